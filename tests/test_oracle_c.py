"""The C + OpenMP oracle (CPU baseline) agrees with the NumPy oracle."""

import numpy as np
import pytest

from oracle import c_oracle
from oracle import fluid as F


def _case(shape, seed, frac=0.05):
    rng = np.random.default_rng(seed)
    w = [rng.uniform(0.6, 1.4, size=n) / n for n in shape]
    obst = (rng.random(shape) < frac).astype(np.int32)
    return rng, w, obst


@pytest.mark.parametrize("outlet", [False, True])
def test_c_vcycle_matches_numpy(outlet):
    rng, w, obst = _case((18, 14, 12), 41)
    r = np.where(obst == 0, rng.normal(size=obst.shape), 0.0)
    z = c_oracle.vcycle(w, obst, r, outlet=outlet)
    want = F.vcycle(r, F.build_hierarchy(w, obst), outlet)
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-6 * np.abs(want).max())


@pytest.mark.parametrize("outlet,multigrid", [(False, True), (True, True), (False, False)])
def test_c_pcg_matches_numpy(outlet, multigrid):
    rng, w, obst = _case((20, 16, 14), 43)
    h = F.build_hierarchy(w, obst)
    rhs0 = np.where(obst == 0, rng.normal(size=obst.shape), 0.0)
    x, it, rel, conv = c_oracle.pcg(w, obst, rhs0, outlet=outlet, remove_mean=not outlet, multigrid=multigrid,
                                    max_iters=600, tolerance=1e-10)
    ref = F.pcg(rhs0, np.zeros(obst.shape), h, outlet=outlet, tolerance=1e-10, max_iters=600,
                preconditioner="fv_multigrid" if multigrid else "jacobi", remove_mean=not outlet)
    assert conv and ref.converged
    assert abs(it - ref.iterations) <= 1, (it, ref.iterations)
    np.testing.assert_allclose(x, ref.pressure, rtol=0, atol=1e-7 * np.abs(ref.pressure).max())
