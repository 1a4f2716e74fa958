"""Host-side checks of the learned-guess feature / target transforms
(easyfsi_amd/learned.py) on CPU tensors: closed forms of the 2x2x2 volume
averaging, normalisation and the scaling identity that maps the network output
back to a pressure."""

import math

import numpy as np
import pytest
import torch

from easyfsi_amd import learned as Lm


def _widths(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.5, 1.5, n) / n for _ in range(3)]


def test_coarsen_widths_are_pairwise_sums():
    w = np.array([0.1, 0.2, 0.3, 0.4])
    np.testing.assert_allclose(Lm.coarsen_widths([w])[0], [0.3, 0.7])


def test_features_fine_grid_channels():
    n = 8
    w = _widths(n, 1)
    rng = np.random.default_rng(2)
    div = torch.tensor(rng.normal(size=(n, n, n)), dtype=torch.float32)
    obst = torch.zeros((n, n, n), dtype=torch.int32)
    obst[2:4, 3:5, 1:6] = 1
    f, rms, wn = Lm.features(div, obst, w, coarsen=False)
    assert f.shape == (n ** 3, 5)
    f = f.reshape(n, n, n, 5).double()
    free = (obst == 0).double()
    V = torch.as_tensor(w[0])[:, None, None] * torch.as_tensor(w[1])[None, :, None] * torch.as_tensor(w[2])[None, None, :]
    # the normalised divergence has unit V-weighted RMS over free cells, zero on solid cells
    assert float((V * free * f[..., 0] ** 2).sum() / (V * free).sum()) == pytest.approx(1.0, rel=1e-5)
    assert float(f[..., 0][obst == 1].abs().max()) == 0.0
    np.testing.assert_array_equal(f[..., 1].numpy(), (obst == 1).numpy().astype(np.float64))
    hmin = min(a.min() for a in w)
    np.testing.assert_allclose(f[:, 0, 0, 2].numpy(), np.log2(w[0] / hmin), rtol=1e-6)
    np.testing.assert_allclose(f[0, 0, :, 4].numpy(), np.log2(w[2] / hmin), rtol=1e-6)


def test_features_coarsened_are_volume_averages():
    n = 8
    w = _widths(n, 3)
    rng = np.random.default_rng(4)
    div = torch.tensor(rng.normal(size=(n, n, n)), dtype=torch.float32)
    obst = torch.zeros((n, n, n), dtype=torch.int32)
    obst[0:2, 0:2, 0:1] = 1  # 2 of the 8 children of coarse cell (0, 0, 0) are solid
    obst[4:6, 4:6, 4:6] = 1  # coarse cell (2, 2, 2) is fully solid
    f, rms, wc = Lm.features(div, obst, w, coarsen=True)
    f = f.reshape(4, 4, 4, 5).double()
    V = (torch.as_tensor(w[0])[:, None, None] * torch.as_tensor(w[1])[None, :, None] *
         torch.as_tensor(w[2])[None, None, :])
    blk = (slice(0, 2), slice(0, 2), slice(0, 2))
    fv = (V[blk] * (obst[blk] == 0)).sum()
    d = float((V[blk] * div.double()[blk] * (obst[blk] == 0)).sum() / fv)
    assert float(f[0, 0, 0, 0]) * rms == pytest.approx(d, rel=1e-5)
    assert float(f[0, 0, 0, 1]) == pytest.approx(1.0 - float(fv / V[blk].sum()), rel=1e-6)
    assert float(f[2, 2, 2, 1]) == pytest.approx(1.0) and float(f[2, 2, 2, 0]) == 0.0
    np.testing.assert_allclose(wc[0], Lm.coarsen_widths(w)[0])


def test_target_scaling_identity():
    """p_hat = -p kappa / ((rho/dt) rms), kappa = 1/h_min^2 of the network grid."""
    n = 8
    w = _widths(n, 5)
    rng = np.random.default_rng(6)
    p = torch.tensor(rng.normal(size=(n, n, n)))
    t = Lm.target(p, 2.5, w, 1.0e6, coarsen=False).double()
    hmin = min(a.min() for a in w)
    np.testing.assert_allclose(t.numpy(), (-p * (1 / hmin ** 2) / (1.0e6 * 2.5)).numpy(), rtol=1e-6)
    tc = Lm.target(p, 2.5, w, 1.0e6, coarsen=True).double()
    hc = min(a.min() for a in Lm.coarsen_widths(w))
    V = (torch.as_tensor(w[0])[:, None, None] * torch.as_tensor(w[1])[None, :, None] *
         torch.as_tensor(w[2])[None, None, :])
    pa = float((V[:2, :2, :2] * p[:2, :2, :2]).sum() / V[:2, :2, :2].sum())
    assert float(tc[0, 0, 0]) == pytest.approx(-pa / hc ** 2 / 2.5e6, rel=1e-5)


def test_bench_axis_is_the_benchmark_grid():
    a = Lm.bench_axis(256)
    assert a.size == 256 and a.sum() == pytest.approx(1.0)
    assert a.max() / a.min() == pytest.approx(1.04 ** 64, rel=1e-9)
