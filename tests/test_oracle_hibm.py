"""Pins the HIBM marker -> particle force scatter oracle (oracle/hibm.py).

The reference's scatter tests (tests/test_hibm.py:605-734, 4040-4070) are
re-expressed; force conservation, the support cut-off and a scalar per-pair
loop pin the rest.
"""

import numpy as np
import pytest

from oracle import hibm as H


def test_one_marker_on_one_particle():  # test_hibm.py:605-642
    ext, rep = H.scatter_marker_forces([[0.5, 0.5, 0.5]], [[1.0, -2.0, 3.0]], [[0.5, 0.5, 0.5]], np.zeros((1, 3)), 0.25)
    assert rep["active_marker_count"] == 1 and rep["invalid_marker_count"] == 0 and rep["active_particle_count"] == 1
    assert tuple(ext[0]) == (1.0, -2.0, 3.0)
    assert tuple(rep["total_mpm_external_force_n"]) == (1.0, -2.0, 3.0)
    assert rep["action_reaction_residual_n"] == 0.0


def test_force_partitions_across_supported_particles():  # test_hibm.py:644-682
    xp = [[0.375, 0.5, 0.5], [0.625, 0.5, 0.5]]
    ext, rep = H.scatter_marker_forces([[0.5, 0.5, 0.5]], [[0.0, 0.0, 4.0]], xp, np.zeros((2, 3)), 0.4)
    assert rep["active_particle_count"] == 2
    assert ext[0, 2] == pytest.approx(2.0, abs=1e-5) and ext[1, 2] == pytest.approx(2.0, abs=1e-5)
    assert rep["action_reaction_residual_n"] == pytest.approx(0.0, abs=1e-5)


def test_non_finite_marker_leaves_forces_untouched():  # test_hibm.py:4040-4070
    ext0 = np.array([[1.0, 2.0, 3.0]], np.float32)
    ext, rep = H.scatter_marker_forces([[0.5, 0.5, 0.5]], [[np.nan, 0.0, 0.0]], [[0.5, 0.5, 0.5]], ext0, 0.5)
    assert rep["active_marker_count"] == 0 and rep["invalid_marker_count"] == 1
    assert tuple(rep["total_marker_force_n"]) == (0.0, 0.0, 0.0)
    np.testing.assert_array_equal(ext, ext0)


def test_conservation_cutoff_and_scalar_loop():
    rng = np.random.default_rng(2)
    xp = rng.uniform(0.0, 1.0, (300, 3)).astype(np.float32)
    xp[7] = np.nan  # skipped particle
    xm = rng.uniform(0.1, 0.9, (25, 3)).astype(np.float32)
    fm = rng.normal(size=(25, 3))
    fm[3, 1] = np.inf  # invalid marker
    xm[9] = (5.0, 5.0, 5.0)  # no particle in reach: invalid
    ext0 = rng.normal(size=(300, 3)).astype(np.float32)
    h = 0.15
    ext, rep = H.scatter_marker_forces(xm, fm, xp, ext0, h)
    assert rep["invalid_marker_count"] == 2 and rep["active_marker_count"] == 23
    # each active marker's force arrives in full (normalised weights sum to one)
    np.testing.assert_allclose(rep["total_mpm_external_force_n"], rep["total_marker_force_n"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose((ext.astype(np.float64) - ext0).sum(axis=0), rep["total_marker_force_n"], rtol=1e-4,
                               atol=1e-4)
    # particles beyond the support along any axis receive nothing
    far = np.ones(300, bool)
    for m in range(25):
        if m in (3, 9):
            continue
        far &= ~(np.abs(xp - xm[m]) < h).all(axis=1)
    far &= np.isfinite(xp).all(axis=1)
    np.testing.assert_array_equal(ext[far], ext0[far])
    loop_ext, counts = H.scatter_scalar_loop(xm, fm, xp, ext0, h)
    assert counts == [rep["active_marker_count"], rep["invalid_marker_count"], rep["active_particle_count"]]
    np.testing.assert_allclose(ext, loop_ext, rtol=0, atol=2e-6 * np.abs(ext).max())


def test_surface_feedback_moves_marker_with_particle_velocity():  # test_hibm.py:736-778
    xm, vm, rep = H.surface_feedback([[0.5, 0.5, 0.5]], [[0.6, 0.45, 0.7]], [[1.0, -2.0, 3.0]], 0.5, 0.1)
    assert rep["updated_marker_count"] == 1 and rep["invalid_marker_count"] == 0
    np.testing.assert_allclose(xm[0], (0.6, 0.3, 0.8), atol=1e-6)
    np.testing.assert_allclose(vm[0], (1.0, -2.0, 3.0), atol=1e-6)


def test_surface_feedback_reproduces_uniform_and_linear_velocity_fields():
    rng = np.random.default_rng(6)
    xp = rng.uniform(0.0, 1.0, (4000, 3)).astype(np.float32)
    xm = rng.uniform(0.2, 0.8, (30, 3)).astype(np.float32)
    u = np.array([0.3, -0.1, 0.25], np.float32)
    xm1, vm1, rep = H.surface_feedback(xm, xp, np.broadcast_to(u, xp.shape), 0.12, 0.01)
    assert rep["updated_marker_count"] == 30
    np.testing.assert_allclose(vm1, np.broadcast_to(u, vm1.shape), rtol=2e-6)  # normalised weights average a constant
    np.testing.assert_allclose(xm1, xm + np.float32(0.01) * u, atol=1e-6)
    # a far marker is invalid and stays put
    xm2, _, rep2 = H.surface_feedback(np.array([[5.0, 5.0, 5.0]], np.float32), xp, np.zeros_like(xp), 0.12, 0.01)
    assert rep2["invalid_marker_count"] == 1 and np.array_equal(xm2, [[5.0, 5.0, 5.0]])
