"""Pins the NumPy MPM oracle (oracle/mpm.py) with the reference's MPM tests
(tests/test_neo_hookean_mpm.py) re-expressed, closed forms of the quadratic
B-spline and numpy.linalg.svd invariants."""

import math

import numpy as np
import pytest

from easyfsi_amd.simulation_core.hyperelastic import ecoflex_0010_material
from oracle import mpm as M

MAT = ecoflex_0010_material()


def _box(counts, lo, hi, density=MAT.density_kgm3):
    return M.initialize_box(counts, lo, hi, density)


def _step(s, bmin, bmax, nodes, dt=1e-5, damping=1.0, mu=MAT.shear_modulus_pa, lam=MAT.lame_lambda_pa):
    return M.step(s, bounds_min=bmin, bounds_max=bmax, grid_nodes=nodes, dt=dt, mu=mu, lam=lam, damping=damping)


def test_bspline_partition_of_unity_and_zero_first_moment():
    fx = np.random.default_rng(0).uniform(0.5, 1.5, size=1000)
    w = M.weights(fx)
    np.testing.assert_allclose(w.sum(axis=0), 1.0, atol=1e-14)
    np.testing.assert_allclose(sum(w[o] * (o - fx) for o in range(3)), 0.0, atol=1e-14)
    assert (w >= 0).all()


def test_svd_rotation_convention():
    rng = np.random.default_rng(1)
    F = rng.normal(size=(200, 3, 3))
    U, s, V = M.svd_rotations(F)
    np.testing.assert_allclose(U @ (s[:, :, None] * np.swapaxes(V, 1, 2)), F, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(V), 1.0, atol=1e-12)
    assert np.all(np.sign(s[:, 2]) == np.sign(np.linalg.det(F)))
    np.testing.assert_allclose(np.abs(s), np.linalg.svd(F, compute_uv=False), atol=1e-12)


def test_uniform_velocity_transfer_is_conservative():  # test_neo_hookean_mpm.py:102-130
    s = _box((4, 4, 4), (-0.005,) * 3, (0.005,) * 3)
    s.v[:] = (0.02, -0.01, 0.03)
    r = _step(s, (-0.02,) * 3, (0.02,) * 3, (12, 12, 12)).values
    assert r["active_grid_nodes"] > 0
    assert r["transfer_relative_error"] < 2e-5
    assert r["total_mass"] > 0
    assert r["max_abs_j"] == pytest.approx(1.0, abs=1e-4)


def test_non_cubic_grid_spacing():  # test_neo_hookean_mpm.py:132-161
    s = _box((2, 2, 2), (-0.002, -0.002, 0.024), (0.002, 0.002, 0.030))
    s.v[:] = (0.0, 0.0, 0.01)
    r = _step(s, (-0.01, -0.01, -0.04), (0.01, 0.01, 0.04), (12, 12, 12)).values
    assert r["active_grid_nodes"] > 0 and r["transfer_relative_error"] < 2e-5
    assert r["max_abs_j"] == pytest.approx(1.0, abs=1e-4)


def test_external_force_changes_grid_momentum():  # test_neo_hookean_mpm.py:163-189
    s = _box((2, 2, 2), (-0.004,) * 3, (0.004,) * 3)
    s.f_ext[:] = (0.0, 0.0, -1e-4)
    r = _step(s, (-0.02,) * 3, (0.02,) * 3, (12, 12, 12)).values
    assert r["external_force"][2] < 0 and r["grid_momentum"][2] < 0
    assert r["transfer_relative_error"] < 2e-5


def test_damping_scales_grid_momentum_only():  # test_neo_hookean_mpm.py:277-309
    s = _box((2, 2, 2), (-0.004,) * 3, (0.004,) * 3)
    s.v[:] = (0.05, 0.0, 0.0)
    r = _step(s, (-0.02,) * 3, (0.02,) * 3, (12, 12, 12), damping=0.9).values
    assert r["max_speed"] > 0 and r["particle_momentum"][0] > 0
    assert r["grid_momentum"][0] == pytest.approx(0.9 * r["particle_momentum"][0], rel=5e-5)
    assert r["transfer_relative_error"] < 2e-5


def test_radial_stretch_translation_invariant():  # test_neo_hookean_mpm.py:348-372
    s = _box((2, 2, 2), (-0.05, -0.05, 1.0), (0.05, 0.05, 1.1))
    s.x = s.rest_x + np.array([0.25, -0.1, 0.05])
    r = _step(s, (-1.0, -1.0, 0.5), (1.0, 1.0, 2.0), (16, 16, 16), dt=0.0).values
    assert r["radial_stretch_sum"] / max(r["radial_stretch_count"], 1) == pytest.approx(1.0, abs=1e-6)
    assert r["max_radial_stretch_error"] == pytest.approx(0.0, abs=1e-6)


def test_out_of_bounds_particle_deposits_nothing():  # test_neo_hookean_mpm.py:374-444
    s = _box((2, 2, 2), (-0.004,) * 3, (0.004,) * 3)
    dx = 0.04 / 8
    s.x[0, 0] = -0.02 + 0.25 * dx
    out = _step(s, (-0.02,) * 3, (0.02,) * 3, (8, 8, 8), dt=0.0)
    assert out.values["oob_count"] == 1
    assert out.grid_mass.sum() == pytest.approx(s.mass[1:].sum(), abs=1e-12)


@pytest.mark.parametrize("F", [np.diag([1.0, 1.0, -1.0]), np.diag([-2.0, 0.5, 0.5])])
def test_svd_clamp_of_inverted_deformation(F):  # test_neo_hookean_mpm.py:543-620
    s = _box((1, 1, 1), (-0.001,) * 3, (0.001,) * 3)
    s.F[0] = F
    r = _step(s, (-0.01,) * 3, (0.01,) * 3, (8, 8, 8), dt=0.0).values
    assert r["clamp_count"] == 1
    assert np.linalg.det(s.F[0]) > 0
    sv = np.linalg.svd(s.F[0], compute_uv=False)
    assert sv.min() >= 1e-2 - 1e-12 and sv.max() <= 1e2
    if np.allclose(F, np.diag([1.0, 1.0, -1.0])):
        assert abs(r["max_abs_j"] - 1.0) < 1e-5


def test_surface_area_and_normal_follow_deformation():  # test_neo_hookean_mpm.py:38-100
    s = _box((1, 1, 1), (-0.001,) * 3, (0.001,) * 3)
    s.rest_area[0] = 0.5
    s.area[0] = 0.5
    s.rest_normal[0] = (0.0, 0.0, 1.0)
    s.normal[0] = (0.0, 0.0, 1.0)
    s.F[0] = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.5, 0.0, 1.0]])
    _step(s, (-1.0,) * 3, (1.0,) * 3, (12, 12, 12), dt=0.0, mu=0.0, lam=0.0)
    n = np.array([-0.5, 0.0, 1.0]) / math.sqrt(1.25)
    assert s.area[0] == pytest.approx(0.5 * math.sqrt(1.25), abs=1e-12)
    np.testing.assert_allclose(s.normal[0], n, atol=1e-12)


def test_rigid_translation_keeps_identity_deformation():
    """A uniform velocity field is transported exactly: C = 0, F = I, x += dt v."""
    s = _box((3, 3, 3), (-0.004,) * 3, (0.004,) * 3)
    s.v[:] = (0.1, -0.2, 0.05)
    x0 = s.x.copy()
    _step(s, (-0.02,) * 3, (0.02,) * 3, (12, 12, 12), dt=1e-4, mu=0.0, lam=0.0)
    np.testing.assert_allclose(s.v, np.tile([0.1, -0.2, 0.05], (27, 1)), atol=1e-12)
    np.testing.assert_allclose(s.C, 0.0, atol=1e-9)
    np.testing.assert_allclose(s.x, x0 + 1e-4 * np.array([0.1, -0.2, 0.05]), atol=1e-15)
