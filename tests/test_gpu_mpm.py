"""GPU parity of the HIP neo-Hookean MPM step (libhfsi hfsi_mpm_step) against
the float64 NumPy oracle (oracle/mpm.py), through NeoHookeanMpmState.

Tolerances: the device state is float32 and P2G scatters with float32
atomics (summation order is not fixed), the oracle is float64.  Fields are
compared with |gpu - oracle| <= rel * max|oracle| per field: positions 1e-6 of
the box size (float32 rounding of ~1e-2 m coordinates), F 1e-5, velocities,
affine C, grid velocity and momenta 2e-4 (C, which vanishes for rigid motion,
against at least 1e-3 of its natural scale 4|v|/dx).  Integer report counts are exact.
"""

import numpy as np
import pytest

from easyfsi_amd.simulation_core.hyperelastic import ecoflex_0010_material
from oracle import mpm as M

pytestmark = pytest.mark.gpu

MAT = ecoflex_0010_material()
MU, LAM = MAT.shear_modulus_pa, MAT.lame_lambda_pa


def _state(nodes, lo, hi, counts, box_lo, box_hi, capacity=None, p2g="binned"):
    from easyfsi_amd.simulation_core.neo_hookean_mpm import NeoHookeanMpmState

    P = int(np.prod(counts))
    s = NeoHookeanMpmState(capacity or P, lo, hi, nodes)
    s.set_p2g_mode(p2g)
    s.initialize_box(particle_counts=counts, box_min_m=box_lo, box_max_m=box_hi, density_kgm3=MAT.density_kgm3)
    return s


def _oracle_from(s) -> M.MpmState:
    n = s.particle_count
    g = lambda f: f.to_numpy()[:n].astype(np.float64)  # noqa: E731
    return M.MpmState(x=g(s.x), v=g(s.v), C=g(s.C), F=g(s.F), mass=g(s.mass_kg), volume=g(s.volume_m3),
                      rest_x=g(s.rest_x), fixed=s.fixed_particle.to_numpy()[:n].copy(),
                      region=s.region_id.to_numpy()[:n].copy(), f_ext=g(s.external_force_n),
                      normal=g(s.surface_normal), rest_normal=g(s.rest_surface_normal), area=g(s.area_weight_m2),
                      rest_area=g(s.rest_area_weight_m2))


def _close(name, got, want, rel, floor=1e-30):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = max(float(np.abs(want).max()) if want.size else 0.0, floor)
    err = float(np.abs(got - want).max()) if want.size else 0.0
    assert err <= rel * scale, f"{name}: max err {err:.3e} > {rel:.1e} * {scale:.3e}"


def _compare(s, o: M.MpmState, rep, ref: M.StepReport, *, x_scale):
    n = s.particle_count
    _close("x", s.x.to_numpy()[:n], o.x, 1e-6 * x_scale / max(np.abs(o.x).max(), 1e-30))
    _close("v", s.v.to_numpy()[:n], o.v, 2e-4, floor=1e-12)
    # C is ~0 for rigid motion: its float32 rounding floor is 1e-3 of |v| * 4 / dx
    c_nat = 4.0 * float(np.abs(o.v).max()) / min(s.dx)
    _close("C", s.C.to_numpy()[:n], o.C, 2e-4, floor=max(1e-3 * c_nat, 1e-9))
    _close("F", s.F.to_numpy()[:n], o.F, 1e-5)
    _close("grid mass", s.grid_mass_kg.to_numpy(), ref.grid_mass, 1e-5)
    _close("grid velocity", s.grid_velocity_mps.to_numpy(), ref.grid_velocity, 2e-4, floor=1e-12)
    _close("grid force", s.grid_force_n.to_numpy(), ref.grid_force, 1e-5, floor=1e-30)
    _close("normal", s.surface_normal.to_numpy()[:n], o.normal, 1e-5, floor=1e-30)
    _close("area", s.area_weight_m2.to_numpy()[:n], o.area, 1e-5, floor=1e-30)
    v = ref.values
    assert rep.active_grid_nodes == v["active_grid_nodes"]
    assert rep.grid_out_of_bounds_particle_count == v["oob_count"]
    assert rep.deformation_clamp_count == v["clamp_count"]
    assert rep.total_mass_kg == pytest.approx(v["total_mass"], rel=1e-6)
    assert rep.total_volume_m3 == pytest.approx(v["total_volume"], rel=1e-6)
    mscale = max(np.abs(v["particle_momentum"]).max(), np.sqrt(v["momentum_square_sum"]), 1e-30)
    _close("particle momentum", rep.particle_momentum_kg_mps, v["particle_momentum"], 1e-5 * mscale
           / max(np.abs(v["particle_momentum"]).max(), 1e-30))
    gscale = max(np.abs(v["grid_momentum"]).max(), mscale)
    assert np.abs(np.asarray(rep.grid_momentum_kg_mps) - v["grid_momentum"]).max() <= 2e-4 * gscale
    _close("external force", rep.external_force_n, v["external_force"], 1e-6, floor=1e-30)
    assert rep.max_abs_j == pytest.approx(v["max_abs_j"], rel=1e-5)
    assert rep.max_speed_mps == pytest.approx(v["max_speed"], rel=2e-4, abs=1e-12)
    assert rep.transfer_relative_error <= max(2.0 * v["transfer_relative_error"], 2e-5)
    pc = max(v["primary_count"], 1)
    sc = max(v["secondary_count"], 1)
    _close("primary displacement", rep.primary_mean_displacement_m, v["primary_displacement_sum"] / pc,
           1e-6 * x_scale / max(np.abs(v["primary_displacement_sum"] / pc).max(), 1e-30), floor=1e-30)
    _close("secondary velocity", rep.secondary_mean_velocity_mps, v["secondary_velocity_sum"] / sc, 2e-4,
           floor=1e-12)
    rc = max(v["radial_stretch_count"], 1)
    assert rep.mean_radial_stretch == pytest.approx(v["radial_stretch_sum"] / rc, rel=1e-5)


@pytest.mark.parametrize("p2g", ["binned", "atomic"])
def test_uniform_velocity_step_matches_oracle(p2g):  # test_neo_hookean_mpm.py:102-130
    s = _state((12, 12, 12), (-0.02,) * 3, (0.02,) * 3, (4, 4, 4), (-0.005,) * 3, (0.005,) * 3, p2g=p2g)
    s.set_uniform_velocity((0.02, -0.01, 0.03))
    o = _oracle_from(s)
    ref = M.step(o, bounds_min=(-0.02,) * 3, bounds_max=(0.02,) * 3, grid_nodes=(12, 12, 12), dt=1e-5, mu=MU,
                 lam=LAM)
    rep = s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    assert s.last_report_host_reads == 1
    _compare(s, o, rep, ref, x_scale=0.04)
    assert rep.transfer_relative_error < 2e-5
    assert rep.max_abs_j == pytest.approx(1.0, abs=1e-4)


def test_non_cubic_grid_spacing_matches_oracle():  # test_neo_hookean_mpm.py:132-161
    lo, hi, nodes = (-0.01, -0.01, -0.04), (0.01, 0.01, 0.04), (12, 12, 12)
    s = _state(nodes, lo, hi, (2, 2, 2), (-0.002, -0.002, 0.024), (0.002, 0.002, 0.030))
    s.set_uniform_velocity((0.0, 0.0, 0.01))
    o = _oracle_from(s)
    ref = M.step(o, bounds_min=lo, bounds_max=hi, grid_nodes=nodes, dt=1e-5, mu=MU, lam=LAM)
    rep = s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    _compare(s, o, rep, ref, x_scale=0.08)
    assert rep.grid_spacing_m == pytest.approx(((0.02 / 12), (0.02 / 12), (0.08 / 12)))


def _loaded_state(seed, p2g="binned"):
    rng = np.random.default_rng(seed)
    lo, hi, nodes = (-0.03, -0.025, -0.02), (0.03, 0.025, 0.04), (20, 18, 24)
    s = _state(nodes, lo, hi, (7, 6, 5), (-0.01, -0.008, -0.005), (0.012, 0.009, 0.02), p2g=p2g)
    n = s.particle_count
    F = np.eye(3)[None] + 0.06 * rng.standard_normal((n, 3, 3))
    s.F.tensor[:n] = _t(F)
    s.v.tensor[:n] = _t(0.05 * rng.standard_normal((n, 3)))
    s.C.tensor[:n] = _t(2.0 * rng.standard_normal((n, 3, 3)))
    s.external_force_n.tensor[:n] = _t(1e-3 * rng.standard_normal((n, 3)))
    s.fixed_particle.tensor[:n] = _t((rng.random(n) < 0.1).astype(np.int32))
    s.region_id.tensor[:n] = _t(rng.integers(0, 3, n).astype(np.int32))
    surf = rng.random(n) < 0.4
    rn = rng.standard_normal((n, 3))
    rn /= np.linalg.norm(rn, axis=1, keepdims=True)
    s.rest_surface_normal.tensor[:n] = _t(np.where(surf[:, None], rn, 0.0))
    s.rest_area_weight_m2.tensor[:n] = _t(np.where(surf, 1e-6 * (1 + rng.random(n)), 0.0))
    s.surface_normal.tensor[:n] = s.rest_surface_normal.tensor[:n]
    s.area_weight_m2.tensor[:n] = s.rest_area_weight_m2.tensor[:n]
    return s, lo, hi, nodes


def _t(a):
    import torch

    a = np.asarray(a)
    dt = torch.int32 if a.dtype.kind == "i" else torch.float32
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dt)


@pytest.mark.parametrize("damping,p2g", [(1.0, "binned"), (0.98, "binned"), (1.0, "atomic")])
def test_deformed_loaded_state_matches_oracle_per_step(damping, p2g):
    s, lo, hi, nodes = _loaded_state(7, p2g)
    for _ in range(3):
        o = _oracle_from(s)
        ref = M.step(o, bounds_min=lo, bounds_max=hi, grid_nodes=nodes, dt=2e-5, mu=MU, lam=LAM, damping=damping,
                     primary_region_id=1, secondary_region_id=2)
        rep = s.step(dt_s=2e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2,
                     velocity_damping=damping)
        _compare(s, o, rep, ref, x_scale=0.06)
        fixed = s.fixed_particle.to_numpy()[: s.particle_count] != 0
        assert fixed.any()
        np.testing.assert_array_equal(s.v.to_numpy()[: s.particle_count][fixed], 0.0)
        np.testing.assert_array_equal(s.F.to_numpy()[: s.particle_count][fixed], np.eye(3)[None].repeat(fixed.sum(), 0))


def test_deformation_clamp_and_reflection_match_oracle():  # test_neo_hookean_mpm.py:163-223
    rng = np.random.default_rng(3)
    s = _state((12, 12, 12), (-0.02,) * 3, (0.02,) * 3, (3, 3, 3), (-0.004,) * 3, (0.004,) * 3)
    n = s.particle_count
    Fs = []
    for p in range(n):
        Q1, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        Q2, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        sig = [(1.5, 0.9, 0.6), (300.0, 1.2, 0.7), (1.1, 0.5, 2e-3), (1.3, 0.8, -0.5), (1.0, 0.97, 0.93)][p % 5]
        Fs.append(Q1 @ np.diag(sig) @ Q2.T)
    s.F.tensor[:n] = _t(np.array(Fs))
    o = _oracle_from(s)
    ref = M.step(o, bounds_min=(-0.02,) * 3, bounds_max=(0.02,) * 3, grid_nodes=(12, 12, 12), dt=1e-6, mu=MU,
                 lam=LAM)
    rep = s.step(dt_s=1e-6, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    assert rep.deformation_clamp_count == ref.values["clamp_count"] > 0
    _close("F", s.F.to_numpy()[:n], o.F, 1e-5)
    assert (np.linalg.det(s.F.to_numpy()[:n].astype(np.float64)) > 0).all()


@pytest.mark.parametrize("p2g", ["binned", "atomic"])
def test_all_particles_outside_grid_raises(p2g):  # test_neo_hookean_mpm.py:225-246
    s = _state((8, 8, 8), (0.0,) * 3, (0.01,) * 3, (2, 2, 2), (0.05,) * 3, (0.06,) * 3, p2g=p2g)
    with pytest.raises(RuntimeError, match="outside the background grid"):
        s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
        print(f"no raise ({p2g}): particle_count {s.particle_count}, report counts {s._report[28:34]}", flush=True)


@pytest.mark.parametrize("p2g", ["binned", "atomic"])
def test_nan_positions_count_as_out_of_bounds(p2g):
    s = _state((8, 8, 8), (-0.01,) * 3, (0.01,) * 3, (2, 2, 2), (-0.002,) * 3, (0.002,) * 3, p2g=p2g)
    s.x.tensor[0, 1] = float("nan")
    rep = s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    assert rep.grid_out_of_bounds_particle_count == 1


def test_read_report_false_and_save_restore():  # test_neo_hookean_mpm.py:248-300
    s, lo, hi, nodes = _loaded_state(11)
    n = s.particle_count
    assert s.step(dt_s=2e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2,
                  read_report=False) is None
    assert s.last_report_host_reads == 0
    s.save_state()
    saved = {k: getattr(s, k).to_numpy()[:n].copy() for k in ("x", "v", "C", "F")}
    for _ in range(2):
        s.step(dt_s=2e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    s.restore_state()
    for k, want in saved.items():
        np.testing.assert_array_equal(getattr(s, k).to_numpy()[:n], want)
    np.testing.assert_array_equal(s.external_force_n.to_numpy()[:n], 0.0)
    o = _oracle_from(s)
    M._update_surface(o, np.arange(n))
    _close("normal", s.surface_normal.to_numpy()[:n], o.normal, 1e-5, floor=1e-30)
    _close("area", s.area_weight_m2.to_numpy()[:n], o.area, 1e-5, floor=1e-30)


def test_step_without_particles_raises():
    from easyfsi_amd.simulation_core.neo_hookean_mpm import NeoHookeanMpmState

    s = NeoHookeanMpmState(8, (0.0,) * 3, (0.01,) * 3, (8, 8, 8))
    with pytest.raises(ValueError, match="initialize particles before stepping"):
        s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)


def test_binned_and_atomic_scatter_agree_at_scale():
    """64^3 particles with a rotation on a 96^3 grid: the LDS-binned P2G and the global-atomic P2G give the
    same grid (float summation order only) and the same particle state."""
    out = {}
    for mode in ("binned", "atomic"):
        s = _state((96, 96, 96), (-0.05,) * 3, (0.05,) * 3, (64, 64, 64), (-0.02,) * 3, (0.02,) * 3, p2g=mode)
        x = s.x.tensor[: s.particle_count]
        v = torch_zeros_like(x)
        v[:, 0] = -3.0 * x[:, 1]
        v[:, 1] = 3.0 * x[:, 0]
        s.v.tensor[: s.particle_count] = v
        rep = s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
        out[mode] = (s.grid_mass_kg.to_numpy(), s.grid_velocity_mps.to_numpy(), s.v.to_numpy(), s.C.to_numpy(), rep)
    (gm_b, gv_b, v_b, c_b, r_b), (gm_a, gv_a, v_a, c_a, r_a) = out["binned"], out["atomic"]
    _close("grid mass", gm_b, gm_a, 1e-5)
    _close("grid velocity", gv_b, gv_a, 1e-4)
    _close("v", v_b, v_a, 1e-4)
    _close("C", c_b, c_a, 1e-4)
    assert r_b.active_grid_nodes == r_a.active_grid_nodes
    assert r_b.transfer_relative_error < 2e-5 and r_a.transfer_relative_error < 2e-5


def torch_zeros_like(t):
    import torch

    return torch.zeros_like(t)


def test_large_block_conserves_mass_and_momentum():
    """64^3 particles on a 96^3 grid: size-independent transfer invariants."""
    s = _state((96, 96, 96), (-0.05,) * 3, (0.05,) * 3, (64, 64, 64), (-0.02,) * 3, (0.02,) * 3)
    s.set_uniform_velocity((0.03, -0.02, 0.01))
    rep = s.step(dt_s=1e-5, mu_pa=MU, lambda_pa=LAM, primary_region_id=1, secondary_region_id=2)
    assert rep.grid_out_of_bounds_particle_count == 0
    assert rep.transfer_relative_error < 2e-5
    gm = s.grid_mass_kg.tensor.double().sum().item()
    assert gm == pytest.approx(rep.total_mass_kg, rel=1e-5)
    np.testing.assert_allclose(rep.grid_momentum_kg_mps, rep.particle_momentum_kg_mps, rtol=1e-4)
    v = s.v.to_numpy()[: s.particle_count].astype(np.float64)
    np.testing.assert_allclose(v, np.broadcast_to((0.03, -0.02, 0.01), v.shape), rtol=1e-5)
    assert rep.max_abs_j == pytest.approx(1.0, abs=1e-4)
