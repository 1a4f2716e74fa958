"""Pins the predictor oracle (oracle/predict.py).

The reference's own predictor tests (tests/test_core_fluid.py:3152-3190,
3738-3846, 3848-3930) are re-expressed against the oracle; the coordinate
lookup is checked against the host grid's float64 linear scan and
numpy.searchsorted; closed forms fix the trilinear sample, the departure
point and the FV Laplacian on graded axes; a scalar per-cell loop checks the
vectorised masks and clamps bit for bit.
"""

import numpy as np
import pytest

from easyfsi_amd.simulation_core.grid import CartesianGrid, _axis_grid_coordinate
from oracle import fluid as F
from oracle import predict as P

F32 = np.float32


def _axes_uniform(shape, length=1.0):
    return [P.AxisMetrics.from_widths(0.0, [length / n] * n) for n in shape]


def _axes_from_grid(g: CartesianGrid):
    return [P.AxisMetrics.from_arrays(g.cell_faces_x_m, g.cell_centers_x_m, g.cell_widths_x_m, g.center_distances_x_m),
            P.AxisMetrics.from_arrays(g.cell_faces_y_m, g.cell_centers_y_m, g.cell_widths_y_m, g.center_distances_y_m),
            P.AxisMetrics.from_arrays(g.cell_faces_z_m, g.cell_centers_z_m, g.cell_widths_z_m, g.center_distances_z_m)]


def _graded8():  # test_core_fluid.py:3168-3174
    return CartesianGrid(bounds_min_m=(0.0, 0.0, 0.0),
                         cell_widths_x_m=(0.08, 0.10, 0.12, 0.16, 0.18, 0.17, 0.11, 0.08),
                         cell_widths_y_m=(0.09, 0.11, 0.15, 0.16, 0.14, 0.13, 0.12, 0.10),
                         cell_widths_z_m=(0.07, 0.09, 0.12, 0.15, 0.17, 0.16, 0.14, 0.10))


# ---------------------------------------------------------------------------
# the reference's predictor tests
# ---------------------------------------------------------------------------

def test_predict_preserves_uniform_velocity():  # test_core_fluid.py:3152-3164
    shape = (12, 12, 12)
    v = np.broadcast_to(np.array([0.12, -0.03, 0.04], F32), shape + (3,)).copy()
    out = P.predict(v, np.zeros(shape, np.int32), _axes_uniform(shape), 1.0e-3, 1.0e-3 / 1000.0)
    assert float(np.max(np.abs(out - v))) < 1.0e-6


def test_nonuniform_predict_preserves_uniform_velocity():  # test_core_fluid.py:3166-3190
    g = _graded8()
    shape = g.grid_nodes
    v = np.broadcast_to(np.array([0.01, -0.015, 0.02], F32), shape + (3,)).copy()
    out = P.predict(v, np.zeros(shape, np.int32), _axes_from_grid(g), 1.0e-3, 1.0e-3 / 1000.0)
    assert float(np.max(np.abs(out - v))) < 1.0e-6


def test_predict_viscosity_damps_shear_mode():  # test_core_fluid.py:3738-3759
    shape = (32, 32, 32)
    v = np.zeros(shape + (3,), F32)
    y = (np.arange(32, dtype=F32) + F32(0.5)) / F32(32.0)
    v[:, :, :, 0] = np.sin(6.0 * np.pi * y)[None, :, None]
    before = float(np.linalg.norm(v[2:-2, 2:-2, 2:-2, 0]))
    out = P.predict(v, np.zeros(shape, np.int32), _axes_uniform(shape), 1.0e-3, 5.0e-2 / 1.0)
    after = float(np.linalg.norm(out[2:-2, 2:-2, 2:-2, 0]))
    assert before * 0.9 < after < before * 0.99


def test_nonuniform_predict_viscosity_damps_shear_mode():  # test_core_fluid.py:3761-3792
    g = CartesianGrid(bounds_min_m=(0.0, 0.0, 0.0), cell_widths_x_m=(0.04,) * 8,
                      cell_widths_y_m=(0.020, 0.024, 0.029, 0.035, 0.042, 0.050, 0.060, 0.072, 0.086, 0.102, 0.122,
                                       0.145),
                      cell_widths_z_m=(0.05,) * 6)
    y = np.asarray(g.cell_centers_y_m, dtype=F32)
    extent = float(g.bounds_max_m[1] - g.bounds_min_m[1])
    mode = np.sin(3.0 * np.pi * (y - float(g.bounds_min_m[1])) / extent).astype(F32)
    v = np.zeros(g.grid_nodes + (3,), F32)
    v[:, :, :, 0] = mode[None, :, None]
    before = float(np.linalg.norm(v[:, 2:-2, :, 0]))
    out = P.predict(v, np.zeros(g.grid_nodes, np.int32), _axes_from_grid(g), 2.0e-4, 5.0e-2)
    after = float(np.linalg.norm(out[:, 2:-2, :, 0]))
    assert before * 0.9 < after < before


def _rotation_case(shape=(32, 32, 8), dt=0.12, omega=1.5):  # test_core_fluid.py:3794-3822
    x = (np.arange(shape[0], dtype=F32) + F32(0.5)) / shape[0]
    y = (np.arange(shape[1], dtype=F32) + F32(0.5)) / shape[1]
    xx, yy = np.meshgrid(x, y, indexing="ij")
    sl = np.zeros((shape[0], shape[1], 3), F32)
    sl[:, :, 0] = -omega * (yy - 0.5)
    sl[:, :, 1] = omega * (xx - 0.5)
    v = np.repeat(sl[:, :, None, :], shape[2], axis=2)
    ang = -omega * dt
    dx, dy = xx - 0.5, yy - 0.5
    bx = 0.5 + np.cos(ang) * dx - np.sin(ang) * dy
    by = 0.5 + np.sin(ang) * dx + np.cos(ang) * dy
    ex = np.zeros_like(sl)
    ex[:, :, 0] = -omega * (by - 0.5)
    ex[:, :, 1] = omega * (bx - 0.5)
    return v, np.repeat(ex[:, :, None, :], shape[2], axis=2)


def test_predict_rk2_backtrace_reduces_rotational_advection_error():  # test_core_fluid.py:3794-3846
    shape, dt = (32, 32, 8), 0.12
    v, exact = _rotation_case(shape, dt)
    interior = (slice(6, -6), slice(6, -6), slice(1, -1), slice(None))
    errs = {}
    for scheme in ("euler", "rk2"):
        out = P.predict(v, np.zeros(shape, np.int32), _axes_uniform(shape), dt, 0.0, scheme)
        errs[scheme] = float(np.linalg.norm(out[interior] - exact[interior]) / np.linalg.norm(exact[interior]))
    assert errs["rk2"] < errs["euler"] * 0.55


def test_taylor_green_predict_project_error_decreases_on_finer_grid():  # test_core_fluid.py:3848-3895
    # The reference projects with its default legacy Jacobi solve; the oracle's FV-CG solve to 1e-6
    # stands in for it (a converged projection), so the refinement claim is tested on the predictor.
    def one(shape):
        nx, ny, nz = shape
        amp, nu, dt = 0.05, 1.0e-2, 5.0e-4
        x = (np.arange(nx, dtype=F32) + F32(0.5)) / nx
        y = (np.arange(ny, dtype=F32) + F32(0.5)) / ny
        xx, yy = np.meshgrid(x, y, indexing="ij")
        sl = np.zeros((nx, ny, 3), F32)
        sl[:, :, 0] = amp * np.sin(2 * np.pi * xx) * np.cos(2 * np.pi * yy)
        sl[:, :, 1] = -amp * np.cos(2 * np.pi * xx) * np.sin(2 * np.pi * yy)
        v = np.repeat(sl[:, :, None, :], nz, axis=2)
        exact = v * F32(np.exp(-8.0 * np.pi * np.pi * nu * dt))
        obst = np.zeros(shape, np.int32)
        pred = P.predict(v, obst, _axes_uniform(shape), dt, nu)
        widths = [np.full(n, 1.0 / n) for n in shape]
        u, _, _, info = F.project_plain(pred, np.zeros(shape), obst, widths, rho=1.0, dt=dt, solver="fv_cg",
                                        iterations=12 * nx, tolerance=1.0e-6)
        assert info["projection"]["l2"] < 1.0e-2
        it = (slice(nx // 4, 3 * nx // 4), slice(ny // 4, 3 * ny // 4), slice(1, nz - 1), slice(None))
        return float(np.linalg.norm(u[it] - exact[it]) / np.linalg.norm(exact[it]))

    coarse, fine = one((24, 24, 8)), one((48, 48, 8))
    assert fine < coarse * 0.85


# ---------------------------------------------------------------------------
# coordinate lookup
# ---------------------------------------------------------------------------

def test_grid_coordinate_matches_host_linear_scan_and_searchsorted():
    g = _graded8()
    ax = _axes_from_grid(g)[0]
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.uniform(-0.3, 1.4, 4000), ax.centers, ax.faces]).astype(F32)
    got = P.grid_coordinate(vals, ax)
    ref = np.array([_axis_grid_coordinate(float(v), g.cell_faces_x_m, g.cell_centers_x_m) for v in vals])
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-5)
    # centres map to their own index exactly
    assert np.array_equal(P.grid_coordinate(ax.centers, ax), np.arange(ax.centers.size, dtype=F32))
    # the binary search brackets like searchsorted on the float32 centres
    inner = vals[(vals > ax.centers[0]) & (vals < ax.centers[-1])]
    lower = np.searchsorted(ax.centers, inner, side="right") - 1
    assert np.array_equal(np.floor(P.grid_coordinate(inner, ax)).astype(np.int64), lower)


def test_grid_coordinate_extrapolates_by_end_half_widths():
    ax = P.AxisMetrics.from_widths(0.0, [0.2, 0.3, 0.5])
    got = P.grid_coordinate(np.array([0.0, -0.1, 1.0, 1.25], F32), ax)
    np.testing.assert_allclose(got, [-0.5, -1.0, 2.5, 3.0], atol=1e-6)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------

def test_departure_point_of_uniform_stream_with_linear_cross_profile():
    # u = (U, s*x, 0): the departure point of cell i is x_i - dt*U and the sampled
    # u_y there is s*(x_i - dt*U) (trilinear reproduces an affine profile).
    shape = (16, 6, 5)
    axes = _axes_uniform(shape)
    U, s, dt = 0.8, 2.0, 0.025
    xc = axes[0].centers
    v = np.zeros(shape + (3,), F32)
    v[..., 0] = U
    v[..., 1] = (s * xc)[:, None, None]
    out = P.predict(v, np.zeros(shape, np.int32), axes, dt, 0.0)
    inner = slice(1, None)  # departure stays right of the first centre for i >= 1 (U*dt = 0.32 cells)
    np.testing.assert_allclose(out[inner, :, :, 1], (s * (xc[inner] - dt * U))[:, None, None] * np.ones((1,) + shape[1:]),
                               rtol=0, atol=2e-6)
    # cell 0 departs left of the first centre: t clamps to 0, the sample is the first cell's value
    np.testing.assert_allclose(out[0, :, :, 1], s * xc[0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(out[..., 0], U, atol=1e-6)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("profile", ["linear", "quadratic"])
def test_fv_laplacian_of_polynomial_profiles_on_graded_axes(profile, axis):
    # component c varies along `axis` only and the other components are zero, so the
    # backtrace runs along c, where the field is constant; predict - u is dt*nu*lap(u)
    # up to the rounding of the trilinear weights.
    g = _graded8()
    axes = _axes_from_grid(g)
    comp = (axis + 2) % 3
    y = axes[axis].centers.astype(np.float64)
    prof = y if profile == "linear" else y * y
    shape = g.grid_nodes
    bshape = [1, 1, 1]
    bshape[axis] = -1
    v = np.zeros(shape + (3,), F32)
    v[..., comp] = prof.astype(F32).reshape(bshape)
    dt, nu = 1.0e-3, 2.0
    out = P.predict(v, np.zeros(shape, np.int32), axes, dt, nu)
    lap = (out[..., comp].astype(np.float64) - v[..., comp]) / (F32(dt) * F32(nu))
    w = axes[axis].widths.astype(np.float64)
    cd = axes[axis].center_distances.astype(np.float64)
    pf = prof.astype(F32).astype(np.float64)
    expect = np.zeros_like(y)
    for j in range(y.size):
        back = (pf[j] - pf[j - 1]) / cd[j] if j > 0 else 0.0
        fwd = (pf[j + 1] - pf[j]) / cd[j + 1] if j < y.size - 1 else 0.0
        expect[j] = (fwd - back) / w[j]
    if profile == "linear":
        assert np.allclose(expect[1:-1], 0.0, atol=1e-3)
        # one-sided boundary cells: the slope over the end cell's width
        np.testing.assert_allclose(expect[[0, -1]], [1.0 / w[0], -1.0 / w[-1]], rtol=1e-4)
    else:
        # flux of y^2 between centres is y_{j+1} + y_j, so lap = (y_{j+1} - y_{j-1}) / w_j
        np.testing.assert_allclose(expect[1:-1], (y[2:] - y[:-2]) / w[1:-1], rtol=1e-4)
    np.testing.assert_allclose(lap, np.broadcast_to(expect.reshape(bshape), shape), rtol=2e-3, atol=2e-3)
    others = [c for c in range(3) if c != comp]
    assert np.all(out[..., others] == 0.0)


@pytest.mark.parametrize("scheme", ["euler", "rk2"])
def test_uniform_flow_beside_obstacle_is_preserved(scheme):
    # Fluid cells hold one uniform velocity, obstacle cells a different one: renormalised
    # samples over the fluid corners and zero diffusive flux through obstacle faces keep
    # every fluid cell at the uniform value; obstacle cells are zeroed.
    g = _graded8()
    shape = g.grid_nodes
    obst = np.zeros(shape, np.int32)
    obst[3:5, 2:6, 3:6] = 1
    u = np.array([0.4, -0.3, 0.25], F32)
    v = np.broadcast_to(u, shape + (3,)).copy()
    v[obst == 1] = (7.0, -7.0, 5.0)
    out = P.predict(v, obst, _axes_from_grid(g), 0.09, 0.5, scheme)
    assert np.all(out[obst == 1] == 0.0)
    np.testing.assert_allclose(out[obst == 0], np.broadcast_to(u, (int((obst == 0).sum()), 3)), rtol=0, atol=2e-6)


def test_obstacle_cells_zero_and_enclosed_sample_falls_back_to_center():
    shape = (6, 6, 6)
    axes = _axes_uniform(shape)
    obst = np.ones(shape, np.int32)
    obst[3, 3, 3] = 0  # one fluid cell walled in by obstacles
    v = np.full(shape + (3,), 0.5, F32)
    v[3, 3, 3] = (0.3, -0.2, 0.1)
    # a backtrace of 1.6 cells leaves every corner of the departure cell inside obstacles
    out = P.predict(v, obst, axes, 1.6 * (1.0 / 6) / 0.3, 0.0)
    assert np.all(out[obst == 1] == 0.0)
    assert np.array_equal(out[3, 3, 3], v[3, 3, 3])


# ---------------------------------------------------------------------------
# vectorised oracle vs scalar loop
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("scheme", ["euler", "rk2"])
@pytest.mark.parametrize("seed", [0, 1])
def test_vectorised_oracle_equals_scalar_loop(scheme, seed):
    rng = np.random.default_rng(seed)
    shape = (5, 6, 7)
    axes = [P.AxisMetrics.from_widths(float(rng.uniform(-1, 1)), rng.uniform(0.05, 0.3, n)) for n in shape]
    obst = (rng.random(shape) < 0.25).astype(np.int32)
    v = rng.normal(0.0, 1.0, shape + (3,)).astype(F32)
    dt, nu = 0.17, 0.031  # departures up to several cells, many leave the domain
    vec = P.predict(v, obst, axes, dt, nu, scheme)
    loop = P.predict_scalar_loop(v, obst, axes, dt, nu, scheme)
    assert np.array_equal(vec, loop)
    cells = rng.choice(np.prod(shape), 40, replace=False)
    assert np.array_equal(P.predict(v, obst, axes, dt, nu, scheme, cells=cells), vec.reshape(-1, 3)[cells])


def test_predict_rejects_thin_axes_and_unknown_scheme():
    axes = _axes_uniform((4, 1, 4))
    with pytest.raises(ValueError):
        P.predict(np.zeros((4, 1, 4, 3), F32), np.zeros((4, 1, 4), np.int32), axes, 1e-3, 0.0)
    axes = _axes_uniform((4, 4, 4))
    with pytest.raises(ValueError):
        P.predict(np.zeros((4, 4, 4, 3), F32), np.zeros((4, 4, 4), np.int32), axes, 1e-3, 0.0, "rk4")
