"""NumPy float32 restatement of the fluid predictor (test oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

Follows ``CartesianFluidSolver._predict_kernel`` (simulation_core/fluid.py:2447-2513)
with its helpers ``_sample_velocity_prev_trilinear`` (2382-2404) and
``_axis_grid_coordinate_device`` (2406-2435): a semi-Lagrangian backtrace
(explicit Euler, or RK2 with a midpoint velocity sample), an obstacle-aware
trilinear sample of the previous velocity at the departure point, and the
explicit FV diffusion term ``dt * nu * lap(u)`` with zero flux through
obstacle and domain faces.  Obstacle cells get zero velocity.

Arithmetic: the reference runs at Taichi ``default_fp=f32``; every quantity
here is float32 and every expression is evaluated in the reference's source
order, one IEEE operation at a time (NumPy neither fuses multiply-adds nor
reassociates).  Divisions are correctly rounded float32 divisions.
``min``/``max`` are ``fmin``/``fmax`` (a NaN operand yields the other one).

Arrays are indexed [i, j, k] with k fastest; velocity is (nx, ny, nz, 3)
float32 and obstacle (nx, ny, nz) int32.  The per-axis metric arrays are the
float32 fields of the solver: faces (n + 1), centres (n), widths (n) and
centre distances (n; entry d > 0 is centre d minus centre d - 1).

Pinning (tests/test_oracle_predict.py): the reference's predictor tests
(uniform velocity preserved on uniform and graded grids, viscous damping of a
shear mode within (0.9, 0.99), RK2 error below 0.55 x Euler for solid-body
rotation, Taylor-Green predict+project error falling with refinement), the
host grid's float64 linear-scan coordinate lookup (grid.py, fluid.py:89-104)
and ``numpy.searchsorted`` for the binary search, closed forms (trilinear
reproduction of affine fields, the FV Laplacian of linear and quadratic
profiles on graded axes, the exact departure point of a uniform flow) and a
scalar per-cell loop on small grids for the vectorised masks and clamps.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

F32 = np.float32
_TINY_WIDTH = F32(1.0e-18)
_TINY_WEIGHT = F32(1.0e-12)


@dataclass
class AxisMetrics:
    """float32 metric arrays of one axis (fluid.py:1013-1024)."""

    faces: np.ndarray
    centers: np.ndarray
    widths: np.ndarray
    center_distances: np.ndarray

    @staticmethod
    def from_arrays(faces, centers, widths, center_distances) -> "AxisMetrics":
        return AxisMetrics(*(np.ascontiguousarray(a, dtype=F32) for a in (faces, centers, widths, center_distances)))

    @staticmethod
    def from_widths(origin: float, widths) -> "AxisMetrics":
        """Metrics the solver derives from float64 widths (fluid.py:66-87), rounded to float32."""
        w = [float(v) for v in widths]
        faces = [float(origin)]
        centers = []
        for v in w:
            centers.append(faces[-1] + 0.5 * v)
            faces.append(faces[-1] + v)
        cdist = [w[0]] + [centers[n] - centers[n - 1] for n in range(1, len(centers))]
        return AxisMetrics.from_arrays(faces, centers, w, cdist)


def grid_coordinate(value, ax: AxisMetrics) -> np.ndarray:
    """Fractional cell coordinate of float32 positions (fluid.py:2406-2435).

    Below the first centre and above the last one the coordinate extrapolates
    with the half width of the end cell; in between, a binary search over the
    centres finds the bracketing pair and the coordinate interpolates linearly.
    """
    v = np.asarray(value, dtype=F32)
    c, f = ax.centers, ax.faces
    n = c.size
    c0, cl = c[0], c[n - 1]
    out = np.zeros(v.shape, dtype=F32)
    low = v <= c0
    high = ~low & (v >= cl)
    mid = ~(low | high)
    if low.any():
        hw = np.fmax(F32(c0 - f[0]), _TINY_WIDTH)
        out[low] = (F32(-0.5) * (c0 - v[low])) / hw
    if high.any():
        hw = np.fmax(F32(f[n] - cl), _TINY_WIDTH)
        out[high] = F32(n - 1) + (F32(0.5) * (v[high] - cl)) / hw
    if mid.any():
        vm = v[mid]
        lower = np.zeros(vm.shape, dtype=np.int64)
        upper = np.full(vm.shape, n - 1, dtype=np.int64)
        while True:
            act = (upper - lower) > 1
            if not act.any():
                break
            m = (lower + upper) // 2
            ge = vm >= c[m]
            lower = np.where(act & ge, m, lower)
            upper = np.where(act & ~ge, m, upper)
        upper = np.minimum(lower + 1, n - 1)
        dist = np.fmax(c[upper] - c[lower], _TINY_WIDTH)
        out[mid] = lower.astype(F32) + (vm - c[lower]) / dist
    return out


def _corner_index(g: np.ndarray, n: int):
    """i0 = clamp(floor(g), 0, n - 2) and t = clamp(g - i0, 0, 1) (fluid.py:2384-2389).

    The clamp is applied to the float floor before the integer conversion,
    which is the reference's result for every finite coordinate and stays
    defined for coordinates beyond the int32 range.
    """
    i0 = np.fmin(np.fmax(np.floor(g), F32(0.0)), F32(n - 2)).astype(np.int64)
    t = np.fmin(np.fmax(g - i0.astype(F32), F32(0.0)), F32(1.0))
    return i0, t


def sample_trilinear(gx, gy, gz, fallback, velocity_prev, obstacle) -> np.ndarray:
    """Obstacle-aware trilinear sample of the previous velocity (fluid.py:2382-2404).

    Corners inside obstacles are dropped and the remaining weights
    renormalised; with no fluid corner (total weight <= 1e-12) the sample is
    ``fallback``.  Corners are visited with the z offset fastest.
    """
    nx, ny, nz = obstacle.shape
    i0, tx = _corner_index(gx, nx)
    j0, ty = _corner_index(gy, ny)
    k0, tz = _corner_index(gz, nz)
    value = np.zeros(np.shape(gx) + (3,), dtype=F32)
    fluid_weight = np.zeros(np.shape(gx), dtype=F32)
    one = F32(1.0)
    for oi in (0, 1):
        wx = one - tx if oi == 0 else tx
        for oj in (0, 1):
            wy = one - ty if oj == 0 else ty
            for ok in (0, 1):
                wz = one - tz if ok == 0 else tz
                weight = (wx * wy) * wz
                ci, cj, ck = i0 + oi, j0 + oj, k0 + ok
                fluid = obstacle[ci, cj, ck] == 0
                contrib = weight[..., None] * velocity_prev[ci, cj, ck]
                value = np.where(fluid[..., None], value + contrib, value)
                fluid_weight = np.where(fluid, fluid_weight + weight, fluid_weight)
    ok = fluid_weight > _TINY_WEIGHT
    safe = np.where(ok, fluid_weight, one)
    return np.where(ok[..., None], value / safe[..., None], np.asarray(fallback, dtype=F32))


def predict(velocity_prev: np.ndarray, obstacle: np.ndarray, axes, dt_s: float, nu_m2_s: float,
            scheme: str = "euler", cells=None) -> np.ndarray:
    """New velocity from ``velocity_prev`` (fluid.py:2447-2513).

    ``axes`` is a sequence of three ``AxisMetrics``.  ``dt_s`` and ``nu_m2_s``
    are rounded to float32 as the Taichi kernel arguments are.  With ``cells``
    (flat indices) only those cells are computed and a (len(cells), 3) array
    is returned; otherwise the full (nx, ny, nz, 3) field.
    """
    vp = np.ascontiguousarray(velocity_prev, dtype=F32)
    ob = np.ascontiguousarray(obstacle, dtype=np.int32)
    nx, ny, nz = ob.shape
    if min(nx, ny, nz) < 2:
        raise ValueError("the predictor needs at least 2 cells along every axis")
    scheme = str(scheme).lower()
    if scheme not in ("euler", "rk2"):
        raise ValueError(f"unsupported advection_scheme: {scheme!r}")
    dt = F32(dt_s)
    nu = F32(nu_m2_s)
    ax, ay, az = axes
    if cells is None:
        flat = np.arange(nx * ny * nz, dtype=np.int64)
    else:
        flat = np.asarray(cells, dtype=np.int64).ravel()
    i, rem = np.divmod(flat, ny * nz)
    j, k = np.divmod(rem, nz)

    center = vp[i, j, k]
    x = ax.centers[i]
    y = ay.centers[j]
    z = az.centers[k]
    trace = center
    if scheme == "rk2":
        half_dt = F32(0.5) * dt
        mgx = grid_coordinate(x - half_dt * center[:, 0], ax)
        mgy = grid_coordinate(y - half_dt * center[:, 1], ay)
        mgz = grid_coordinate(z - half_dt * center[:, 2], az)
        trace = sample_trilinear(mgx, mgy, mgz, center, vp, ob)
    gx = grid_coordinate(x - dt * trace[:, 0], ax)
    gy = grid_coordinate(y - dt * trace[:, 1], ay)
    gz = grid_coordinate(z - dt * trace[:, 2], az)
    advected = sample_trilinear(gx, gy, gz, center, vp, ob)

    zero = np.zeros_like(center)

    def flux(cond, nb_idx, dist, forward):
        nb = vp[nb_idx]
        use = cond & (ob[nb_idx] == 0)
        d = dist[:, None]
        val = (nb - center) / d if forward else (center - nb) / d
        return np.where(use[:, None], val, zero)

    im, ip = np.maximum(i - 1, 0), np.minimum(i + 1, nx - 1)
    jm, jp = np.maximum(j - 1, 0), np.minimum(j + 1, ny - 1)
    km, kp = np.maximum(k - 1, 0), np.minimum(k + 1, nz - 1)
    cdx, cdy, cdz = ax.center_distances, ay.center_distances, az.center_distances
    fxb = flux(i > 0, (im, j, k), cdx[i], False)
    fxf = flux(i < nx - 1, (ip, j, k), cdx[np.minimum(i + 1, nx - 1)], True)
    fyb = flux(j > 0, (i, jm, k), cdy[j], False)
    fyf = flux(j < ny - 1, (i, jp, k), cdy[np.minimum(j + 1, ny - 1)], True)
    fzb = flux(k > 0, (i, j, km), cdz[k], False)
    fzf = flux(k < nz - 1, (i, j, kp), cdz[np.minimum(k + 1, nz - 1)], True)
    lap = ((fxf - fxb) / ax.widths[i][:, None] + (fyf - fyb) / ay.widths[j][:, None]) \
        + (fzf - fzb) / az.widths[k][:, None]
    out = advected + (dt * nu) * lap
    out = np.where((ob[i, j, k] == 1)[:, None], zero, out)
    if cells is None:
        return out.reshape(nx, ny, nz, 3)
    return out


def predict_scalar_loop(velocity_prev, obstacle, axes, dt_s, nu_m2_s, scheme="euler") -> np.ndarray:
    """Cell-by-cell loop over the same kernel with float32 scalars (small grids only).

    A second, independent statement of the masks, clamps and branch structure
    of fluid.py:2447-2513, used to check the vectorised ``predict``.
    """
    vp = np.asarray(velocity_prev, dtype=F32)
    ob = np.asarray(obstacle)
    nx, ny, nz = ob.shape
    dt, nu = F32(dt_s), F32(nu_m2_s)
    ax, ay, az = axes
    out = np.zeros_like(vp)

    def coord(v, a):
        c, f = a.centers, a.faces
        n = c.size
        if v <= c[0]:
            return (F32(-0.5) * (c[0] - v)) / max(F32(c[0] - f[0]), _TINY_WIDTH)
        if v >= c[n - 1]:
            return F32(n - 1) + (F32(0.5) * (v - c[n - 1])) / max(F32(f[n] - c[n - 1]), _TINY_WIDTH)
        lo, hi = 0, n - 1
        while hi - lo > 1:
            m = (lo + hi) // 2
            if v >= c[m]:
                lo = m
            else:
                hi = m
        hi = min(lo + 1, n - 1)
        return F32(lo) + (v - c[lo]) / max(F32(c[hi] - c[lo]), _TINY_WIDTH)

    def sample(g, fallback):
        idx, ts = [], []
        for gv, n in zip(g, (nx, ny, nz)):
            i0 = int(min(max(np.floor(gv), F32(0.0)), F32(n - 2)))
            idx.append(i0)
            ts.append(min(max(F32(gv - F32(i0)), F32(0.0)), F32(1.0)))
        value = np.zeros(3, dtype=F32)
        fw = F32(0.0)
        for oi in (0, 1):
            for oj in (0, 1):
                for ok in (0, 1):
                    wx = F32(1.0) - ts[0] if oi == 0 else ts[0]
                    wy = F32(1.0) - ts[1] if oj == 0 else ts[1]
                    wz = F32(1.0) - ts[2] if ok == 0 else ts[2]
                    w = F32(F32(wx * wy) * wz)
                    if ob[idx[0] + oi, idx[1] + oj, idx[2] + ok] == 0:
                        value = (value + w * vp[idx[0] + oi, idx[1] + oj, idx[2] + ok]).astype(F32)
                        fw = F32(fw + w)
        if fw > _TINY_WEIGHT:
            return (value / fw).astype(F32)
        return np.asarray(fallback, dtype=F32)

    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if ob[i, j, k] == 1:
                    continue
                c = vp[i, j, k]
                pos = (ax.centers[i], ay.centers[j], az.centers[k])
                tv = c
                if scheme == "rk2":
                    hd = F32(F32(0.5) * dt)
                    tv = sample([coord(F32(p - hd * cv), a) for p, cv, a in zip(pos, c, axes)], c)
                adv = sample([coord(F32(p - dt * tvv), a) for p, tvv, a in zip(pos, tv, axes)], c)
                lap = np.zeros(3, dtype=F32)
                for d, (n, a) in enumerate(zip((nx, ny, nz), axes)):
                    cell = [i, j, k]
                    pos_d = cell[d]
                    back = np.zeros(3, dtype=F32)
                    fwd = np.zeros(3, dtype=F32)
                    if pos_d > 0:
                        nb = list(cell)
                        nb[d] -= 1
                        if ob[tuple(nb)] == 0:
                            back = ((c - vp[tuple(nb)]) / a.center_distances[pos_d]).astype(F32)
                    if pos_d < n - 1:
                        nb = list(cell)
                        nb[d] += 1
                        if ob[tuple(nb)] == 0:
                            fwd = ((vp[tuple(nb)] - c) / a.center_distances[pos_d + 1]).astype(F32)
                    lap = (lap + (fwd - back) / a.widths[pos_d]).astype(F32)
                out[i, j, k] = (adv + F32(dt * nu) * lap).astype(F32)
    return out
