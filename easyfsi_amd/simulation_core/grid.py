"""Cartesian and graded MAC-grid geometry (host side).

Widths, faces, centres and centre distances of the cell-centred grid that the
pressure projection runs on.  Everything here is plain Python float64 work done
once per grid; the device kernels only ever see the per-axis 1D metric arrays
built from it.

Reference: ``simulation_core/fluid.py:20-647`` (``CartesianGrid``,
``RefinementRegion``, ``GradedGridSpec``, ``build_graded_grid`` and its
helpers, ``FluidDomainSpec``).  The graded-axis construction has to reproduce
the reference's widths bit for bit, so the floating-point operation order of
each step (geometric bounds, interpolation between them, left-to-right sums)
follows the reference; the code organisation is this package's own.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

_REL_EPS = 1.0e-10  # relative slack used by every feasibility test of the builder


def _as_triple(value, name: str) -> tuple[float, float, float]:
    if isinstance(value, tuple):
        out = tuple(float(v) for v in value)
        if len(out) != 3:
            raise ValueError(f"{name} tuple must have three values")
        return out  # type: ignore[return-value]
    return (float(value),) * 3


# ---------------------------------------------------------------------------
# 1D axis helpers (fluid.py:66-104)
# ---------------------------------------------------------------------------

def axis_faces(origin: float, widths: Sequence[float]) -> tuple[float, ...]:
    """Face coordinates: origin followed by running sums (fluid.py:75-80)."""
    out = [float(origin)]
    for w in widths:
        out.append(out[-1] + w)
    return tuple(out)


def axis_centers(origin: float, widths: Sequence[float]) -> tuple[float, ...]:
    """Cell centres: running face + half width (fluid.py:66-73)."""
    out = []
    face = float(origin)
    for w in widths:
        out.append(face + 0.5 * w)
        face += w
    return tuple(out)


def axis_center_distances(centers: Sequence[float], widths: Sequence[float]) -> tuple[float, ...]:
    """Distance to the previous centre; entry 0 is the first width (fluid.py:82-87)."""
    out = [float(widths[0])]
    out.extend(float(centers[n] - centers[n - 1]) for n in range(1, len(centers)))
    return tuple(out)


def _axis_grid_coordinate(value: float, faces: Sequence[float], centers: Sequence[float]) -> float:
    """Fractional cell coordinate of ``value``; centres map to integers (fluid.py:89-104)."""
    if value <= centers[0]:
        return -0.5 * (centers[0] - value) / max(centers[0] - faces[0], 1.0e-18)
    if value >= centers[-1]:
        return float(len(centers) - 1) + 0.5 * (value - centers[-1]) / max(faces[-1] - centers[-1], 1.0e-18)
    for n in range(len(centers) - 1):
        lo, hi = centers[n], centers[n + 1]
        if lo <= value <= hi:
            return float(n) + (value - lo) / max(hi - lo, 1.0e-18)
    return 0.0


def _axis_is_uniform(widths: Sequence[float]) -> bool:
    w0 = widths[0]
    tol = max(abs(w0), 1.0) * 1.0e-12
    return all(abs(w - w0) <= tol for w in widths)


# ---------------------------------------------------------------------------
# Grid descriptions
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class CartesianGrid:
    """Tensor-product grid given by per-axis cell widths (fluid.py:20-182)."""

    bounds_min_m: tuple[float, float, float]
    cell_widths_x_m: tuple[float, ...]
    cell_widths_y_m: tuple[float, ...]
    cell_widths_z_m: tuple[float, ...]

    @classmethod
    def uniform(cls, *, bounds_min_m, bounds_max_m, grid_nodes) -> "CartesianGrid":
        if any(n <= 0 for n in grid_nodes):
            raise ValueError("grid_nodes must be positive in every dimension")
        if len(grid_nodes) != 3 or len(bounds_min_m) != 3 or len(bounds_max_m) != 3:
            raise ValueError("bounds_min_m must have three coordinates")
        per_axis = [((hi - lo) / n, int(n)) for lo, hi, n in zip(bounds_min_m, bounds_max_m, grid_nodes)]
        return cls(
            bounds_min_m=tuple(float(v) for v in bounds_min_m),
            cell_widths_x_m=(float(per_axis[0][0]),) * per_axis[0][1],
            cell_widths_y_m=(float(per_axis[1][0]),) * per_axis[1][1],
            cell_widths_z_m=(float(per_axis[2][0]),) * per_axis[2][1],
        )

    def __post_init__(self) -> None:
        lo = tuple(float(v) for v in self.bounds_min_m)
        axes = [tuple(float(v) for v in w) for w in (self.cell_widths_x_m, self.cell_widths_y_m, self.cell_widths_z_m)]
        if len(lo) != 3:
            raise ValueError("bounds_min_m must have three coordinates")
        if any(len(a) == 0 for a in axes):
            raise ValueError("CartesianGrid axes must contain at least one cell")
        if any(w <= 0.0 for a in axes for w in a):
            raise ValueError("CartesianGrid cell widths must be positive")
        object.__setattr__(self, "bounds_min_m", lo)
        object.__setattr__(self, "cell_widths_x_m", axes[0])
        object.__setattr__(self, "cell_widths_y_m", axes[1])
        object.__setattr__(self, "cell_widths_z_m", axes[2])

    # -- per-axis views ---------------------------------------------------
    def axis_widths(self, axis: int) -> tuple[float, ...]:
        return (self.cell_widths_x_m, self.cell_widths_y_m, self.cell_widths_z_m)[axis]

    @property
    def grid_nodes(self) -> tuple[int, int, int]:
        return (len(self.cell_widths_x_m), len(self.cell_widths_y_m), len(self.cell_widths_z_m))

    @property
    def bounds_max_m(self) -> tuple[float, float, float]:
        return tuple(self.bounds_min_m[a] + sum(self.axis_widths(a)) for a in range(3))  # type: ignore[return-value]

    @property
    def cell_centers_x_m(self):
        return axis_centers(self.bounds_min_m[0], self.cell_widths_x_m)

    @property
    def cell_centers_y_m(self):
        return axis_centers(self.bounds_min_m[1], self.cell_widths_y_m)

    @property
    def cell_centers_z_m(self):
        return axis_centers(self.bounds_min_m[2], self.cell_widths_z_m)

    @property
    def cell_faces_x_m(self):
        return axis_faces(self.bounds_min_m[0], self.cell_widths_x_m)

    @property
    def cell_faces_y_m(self):
        return axis_faces(self.bounds_min_m[1], self.cell_widths_y_m)

    @property
    def cell_faces_z_m(self):
        return axis_faces(self.bounds_min_m[2], self.cell_widths_z_m)

    @property
    def center_distances_x_m(self):
        return axis_center_distances(self.cell_centers_x_m, self.cell_widths_x_m)

    @property
    def center_distances_y_m(self):
        return axis_center_distances(self.cell_centers_y_m, self.cell_widths_y_m)

    @property
    def center_distances_z_m(self):
        return axis_center_distances(self.cell_centers_z_m, self.cell_widths_z_m)

    @property
    def is_uniform(self) -> bool:
        return all(_axis_is_uniform(self.axis_widths(a)) for a in range(3))

    @property
    def uniform_spacing_m(self) -> tuple[float, float, float]:
        if not self.is_uniform:
            raise ValueError("uniform_spacing_m is only available for a uniform CartesianGrid")
        return (self.cell_widths_x_m[0], self.cell_widths_y_m[0], self.cell_widths_z_m[0])

    def grid_coordinate_x(self, x_m: float) -> float:
        return _axis_grid_coordinate(float(x_m), self.cell_faces_x_m, self.cell_centers_x_m)

    def grid_coordinate_y(self, y_m: float) -> float:
        return _axis_grid_coordinate(float(y_m), self.cell_faces_y_m, self.cell_centers_y_m)

    def grid_coordinate_z(self, z_m: float) -> float:
        return _axis_grid_coordinate(float(z_m), self.cell_faces_z_m, self.cell_centers_z_m)


@dataclass(frozen=True)
class RefinementRegion:
    """Box whose cells must be no wider than ``target_spacing_m`` (fluid.py:185-209)."""

    bounds_min_m: tuple[float, float, float]
    bounds_max_m: tuple[float, float, float]
    target_spacing_m: float | tuple[float, float, float]

    def __post_init__(self) -> None:
        lo = tuple(float(v) for v in self.bounds_min_m)
        hi = tuple(float(v) for v in self.bounds_max_m)
        if len(lo) != 3 or len(hi) != 3:
            raise ValueError("RefinementRegion bounds must have three coordinates")
        if any(b <= a for a, b in zip(lo, hi)):
            raise ValueError("RefinementRegion bounds_max_m must be greater than bounds_min_m")
        target = _as_triple(self.target_spacing_m, "target_spacing_m")
        if any(v <= 0.0 for v in target):
            raise ValueError("target_spacing_m must be positive")
        object.__setattr__(self, "bounds_min_m", lo)
        object.__setattr__(self, "bounds_max_m", hi)
        object.__setattr__(self, "target_spacing_m", target)


@dataclass(frozen=True)
class GradedGridSpec:
    """Farfield spacing, growth limit and refinement boxes of a graded grid (fluid.py:212-246)."""

    bounds_min_m: tuple[float, float, float]
    bounds_max_m: tuple[float, float, float]
    farfield_spacing_m: float | tuple[float, float, float]
    max_growth_ratio: float
    refinement_regions: tuple[RefinementRegion, ...] = ()
    max_cells: int | None = None

    def __post_init__(self) -> None:
        lo = tuple(float(v) for v in self.bounds_min_m)
        hi = tuple(float(v) for v in self.bounds_max_m)
        if len(lo) != 3 or len(hi) != 3:
            raise ValueError("GradedGridSpec bounds must have three coordinates")
        if any(b <= a for a, b in zip(lo, hi)):
            raise ValueError("bounds_max_m must be greater than bounds_min_m")
        far = _as_triple(self.farfield_spacing_m, "farfield_spacing_m")
        if any(v <= 0.0 for v in far):
            raise ValueError("farfield_spacing_m must be positive")
        if self.max_growth_ratio <= 1.0:
            raise ValueError("max_growth_ratio must be greater than 1")
        cap = None if self.max_cells is None else int(self.max_cells)
        if cap is not None and cap <= 0:
            raise ValueError("max_cells must be positive when provided")
        object.__setattr__(self, "bounds_min_m", lo)
        object.__setattr__(self, "bounds_max_m", hi)
        object.__setattr__(self, "farfield_spacing_m", far)
        object.__setattr__(self, "refinement_regions", tuple(self.refinement_regions))
        object.__setattr__(self, "max_cells", cap)


# ---------------------------------------------------------------------------
# Graded-axis construction (fluid.py:249-536)
# ---------------------------------------------------------------------------

def _uniform_fill_faces(left: float, right: float, spacing: float) -> list[float]:
    """Interior faces of the fewest equal cells no wider than ``spacing`` (fluid.py:249-257)."""
    length = right - left
    if length <= 0.0:
        return []
    n = max(1, int(round(length / spacing)))
    while length / n > spacing * (1.0 + _REL_EPS):
        n += 1
    w = length / n
    return [left + w * m for m in range(1, n)]


def _uniform_axis(lo: float, hi: float, spacing: float) -> tuple[float, ...]:
    faces = [lo] + _uniform_fill_faces(lo, hi, spacing) + [hi]
    return tuple(faces[m + 1] - faces[m] for m in range(len(faces) - 1))


def _growth_ok(widths: Sequence[float], growth: float, left: float | None = None,
               right: float | None = None) -> bool:
    """Every neighbour ratio (including the boundary widths) within ``growth`` (fluid.py:315-335)."""
    seq: list[float] = ([] if left is None else [float(left)]) + [float(w) for w in widths]
    if right is not None:
        seq.append(float(right))
    limit = float(growth) * (1.0 + _REL_EPS)
    for a, b in zip(seq, seq[1:]):
        if a <= 0.0 or b <= 0.0 or max(a / b, b / a) > limit:
            return False
    return True


def _fit_between_bounds(distance: float, lower: list[float], upper: list[float]) -> tuple[float, ...] | None:
    """Blend per-cell lower/upper width bounds so the widths sum to ``distance``.

    Returns None when ``distance`` lies outside [sum(lower), sum(upper)] (with the
    builder's relative slack).  Blend weight and sums follow fluid.py:289-296.
    """
    s_lo = sum(lower)
    s_hi = sum(upper)
    if distance < s_lo * (1.0 - _REL_EPS) or distance > s_hi * (1.0 + _REL_EPS):
        return None
    if s_hi <= s_lo:
        return tuple(lower)
    t = (float(distance) - s_lo) / (s_hi - s_lo)
    return tuple(a + t * (b - a) for a, b in zip(lower, upper))


def _search_limit(distance: float, *anchors: float) -> int:
    smallest = max(min(anchors), 1.0e-12)
    return max(64, int(distance / smallest) * 8 + 64)


def graded_side_widths(*, distance: float, inner_width: float, farfield_spacing: float,
                       max_growth_ratio: float) -> tuple[float, ...]:
    """Widths growing away from a refined segment towards a wall (fluid.py:260-304).

    Tries 1, 2, ... cells; cell m (1-based) is bounded by inner/g^m below and by
    min(farfield, inner*g^m) above.  The first count whose bounds bracket the
    distance and whose blended widths respect the growth limit wins.
    """
    if distance <= 0.0:
        return ()
    g = float(max_growth_ratio)
    inner = float(inner_width)
    far = float(farfield_spacing)
    for n in range(1, _search_limit(distance, inner_width, farfield_spacing) + 1):
        lower: list[float] = []
        upper: list[float] = []
        for m in range(1, n + 1):
            lo_w = inner / (g ** m)
            hi_w = min(far, inner * (g ** m))
            if lo_w > hi_w * (1.0 + _REL_EPS):
                break
            lower.append(lo_w)
            upper.append(hi_w)
        if len(lower) != n:
            continue
        widths = _fit_between_bounds(distance, lower, upper)
        if widths is not None and _growth_ok(widths, max_growth_ratio, left=inner_width):
            return widths
    raise ValueError("could not build graded side satisfying max_growth_ratio")


def graded_bridge_widths(*, distance: float, left_width: float, right_width: float,
                         farfield_spacing: float, max_growth_ratio: float) -> tuple[float, ...]:
    """Widths joining two refined segments of different fine widths (fluid.py:338-403)."""
    if distance <= 0.0:
        return ()
    g = float(max_growth_ratio)
    far = float(farfield_spacing)

    def capped_growth(start: float, n: int) -> list[float]:
        out, w = [], float(start)
        for _ in range(n):
            w = min(far, w * g)
            out.append(w)
        return out

    for n in range(1, _search_limit(distance, left_width, right_width, farfield_spacing) + 1):
        from_left = capped_growth(left_width, n)
        from_right = capped_growth(right_width, n)
        lower: list[float] = []
        upper: list[float] = []
        ok = True
        for m in range(n):
            lo_w = max(float(left_width) / (g ** (m + 1)), float(right_width) / (g ** (n - m)))
            hi_w = min(far, from_left[m], from_right[n - m - 1])
            if lo_w > hi_w * (1.0 + _REL_EPS):
                ok = False
                break
            lower.append(lo_w)
            upper.append(hi_w)
        if not ok:
            continue
        widths = _fit_between_bounds(distance, lower, upper)
        if widths is not None and _growth_ok(widths, max_growth_ratio, left=left_width, right=right_width):
            return widths
    raise ValueError("could not build graded bridge satisfying max_growth_ratio")


@dataclass(frozen=True)
class _FineSegment:
    left: float
    right: float
    fine_width: float
    fine_cells: int


def _merge_refinement_intervals(*, lo: float, hi: float, farfield: float, growth: float,
                                intervals, tol: float) -> tuple[_FineSegment, ...]:
    """Clip, snap, sort and merge refinement intervals on one axis (fluid.py:406-472).

    Intervals ending closer to a wall than one coarsened cell snap to the wall;
    neighbours closer than the smaller width/growth, or whose gap no bridge can
    span, merge (taking the finer spacing).
    """
    clipped: list[tuple[float, float, float]] = []
    for a, b, h in intervals:
        left = max(lo, float(a))
        right = min(hi, float(b))
        h = float(h)
        if right <= lo or left >= hi or right <= left:
            continue
        snap = h / float(growth)
        if left - lo < snap:
            left = lo
        if hi - right < snap:
            right = hi
        clipped.append((left, right, h))
    if not clipped:
        return ()
    clipped.sort(key=lambda iv: (iv[0], iv[1]))

    merged: list[list[float]] = []
    for left, right, h in clipped:
        if merged:
            prev = merged[-1]
            gap = left - prev[1]
            join = left <= prev[1] + max(tol, min(prev[2], h) / float(growth))
            if not join and gap > 0.0:
                try:
                    graded_bridge_widths(distance=gap, left_width=prev[2], right_width=h,
                                         farfield_spacing=farfield, max_growth_ratio=growth)
                except ValueError:
                    join = True
            if join:
                prev[1] = max(prev[1], right)
                prev[2] = min(prev[2], h)
                continue
        merged.append([left, right, h])

    out = []
    for left, right, h in merged:
        n = max(1, int(round((right - left) / h)))
        while (right - left) / n > h:
            n += 1
        out.append(_FineSegment(left=left, right=right, fine_width=(right - left) / n, fine_cells=n))
    return tuple(out)


def build_graded_axis(*, bounds_min: float, bounds_max: float, farfield_spacing: float,
                      max_growth_ratio: float, intervals) -> tuple[float, ...]:
    """One axis: wall side, fine segments with bridges, far side (fluid.py:475-536)."""
    tol = max(abs(bounds_max - bounds_min), 1.0) * 1.0e-12
    segments = ()
    if intervals:
        segments = _merge_refinement_intervals(lo=bounds_min, hi=bounds_max, farfield=farfield_spacing,
                                               growth=max_growth_ratio, intervals=intervals, tol=tol)
    if not segments:
        return _uniform_axis(bounds_min, bounds_max, farfield_spacing)

    first = segments[0]
    widths: list[float] = list(reversed(graded_side_widths(
        distance=first.left - bounds_min, inner_width=first.fine_width,
        farfield_spacing=farfield_spacing, max_growth_ratio=max_growth_ratio)))
    widths += [first.fine_width] * first.fine_cells
    prev = first
    for seg in segments[1:]:
        widths += graded_bridge_widths(distance=seg.left - prev.right, left_width=prev.fine_width,
                                       right_width=seg.fine_width, farfield_spacing=farfield_spacing,
                                       max_growth_ratio=max_growth_ratio)
        widths += [seg.fine_width] * seg.fine_cells
        prev = seg
    widths += graded_side_widths(distance=bounds_max - prev.right, inner_width=prev.fine_width,
                                 farfield_spacing=farfield_spacing, max_growth_ratio=max_growth_ratio)
    return tuple(widths)


def build_graded_grid(spec: GradedGridSpec) -> CartesianGrid:
    """Graded tensor grid of ``spec`` (fluid.py:539-570)."""
    per_axis = [[(r.bounds_min_m[a], r.bounds_max_m[a], r.target_spacing_m[a]) for r in spec.refinement_regions]
                for a in range(3)]
    widths = [build_graded_axis(bounds_min=spec.bounds_min_m[a], bounds_max=spec.bounds_max_m[a],
                                farfield_spacing=spec.farfield_spacing_m[a],
                                max_growth_ratio=float(spec.max_growth_ratio), intervals=tuple(per_axis[a]))
              for a in range(3)]
    count = len(widths[0]) * len(widths[1]) * len(widths[2])
    if spec.max_cells is not None and count > int(spec.max_cells):
        raise ValueError(f"graded grid cell count {count} exceeds max_cells {int(spec.max_cells)}")
    return CartesianGrid(bounds_min_m=spec.bounds_min_m, cell_widths_x_m=widths[0],
                         cell_widths_y_m=widths[1], cell_widths_z_m=widths[2])


def explicit_graded_axis(core_cells: int, core_width: float, side_cells: int, growth: float) -> tuple[float, ...]:
    """Core of equal cells with geometric sides: w*g^k, k=1..side_cells on each side.

    The 256-cell graded benchmark axis is ``explicit_graded_axis(128, h, 64, 1.04)``.
    """
    side = [core_width * growth ** k for k in range(1, side_cells + 1)]
    return tuple(reversed(side)) + (float(core_width),) * core_cells + tuple(side)


@dataclass(frozen=True)
class FluidDomainSpec:
    """Fluid box, material constants and its grid (fluid.py:573-647)."""

    bounds_min_m: tuple[float, float, float]
    bounds_max_m: tuple[float, float, float]
    grid_nodes: tuple[int, int, int] | None
    density_kgm3: float
    viscosity_pa_s: float
    dt_s: float
    cartesian_grid: CartesianGrid | None = None
    graded_grid: GradedGridSpec | None = None

    @classmethod
    def unit_box(cls, grid_nodes=(32, 32, 32), density_kgm3=1000.0, viscosity_pa_s=1.0e-3, dt_s=1.0e-3):
        return cls(bounds_min_m=(0.0, 0.0, 0.0), bounds_max_m=(1.0, 1.0, 1.0), grid_nodes=grid_nodes,
                   density_kgm3=density_kgm3, viscosity_pa_s=viscosity_pa_s, dt_s=dt_s)

    def __post_init__(self) -> None:
        if any(b <= a for a, b in zip(self.bounds_min_m, self.bounds_max_m)):
            raise ValueError("bounds_max_m must be greater than bounds_min_m")
        if self.density_kgm3 <= 0.0:
            raise ValueError("density_kgm3 must be positive")
        if self.viscosity_pa_s < 0.0:
            raise ValueError("viscosity_pa_s must be non-negative")
        if self.dt_s <= 0.0:
            raise ValueError("dt_s must be positive")
        if self.cartesian_grid is not None and self.graded_grid is not None:
            raise ValueError("cartesian_grid and graded_grid are mutually exclusive")
        if self.graded_grid is not None:
            grid = build_graded_grid(self.graded_grid)
        elif self.cartesian_grid is not None:
            grid = self.cartesian_grid
        else:
            if self.grid_nodes is None:
                raise ValueError("grid_nodes is required when no cartesian_grid or graded_grid is provided")
            grid = CartesianGrid.uniform(bounds_min_m=self.bounds_min_m, bounds_max_m=self.bounds_max_m,
                                         grid_nodes=self.grid_nodes)
        nodes = grid.grid_nodes if self.grid_nodes is None else tuple(int(n) for n in self.grid_nodes)
        if any(n < 4 for n in nodes):
            raise ValueError("grid_nodes must be at least 4 in every dimension")
        if grid.grid_nodes != nodes:
            raise ValueError("cartesian_grid grid_nodes must match FluidDomainSpec.grid_nodes")
        for got, want in zip(grid.bounds_min_m, self.bounds_min_m):
            if abs(got - want) > max(abs(want), 1.0) * 1.0e-12:
                raise ValueError("cartesian_grid bounds_min_m must match FluidDomainSpec.bounds_min_m")
        for got, want in zip(grid.bounds_max_m, self.bounds_max_m):
            if abs(got - want) > max(abs(want), 1.0) * 1.0e-12:
                raise ValueError("cartesian_grid bounds_max_m must match FluidDomainSpec.bounds_max_m")
        object.__setattr__(self, "grid_nodes", nodes)
        object.__setattr__(self, "cartesian_grid", grid)

    @property
    def spacing_m(self) -> tuple[float, float, float]:
        return self.cartesian_grid.uniform_spacing_m

    @property
    def cell_volume_m3(self) -> float:
        dx, dy, dz = self.spacing_m
        return dx * dy * dz


# ---------------------------------------------------------------------------
# Multigrid hierarchy geometry (fluid.py:692-734)
# ---------------------------------------------------------------------------

def multigrid_shapes(shape: tuple[int, int, int]) -> tuple[tuple[int, int, int], ...]:
    """Halve (rounding up, at least 2) while every axis is above 3 (fluid.py:692-701)."""
    shapes = [tuple(int(n) for n in shape)]
    while min(shapes[-1]) > 3:
        nxt = tuple(max(2, (n + 1) // 2) for n in shapes[-1])
        if nxt == shapes[-1]:
            break
        shapes.append(nxt)
    return tuple(shapes)  # type: ignore[return-value]


def coarsen_axis_widths(widths: Sequence[float]) -> tuple[float, ...]:
    """Pairwise sums; an odd last cell stays alone (fluid.py:704-708)."""
    return tuple(sum(widths[n:n + 2]) for n in range(0, len(widths), 2))


def multigrid_axis_widths(grid: CartesianGrid, shapes) -> tuple[tuple[tuple[float, ...], ...], ...]:
    """Per-level per-axis widths, checked against the level shapes (fluid.py:711-729)."""
    levels = []
    cur = tuple(grid.axis_widths(a) for a in range(3))
    for shp in shapes:
        if tuple(len(a) for a in cur) != tuple(shp):
            raise ValueError("multigrid axis widths do not match multigrid shapes")
        levels.append(cur)
        cur = tuple(coarsen_axis_widths(a) for a in cur)
    return tuple(levels)
