"""Host-side grid geometry: re-expression of the reference's grid tests
(tests/test_core_fluid.py:72-366, 704-743) plus the benchmark grid."""

import numpy as np
import pytest

from easyfsi_amd.simulation_core.grid import (
    CartesianGrid,
    FluidDomainSpec,
    GradedGridSpec,
    RefinementRegion,
    build_graded_grid,
    coarsen_axis_widths,
    explicit_graded_axis,
    multigrid_axis_widths,
    multigrid_shapes,
)


def _max_ratio(widths):
    return max(max(a, b) / min(a, b) for a, b in zip(widths, widths[1:]))


def _count_between(centers, lo, hi):
    return sum(1 for c in centers if lo <= c <= hi)


def test_uniform_spec_builds_grid():  # test_core_fluid.py:72-88
    spec = FluidDomainSpec(bounds_min_m=(1.0, 2.0, 3.0), bounds_max_m=(2.0, 4.0, 6.0), grid_nodes=(4, 5, 6),
                           density_kgm3=1000.0, viscosity_pa_s=1.0e-3, dt_s=1.0e-3)
    g = spec.cartesian_grid
    assert g.grid_nodes == (4, 5, 6)
    assert g.uniform_spacing_m == (0.25, 0.4, 0.5)
    np.testing.assert_allclose(g.cell_centers_x_m, (1.125, 1.375, 1.625, 1.875))
    np.testing.assert_allclose(g.center_distances_z_m, (0.5,) * 6)


def test_nonuniform_coordinate_lookup():  # test_core_fluid.py:112-128
    g = CartesianGrid(bounds_min_m=(0.0, 0.0, 0.0), cell_widths_x_m=(0.5, 1.0, 1.5, 1.0),
                      cell_widths_y_m=(1.0,) * 4, cell_widths_z_m=(1.0,) * 4)
    np.testing.assert_allclose(g.cell_faces_x_m, (0.0, 0.5, 1.5, 3.0, 4.0))
    np.testing.assert_allclose(g.cell_centers_x_m, (0.25, 1.0, 2.25, 3.5))
    assert g.grid_coordinate_x(0.25) == pytest.approx(0.0)
    assert g.grid_coordinate_x(1.0) == pytest.approx(1.0)
    assert g.grid_coordinate_x(2.25) == pytest.approx(2.0)
    assert g.grid_coordinate_x(0.0) == pytest.approx(-0.5)
    assert g.grid_coordinate_x(4.0) == pytest.approx(3.5)
    assert g.grid_coordinate_x(1.5) == pytest.approx(1.4)


def test_uniform_coordinate_lookup_matches_formula():  # test_core_fluid.py:130-137
    g = CartesianGrid.uniform(bounds_min_m=(-1.0, 0.0, 0.0), bounds_max_m=(1.0, 1.0, 1.0), grid_nodes=(8, 4, 4))
    for x in (-1.0, -0.875, -0.125, 0.0, 0.875, 1.0):
        assert g.grid_coordinate_x(x) == pytest.approx((x + 1.0) / 0.25 - 0.5)


def test_graded_without_regions_is_uniform():  # test_core_fluid.py:139-150
    g = build_graded_grid(GradedGridSpec(bounds_min_m=(0, 0, 0), bounds_max_m=(0.012,) * 3,
                                         farfield_spacing_m=0.003, max_growth_ratio=1.2))
    assert g.is_uniform
    assert g.grid_nodes == (4, 4, 4)
    np.testing.assert_allclose(g.uniform_spacing_m, (0.003,) * 3)


def test_graded_rejects_cell_count_above_limit():  # test_core_fluid.py:152-162
    with pytest.raises(ValueError, match="graded grid cell count 64 exceeds max_cells 63"):
        build_graded_grid(GradedGridSpec(bounds_min_m=(0, 0, 0), bounds_max_m=(0.012,) * 3,
                                         farfield_spacing_m=0.003, max_growth_ratio=1.2, max_cells=63))


def test_graded_resolves_nozzle_region():  # test_core_fluid.py:164-206
    r = 0.003
    h = r / 5.0
    g = build_graded_grid(GradedGridSpec(
        bounds_min_m=(-0.018, -0.018, 0.0), bounds_max_m=(0.018, 0.018, 0.060), farfield_spacing_m=0.003,
        max_growth_ratio=1.2,
        refinement_regions=(RefinementRegion(bounds_min_m=(-r, -r, 0.020), bounds_max_m=(r, r, 0.040),
                                             target_spacing_m=h),)))
    assert _count_between(g.cell_centers_x_m, -r, r) >= 10
    assert _count_between(g.cell_centers_y_m, -r, r) >= 10
    assert _count_between(g.cell_centers_z_m, 0.020, 0.040) >= 33
    refined = [w for c, w in zip(g.cell_centers_x_m, g.cell_widths_x_m) if -r <= c <= r]
    assert max(refined) <= h * 1.0000001
    for a in (g.cell_widths_x_m, g.cell_widths_y_m, g.cell_widths_z_m):
        assert _max_ratio(a) <= 1.2 + 1.0e-9
    assert max(g.cell_widths_x_m) <= 0.003 * 1.0000001


def test_graded_extends_refinement_over_boundary_sliver():  # test_core_fluid.py:208-224
    g = build_graded_grid(GradedGridSpec(
        bounds_min_m=(0, 0, 0), bounds_max_m=(1, 1, 1), farfield_spacing_m=0.30, max_growth_ratio=1.2,
        refinement_regions=(RefinementRegion(bounds_min_m=(0.05, 0.2, 0.2), bounds_max_m=(0.4, 0.8, 0.8),
                                             target_spacing_m=0.10),)))
    assert _max_ratio(g.cell_widths_x_m) <= 1.2 + 1.0e-12
    assert min(g.cell_widths_x_m) >= 0.10 / 1.2


def test_graded_side_transition_preserves_growth():  # test_core_fluid.py:226-246
    g = build_graded_grid(GradedGridSpec(
        bounds_min_m=(0, 0, 0), bounds_max_m=(0.0057, 1.0, 1.0), farfield_spacing_m=(0.001, 1.0, 1.0),
        max_growth_ratio=1.2,
        refinement_regions=(RefinementRegion(bounds_min_m=(0.0012, 0, 0), bounds_max_m=(0.0037, 1, 1),
                                             target_spacing_m=(0.0005, 1.0, 1.0)),)))
    assert _max_ratio(g.cell_widths_x_m) <= 1.2 + 1.0e-12


def test_graded_growth_between_regions():  # test_core_fluid.py:248-276
    h = 1.0e-3
    regions = tuple(RefinementRegion(bounds_min_m=(a, 0, 0), bounds_max_m=(b, 0.01, 0.01),
                                     target_spacing_m=(h, 0.01, 0.01)) for a, b in ((0.020, 0.025), (0.040, 0.045)))
    g = build_graded_grid(GradedGridSpec(bounds_min_m=(0, 0, 0), bounds_max_m=(0.10, 0.01, 0.01),
                                         farfield_spacing_m=(0.01,) * 3, max_growth_ratio=1.2,
                                         refinement_regions=regions))
    assert min(g.cell_widths_x_m) >= h / 1.2
    assert _max_ratio(g.cell_widths_x_m) <= 1.2 + 1.0e-12


def test_graded_merges_unbridgeable_gaps():  # test_core_fluid.py:278-322
    fine, coarse = 1.0e-3, 1.0e-2
    right1 = 0.025
    left2 = right1 + fine / 1.2 * 1.1
    g = build_graded_grid(GradedGridSpec(
        bounds_min_m=(0, 0, 0), bounds_max_m=(0.10, 0.01, 0.01), farfield_spacing_m=(coarse, 0.01, 0.01),
        max_growth_ratio=1.2,
        refinement_regions=(
            RefinementRegion(bounds_min_m=(0.020, 0, 0), bounds_max_m=(right1, 0.01, 0.01),
                             target_spacing_m=(fine, 0.01, 0.01)),
            RefinementRegion(bounds_min_m=(left2, 0, 0), bounds_max_m=(left2 + coarse, 0.01, 0.01),
                             target_spacing_m=(coarse, 0.01, 0.01)))))
    assert _max_ratio(g.cell_widths_x_m) <= 1.2 + 1.0e-12
    assert _count_between(g.cell_centers_x_m, right1, left2) >= 1


def test_multigrid_coarse_widths_are_pairwise_sums():  # test_core_fluid.py:324-366
    g = CartesianGrid(bounds_min_m=(0, 0, 0), cell_widths_x_m=(0.1, 0.2, 0.3, 0.4),
                      cell_widths_y_m=(0.25,) * 4, cell_widths_z_m=(0.25,) * 4)
    shapes = multigrid_shapes(g.grid_nodes)
    levels = multigrid_axis_widths(g, shapes)
    np.testing.assert_allclose(levels[1][0], (0.3, 0.7))
    assert coarsen_axis_widths((1.0, 2.0, 3.0)) == (3.0, 3.0)


def test_multigrid_shapes_halve_to_two():
    assert multigrid_shapes((256, 256, 256)) == tuple((n, n, n) for n in (256, 128, 64, 32, 16, 8, 4, 2))
    assert multigrid_shapes((5, 4, 4)) == ((5, 4, 4), (3, 2, 2))
    assert multigrid_shapes((3, 8, 8)) == ((3, 8, 8),)


def test_spec_validation_messages():  # fluid.py:596-647
    with pytest.raises(ValueError, match="grid_nodes must be at least 4"):
        FluidDomainSpec.unit_box(grid_nodes=(3, 4, 4))
    with pytest.raises(ValueError, match="mutually exclusive"):
        g = CartesianGrid.uniform(bounds_min_m=(0, 0, 0), bounds_max_m=(1, 1, 1), grid_nodes=(4, 4, 4))
        FluidDomainSpec(bounds_min_m=(0, 0, 0), bounds_max_m=(1, 1, 1), grid_nodes=None, density_kgm3=1.0,
                        viscosity_pa_s=0.0, dt_s=1.0, cartesian_grid=g,
                        graded_grid=GradedGridSpec(bounds_min_m=(0, 0, 0), bounds_max_m=(1, 1, 1),
                                                   farfield_spacing_m=0.25, max_growth_ratio=1.2))


def test_benchmark_graded_axis():
    """256-cell axis: 128-cell core of width h, 64 cells each side at h*1.04^k."""
    w = explicit_graded_axis(128, 1.0, 64, 1.04)
    assert len(w) == 256
    assert sum(w) == pytest.approx(128 + 2 * sum(1.04 ** k for k in range(1, 65)))
    assert _max_ratio(w) == pytest.approx(1.04)
    assert max(w) / min(w) == pytest.approx(1.04 ** 64)
