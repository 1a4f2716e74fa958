"""Pins the sharp-interface coupling part of the fluid oracle (oracle/fluid.py:
interface_couplings, coupling_apply, interface_terms_report,
interface_projection_divergence and pcg(couplings=...)).

The reference's tests (tests/test_core_fluid.py:1756-1896, 2139-2179) are
re-expressed against the oracle; V-weighted self-adjointness of the coupled
operator, scipy.sparse spsolve for the coupled system and the report's
counting rules pin the rest.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fluid as F


def _unit(n):
    return [np.full(n, 1.0 / n)] * 3


def _lvl(widths, obst, diag=None):
    return F.make_level(widths, obst, iface_diag=diag)


def _slots(shape):
    return dict(slot_active=np.zeros(shape, np.int32), slot_neighbor=np.zeros(shape + (3,), np.int32),
                slot_coef=np.zeros(shape), extra_neighbor=np.zeros(shape + (F.COUPLING_SLOTS - 1, 3), np.int32),
                extra_coef=np.zeros(shape + (F.COUPLING_SLOTS - 1,)))


def _coupled_apply(p, widths, obst, diag, fields):
    c = F.interface_couplings(obst, widths, **fields)
    return F.fv_apply(p, _lvl(widths, obst, diag), False) + F.coupling_apply(p, c)


def test_coupling_enters_operator_symmetrically():  # test_core_fluid.py:1756-1790
    shape = (5, 5, 5)
    w = _unit(5)
    obst = np.zeros(shape, np.int32)
    p = np.ones(shape)
    p[2, 2, 2], p[2, 2, 3] = 3.0, 1.0
    base = F.fv_apply(p, _lvl(w, obst), False)
    diag = np.zeros(shape)
    diag[2, 2, 2] = diag[2, 2, 3] = 4.0
    f = _slots(shape)
    f["slot_active"][2, 2, 2] = 1
    f["slot_neighbor"][2, 2, 2] = (2, 2, 3)
    f["slot_coef"][2, 2, 2] = 4.0 * 0.2 ** 3
    coupled = _coupled_apply(p, w, obst, diag, f)
    assert coupled[2, 2, 2] - base[2, 2, 2] == pytest.approx(8.0, abs=1e-5)
    assert coupled[2, 2, 3] - base[2, 2, 3] == pytest.approx(-8.0, abs=1e-5)


def test_coupling_uses_transmissibility_on_nonuniform_cells():  # test_core_fluid.py:1792-1837
    w = [np.array([0.1, 0.2, 0.35, 0.35]), np.full(4, 0.25), np.full(4, 0.25)]
    shape = (4, 4, 4)
    obst = np.zeros(shape, np.int32)
    rng = np.random.default_rng(41)
    x = rng.normal(size=shape).astype(np.float32).astype(np.float64)
    y = rng.normal(size=shape).astype(np.float32).astype(np.float64)
    T = 0.375
    left, right = (1, 2, 2), (2, 2, 2)
    diag = np.zeros(shape)
    diag[left] = np.float32(T / (0.2 * 0.25 * 0.25))
    diag[right] = np.float32(T / (0.35 * 0.25 * 0.25))
    f = _slots(shape)
    f["slot_active"][left] = 1
    f["slot_neighbor"][left] = right
    f["slot_coef"][left] = T
    lvl = _lvl(w, obst, diag)
    ax, ay = _coupled_apply(x, w, obst, diag, f), _coupled_apply(y, w, obst, diag, f)
    lhs, rhs = F.weighted_dot(ax, y, lvl), F.weighted_dot(x, ay, lvl)
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-30) < 1e-5


def test_coupling_applies_multiple_owner_slots():  # test_core_fluid.py:1839-1896
    shape = (5, 5, 5)
    w = _unit(5)
    obst = np.zeros(shape, np.int32)
    owner, nb0, nb1 = (2, 2, 2), (2, 2, 3), (2, 3, 2)
    p = np.ones(shape)
    p[owner], p[nb0], p[nb1] = 3.0, 1.0, 2.0
    base = F.fv_apply(p, _lvl(w, obst), False)
    V = 0.2 ** 3
    c0, c1 = 4.0 * V, 7.0 * V
    diag = np.zeros(shape, np.float32)
    diag[owner] = (c0 + c1) / V
    diag[nb0] = c0 / V
    diag[nb1] = c1 / V
    f = _slots(shape)
    f["slot_active"][owner] = 2
    f["slot_neighbor"][owner] = nb0
    f["slot_coef"][owner] = c0
    f["extra_neighbor"][owner + (0,)] = nb1
    f["extra_coef"][owner + (0,)] = c1
    op = _coupled_apply(p, w, obst, diag.astype(np.float64), f)
    assert op[owner] - base[owner] == pytest.approx(15.0, abs=5e-5)
    assert op[nb0] - base[nb0] == pytest.approx(-8.0, abs=5e-5)
    assert op[nb1] - base[nb1] == pytest.approx(-7.0, abs=5e-5)


def _rows(entries, n):
    own = np.zeros((n, 3), np.int32)
    nb = np.zeros((n, 3), np.int32)
    T = np.zeros(n)
    for r, (o, b, t) in enumerate(entries):
        own[r], nb[r], T[r] = o, b, t
    return dict(row_owner=own, row_neighbor=nb, row_T=T, row_count=len(entries))


def test_row_list_projection_flux_enters_divergence_report():  # test_core_fluid.py:2139-2179
    shape = (5, 5, 5)
    w = _unit(5)
    obst = np.zeros(shape, np.int32)
    p = np.zeros(shape)
    p[2, 2, 2], p[2, 2, 1] = 3.0, 1.0
    V = 0.2 ** 3
    T = 4.0 * V
    diag = np.zeros(shape)
    diag[2, 2, 2] = diag[2, 2, 1] = np.float32(T / V)
    fields = _rows([((2, 2, 2), (2, 2, 1), T)], 125)
    c = F.interface_couplings(obst, w, **fields)
    operator = F.fv_apply(p, _lvl(w, obst, diag), False) + F.coupling_apply(p, c)
    u = F.subtract_pressure_gradient(np.zeros(shape + (3,), np.float32), p, obst, w, 1.0, False)
    total = F.divergence(u, obst, w, False) + F.interface_projection_divergence(p, obst, diag, np.zeros(shape), c,
                                                                                  1.0)
    np.testing.assert_allclose(total, operator, rtol=1e-6, atol=5e-5)


def _random_couplings(shape, rng, obst, rows: bool, count=12):
    n = int(np.prod(shape))
    entries = []
    while len(entries) < count:
        a = tuple(int(v) for v in rng.integers(0, shape))
        b = tuple(int(v) for v in rng.integers(0, shape))
        if a != b and obst[a] == 0 and obst[b] == 0:
            entries.append((a, b, float(rng.uniform(0.01, 0.2))))
    if rows:
        return _rows(entries, n)
    f = _slots(shape)
    for a, b, t in entries:
        s = int(f["slot_active"][a])
        if s == 0:
            f["slot_neighbor"][a], f["slot_coef"][a] = b, t
        elif s < F.COUPLING_SLOTS:
            f["extra_neighbor"][a + (s - 1,)], f["extra_coef"][a + (s - 1,)] = b, t
        f["slot_active"][a] = s + 1
    return f


@pytest.mark.parametrize("rows", [False, True])
def test_coupled_operator_is_v_self_adjoint_on_graded_grid(rows):
    rng = np.random.default_rng(7)
    shape = (6, 5, 7)
    w = [rng.uniform(0.5, 1.5, n) / n for n in shape]
    obst = (rng.random(shape) < 0.1).astype(np.int32)
    f = _random_couplings(shape, rng, obst, rows)
    c = F.interface_couplings(obst, w, **f)
    assert len(c.a) == 12
    lvl = _lvl(w, obst)
    x, y = rng.normal(size=shape), rng.normal(size=shape)
    lhs = F.weighted_dot(F.coupling_apply(x, c), y, lvl)
    rhs = F.weighted_dot(x, F.coupling_apply(y, c), lvl)
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-30)
    # C p of a constant is the negative row sums -T/V_a: mass balance sum_V (C 1) = -2 sum T
    one = np.ones(shape)
    total = float((F.coupling_apply(one, c) * F.cell_volumes_f32(w)).sum())
    tsum = float(np.sum(f["row_T"][:f["row_count"]]) if rows else
                 np.sum(f["slot_coef"][f["slot_active"] > 0]) + np.sum(f["extra_coef"]))
    assert total == pytest.approx(-2.0 * tsum, rel=1e-5)


def test_coupled_pcg_converges_to_spsolve():
    rng = np.random.default_rng(3)
    shape = (6, 6, 5)
    w = [rng.uniform(0.6, 1.4, n) / n for n in shape]
    obst = np.zeros(shape, np.int32)
    f = _random_couplings(shape, rng, obst, rows=True)
    c = F.interface_couplings(obst, w, **f)
    # the interface diagonal is the row diagonal density (the report's zero-mismatch state)
    rep0 = F.interface_terms_report(obst, w, np.zeros(shape), np.zeros(shape), **f)
    dens = np.zeros(shape)
    flat = dens.ravel()
    np.add.at(flat, c.a, c.ca)
    np.add.at(flat, c.b, c.cb)
    diag = flat.reshape(shape)
    rep = F.interface_terms_report(obst, w, diag, np.zeros(shape), **f)
    assert rep["row_diagonal_max_abs_density_mismatch"] < 1e-12 < rep0["row_diagonal_max_abs_density_mismatch"]
    # D + C annihilates constants like the closed FV Laplacian; a diagonal above the density
    # anchors the system (symmetric positive definite)
    diag = 1.5 * diag
    hier = F.build_hierarchy(w, obst, iface_diag=diag)
    lvl = hier.levels[0]
    n = int(np.prod(shape))
    cols = []
    for m in range(n):
        e = np.zeros(n)
        e[m] = 1.0
        cols.append((F.fv_apply(e.reshape(shape), lvl, False) + F.coupling_apply(e.reshape(shape), c)).ravel())
    A = sp.csr_matrix(np.array(cols).T)
    rhs0 = rng.normal(size=shape)
    res = F.pcg(rhs0, np.zeros(shape), hier, outlet=False, tolerance=1e-12, max_iters=400,
                preconditioner="jacobi", remove_mean=False, interface_active=True, couplings=c)
    assert res.converged
    want = spla.spsolve(A, -rhs0.ravel()).reshape(shape)
    np.testing.assert_allclose(res.pressure, want, rtol=0, atol=1e-8 * np.abs(want).max())
    # without the couplings the same solve lands elsewhere
    res0 = F.pcg(rhs0, np.zeros(shape), hier, outlet=False, tolerance=1e-12, max_iters=400,
                 preconditioner="jacobi", remove_mean=False, interface_active=True)
    assert np.abs(res0.pressure - want).max() > 1e-3 * np.abs(want).max()


def test_terms_report_counts_rows_like_the_reference():  # fluid.py:2864-3078
    shape = (4, 4, 4)
    w = _unit(4)
    obst = np.zeros(shape, np.int32)
    obst[3, 3, 3] = 1
    T = 0.01
    entries = [((0, 0, 0), (0, 0, 1), T),      # valid
               ((1, 1, 1), (1, 1, 1), T),      # owner == neighbour: invalid in the report, applied
               ((0, 0, 0), (3, 3, 3), T),      # neighbour is an obstacle
               ((0, 0, 0), (4, 0, 0), T),      # out of range
               ((0, 1, 0), (0, 1, 1), 0.0),    # zero transmissibility
               ((2, 2, 2), (2, 1, 2), 2 * T)]  # valid
    f = _rows(entries, 64)
    rep = F.interface_terms_report(obst, w, np.zeros(shape), np.zeros(shape), **f)
    assert rep["row_count"] == 6 and rep["row_active_count"] == 2 and rep["row_invalid_count"] == 4
    assert rep["row_overflow_count"] == 0
    assert rep["row_diagonal_integral"] == pytest.approx(2 * T + 4 * T)
    assert rep["row_max_transmissibility"] == pytest.approx(2 * T)
    c = F.interface_couplings(obst, w, **f)
    assert len(c.a) == 3  # the self row is applied
    p = np.zeros(shape)
    p[1, 1, 1] = 1.0
    assert F.coupling_apply(p, c)[1, 1, 1] == pytest.approx(-2 * T / (0.25 ** 3))
    # row_count beyond the capacity (one row per cell) is reported as overflow
    f2 = dict(f, row_count=70)
    rep2 = F.interface_terms_report(obst, w, np.zeros(shape), np.zeros(shape),
                                    **{**f2, "row_owner": np.zeros((70, 3), np.int32),
                                       "row_neighbor": np.zeros((70, 3), np.int32), "row_T": np.zeros(70)})
    assert rep2["row_overflow_count"] == 6 and rep2["row_invalid_count"] == 64
    # rows take precedence over slots
    fs = _slots(shape)
    fs["slot_active"][0, 0, 0] = 1
    fs["slot_neighbor"][0, 0, 0] = (1, 0, 0)
    fs["slot_coef"][0, 0, 0] = T
    assert len(F.interface_couplings(obst, w, **fs, **{k: f[k] for k in f}).a) == 3
    assert len(F.interface_couplings(obst, w, **fs).a) == 1


# ---------------------------------------------------------------------------
# BiCGSTAB fallback (fluid.py:6264-6533)
# ---------------------------------------------------------------------------

def test_bicgstab_solves_generic_fv_rhs():  # test_core_fluid.py:3713-3736
    shape = (4, 4, 4)
    w = _unit(4)
    obst = np.zeros(shape, np.int32)
    lvl = _lvl(w, obst)
    rhs = np.zeros(shape)
    rhs[2, 2, 2] = 1.0
    b_norm = float(np.sqrt(max(F.weighted_dot(rhs, rhs, lvl), 0.0)))
    res = F.bicgstab(rhs, np.zeros(shape), lvl, outlet=True, b_norm=b_norm, tolerance=1e-8, max_iters=80)
    assert res["converged"], res["breakdown"]
    assert res["relative_residual"] <= 1e-8
    # the outlet system is nonsingular: the iterate is spsolve's solution
    n = 64
    cols = []
    for m in range(n):
        e = np.zeros(n)
        e[m] = 1.0
        cols.append(F.fv_apply(e.reshape(shape), lvl, True).ravel())
    want = spla.spsolve(sp.csr_matrix(np.array(cols).T), rhs.ravel()).reshape(shape)
    np.testing.assert_allclose(res["pressure"], want, rtol=0, atol=1e-7 * np.abs(want).max())


def test_bicgstab_with_couplings_matches_spsolve_and_counts():
    rng = np.random.default_rng(11)
    shape = (6, 5, 6)
    w = [rng.uniform(0.6, 1.4, n) / n for n in shape]
    obst = (rng.random(shape) < 0.08).astype(np.int32)
    f = _random_couplings(shape, rng, obst, rows=True, count=15)
    c = F.interface_couplings(obst, w, **f)
    dens = np.zeros(shape).ravel()
    np.add.at(dens, c.a, c.ca)
    np.add.at(dens, c.b, c.cb)
    diag = 1.5 * dens.reshape(shape)
    lvl = _lvl(w, obst, diag)
    rhs = np.where(obst == 0, rng.normal(size=shape), 0.0)
    b_norm = float(np.sqrt(F.weighted_dot(rhs, rhs, lvl)))
    res = F.bicgstab(rhs, np.zeros(shape), lvl, outlet=True, b_norm=b_norm, tolerance=1e-11, max_iters=400,
                     start_iteration=7, couplings=c)
    assert res["converged"] and res["iterations"] > 7
    free = (obst == 0).ravel()
    idx = np.nonzero(free)[0]
    cols = []
    for m in idx:
        e = np.zeros(int(np.prod(shape)))
        e[m] = 1.0
        q = F.fv_apply(e.reshape(shape), lvl, True) + F.coupling_apply(e.reshape(shape), c)
        cols.append(q.ravel()[idx])
    want = np.zeros(int(np.prod(shape)))
    want[idx] = spla.spsolve(sp.csr_matrix(np.array(cols).T), rhs.ravel()[idx])
    np.testing.assert_allclose(res["pressure"].ravel(), want, rtol=0, atol=1e-8 * np.abs(want).max())
    # a budget too small to converge ends with the reference's message
    short = F.bicgstab(rhs, np.zeros(shape), lvl, outlet=True, b_norm=b_norm, tolerance=1e-14, max_iters=3,
                       couplings=c)
    assert not short["converged"] and short["breakdown"] == "BiCGSTAB fallback did not converge"
    assert short["iterations"] == 3 and short["host_residual_checks"] >= 4
