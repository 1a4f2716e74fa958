"""NumPy float64 restatement of the mixed-precision V-cycle (test oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

The MIXED solve has no counterpart in the reference: it keeps the reference's
flexible PCG (simulation_core/fluid.py:6535-6811) and replaces the V-cycle of
fluid.py:6085-6262 by a red-black Gauss-Seidel V-cycle on a semi-coarsened
hierarchy.  This module states that V-cycle in float64 so that the device
implementation (float32 arithmetic, fp16 or float32 storage) can be checked
cell by cell:

  hierarchy   with T_l = 2^l h_min (h_min: smallest level-0 width over all
              axes), along each axis cells i, i+1 of level l merge when
              w_i + w_{i+1} <= 2.5 T_l, or always once max width <= 2.5 T_l;
              stop when an axis has <= 3 cells or nothing merges.  Level
              operators are the reference's FV operator (fluid.py:4079-4221)
              rediscretised on the merged widths (level 0: the level-0
              metrics rounded through float32, as oracle.fluid.make_level).
  masks       a coarse cell is solid only when none of its children is free
              (the reference's 2:1 hierarchy uses "any child is an obstacle
              row", fluid.py:5334-5388, which thickens the solid per level).
  smoother    RB-GS, omega = 1: red (i+j+k even) then black before the coarse
              correction, black then red after it; the first pre-sweep starts
              from zero.
  restriction volume-weighted mean of the free children's residuals.
  prolongation per axis, linear between the parent and the coarse neighbour
              on the side of the fine centre (weight from the centres),
              product over axes, renormalised over free coarse cells.
  coarsest    `coarse` RB sweeps.

The restriction is not the adjoint of the prolongation, so M is not
symmetric; the flexible (Polak-Ribiere) beta of the PCG driver tolerates that.

Pinning: the hierarchy rule is checked on closed cases (uniform grids give
the reference's 2:1 hierarchy; child tables partition every axis and sum the
widths), the level-0 operator against oracle.fluid.fv_apply, the smoother
against scipy's exact solve on tiny grids (RB-GS converges to it) and the
V-cycle by its contraction of the error in the energy norm when used as a
stationary iteration (tests/test_oracle_mixed.py).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import fluid as F

MERGE_RATIO = 2.5   # merge a pair while its width <= MERGE_RATIO x the nominal width T_l
FULL_RATIO = 2.5    # plain 2:1 once every width <= FULL_RATIO x T_l


@dataclass
class MixLevel:
    w: list            # per-axis widths (float64)
    ctr: list          # per-axis centres
    cl: list           # per-axis low-face coefficients (unscaled)
    ch: list
    obst_row: np.ndarray   # bool: obstacle row (obstacle == 1)
    notfree: np.ndarray    # bool: obstacle != 0
    outlet_add: float
    par: list = field(default_factory=list)    # fine index -> parent (to the next level)
    first: list = field(default_factory=list)  # this level's cell -> first child (previous level)
    cnt: list = field(default_factory=list)

    @property
    def shape(self):
        return self.notfree.shape


def _centres(w):
    return np.cumsum(w) - 0.5 * np.asarray(w)


def _coefficients(w, ctr):
    n = len(w)
    cl = np.zeros(n)
    ch = np.zeros(n)
    for i in range(n):
        if i > 0:
            cl[i] = 1.0 / (w[i] * (ctr[i] - ctr[i - 1]))
        if i + 1 < n:
            ch[i] = 1.0 / (w[i] * (ctr[i + 1] - ctr[i]))
    return cl, ch


def build_hierarchy(widths, obstacle) -> list[MixLevel]:
    """Semi-coarsened levels from the level-0 widths and obstacle array."""
    lvl0 = F.make_level(widths, obstacle)
    obstacle = np.asarray(obstacle, np.int32)
    w0 = [np.asarray(lvl0.w[a], np.float64) for a in range(3)]
    cl0, ch0 = [], []
    for a in range(3):
        d = lvl0.d[a]
        n = len(w0[a])
        cl = np.zeros(n)
        ch = np.zeros(n)
        cl[1:] = 1.0 / (w0[a][1:] * d[1:])
        ch[:-1] = 1.0 / (w0[a][:-1] * d[1:])
        cl0.append(cl)
        ch0.append(ch)
    levels = [MixLevel(w=w0, ctr=[_centres(w) for w in w0], cl=cl0, ch=ch0, obst_row=obstacle == 1,
                       notfree=obstacle != 0, outlet_add=2.0 / (w0[2][0] * w0[2][0]))]
    while len(levels) < 24:
        f = levels[-1]
        if min(len(w) for w in f.w) <= 3:
            break
        hmin = min(float(w.min()) for w in levels[0].w) * 2.0 ** (len(levels) - 1)
        hmax = float(np.concatenate(f.w).max())
        full = hmax <= FULL_RATIO * hmin
        nw, par, first, cnt = [], [], [], []
        any_merge = False
        for a in range(3):
            wa = f.w[a]
            n = len(wa)
            pa = np.zeros(n, np.int64)
            cw, fi, ci = [], [], []
            i = 0
            while i < n:
                fi.append(i)
                if i + 1 < n and (full or wa[i] + wa[i + 1] <= MERGE_RATIO * hmin):
                    cw.append(wa[i] + wa[i + 1])
                    ci.append(2)
                    pa[i] = pa[i + 1] = len(cw) - 1
                    i += 2
                    any_merge = True
                else:
                    cw.append(wa[i])
                    ci.append(1)
                    pa[i] = len(cw) - 1
                    i += 1
            nw.append(np.asarray(cw))
            par.append(pa)
            first.append(np.asarray(fi, np.int64))
            cnt.append(np.asarray(ci, np.int64))
        if not any_merge:
            break
        f.par = par
        ctr = [_centres(w) for w in nw]
        coef = [_coefficients(nw[a], ctr[a]) for a in range(3)]
        shape = tuple(len(w) for w in nw)
        nfree = np.zeros(shape)
        pi, pj, pk = par
        np.add.at(nfree, _broadcast_index(pi, pj, pk), (~f.notfree).astype(np.float64))
        obst = nfree == 0
        levels.append(MixLevel(w=nw, ctr=ctr, cl=[c[0] for c in coef], ch=[c[1] for c in coef], obst_row=obst,
                               notfree=obst.copy(), outlet_add=2.0 / (nw[2][0] * nw[2][0]), first=first, cnt=cnt))
    return levels


def _broadcast_index(pi, pj, pk):
    """Index arrays (i, j, k) -> (pi[i], pj[j], pk[k]) broadcast to the fine shape."""
    shape = (len(pi), len(pj), len(pk))
    return (np.broadcast_to(pi[:, None, None], shape), np.broadcast_to(pj[None, :, None], shape),
            np.broadcast_to(pk[None, None, :], shape))


def _shift(a, axis, step):
    """b[c] = a[c - step e_axis] with zero fill (step = +1: value of the lower neighbour)."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if step > 0:
        src[axis] = slice(0, n - 1)
        dst[axis] = slice(1, n)
    else:
        src[axis] = slice(1, n)
        dst[axis] = slice(0, n - 1)
    out[tuple(dst)] = a[tuple(src)]
    return out


def _open_faces(lvl: MixLevel):
    free = ~lvl.notfree
    opens = []
    for a in range(3):
        lo = _shift(free.astype(np.int8), a, +1).astype(bool) & free
        hi = _shift(free.astype(np.int8), a, -1).astype(bool) & free
        opens.append((lo, hi))
    return opens


def _nb_den(lvl: MixLevel, e, opens, outlet, with_nb=True):
    shape = lvl.shape
    nb = np.zeros(shape)
    den = np.zeros(shape)
    for a in range(3):
        bshape = [1, 1, 1]
        bshape[a] = shape[a]
        cl = lvl.cl[a].reshape(bshape)
        ch = lvl.ch[a].reshape(bshape)
        lo, hi = opens[a]
        den += np.where(lo, cl, 0.0) + np.where(hi, ch, 0.0)
        if with_nb:
            nb += np.where(lo, cl * _shift(e, a, +1), 0.0) + np.where(hi, ch * _shift(e, a, -1), 0.0)
    if outlet:
        den[:, :, 0] += lvl.outlet_add
    return nb, den


def apply(lvl: MixLevel, e, outlet=False):
    """A e on free cells (0 elsewhere), unscaled operator."""
    opens = _open_faces(lvl)
    nb, den = _nb_den(lvl, e, opens, outlet)
    return np.where(lvl.notfree, 0.0, den * e - nb)


def rb_pass(lvl: MixLevel, e, rhs, color, outlet, from_zero=False):
    shape = lvl.shape
    if from_zero:
        e = np.zeros(shape)
    ii, jj, kk = np.indices(shape)
    mask = ((ii + jj + kk) & 1) == color
    opens = _open_faces(lvl)
    nb, den = _nb_den(lvl, e, opens, outlet, with_nb=not from_zero)
    v = np.where(den > 0, (nb + rhs) / np.where(den > 0, den, 1.0), 0.0)
    v = np.where(lvl.notfree, 0.0, v)
    return np.where(mask, v, e)


def restrict(f: MixLevel, c: MixLevel, e, rhs, outlet):
    res = np.where(f.notfree, 0.0, rhs - apply(f, e, outlet))
    V = f.w[0][:, None, None] * f.w[1][None, :, None] * f.w[2][None, None, :]
    Vf = np.where(f.notfree, 0.0, V)
    s = np.zeros(c.shape)
    vs = np.zeros(c.shape)
    idx = _broadcast_index(*f.par)
    np.add.at(s, idx, Vf * res)
    np.add.at(vs, idx, Vf)
    out = np.where(vs > 0, s / np.where(vs > 0, vs, 1.0), 0.0)
    return np.where(c.notfree, 0.0, out)


def interp_tables(f: MixLevel, c: MixLevel, a: int):
    """Per fine index on axis a: (parent, partner, weight of parent)."""
    par = f.par[a]
    n = len(par)
    nb = par.copy()
    w = np.ones(n)
    for i in range(n):
        I = par[i]
        x, X = f.ctr[a][i], c.ctr[a][I]
        if abs(x - X) > 1e-9 * f.w[a][i]:
            cand = I - 1 if x < X else I + 1
            if 0 <= cand < len(c.w[a]):
                nb[i] = cand
                Xn = c.ctr[a][cand]
                w[i] = (Xn - x) / (Xn - X)
    return par, nb, w


def prolong_add(f: MixLevel, c: MixLevel, ef, ec):
    tabs = [interp_tables(f, c, a) for a in range(3)]
    s = np.zeros(f.shape)
    ws = np.zeros(f.shape)
    cfree = ~c.notfree
    for a in range(2):
        for b in range(2):
            for d in range(2):
                I = tabs[0][0] if a == 0 else tabs[0][1]
                J = tabs[1][0] if b == 0 else tabs[1][1]
                K = tabs[2][0] if d == 0 else tabs[2][1]
                wi = tabs[0][2] if a == 0 else 1.0 - tabs[0][2]
                wj = tabs[1][2] if b == 0 else 1.0 - tabs[1][2]
                wk = tabs[2][2] if d == 0 else 1.0 - tabs[2][2]
                w = wi[:, None, None] * wj[None, :, None] * wk[None, None, :]
                sel = np.ix_(I, J, K)
                ok = cfree[sel] & (w != 0)
                s += np.where(ok, w * ec[sel], 0.0)
                ws += np.where(ok, w, 0.0)
    add = np.where(ws > 0, s / np.where(ws > 0, ws, 1.0), 0.0)
    return np.where(f.notfree, ef, ef + add)


def dense_solve(lvl: MixLevel, rhs, outlet=False):
    """Direct solve of L e = rhs on a (small) level, e = 0 on cells that are not
    free: S = V L over the free cells is symmetric positive semidefinite, with
    the constants of every connected set of free cells without an outlet row
    as null space; adding alpha 1 1^T on such a set (alpha = mean diagonal of
    the set / its size) makes it definite and leaves the solution of a
    consistent right-hand side (the mean-free one) unchanged."""
    shape = lvl.shape
    free = ~lvl.notfree
    cells = np.flatnonzero(free.ravel())
    m = cells.size
    e = np.zeros(shape)
    if m == 0:
        return e
    V = (lvl.w[0][:, None, None] * lvl.w[1][None, :, None] * lvl.w[2][None, None, :]).ravel()[cells]
    S = np.zeros((m, m))
    unit = np.zeros(int(np.prod(shape)))
    for t, c in enumerate(cells):
        unit[c] = 1.0
        S[:, t] = V * apply(lvl, unit.reshape(shape), outlet).ravel()[cells]
        unit[c] = 0.0
    S = 0.5 * (S + S.T)
    # connected sets of free cells through open faces
    comp = -np.ones(m, np.int64)
    ncomp = 0
    for t0 in range(m):
        if comp[t0] >= 0:
            continue
        stack, members = [t0], []
        comp[t0] = ncomp
        while stack:
            t = stack.pop()
            members.append(t)
            for u in np.flatnonzero((S[t] != 0.0) & (comp < 0)):
                comp[u] = ncomp
                stack.append(int(u))
        members = np.asarray(members)
        has_outlet = outlet and bool(np.any(cells[members] % shape[2] == 0))
        if not has_outlet:
            alpha = S[members, members].sum() / members.size / members.size
            S[np.ix_(members, members)] += alpha
        ncomp += 1
    b = V * np.asarray(rhs, np.float64).ravel()[cells]
    e.ravel()[cells] = np.linalg.solve(S, b)
    return e


def vcycle(levels: list[MixLevel], r, *, pre=2, coarse=24, post=2, outlet=False):
    """z = M r (float64), the V-cycle of the mixed solve; coarse <= 0 solves
    the coarsest level directly (dense_solve), coarse > 0 gives its RB sweeps."""
    L = len(levels)
    rhs = [np.where(levels[0].notfree, 0.0, np.asarray(r, np.float64))] + [None] * (L - 1)
    e = [None] * L
    if L == 1:
        x = np.zeros(levels[0].shape)
        for s in range(coarse):
            x = rb_pass(levels[0], x, rhs[0], 0, outlet, from_zero=(s == 0))
            x = rb_pass(levels[0], x, rhs[0], 1, outlet)
        return x
    for l in range(L - 1):
        x = np.zeros(levels[l].shape)
        for s in range(pre):
            x = rb_pass(levels[l], x, rhs[l], 0, outlet, from_zero=(s == 0))
            x = rb_pass(levels[l], x, rhs[l], 1, outlet)
        e[l] = x
        rhs[l + 1] = restrict(levels[l], levels[l + 1], x, rhs[l], outlet)
    x = np.zeros(levels[-1].shape)
    if coarse <= 0:
        x = dense_solve(levels[-1], rhs[-1], outlet)
    for s in range(coarse):
        x = rb_pass(levels[-1], x, rhs[-1], 0, outlet, from_zero=(s == 0))
        x = rb_pass(levels[-1], x, rhs[-1], 1, outlet)
    e[-1] = x
    for l in range(L - 2, -1, -1):
        x = prolong_add(levels[l], levels[l + 1], e[l], e[l + 1])
        for s in range(post):
            x = rb_pass(levels[l], x, rhs[l], 1, outlet)
            x = rb_pass(levels[l], x, rhs[l], 0, outlet)
        e[l] = x
    return np.where(levels[0].notfree, 0.0, e[0])
