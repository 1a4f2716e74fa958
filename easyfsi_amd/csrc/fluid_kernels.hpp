// Device kernels of the FV pressure projection (gfx950).
//
// Operator (simulation_core/fluid.py:4079-4221): for a free cell c,
//   nb  = sum_faces  coef_f * w_f * p_nb,   den = sum_faces coef_f * w_f
//   (A p)_c = (den + iface_diag_c) p_c - nb
// with coef = 1/(w_a[i] d_a[i]) for the low face and 1/(w_a[i] d_a[i+1]) for the
// high face of axis a, w_f the velocity-Dirichlet face weight (1 when the
// face's owner cell has no active row), and the z-min outlet adding the ghost
// term 2/w_z[0]^2 to den.  A face is open when the neighbour exists and is
// free; the per-cell byte `flags` caches that:
//   bit0  row is an obstacle row (obstacle == 1)
//   bit1..6 face open towards -x, +x, -y, +y, -z, +z
//   bit7  cell is not free (obstacle != 0); excluded from dots and means
//
// Template parameters: C = compute type, loaders convert storage to C.
#pragma once

#include "hfsi_common.hpp"

namespace hfsi {

constexpr uint8_t kObst = 1, kXm = 2, kXp = 4, kYm = 8, kYp = 16, kZm = 32, kZp = 64, kNotFree = 128;

template <typename C>
struct OpGeo {
  uint32_t nx, ny, nz, n, sx, sy;  // sx = ny*nz (i stride), sy = nz (j stride)
  const uint8_t* flags;
  const float* fw;     // low-face weight per cell, or nullptr when no velocity-Dirichlet rows
  const double* diag;  // interface diagonal, or nullptr
  const C* cl[3];      // low-face coefficients per axis
  const C* ch[3];      // high-face coefficients per axis
  const float* wf[3];  // float widths (restriction volumes)
  const double* wd[3]; // widths as float64 (dot volumes)
  C outlet_add;
  int outlet;
};

// ---------------------------------------------------------------------------
// Storage loaders
// ---------------------------------------------------------------------------

template <typename T, typename C>
struct Ld {
  using Storage = T;
  static constexpr int kSign = 1;
  const T* p;
  __device__ C operator()(uint32_t c, uint32_t, uint32_t, uint32_t) const { return (C)p[c]; }
  __device__ C at(uint32_t c) const { return (C)p[c]; }
};

template <typename T, typename C>
struct LdNeg {
  using Storage = T;
  static constexpr int kSign = -1;
  const T* p;
  __device__ C at(uint32_t c) const { return -(C)p[c]; }
};

template <typename C>
struct LdZero {
  __device__ C operator()(uint32_t, uint32_t, uint32_t, uint32_t) const { return (C)0; }
  __device__ C at(uint32_t) const { return (C)0; }
};

// Fine value plus s * coarse parent value for cells that are free
// (_mg_prolongate_add_kernel, fluid.py:5390-5406).
template <typename TF, typename TCo, typename C>
struct LdProlong {
  const TF* pf;
  const TCo* pc;
  const uint8_t* flags;
  uint32_t cnx, cny, cnz, csx, csy;
  C s;
  __device__ C operator()(uint32_t c, uint32_t i, uint32_t j, uint32_t k) const {
    C v = (C)pf[c];
    if (!(flags[c] & kNotFree)) {
      uint32_t ci = min(i >> 1, cnx - 1), cj = min(j >> 1, cny - 1), ck = min(k >> 1, cnz - 1);
      v += s * (C)pc[ci * csx + cj * csy + ck];
    }
    return v;
  }
};

__device__ inline void decompose(uint32_t c, uint32_t ny, uint32_t nz, uint32_t& i, uint32_t& j, uint32_t& k) {
  k = c % nz;
  uint32_t t = c / nz;
  j = t % ny;
  i = t / ny;
}

// Neighbour sum and denominator (without interface diagonal) of cell c.
template <typename C, typename P>
__device__ inline void nb_den(const OpGeo<C>& g, uint32_t c, uint32_t i, uint32_t j, uint32_t k, uint8_t f, const P& p,
                             C& nb, C& den) {
  nb = (C)0;
  den = (C)0;
  const float* fw = g.fw;
  if (f & kXm) {
    C cf = g.cl[0][i];
    if (fw) cf *= (C)fw[c];
    nb += cf * p(c - g.sx, i - 1, j, k);
    den += cf;
  }
  if (f & kXp) {
    C cf = g.ch[0][i];
    if (fw) cf *= (C)fw[c + g.sx];
    nb += cf * p(c + g.sx, i + 1, j, k);
    den += cf;
  }
  if (f & kYm) {
    C cf = g.cl[1][j];
    if (fw) cf *= (C)fw[c];
    nb += cf * p(c - g.sy, i, j - 1, k);
    den += cf;
  }
  if (f & kYp) {
    C cf = g.ch[1][j];
    if (fw) cf *= (C)fw[c + g.sy];
    nb += cf * p(c + g.sy, i, j + 1, k);
    den += cf;
  }
  if (f & kZm) {
    C cf = g.cl[2][k];
    if (fw) cf *= (C)fw[c];
    nb += cf * p(c - 1, i, j, k - 1);
    den += cf;
  }
  if (f & kZp) {
    C cf = g.ch[2][k];
    if (fw) cf *= (C)fw[c + 1];
    nb += cf * p(c + 1, i, j, k + 1);
    den += cf;
  }
  if (g.outlet && k == 0) den += g.outlet_add;
}

// One weighted-Jacobi update of cell c on L p = rhs, L = nb - den p
// (_pressure_fv_jacobi_kernel, fluid.py:4223-4273).
template <typename C, typename P, typename R>
__device__ inline C jacobi_cell(const OpGeo<C>& g, uint32_t c, uint32_t i, uint32_t j, uint32_t k, const P& p,
                                const R& rhs, C omega) {
  const uint8_t f = g.flags[c];
  if (f & kObst) return (C)0;
  C nb, den;
  nb_den(g, c, i, j, k, f, p, nb, den);
  if (g.diag) den += (C)g.diag[c];
  if (!(den > (C)0)) return (C)0;
  C pc = p(c, i, j, k);
  C jac = (nb - rhs.at(c)) / den;
  return pc + omega * (jac - pc);
}

// Legacy residual rhs - (nb - den p) of cell c (_mg_compute_residual_kernel, fluid.py:5261-5307).
template <typename C, typename P, typename R>
__device__ inline C residual_cell(const OpGeo<C>& g, uint32_t c, uint32_t i, uint32_t j, uint32_t k, const P& p,
                                  const R& rhs) {
  const uint8_t f = g.flags[c];
  if (f & kObst) return (C)0;
  C nb, den;
  nb_den(g, c, i, j, k, f, p, nb, den);
  if (g.diag) den += (C)g.diag[c];
  return rhs.at(c) - (nb - den * p(c, i, j, k));
}

// (A p)_c (_fv_laplacian_apply_kernel, fluid.py:4311-4340).
template <typename C, typename P>
__device__ inline C apply_cell(const OpGeo<C>& g, uint32_t c, uint32_t i, uint32_t j, uint32_t k, const P& p) {
  const uint8_t f = g.flags[c];
  if (f & kObst) return (C)0;
  C nb, den;
  nb_den(g, c, i, j, k, f, p, nb, den);
  if (g.diag) den += (C)g.diag[c];
  return den * p(c, i, j, k) - nb;
}

template <typename C>
__device__ inline double volume64(const OpGeo<C>& g, uint32_t i, uint32_t j, uint32_t k) {
  return g.wd[0][i] * g.wd[1][j] * g.wd[2][k];
}

// Coarse cell (ci,cj,ck): V-weighted mean of the free children's legacy
// residual (_mg_restrict_residual_kernel, fluid.py:5334-5379).  Children
// visited in (oi, oj, ok) order with ok fastest; the residual is rounded to
// the fine level's storage type TRes like the reference's residual field.
template <typename TRes, typename C, typename P, typename R>
__device__ inline C restrict_cell(const OpGeo<C>& fine, uint32_t ci, uint32_t cj, uint32_t ck, const P& p,
                                  const R& rhs) {
  C rsum = (C)0, vsum = (C)0;
#pragma unroll
  for (int oi = 0; oi < 2; ++oi) {
#pragma unroll
    for (int oj = 0; oj < 2; ++oj) {
#pragma unroll
      for (int ok = 0; ok < 2; ++ok) {
        uint32_t fi = 2 * ci + oi, fj = 2 * cj + oj, fk = 2 * ck + ok;
        if (fi < fine.nx && fj < fine.ny && fk < fine.nz) {
          uint32_t fc = fi * fine.sx + fj * fine.sy + fk;
          if (!(fine.flags[fc] & kObst)) {
            C vol = (C)(fine.wf[0][fi] * fine.wf[1][fj] * fine.wf[2][fk]);
            C res = (C)(TRes)residual_cell(fine, fc, fi, fj, fk, p, rhs);
            rsum += res * vol;
            vsum += vol;
          }
        }
      }
    }
  }
  return rsum / (vsum > (C)1.0e-30 ? vsum : (C)1.0e-30);
}

// ---------------------------------------------------------------------------
// Full-grid kernels (XCD-aware chunk loop)
// ---------------------------------------------------------------------------


template <typename C, typename P, typename R, typename TOut>
__global__ void __launch_bounds__(kBlock) k_jacobi(OpGeo<C> g, P p, R rhs, TOut* out, C omega) {
  HFSI_CELL_LOOP(g.n) {
    uint32_t i, j, k;
    decompose(c, g.ny, g.nz, i, j, k);
    out[c] = (TOut)jacobi_cell(g, c, i, j, k, p, rhs, omega);
  }
}

// Thread per coarse cell: fine residual + restriction.
template <typename TRes, typename C, typename P, typename R, typename TOut>
__global__ void __launch_bounds__(kBlock) k_resid_restrict(OpGeo<C> fine, P p, R rhs, OpGeo<C> coarse, TOut* out) {
  HFSI_CELL_LOOP(coarse.n) {
    uint32_t ci, cj, ck;
    decompose(c, coarse.ny, coarse.nz, ci, cj, ck);
    C v = (C)0;
    if (!(coarse.flags[c] & kObst)) v = restrict_cell<TRes>(fine, ci, cj, ck, p, rhs);
    out[c] = (TOut)v;
  }
}

// q = A d with the per-workgroup partials of <d, q>_V (combined by k_finish_sums).
template <typename C, typename TD, typename TQ>
__global__ void __launch_bounds__(kBlock) k_spmv_dot(OpGeo<C> g, const TD* d, TQ* q, double* partials) {
  __shared__ double lds[kBlock / 64];
  double acc[1] = {0.0};
  Ld<TD, C> P{d};
  HFSI_CELL_LOOP(g.n) {
    uint32_t i, j, k;
    decompose(c, g.ny, g.nz, i, j, k);
    C v = apply_cell(g, c, i, j, k, P);
    TQ vs = (TQ)v;
    q[c] = vs;
    if (!(g.flags[c] & kNotFree)) acc[0] += (double)d[c] * (double)vs * volume64(g, i, j, k);
  }
  block_partials<1>(acc, partials, lds);
}

// Plain apply (no dot), any storage.
template <typename C, typename TP, typename TQ>
__global__ void __launch_bounds__(kBlock) k_apply(OpGeo<C> g, const TP* p, TQ* q) {
  Ld<TP, C> P{p};
  HFSI_CELL_LOOP(g.n) {
    uint32_t i, j, k;
    decompose(c, g.ny, g.nz, i, j, k);
    q[c] = (TQ)apply_cell(g, c, i, j, k, P);
  }
}

template <typename C>
__global__ void __launch_bounds__(kBlock) k_diagonal(OpGeo<C> g, double* out) {
  LdZero<C> P;
  HFSI_CELL_LOOP(g.n) {
    uint32_t i, j, k;
    decompose(c, g.ny, g.nz, i, j, k);
    uint8_t f = g.flags[c];
    double v = 0.0;
    if (!(f & kObst)) {
      C nb, den;
      nb_den(g, c, i, j, k, f, P, nb, den);
      if (g.diag) den += (C)g.diag[c];
      v = (double)den;
    }
    out[c] = v;
  }
}

// Generic weighted sums: up to NS sums of V * (products chosen by Op), as
// per-workgroup partials (combined by k_finish_sums).
template <int NS, typename C, typename Op>
__global__ void __launch_bounds__(kBlock) k_sums(OpGeo<C> g, Op op, double* partials) {
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  HFSI_CELL_LOOP(g.n) {
    uint32_t i, j, k;
    decompose(c, g.ny, g.nz, i, j, k);
    if (!(g.flags[c] & kNotFree)) op(c, volume64(g, i, j, k), acc);
  }
  block_partials<NS>(acc, partials, lds);
}

// Generic pointwise pass over all cells: op(c, free).
template <typename Op>
__global__ void __launch_bounds__(kBlock) k_pointwise(uint32_t n, const uint8_t* flags, Op op) {
  HFSI_CELL_LOOP(n) { op(c, !(flags[c] & kNotFree), flags[c]); }
}

}  // namespace hfsi
