// Sharp-interface pressure couplings of the FV operator (gfx950):
// the off-diagonal terms of CartesianFluidSolver._fv_laplacian_apply_kernel
// (simulation_core/fluid.py:4346-4456), their report (2864-3078) and the
// projection-divergence accumulation (5584-5759).
//
// A coupling joins two fluid cells a and b with transmissibility T:
//   (A p)[a] -= (T / V_a) p[b],   (A p)[b] -= (T / V_b) p[a],
// V the float32 cell volume (wx * wy) * wz floored at 1e-30.  The couplings come either
// from the row list (owner, neighbour, T; used when its count is positive, at most one
// row per cell) or from per-owner slots (a count, one primary neighbour / coefficient
// and 15 extra ones).  The reference walks those fields at every operator application;
// here they are compiled once into a compact list (a, b, T/V_a, T/V_b) after the
// caller sets them, and every application is one pass over the list with float64
// atomics (the reference also accumulates with atomics, so the summation order of a
// cell's coupling terms is unspecified in both).
#pragma once

#include "hfsi_common.hpp"

namespace hfsi {

constexpr int kCouplingSlots = 16;  // primary + 15 extra per owner

struct CoupleSrc {
  uint32_t nx, ny, nz, n;
  const int32_t* obst;
  const float* w[3];  // float32 widths
  const int32_t* slot_active;
  const int32_t* slot_nb;    // [n][3]
  const double* slot_coef;   // [n]
  const int32_t* extra_nb;   // [n][15][3] or null
  const double* extra_coef;  // [n][15] or null
  const int32_t* row_owner;  // [rows][3]
  const int32_t* row_nb;     // [rows][3]
  const double* row_T;       // [rows]
  uint32_t rows;             // rows walked (min(row_count, n)); 0: slots
};

struct CoupleList {
  uint32_t* a;
  uint32_t* b;
  double* ca;
  double* cb;
  uint32_t cap;  // entries allocated
};

// Build-pass statistics (device): counters and sums of the terms report.
struct CoupleStats {
  uint32_t applied;      // couplings the operator applies
  uint32_t active;       // report-valid rows (applied and owner != neighbour)
  uint32_t invalid;      // rows or slots that are not report-valid
  uint32_t pad;
  double diag_integral;  // sum of 2 T over report-valid rows
  uint64_t tmax_bits;    // max T over report-valid rows (T > 0: bit order = value order)
};

__device__ inline float cell_volume_f32(const CoupleSrc& s, uint32_t i, uint32_t j, uint32_t k) {
#pragma clang fp contract(off)
  return fmaxf(s.w[0][i] * s.w[1][j] * s.w[2][k], 1.0e-30f);
}

// One candidate coupling: apply-validity (fluid.py:4361-4371, 4427-4433) and the
// report's validity, which also needs owner != neighbour (fluid.py:2915, 2996).
template <bool kFill>
__device__ inline void couple_candidate(const CoupleSrc& s, int oi, int oj, int ok, int ni, int nj, int nk, double T,
                                        bool check_owner, CoupleList L, CoupleStats* st, double* density) {
  const bool own_in = oi >= 0 && oi < (int)s.nx && oj >= 0 && oj < (int)s.ny && ok >= 0 && ok < (int)s.nz;
  const bool nb_in = ni >= 0 && ni < (int)s.nx && nj >= 0 && nj < (int)s.ny && nk >= 0 && nk < (int)s.nz;
  bool valid = nb_in && (!check_owner || own_in) && T > 0.0;
  uint32_t a = 0, b = 0;
  if (valid) {
    a = ((uint32_t)oi * s.ny + (uint32_t)oj) * s.nz + (uint32_t)ok;
    b = ((uint32_t)ni * s.ny + (uint32_t)nj) * s.nz + (uint32_t)nk;
    valid = s.obst[b] == 0 && (!check_owner || s.obst[a] == 0);
  }
  const bool report_valid = valid && a != b;
  double ca = 0.0, cb = 0.0;
  if (valid) {
    ca = T / (double)cell_volume_f32(s, (uint32_t)oi, (uint32_t)oj, (uint32_t)ok);
    cb = T / (double)cell_volume_f32(s, (uint32_t)ni, (uint32_t)nj, (uint32_t)nk);
  }
  if (kFill) {
    if (valid) {
      const uint32_t e = atomicAdd(&st->applied, 1u);
      if (e < L.cap) {  // a count above the first pass's is reported by the caller, never written
        L.a[e] = a;
        L.b[e] = b;
        L.ca[e] = ca;
        L.cb[e] = cb;
      }
    }
  } else {
    if (valid) atomicAdd(&st->applied, 1u);
    if (report_valid) {
      atomicAdd(&st->active, 1u);
      atomicAdd(&st->diag_integral, 2.0 * T);
      atomicMax((unsigned long long*)&st->tmax_bits, (unsigned long long)__double_as_longlong(T));
      if (density) {
        atomicAdd(&density[a], ca);
        atomicAdd(&density[b], cb);
      }
    } else {
      atomicAdd(&st->invalid, 1u);
    }
  }
}

// kFill = false: count the couplings and gather the report; kFill = true: write the list
// (positions from a counter reset between the passes).
template <bool kFill>
__global__ void __launch_bounds__(kBlock) k_couple_build(CoupleSrc s, CoupleList L, CoupleStats* st,
                                                         double* density) {
  const uint32_t stride = gridDim.x * blockDim.x;
  if (s.rows > 0) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < s.rows; r += stride) {
      const int32_t* o = s.row_owner + 3 * (size_t)r;
      const int32_t* b = s.row_nb + 3 * (size_t)r;
      couple_candidate<kFill>(s, o[0], o[1], o[2], b[0], b[1], b[2], s.row_T[r], true, L, st, density);
    }
    return;
  }
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < s.n; c += stride) {
    const int32_t act = s.slot_active[c];
    if (act <= 0 || s.obst[c] != 0) continue;
    const uint32_t k = c % s.nz, t = c / s.nz, j = t % s.ny, i = t / s.ny;
    const int cnt = act < kCouplingSlots ? act : kCouplingSlots;
    for (int q = 0; q < cnt; ++q) {
      // absent extra-slot arrays read as the zero-initialised fields they stand for (T = 0: invalid)
      int ni = 0, nj = 0, nk = 0;
      double T = 0.0;
      if (q == 0) {
        const int32_t* nb = s.slot_nb + 3 * (size_t)c;
        ni = nb[0], nj = nb[1], nk = nb[2];
        T = s.slot_coef[c];
      } else if (s.extra_nb && s.extra_coef) {
        const size_t e = (size_t)c * (kCouplingSlots - 1) + q - 1;
        ni = s.extra_nb[3 * e], nj = s.extra_nb[3 * e + 1], nk = s.extra_nb[3 * e + 2];
        T = s.extra_coef[e];
      }
      couple_candidate<kFill>(s, (int)i, (int)j, (int)k, ni, nj, nk, T, false, L, st, density);
    }
  }
}

// out += scale * (C p): (C p)[a] = -ca p[b], (C p)[b] = -cb p[a].
template <typename TP, typename TO>
__global__ void __launch_bounds__(kBlock) k_couple_apply(CoupleList L, uint32_t count, const TP* p, TO* out,
                                                         double scale) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride) {
    const uint32_t a = L.a[e], b = L.b[e];
    const double va = scale * (-L.ca[e] * (double)p[b]);
    const double vb = scale * (-L.cb[e] * (double)p[a]);
    unsafeAtomicAdd(&out[a], (TO)va);
    unsafeAtomicAdd(&out[b], (TO)vb);
  }
}

}  // namespace hfsi
