// Velocity predictor of the MAC-grid fluid (gfx950):
// CartesianFluidSolver._predict_kernel (simulation_core/fluid.py:2447-2513).
//
// For every cell that is not an obstacle (obstacle != 1):
//   departure point  x_d = x_c - dt * u_trace, with u_trace = u_prev(c) (Euler) or the
//                    sample of u_prev at x_c - (0.5 dt) u_prev(c) (RK2, fluid.py:2462-2474)
//   advected         obstacle-aware trilinear sample of u_prev at x_d (fluid.py:2382-2404):
//                    corners with obstacle != 0 are dropped and the weights renormalised;
//                    with total weight <= 1e-12 the sample is u_prev(c)
//   diffusion        FV Laplacian with zero flux through domain faces and faces shared
//                    with obstacle != 0 cells (fluid.py:2484-2512)
//   u(c)           = advected + (dt * nu) * lap
// and u(c) = 0 for obstacle == 1.  Positions map to fractional cell coordinates by a
// binary search over the cell centres with half-width extrapolation past the end
// centres (fluid.py:2406-2435).
//
// Arithmetic is float32 in the reference's operation order with fused multiply-add
// contraction disabled and correctly rounded division, so the result is the IEEE
// evaluation of the reference's expressions, operation by operation.
//
// Layout: velocity [n][3] float32 (component fastest), obstacle [n] int32, cell
// c = (i * ny + j) * nz + k.  One thread per cell under the XCD-aware chunk loop.
// Every divisor but the sample weight is a metric value, so each workgroup stages the
// centres and the correctly rounded reciprocals of the metric divisors in LDS and forms
// each quotient from its reciprocal with two residual corrections (div_ry), which
// yields the correctly rounded quotient in five operations instead of the dozen of a
// general division.
#pragma once

#include "hfsi_common.hpp"

namespace hfsi {

// LDS tables per workgroup (kLds): for each axis the centres, widths, centre distances
// and the correctly rounded reciprocals of the widths, centre distances and centre gaps,
// 6 floats per cell.
// With 32-bit offsets the x axis also keeps its plane offsets i * ny * nz in LDS.
constexpr uint32_t kPredictLdsTables = 6;
constexpr uint32_t kPredictLdsMaxCells = 2048;  // nx + ny + nz; at most 56 KiB of LDS

struct PredictArgs {
  uint32_t nx, ny, nz, n;  // n < 2^31, every axis < 2^24
  FastDiv fz, fy;          // division by nz and by ny
  const float* face[3];    // n_a + 1 face coordinates
  const float* centre[3];  // n_a centres
  const float* width[3];   // n_a widths
  const float* cdist[3];   // n_a centre distances (entry d > 0: centre d - centre d-1)
  const float* vp;         // previous velocity [n][3]
  const int32_t* obst;     // obstacle [n]
  float* v;                // new velocity [n][3]
  float dt, nu;
};

struct Vec3 {
  float x, y, z;
};

// Loads and stores at a byte offset from a uniform base.  With Off = uint32_t the
// offset stays a 32-bit register operand of the memory instruction.
template <typename Off>
__device__ inline Vec3 ld3(const float* p, Off byte_off) {
  const float* q = (const float*)((const char*)p + byte_off);
  return Vec3{q[0], q[1], q[2]};
}

template <typename Off>
__device__ inline void st3(float* p, Off byte_off, Vec3 v) {
  float* q = (float*)((char*)p + byte_off);
  q[0] = v.x;
  q[1] = v.y;
  q[2] = v.z;
}

template <typename Off>
__device__ inline int32_t ld_obst(const int32_t* p, Off byte_off) {
  return *(const int32_t*)((const char*)p + byte_off);
}

// 3 x with one full-rate shift-add (a 32-bit multiply would be a quarter-rate instruction).
__device__ inline uint32_t times3(uint32_t x) {
  uint32_t r;
  asm("v_lshl_add_u32 %0, %1, 1, %1" : "=v"(r) : "v"(x));
  return r;
}
__device__ inline uint64_t times3(uint64_t x) { return 3 * x; }

// RN(n / d) for d > 0 from y = RN(1 / d).  q = RN(n y) is within 2 ulps of n / d; one
// residual correction brings it within 1 ulp, and a second one, with the residual
// n - d q exact under the fused multiply-add, rounds to RN(n / d) (P. Markstein, "Computation
// of elementary functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990,
// theorem for y within half an ulp of 1/d and q within one ulp of n/d).  Valid for finite
// normal operands and quotients away from the overflow and underflow thresholds; a zero
// numerator yields +0.
__device__ inline float div_ry(float n, float d, float y) {
  float q = n * y;
  float r = __builtin_fmaf(-d, q, n);
  q = __builtin_fmaf(r, y, q);
  r = __builtin_fmaf(-d, q, n);
  return __builtin_fmaf(r, y, q);
}

// Per-axis data of the coordinate lookup and the diffusion term.
struct PredictAxis {
  const float* cen;   // centres
  const float* rw;    // RN(1 / width)            (LDS tables) or nullptr
  const float* rcd;   // RN(1 / centre distance)  (LDS tables) or nullptr
  const float* rgap;  // RN(1 / max(c[l+1] - c[l], 1e-18))
  const float* w;     // widths
  const float* cd;    // centre distances
  float c0, cl, hwlo, hwhi, rhwlo, rhwhi;
  uint32_t n;
};

// Fractional coordinate of `pos` (fluid.py:2406-2435).  The bracketing pair of centres is
// looked for next to the cell's own index `own` first and by the reference's binary search
// over the remaining range otherwise; for non-decreasing centres the bracket
// c[lo] <= pos < c[lo + 1] is unique, so both find the same lo.
template <bool kTables>
__device__ inline float predict_coord(float pos, uint32_t own, const PredictAxis& A) {
#pragma clang fp contract(off)
  if (pos <= A.c0) {
    const float num = -0.5f * (A.c0 - pos);
    return kTables ? div_ry(num, A.hwlo, A.rhwlo) : num / A.hwlo;
  }
  if (pos >= A.cl) {
    const float num = 0.5f * (pos - A.cl);
    return (float)(A.n - 1) + (kTables ? div_ry(num, A.hwhi, A.rhwhi) : num / A.hwhi);
  }
  const float* cen = A.cen;
  uint32_t lo, hi;
  const uint32_t l = min(own, A.n - 2);
  if (pos >= cen[l]) {
    // l + 1 == n - 1 implies pos < c[n - 1] = c[l + 1]; otherwise l + 2 <= n - 1
    if (pos < cen[l + 1]) {
      lo = l;
      hi = l;
    } else if (pos < cen[l + 2]) {
      lo = l + 1;
      hi = l + 1;
    } else {
      lo = l + 2;  // c[l + 2] <= pos < c[n - 1]
      hi = A.n - 1;
    }
  } else {
    // pos > c[0] here, so l >= 1
    if (pos >= cen[l - 1]) {
      lo = l - 1;
      hi = l - 1;
    } else {
      lo = 0;
      hi = l - 1;
    }
  }
  while (hi > lo + 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (pos >= cen[m])
      lo = m;
    else
      hi = m;
  }
  const float gap = fmaxf(cen[lo + 1] - cen[lo], 1.0e-18f);
  const float num = pos - cen[lo];
  return (float)lo + (kTables ? div_ry(num, gap, A.rgap[lo]) : num / gap);
}

// Lower corner index clamp(floor(g), 0, n - 2) and its fraction clamp(g - i0, 0, 1)
// (fluid.py:2384-2389); the clamp acts on the float floor, which equals the reference
// for every finite coordinate.
__device__ inline uint32_t predict_corner(float g, uint32_t n, float& t) {
#pragma clang fp contract(off)
  const float f = fminf(fmaxf(floorf(g), 0.0f), (float)(n - 2));
  const uint32_t i0 = (uint32_t)f;
  t = fminf(fmaxf(g - (float)i0, 0.0f), 1.0f);
  return i0;
}

// Obstacle-aware trilinear sample of a.vp at grid coordinates (gx, gy, gz) (fluid.py:2382-2404).
// `plane` (LDS; read with kTables and 32-bit offsets only): i * ny * nz for each plane i.
template <bool kTables, typename Off>
__device__ inline Vec3 predict_sample(const PredictArgs& a, const uint32_t* plane, float gx, float gy, float gz,
                                      Vec3 fallback) {
#pragma clang fp contract(off)
  float tx, ty, tz;
  const uint32_t i0 = predict_corner(gx, a.nx, tx);
  const uint32_t j0 = predict_corner(gy, a.ny, ty);
  const uint32_t k0 = predict_corner(gz, a.nz, tz);
  const Off sx = (Off)a.ny * a.nz, sy = a.nz;
  Off base;
  if constexpr (sizeof(Off) == 4 && kTables) {
    base = plane[i0] + __umul24(j0, a.nz) + k0;
  } else {
    base = ((Off)i0 * a.ny + j0) * a.nz + k0;
  }
  int32_t ob[8];
  Vec3 u[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const Off c4 = (base + ((q >> 2) & 1) * sx + ((q >> 1) & 1) * sy + (q & 1)) * 4;
    ob[q] = ld_obst(a.obst, c4);
    u[q] = ld3(a.vp, times3(c4));
  }
  Vec3 val{0.0f, 0.0f, 0.0f};
  float fw = 0.0f;
  const int32_t any_solid = ob[0] | ob[1] | ob[2] | ob[3] | ob[4] | ob[5] | ob[6] | ob[7];
  if (__builtin_amdgcn_ballot_w64(any_solid != 0) == 0) {
    // every corner of every lane of the wave is fluid: the same sums without the masks
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float wx = (q & 4) ? tx : 1.0f - tx;
      const float wy = (q & 2) ? ty : 1.0f - ty;
      const float wz = (q & 1) ? tz : 1.0f - tz;
      const float w = wx * wy * wz;
      val.x = val.x + w * u[q].x;
      val.y = val.y + w * u[q].y;
      val.z = val.z + w * u[q].z;
      fw = fw + w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float wx = (q & 4) ? tx : 1.0f - tx;
      const float wy = (q & 2) ? ty : 1.0f - ty;
      const float wz = (q & 1) ? tz : 1.0f - tz;
      const float w = wx * wy * wz;
      if (ob[q] == 0) {
        val.x = val.x + w * u[q].x;
        val.y = val.y + w * u[q].y;
        val.z = val.z + w * u[q].z;
        fw = fw + w;
      }
    }
  }
  if (fw > 1.0e-12f) {
    if (kTables) {
      const float y = 1.0f / fw;
      return Vec3{div_ry(val.x, fw, y), div_ry(val.y, fw, y), div_ry(val.z, fw, y)};
    }
    return Vec3{val.x / fw, val.y / fw, val.z / fw};
  }
  return fallback;
}

// (hi - lo) / d when `open`, else 0.
template <bool kTables>
__device__ inline Vec3 predict_flux(bool open, Vec3 lo, Vec3 hi, float d, float rd) {
#pragma clang fp contract(off)
  if (!open) return Vec3{0.0f, 0.0f, 0.0f};
  if (kTables) return Vec3{div_ry(hi.x - lo.x, d, rd), div_ry(hi.y - lo.y, d, rd), div_ry(hi.z - lo.z, d, rd)};
  return Vec3{(hi.x - lo.x) / d, (hi.y - lo.y) / d, (hi.z - lo.z) / d};
}

template <bool kTables>
__device__ inline float predict_lap(float f, float b, float w, float rw) {
#pragma clang fp contract(off)
  return kTables ? div_ry(f - b, w, rw) : (f - b) / w;
}

// Neighbour data of the FV diffusion term of one cell.
struct PredictNb {
  Vec3 ctr, uxm, uxp, uym, uyp, uzm, uzp;
  float cd[6], rcd[6];  // centre distances (and reciprocals) of the -x, +x, -y, +y, -z, +z faces
  float w[3], rw[3];    // widths (and reciprocals) along x, y, z
};

// lap(u) = ((fxf - fxb) / wx + (fyf - fyb) / wy) + (fzf - fzb) / wz with face fluxes
// (hi - lo) / d; kMasked zeroes the fluxes of closed faces (fluid.py:2484-2512).
template <bool kTables, bool kMasked>
__device__ inline Vec3 predict_laplacian(const PredictNb& b, const bool (&open)[6]) {
#pragma clang fp contract(off)
  const Vec3 fxb = predict_flux<kTables>(!kMasked || open[0], b.uxm, b.ctr, b.cd[0], b.rcd[0]);
  const Vec3 fxf = predict_flux<kTables>(!kMasked || open[1], b.ctr, b.uxp, b.cd[1], b.rcd[1]);
  const Vec3 fyb = predict_flux<kTables>(!kMasked || open[2], b.uym, b.ctr, b.cd[2], b.rcd[2]);
  const Vec3 fyf = predict_flux<kTables>(!kMasked || open[3], b.ctr, b.uyp, b.cd[3], b.rcd[3]);
  const Vec3 fzb = predict_flux<kTables>(!kMasked || open[4], b.uzm, b.ctr, b.cd[4], b.rcd[4]);
  const Vec3 fzf = predict_flux<kTables>(!kMasked || open[5], b.ctr, b.uzp, b.cd[5], b.rcd[5]);
  Vec3 lap;
  lap.x = (predict_lap<kTables>(fxf.x, fxb.x, b.w[0], b.rw[0]) + predict_lap<kTables>(fyf.x, fyb.x, b.w[1], b.rw[1])) +
          predict_lap<kTables>(fzf.x, fzb.x, b.w[2], b.rw[2]);
  lap.y = (predict_lap<kTables>(fxf.y, fxb.y, b.w[0], b.rw[0]) + predict_lap<kTables>(fyf.y, fyb.y, b.w[1], b.rw[1])) +
          predict_lap<kTables>(fzf.y, fzb.y, b.w[2], b.rw[2]);
  lap.z = (predict_lap<kTables>(fxf.z, fxb.z, b.w[0], b.rw[0]) + predict_lap<kTables>(fyf.z, fyb.z, b.w[1], b.rw[1])) +
          predict_lap<kTables>(fzf.z, fzb.z, b.w[2], b.rw[2]);
  return lap;
}

// kLds: centres and reciprocal tables in LDS (nx + ny + nz <= kPredictLdsMaxCells), all
// quotients through div_ry; otherwise centres from global memory and IEEE division.
// Off: byte-offset type (uint32_t when 12 n < 2^32).
template <bool kRk2, bool kLds, typename Off>
__global__ void __launch_bounds__(kBlock) k_predict(PredictArgs a) {
#pragma clang fp contract(off)
  extern __shared__ float s_tab[];
  const uint32_t tot = a.nx + a.ny + a.nz;
  const uint32_t nn[3] = {a.nx, a.ny, a.nz};
  PredictAxis ax[3];
  uint32_t off = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    PredictAxis& A = ax[d];
    const uint32_t n = nn[d];
    A.n = n;
    A.w = a.width[d];
    A.cd = a.cdist[d];
    A.c0 = a.centre[d][0];
    A.cl = a.centre[d][n - 1];
    A.hwlo = fmaxf(A.c0 - a.face[d][0], 1.0e-18f);
    A.hwhi = fmaxf(a.face[d][n] - A.cl, 1.0e-18f);
    A.rhwlo = 1.0f / A.hwlo;
    A.rhwhi = 1.0f / A.hwhi;
    if (kLds) {
      A.cen = s_tab + off;
      A.rw = s_tab + tot + off;
      A.rcd = s_tab + 2 * tot + off;
      A.rgap = s_tab + 3 * tot + off;
      A.w = s_tab + 4 * tot + off;
      A.cd = s_tab + 5 * tot + off;
      if (sizeof(Off) == 4 && d == 0) {
        uint32_t* plane = (uint32_t*)(s_tab + kPredictLdsTables * tot);
        for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) plane[t] = t * (a.ny * a.nz);
      }
      for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        const float c = a.centre[d][t];
        const float w = a.width[d][t];
        const float cd = a.cdist[d][t];
        s_tab[off + t] = c;
        s_tab[tot + off + t] = 1.0f / w;
        s_tab[2 * tot + off + t] = 1.0f / cd;
        s_tab[3 * tot + off + t] = t + 1 < n ? 1.0f / fmaxf(a.centre[d][t + 1] - c, 1.0e-18f) : 0.0f;
        s_tab[4 * tot + off + t] = w;
        s_tab[5 * tot + off + t] = cd;
      }
    } else {
      A.cen = a.centre[d];
      A.rw = A.rcd = A.rgap = nullptr;
    }
    off += n;
  }
  if (kLds) __syncthreads();
  const uint32_t* plane = (kLds && sizeof(Off) == 4) ? (const uint32_t*)(s_tab + kPredictLdsTables * tot) : nullptr;

  const float dt = a.dt;
  const float dtnu = a.dt * a.nu;
  const Off sx = (Off)a.ny * a.nz, sy = a.nz;
  HFSI_CELL_LOOP(a.n) {
    const uint32_t t = a.fz.div(c);
    const uint32_t k = c - t * a.nz;
    const uint32_t i = a.fy.div(t);
    const uint32_t j = t - i * a.ny;
    // every load below is at a clamped, in-range index; masks select what counts
    const Off c4 = (Off)c * 4, sx4 = sx * 4, sy4 = sy * 4;
    const Off c12 = times3(c4);
    const Off xm4 = i > 0 ? c4 - sx4 : c4, xp4 = i + 1 < a.nx ? c4 + sx4 : c4;
    const Off ym4 = j > 0 ? c4 - sy4 : c4, yp4 = j + 1 < a.ny ? c4 + sy4 : c4;
    const Off zm4 = k > 0 ? c4 - 4 : c4, zp4 = k + 1 < a.nz ? c4 + 4 : c4;
    const int32_t oc = ld_obst(a.obst, c4);
    const Vec3 ctr = ld3(a.vp, c12);
    const int32_t oxm = ld_obst(a.obst, xm4), oxp = ld_obst(a.obst, xp4);
    const int32_t oym = ld_obst(a.obst, ym4), oyp = ld_obst(a.obst, yp4);
    const int32_t ozm = ld_obst(a.obst, zm4), ozp = ld_obst(a.obst, zp4);
    const Vec3 uxm = ld3(a.vp, times3(xm4)), uxp = ld3(a.vp, times3(xp4));
    const Vec3 uym = ld3(a.vp, times3(ym4)), uyp = ld3(a.vp, times3(yp4));
    const Vec3 uzm = ld3(a.vp, times3(zm4)), uzp = ld3(a.vp, times3(zp4));
    if (oc == 1) {
      st3(a.v, c12, Vec3{0.0f, 0.0f, 0.0f});
      continue;
    }
    const float x = ax[0].cen[i], y = ax[1].cen[j], z = ax[2].cen[k];
    Vec3 tr = ctr;
    if (kRk2) {
      const float hd = 0.5f * dt;
      const float mgx = predict_coord<kLds>(x - hd * ctr.x, i, ax[0]);
      const float mgy = predict_coord<kLds>(y - hd * ctr.y, j, ax[1]);
      const float mgz = predict_coord<kLds>(z - hd * ctr.z, k, ax[2]);
      tr = predict_sample<kLds, Off>(a, plane, mgx, mgy, mgz, ctr);
    }
    const float gx = predict_coord<kLds>(x - dt * tr.x, i, ax[0]);
    const float gy = predict_coord<kLds>(y - dt * tr.y, j, ax[1]);
    const float gz = predict_coord<kLds>(z - dt * tr.z, k, ax[2]);
    const Vec3 adv = predict_sample<kLds, Off>(a, plane, gx, gy, gz, ctr);

    const uint32_t ip = i + 1 < a.nx ? i + 1 : i, jp = j + 1 < a.ny ? j + 1 : j, kp = k + 1 < a.nz ? k + 1 : k;
    PredictNb b;
    b.ctr = ctr;
    b.uxm = uxm;
    b.uxp = uxp;
    b.uym = uym;
    b.uyp = uyp;
    b.uzm = uzm;
    b.uzp = uzp;
    b.cd[0] = ax[0].cd[i];
    b.cd[1] = ax[0].cd[ip];
    b.cd[2] = ax[1].cd[j];
    b.cd[3] = ax[1].cd[jp];
    b.cd[4] = ax[2].cd[k];
    b.cd[5] = ax[2].cd[kp];
    b.w[0] = ax[0].w[i];
    b.w[1] = ax[1].w[j];
    b.w[2] = ax[2].w[k];
    if (kLds) {
      b.rcd[0] = ax[0].rcd[i];
      b.rcd[1] = ax[0].rcd[ip];
      b.rcd[2] = ax[1].rcd[j];
      b.rcd[3] = ax[1].rcd[jp];
      b.rcd[4] = ax[2].rcd[k];
      b.rcd[5] = ax[2].rcd[kp];
      b.rw[0] = ax[0].rw[i];
      b.rw[1] = ax[1].rw[j];
      b.rw[2] = ax[2].rw[k];
    } else {
      for (int q = 0; q < 6; ++q) b.rcd[q] = 0.0f;
      for (int q = 0; q < 3; ++q) b.rw[q] = 0.0f;
    }
    const bool open[6] = {i > 0 && oxm == 0, i + 1 < a.nx && oxp == 0, j > 0 && oym == 0,
                          j + 1 < a.ny && oyp == 0, k > 0 && ozm == 0, k + 1 < a.nz && ozp == 0};
    const bool shut = !(open[0] && open[1] && open[2] && open[3] && open[4] && open[5]);
    // every face of every lane of the wave open: the same fluxes without the masks
    const Vec3 lap = __builtin_amdgcn_ballot_w64(shut) == 0 ? predict_laplacian<kLds, false>(b, open)
                                                            : predict_laplacian<kLds, true>(b, open);
    st3(a.v, c12, Vec3{adv.x + dtnu * lap.x, adv.y + dtnu * lap.y, adv.z + dtnu * lap.z});
  }
}

}  // namespace hfsi
