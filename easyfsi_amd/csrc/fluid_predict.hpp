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
// c = (i * ny + j) * nz + k.  One thread per cell under the XCD-aware chunk loop; the
// cell centres of the three axes are staged in LDS for the binary searches when they
// fit (kLds), otherwise read from global memory.
#pragma once

#include "hfsi_common.hpp"

namespace hfsi {

constexpr uint32_t kPredictLdsMaxCentres = 12288;  // 48 KiB of LDS

struct PredictArgs {
  uint32_t nx, ny, nz, n;
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

__device__ inline Vec3 ld3(const float* p, uint32_t c) {
  const float* q = p + 3 * (size_t)c;
  return Vec3{q[0], q[1], q[2]};
}

__device__ inline void st3(float* p, uint32_t c, Vec3 v) {
  float* q = p + 3 * (size_t)c;
  q[0] = v.x;
  q[1] = v.y;
  q[2] = v.z;
}

// Fractional coordinate of position `pos` on an axis with `n` centres `cen` (fluid.py:2406-2435).
__device__ inline float predict_coord(float pos, const float* cen, const float* face, uint32_t n) {
#pragma clang fp contract(off)
  const float c0 = cen[0];
  const float cl = cen[n - 1];
  if (pos <= c0) {
    const float hw = fmaxf(c0 - face[0], 1.0e-18f);
    return (-0.5f * (c0 - pos)) / hw;
  }
  if (pos >= cl) {
    const float hw = fmaxf(face[n] - cl, 1.0e-18f);
    return (float)(n - 1) + (0.5f * (pos - cl)) / hw;
  }
  uint32_t lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (pos >= cen[m])
      lo = m;
    else
      hi = m;
  }
  hi = min(lo + 1, n - 1);
  const float dist = fmaxf(cen[hi] - cen[lo], 1.0e-18f);
  return (float)lo + (pos - cen[lo]) / dist;
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
__device__ inline Vec3 predict_sample(const PredictArgs& a, float gx, float gy, float gz, Vec3 fallback) {
#pragma clang fp contract(off)
  float tx, ty, tz;
  const uint32_t i0 = predict_corner(gx, a.nx, tx);
  const uint32_t j0 = predict_corner(gy, a.ny, ty);
  const uint32_t k0 = predict_corner(gz, a.nz, tz);
  const uint32_t sx = a.ny * a.nz, sy = a.nz;
  const uint32_t base = (i0 * a.ny + j0) * a.nz + k0;
  int32_t ob[8];
  Vec3 u[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t c = base + ((q >> 2) & 1) * sx + ((q >> 1) & 1) * sy + (q & 1);
    ob[q] = a.obst[c];
    u[q] = ld3(a.vp, c);
  }
  Vec3 val{0.0f, 0.0f, 0.0f};
  float fw = 0.0f;
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
  if (fw > 1.0e-12f) return Vec3{val.x / fw, val.y / fw, val.z / fw};
  return fallback;
}

// Diffusive flux across one face: (hi - lo) / d when `open`, else 0.
__device__ inline Vec3 predict_flux(bool open, Vec3 lo, Vec3 hi, float d) {
#pragma clang fp contract(off)
  if (!open) return Vec3{0.0f, 0.0f, 0.0f};
  return Vec3{(hi.x - lo.x) / d, (hi.y - lo.y) / d, (hi.z - lo.z) / d};
}

template <bool kRk2, bool kLds>
__global__ void __launch_bounds__(kBlock) k_predict(PredictArgs a) {
#pragma clang fp contract(off)
  extern __shared__ float s_cen[];
  const float* cen[3] = {a.centre[0], a.centre[1], a.centre[2]};
  if (kLds) {
    const uint32_t tot = a.nx + a.ny + a.nz;
    for (uint32_t t = threadIdx.x; t < tot; t += blockDim.x)
      s_cen[t] = t < a.nx ? a.centre[0][t] : (t < a.nx + a.ny ? a.centre[1][t - a.nx] : a.centre[2][t - a.nx - a.ny]);
    __syncthreads();
    cen[0] = s_cen;
    cen[1] = s_cen + a.nx;
    cen[2] = s_cen + a.nx + a.ny;
  }
  const float dt = a.dt;
  const float dtnu = a.dt * a.nu;
  const uint32_t sx = a.ny * a.nz, sy = a.nz;
  HFSI_CELL_LOOP(a.n) {
    const uint32_t k = c % a.nz;
    const uint32_t t = c / a.nz;
    const uint32_t j = t % a.ny;
    const uint32_t i = t / a.ny;
    // every load below is at a clamped, in-range index; masks select what counts
    const int32_t oc = a.obst[c];
    const Vec3 ctr = ld3(a.vp, c);
    const uint32_t cxm = i > 0 ? c - sx : c, cxp = i + 1 < a.nx ? c + sx : c;
    const uint32_t cym = j > 0 ? c - sy : c, cyp = j + 1 < a.ny ? c + sy : c;
    const uint32_t czm = k > 0 ? c - 1 : c, czp = k + 1 < a.nz ? c + 1 : c;
    const int32_t oxm = a.obst[cxm], oxp = a.obst[cxp], oym = a.obst[cym], oyp = a.obst[cyp];
    const int32_t ozm = a.obst[czm], ozp = a.obst[czp];
    const Vec3 uxm = ld3(a.vp, cxm), uxp = ld3(a.vp, cxp), uym = ld3(a.vp, cym), uyp = ld3(a.vp, cyp);
    const Vec3 uzm = ld3(a.vp, czm), uzp = ld3(a.vp, czp);
    if (oc == 1) {
      st3(a.v, c, Vec3{0.0f, 0.0f, 0.0f});
      continue;
    }
    const float x = cen[0][i], y = cen[1][j], z = cen[2][k];
    Vec3 tr = ctr;
    if (kRk2) {
      const float hd = 0.5f * dt;
      const float mgx = predict_coord(x - hd * ctr.x, cen[0], a.face[0], a.nx);
      const float mgy = predict_coord(y - hd * ctr.y, cen[1], a.face[1], a.ny);
      const float mgz = predict_coord(z - hd * ctr.z, cen[2], a.face[2], a.nz);
      tr = predict_sample(a, mgx, mgy, mgz, ctr);
    }
    const float gx = predict_coord(x - dt * tr.x, cen[0], a.face[0], a.nx);
    const float gy = predict_coord(y - dt * tr.y, cen[1], a.face[1], a.ny);
    const float gz = predict_coord(z - dt * tr.z, cen[2], a.face[2], a.nz);
    const Vec3 adv = predict_sample(a, gx, gy, gz, ctr);

    const Vec3 fxb = predict_flux(i > 0 && oxm == 0, uxm, ctr, a.cdist[0][i]);
    const Vec3 fxf = predict_flux(i + 1 < a.nx && oxp == 0, ctr, uxp, a.cdist[0][i + 1 < a.nx ? i + 1 : i]);
    const Vec3 fyb = predict_flux(j > 0 && oym == 0, uym, ctr, a.cdist[1][j]);
    const Vec3 fyf = predict_flux(j + 1 < a.ny && oyp == 0, ctr, uyp, a.cdist[1][j + 1 < a.ny ? j + 1 : j]);
    const Vec3 fzb = predict_flux(k > 0 && ozm == 0, uzm, ctr, a.cdist[2][k]);
    const Vec3 fzf = predict_flux(k + 1 < a.nz && ozp == 0, ctr, uzp, a.cdist[2][k + 1 < a.nz ? k + 1 : k]);
    const float wx = a.width[0][i], wy = a.width[1][j], wz = a.width[2][k];
    Vec3 lap;
    lap.x = ((fxf.x - fxb.x) / wx + (fyf.x - fyb.x) / wy) + (fzf.x - fzb.x) / wz;
    lap.y = ((fxf.y - fxb.y) / wx + (fyf.y - fyb.y) / wy) + (fzf.y - fzb.y) / wz;
    lap.z = ((fxf.z - fxb.z) / wx + (fyf.z - fyb.z) / wy) + (fzf.z - fzb.z) / wz;
    st3(a.v, c, Vec3{adv.x + dtnu * lap.x, adv.y + dtnu * lap.y, adv.z + dtnu * lap.z});
  }
}

}  // namespace hfsi
