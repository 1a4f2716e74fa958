// 2.5D marching stencil kernels for the large multigrid levels (gfx950).
//
// Layout: a thread owns 4 consecutive k-cells (k0..k0+3, k0 % 4 == 0) of one
// j-row and marches along i over a chunk of planes; the +-i neighbours stay in
// registers (previous / prefetched next plane), the +-j rows are loaded as
// 4-vectors, the two k-halo values as scalars.  Per-axis coefficients are
// per-thread constants (y, z) or uniform per plane (x), so no per-cell index
// arithmetic is needed.  Workgroups are TK x TJ threads (TK along k); the
// logical block order is (i-chunk, j-block, k-block) with k fastest and the
// physical -> logical remap gives every XCD a contiguous slab of i-chunks.
//
// Requires nz % 4 == 0 and no velocity-Dirichlet face weights (those grids use
// the flat kernels of fluid_kernels.hpp).
#pragma once

#include <hip/hip_fp16.h>

#include "fluid_kernels.hpp"

namespace hfsi {

// ---------------------------------------------------------------------------
// 4-wide storage access
// ---------------------------------------------------------------------------

template <typename T>
struct Pack4;

template <>
struct Pack4<float> {
  template <typename C>
  __device__ static inline void load(const float* p, C v[4]) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = (C)q.x; v[1] = (C)q.y; v[2] = (C)q.z; v[3] = (C)q.w;
  }
  template <typename C>
  __device__ static inline void store(float* p, const C v[4]) {
    *reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
  }
};

template <>
struct Pack4<double> {
  template <typename C>
  __device__ static inline void load(const double* p, C v[4]) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    v[0] = (C)a.x; v[1] = (C)a.y; v[2] = (C)b.x; v[3] = (C)b.y;
  }
  template <typename C>
  __device__ static inline void store(double* p, const C v[4]) {
    *reinterpret_cast<double2*>(p) = make_double2((double)v[0], (double)v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2((double)v[2], (double)v[3]);
  }
};

template <>
struct Pack4<__half> {
  template <typename C>
  __device__ static inline void load(const __half* p, C v[4]) {
    const uint2 raw = *reinterpret_cast<const uint2*>(p);
    const __half2 a = *reinterpret_cast<const __half2*>(&raw.x);
    const __half2 b = *reinterpret_cast<const __half2*>(&raw.y);
    const float2 fa = __half22float2(a), fb = __half22float2(b);
    v[0] = (C)fa.x; v[1] = (C)fa.y; v[2] = (C)fb.x; v[3] = (C)fb.y;
  }
  template <typename C>
  __device__ static inline void store(__half* p, const C v[4]) {
    const __half2 a = __floats2half2_rn((float)v[0], (float)v[1]);
    const __half2 b = __floats2half2_rn((float)v[2], (float)v[3]);
    uint2 raw;
    raw.x = *reinterpret_cast<const uint32_t*>(&a);
    raw.y = *reinterpret_cast<const uint32_t*>(&b);
    *reinterpret_cast<uint2*>(p) = raw;
  }
};

// Value of v after a round trip through storage type T (register-only).
template <typename T, typename C>
__device__ inline C storage_round(C v) {
  return (C)(T)v;
}
template <>
__device__ inline float storage_round<__half, float>(float v) {
  return __half2float(__float2half_rn(v));
}

__device__ inline double to_c_exact(double v) { return v; }
__device__ inline float to_c_exact(float v) { return v; }
__device__ inline float to_c_exact(__half v) { return __half2float(v); }

// ---------------------------------------------------------------------------
// Launch geometry
// ---------------------------------------------------------------------------

struct MarchGeo {
  uint32_t nx, ny, nz, sx, sy;
  uint32_t tk, tj;               // threads along k-quads and along j
  uint32_t kblocks, jblocks, iblocks, ichunk;
  uint32_t nblocks;              // real logical blocks
};

inline MarchGeo march_geo(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t ichunk) {
  MarchGeo m;
  m.nx = nx; m.ny = ny; m.nz = nz; m.sx = ny * nz; m.sy = nz;
  const uint32_t quads = nz / 4;
  uint32_t tk = 1;
  while (tk < quads && tk < 64) tk <<= 1;  // power of two: workgroups are always 256 threads
  m.tk = tk;
  m.tj = 256 / m.tk;
  m.kblocks = (quads + m.tk - 1) / m.tk;
  m.jblocks = (ny + m.tj - 1) / m.tj;
  m.ichunk = ichunk < nx ? ichunk : nx;
  m.iblocks = (nx + m.ichunk - 1) / m.ichunk;
  m.nblocks = m.kblocks * m.jblocks * m.iblocks;
  return m;
}

inline uint32_t march_grid(const MarchGeo& m) { return (m.nblocks + HFSI_NXCD - 1) / HFSI_NXCD * HFSI_NXCD; }
inline uint32_t march_threads(const MarchGeo&) { return 256; }

struct MarchPos {
  bool valid;   // the workgroup has work
  bool active;  // this thread owns cells
  uint32_t k0, j, i0, i1;
};

__device__ inline MarchPos march_pos(const MarchGeo& m) {
  MarchPos p;
  const uint32_t G = gridDim.x;
  const uint32_t b = blockIdx.x;
  const uint32_t L = (b % HFSI_NXCD) * (G / HFSI_NXCD) + b / HFSI_NXCD;
  p.valid = L < m.nblocks;
  const uint32_t kb = L % m.kblocks;
  const uint32_t t = L / m.kblocks;
  const uint32_t jb = t % m.jblocks;
  const uint32_t ib = t / m.jblocks;
  const uint32_t tk = threadIdx.x % m.tk;
  const uint32_t tj = threadIdx.x / m.tk;
  p.k0 = (kb * m.tk + tk) * 4;
  p.j = jb * m.tj + tj;
  p.i0 = ib * m.ichunk;
  p.i1 = min(m.nx, p.i0 + m.ichunk);
  p.active = p.valid && p.k0 < m.nz && p.j < m.ny;
  return p;
}

// ---------------------------------------------------------------------------
// Weighted-Jacobi sweep / apply / residual in marching form
// ---------------------------------------------------------------------------

enum MarchMode { kModeNormal = 0, kModeFromZero = 1, kModeProlong = 2, kModeApply = 3, kModeResidual = 4 };

// partial sums per mode: <d, A d>_V (apply); {<r,r>_V, sum V r} of r = rhs - A p (residual, nothing
// stored); the five CG sums (sweeps)
template <int MODE>
constexpr int march_sums() {
  return MODE == kModeApply ? 1 : (MODE == kModeResidual ? 2 : 5);
}

template <typename C, typename TP, typename TR, typename TO, typename TQ>
struct MarchArgs {
  MarchGeo m;
  const uint8_t* flags;
  const double* diag;
  const C* cl[3];
  const C* ch[3];
  const double* wd[3];
  C outlet_add;
  int outlet;
  C omega;
  C rhs_sign;     // rhs value = rhs_sign * stored value
  const TP* p;    // input iterate (unused for kFromZero)
  const TR* rhs;  // right-hand side (unused for kApply)
  TO* out;
  // prolongation (kModeProlong): p_eff = p + s * coarse[parent] on free cells
  const float* pc;
  uint32_t cnx, cny, cnz, csx, csy;
  C scale;
  // fused reductions
  const TQ* q;  // kNormal/kProlong with DOTS: CG search-direction product
  double* partials;
  unsigned int* counter;
  double* totals;
};

// load 4 values at (i, j, k0..k0+3) with the prolongation added (raw, add split)
template <typename C, typename TP, typename TR, typename TO, typename TQ, int MODE>
__device__ inline void march_load(const MarchArgs<C, TP, TR, TO, TQ>& a, uint32_t i, uint32_t j, uint32_t k0,
                                  C raw[4], C add[4]) {
  const uint32_t c = i * a.m.sx + j * a.m.sy + k0;
  Pack4<TP>::load(a.p + c, raw);
  if (MODE == kModeProlong) {
    const uint32_t ci = min(i >> 1, a.cnx - 1), cj = min(j >> 1, a.cny - 1);
    const uint32_t ck0 = min(k0 >> 1, a.cnz - 1), ck1 = min((k0 >> 1) + 1, a.cnz - 1);
    const float* row = a.pc + ci * a.csx + cj * a.csy;
    const C v0 = a.scale * (C)row[ck0], v1 = a.scale * (C)row[ck1];
    add[0] = v0; add[1] = v0; add[2] = v1; add[3] = v1;
  } else {
    add[0] = add[1] = add[2] = add[3] = (C)0;
  }
}

// One marching kernel for: weighted-Jacobi sweep (normal / from zero /
// prolongation fused), operator apply with <d, A d>_V (kModeApply + DOTS),
// the norm of the residual rhs_sign * rhs - A p (kModeResidual + DOTS, no
// store) and the final level-0 post-sweep with the five CG sums (DOTS).
template <typename C, typename TP, typename TR, typename TO, typename TQ, int MODE, bool DOTS>
__global__ void __launch_bounds__(256) k_march(MarchArgs<C, TP, TR, TO, TQ> a) {
  constexpr int NS = march_sums<MODE>();
  __shared__ double lds[DOTS ? NS * 4 : 1];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const MarchPos P = march_pos(a.m);
  if (P.active) {
    const uint32_t nx = a.m.nx, ny = a.m.ny, nz = a.m.nz, sx = a.m.sx, sy = a.m.sy;
    const uint32_t j = P.j, k0 = P.k0;
    const bool has_jm = j > 0, has_jp = j + 1 < ny;
    const bool has_km = k0 > 0, has_kp = k0 + 4 < nz;
    // per-thread constant coefficients
    const C cyl = a.cl[1][j], cyh = a.ch[1][j];
    C czl[4], czh[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      czl[m] = a.cl[2][k0 + m];
      czh[m] = a.ch[2][k0 + m];
    }
    const bool outlet_row = a.outlet && k0 == 0;
    double wyz[4];
    if (DOTS) {
#pragma unroll
      for (int m = 0; m < 4; ++m) wyz[m] = a.wd[1][j] * a.wd[2][k0 + m];
    }
    C pm_r[4], pm_a[4], pc_r[4], pc_a[4], pn_r[4], pn_a[4];
    const bool use_p = MODE != kModeFromZero;
    if (use_p) {
      if (P.i0 > 0)
        march_load<C, TP, TR, TO, TQ, MODE>(a, P.i0 - 1, j, k0, pm_r, pm_a);
      else
#pragma unroll
        for (int m = 0; m < 4; ++m) pm_r[m] = pm_a[m] = (C)0;
      march_load<C, TP, TR, TO, TQ, MODE>(a, P.i0, j, k0, pc_r, pc_a);
    }
    for (uint32_t i = P.i0; i < P.i1; ++i) {
      const uint32_t c = i * sx + j * sy + k0;
      C jm_r[4], jm_a[4], jp_r[4], jp_a[4], km = (C)0, kp = (C)0;
      if (use_p) {
        if (i + 1 < nx)
          march_load<C, TP, TR, TO, TQ, MODE>(a, i + 1, j, k0, pn_r, pn_a);
        else
#pragma unroll
          for (int m = 0; m < 4; ++m) pn_r[m] = pn_a[m] = (C)0;
        march_load<C, TP, TR, TO, TQ, MODE>(a, i, has_jm ? j - 1 : j, k0, jm_r, jm_a);
        march_load<C, TP, TR, TO, TQ, MODE>(a, i, has_jp ? j + 1 : j, k0, jp_r, jp_a);
        const C km_raw = (C)to_c_exact(a.p[has_km ? c - 1 : c]);
        const C kp_raw = (C)to_c_exact(a.p[has_kp ? c + 4 : c]);
        km = km_raw;
        kp = kp_raw;
        if (MODE == kModeProlong) {
          const uint32_t ci = min(i >> 1, a.cnx - 1), cj = min(j >> 1, a.cny - 1);
          const float* row = a.pc + ci * a.csx + cj * a.csy;
          if (has_km) km += a.scale * (C)row[min((k0 - 1) >> 1, a.cnz - 1)];
          if (has_kp) kp += a.scale * (C)row[min((k0 + 4) >> 1, a.cnz - 1)];
        }
      }
      const uint32_t f4 = *reinterpret_cast<const uint32_t*>(a.flags + c);
      C rhs[4];
      if (MODE != kModeApply) Pack4<TR>::load(a.rhs + c, rhs);
      C qv[4];
      if (DOTS && MODE != kModeApply && MODE != kModeResidual) Pack4<TQ>::load(a.q + c, qv);
      const C cxl = a.cl[0][i], cxh = a.ch[0][i];
      C out[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const uint32_t f = (f4 >> (8 * m)) & 0xFFu;
        const C center = use_p ? pc_r[m] + ((f & kNotFree) ? (C)0 : pc_a[m]) : (C)0;
        C nb = (C)0, den = (C)0;
        if (use_p) {
          if (f & kXm) { nb += cxl * (pm_r[m] + pm_a[m]); den += cxl; }
          if (f & kXp) { nb += cxh * (pn_r[m] + pn_a[m]); den += cxh; }
          if (f & kYm) { nb += cyl * (jm_r[m] + jm_a[m]); den += cyl; }
          if (f & kYp) { nb += cyh * (jp_r[m] + jp_a[m]); den += cyh; }
          const C left = m == 0 ? km : pc_r[m - 1] + pc_a[m - 1];
          const C right = m == 3 ? kp : pc_r[m + 1] + pc_a[m + 1];
          if (f & kZm) { nb += czl[m] * left; den += czl[m]; }
          if (f & kZp) { nb += czh[m] * right; den += czh[m]; }
        } else {
          if (f & kXm) den += cxl;
          if (f & kXp) den += cxh;
          if (f & kYm) den += cyl;
          if (f & kYp) den += cyh;
          if (f & kZm) den += czl[m];
          if (f & kZp) den += czh[m];
        }
        if (outlet_row && m == 0) den += a.outlet_add;
        if (a.diag) den += (C)a.diag[c + m];
        C v;
        if (MODE == kModeApply) {
          v = (f & kObst) ? (C)0 : den * center - nb;
        } else if (MODE == kModeResidual) {
          v = (f & kObst) ? (C)0 : a.rhs_sign * rhs[m] - (den * center - nb);
        } else if (f & kObst) {
          v = (C)0;
        } else if (!(den > (C)0)) {
          v = (C)0;
        } else {
          const C jac = (nb - a.rhs_sign * rhs[m]) / den;
          v = center + a.omega * (jac - center);
        }
        out[m] = v;
      }
      if (MODE != kModeResidual) Pack4<TO>::store(a.out + c, out);
      if (DOTS) {
        const double wx = a.wd[0][i];
        C stored[4];  // the rounded values actually kept
#pragma unroll
        for (int m = 0; m < 4; ++m) stored[m] = storage_round<TO, C>(out[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const uint32_t f = (f4 >> (8 * m)) & 0xFFu;
          if (f & kNotFree) continue;
          const double V = wx * wyz[m];
          if (MODE == kModeApply) {
            acc[0] += (double)pc_r[m] * (double)stored[m] * V;
          } else if (MODE == kModeResidual) {
            const double rv = (double)out[m];
            acc[0] += rv * rv * V;
            acc[1] += rv * V;
          } else {
            const double zv = (double)stored[m], rv = -(double)(a.rhs_sign * rhs[m]), qd = (double)qv[m];
            acc[0] += rv * zv * V;
            acc[1] += zv * qd * V;
            acc[2] += zv * V;
            acc[3] += rv * V;
            acc[4] += qd * V;
          }
        }
      }
      if (use_p) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          pm_r[m] = pc_r[m]; pm_a[m] = pc_a[m];
          pc_r[m] = pn_r[m]; pc_a[m] = pn_a[m];
        }
      }
    }
  }
  if (DOTS) {
    block_partials<NS>(acc, a.partials, lds);
  }
}

// ---------------------------------------------------------------------------
// Residual + restriction in marching form
// ---------------------------------------------------------------------------
//
// A thread owns the fine k-quad k0..k0+3 (coarse ck0 = k0/2, ck0 + 1) of the
// fine rows r0 = 2 cj, r1 = 2 cj + 1 and marches over coarse planes ci (fine
// planes 2 ci, 2 ci + 1), carrying the fine planes 2 ci - 1 and 2 ci + 2 in
// registers.  Geometry: the fine level's MarchGeo with ny -> ceil(ny/2) rows
// (m.ny is the COARSE row count here) and ichunk counted in coarse planes.

template <typename C, typename TP, typename TR, typename TRes>
struct RestrictArgs {
  MarchGeo m;                  // tk/tj/blocks over (coarse planes, coarse rows, fine k-quads)
  uint32_t fnx, fny, fnz, fsx, fsy;
  uint32_t cnx, cny, cnz, csx, csy;
  const uint8_t* flags;        // fine
  const uint8_t* cflags;       // coarse
  const double* diag;          // fine
  const C* cl[3];
  const C* ch[3];
  const float* wf[3];          // fine widths (volumes in float like the reference)
  C outlet_add;
  int outlet;
  C rhs_sign;
  const TP* p;
  const TR* rhs;
  float* out;                  // coarse rhs
};

template <typename C, typename TP>
__device__ inline void rr_load_row(const TP* p, uint32_t c, C v[4]) {
  Pack4<TP>::load(p + c, v);
}

template <typename C, typename TP, typename TR, typename TRes>
__global__ void __launch_bounds__(256) k_march_restrict(RestrictArgs<C, TP, TR, TRes> a) {
  const MarchPos P = march_pos(a.m);  // P.j = coarse row, P.i0/i1 = coarse planes, P.k0 = fine k-quad start
  if (!P.active) return;
  const uint32_t k0 = P.k0;
  const uint32_t cj = P.j;
  const uint32_t r0 = 2 * cj, r1 = 2 * cj + 1;
  const bool has_r1 = r1 < a.fny;
  const bool has_rm = r0 > 0;                 // row r0 - 1
  const bool has_rp = r1 + 1 < a.fny;         // row r1 + 1
  const bool has_km = k0 > 0, has_kp = k0 + 4 < a.fnz;
  const uint32_t rows[2] = {r0, has_r1 ? r1 : r0};
  C cyl[2], cyh[2];
  float wy[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    cyl[r] = a.cl[1][rows[r]];
    cyh[r] = a.ch[1][rows[r]];
    wy[r] = a.wf[1][rows[r]];
  }
  C czl[4], czh[4];
  float wz[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    czl[m] = a.cl[2][k0 + m];
    czh[m] = a.ch[2][k0 + m];
    wz[m] = a.wf[2][k0 + m];
  }
  const bool outlet_row = a.outlet && k0 == 0;
  // plane registers: [plane][row][4]; planes: prev (2ci-1), A (2ci), B (2ci+1), next (2ci+2)
  C pv[2][4], pa[2][4], pb[2][4], pn[2][4];
  auto load_plane = [&](uint32_t fi, C dst[2][4]) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (fi < a.fnx)
        rr_load_row<C, TP>(a.p, fi * a.fsx + rows[r] * a.fsy + k0, dst[r]);
      else
#pragma unroll
        for (int m = 0; m < 4; ++m) dst[r][m] = (C)0;
    }
  };
  const uint32_t fi_start = 2 * P.i0;
  if (fi_start > 0)
    load_plane(fi_start - 1, pv);
  else
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int m = 0; m < 4; ++m) pv[r][m] = (C)0;
  load_plane(fi_start, pa);
  load_plane(fi_start + 1, pb);
  for (uint32_t ci = P.i0; ci < P.i1; ++ci) {
    const uint32_t fa = 2 * ci, fb = 2 * ci + 1;
    load_plane(fb + 1, pn);
    C rsum[2] = {(C)0, (C)0}, vsum[2] = {(C)0, (C)0};
#pragma unroll
    for (int pl = 0; pl < 2; ++pl) {
      const uint32_t fi = pl == 0 ? fa : fb;
      if (fi >= a.fnx) continue;
      C(&cur)[2][4] = pl == 0 ? pa : pb;
      C(&lo)[2][4] = pl == 0 ? pv : pa;
      C(&hi)[2][4] = pl == 0 ? pb : pn;
      const C cxl = a.cl[0][fi], cxh = a.ch[0][fi];
      const float wx = a.wf[0][fi];
      // j-halo rows of this plane
      C hm[4], hp[4];
      rr_load_row<C, TP>(a.p, fi * a.fsx + (has_rm ? r0 - 1 : r0) * a.fsy + k0, hm);
      rr_load_row<C, TP>(a.p, fi * a.fsx + (has_rp ? r1 + 1 : rows[1]) * a.fsy + k0, hp);
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        if (r == 1 && !has_r1) continue;
        const uint32_t c = fi * a.fsx + rows[r] * a.fsy + k0;
        const uint32_t f4 = *reinterpret_cast<const uint32_t*>(a.flags + c);
        C rhs[4];
        Pack4<TR>::load(a.rhs + c, rhs);
        const C km = (C)to_c_exact(a.p[has_km ? c - 1 : c]);
        const C kp = (C)to_c_exact(a.p[has_kp ? c + 4 : c]);
        const C(&ym)[4] = r == 0 ? hm : cur[0];
        const C(&yp)[4] = r == 0 ? cur[1] : hp;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const uint32_t f = (f4 >> (8 * m)) & 0xFFu;
          if (f & kObst) continue;
          C nb = (C)0, den = (C)0;
          if (f & kXm) { nb += cxl * lo[r][m]; den += cxl; }
          if (f & kXp) { nb += cxh * hi[r][m]; den += cxh; }
          if (f & kYm) { nb += cyl[r] * ym[m]; den += cyl[r]; }
          if (f & kYp) { nb += cyh[r] * yp[m]; den += cyh[r]; }
          const C left = m == 0 ? km : cur[r][m - 1];
          const C right = m == 3 ? kp : cur[r][m + 1];
          if (f & kZm) { nb += czl[m] * left; den += czl[m]; }
          if (f & kZp) { nb += czh[m] * right; den += czh[m]; }
          if (outlet_row && m == 0) den += a.outlet_add;
          if (a.diag) den += (C)a.diag[c + m];
          const C res = (C)(TRes)(a.rhs_sign * rhs[m] - (nb - den * cur[r][m]));
          const C vol = (C)(wx * wy[r] * wz[m]);
          rsum[m >> 1] += res * vol;
          vsum[m >> 1] += vol;
        }
      }
    }
    // coarse outputs ck0, ck0 + 1
    const uint32_t ck0 = k0 >> 1;
    const uint32_t cc = ci * a.csx + cj * a.csy + ck0;
    float o[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool obst = (a.cflags[cc + h] & kObst) != 0;
      o[h] = obst ? 0.0f : (float)(rsum[h] / (vsum[h] > (C)1.0e-30 ? vsum[h] : (C)1.0e-30));
    }
    *reinterpret_cast<float2*>(a.out + cc) = make_float2(o[0], o[1]);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        pv[r][m] = pb[r][m];
        pa[r][m] = pn[r][m];
      }
    load_plane(fb + 2, pb);
  }
}

}  // namespace hfsi
