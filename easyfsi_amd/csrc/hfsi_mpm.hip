// libhfsi: neo-Hookean MLS/APIC MPM particle <-> grid transfers for gfx950.
//
// One step (simulation_core/neo_hookean_mpm.py:695-959):
//   k_mpm_clear      grid mass / momentum / force := 0
//   k_mpm_p2g        SVD clamp of F, first Piola stress, affine momentum,
//                    27-node quadratic B-spline scatter (hardware fp32 atomics)
//   k_mpm_grid       v = p/m + dt f/m, transfer snapshot, damping
//   k_mpm_g2p        27-node gather, APIC C, x += dt v, F <- (I + dt C) F,
//                    surface normal / area update
//   k_mpm_radial     centroid-relative radial stretch diagnostics
// Every launch also reduces its report quantities deterministically
// (workgroup partials, combined in workgroup order by a one-workgroup finish
// launch); the final 34-value report snapshot is assembled on the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hfsi.h"
#include "hfsi_common.hpp"

namespace hfsi {

constexpr double kMinSigma = 1.0e-2;
constexpr double kMaxSigma = 1.0e2;

// ---------------------------------------------------------------------------
// 3x3 helpers (float64)
// ---------------------------------------------------------------------------

struct M3 {
  double a[3][3];
};

__device__ inline M3 mmul(const M3& A, const M3& B) {
  M3 C;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
  return C;
}

__device__ inline M3 mtrans(const M3& A) {
  M3 T;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) T.a[i][j] = A.a[j][i];
  return T;
}

__device__ inline double mdet(const M3& A) {
  return A.a[0][0] * (A.a[1][1] * A.a[2][2] - A.a[1][2] * A.a[2][1]) -
         A.a[0][1] * (A.a[1][0] * A.a[2][2] - A.a[1][2] * A.a[2][0]) +
         A.a[0][2] * (A.a[1][0] * A.a[2][1] - A.a[1][1] * A.a[2][0]);
}

// inverse transpose (cofactor matrix / det)
__device__ inline M3 minv_t(const M3& A, double det) {
  M3 R;
  const double id = 1.0 / det;
  R.a[0][0] = (A.a[1][1] * A.a[2][2] - A.a[1][2] * A.a[2][1]) * id;
  R.a[0][1] = (A.a[1][2] * A.a[2][0] - A.a[1][0] * A.a[2][2]) * id;
  R.a[0][2] = (A.a[1][0] * A.a[2][1] - A.a[1][1] * A.a[2][0]) * id;
  R.a[1][0] = (A.a[0][2] * A.a[2][1] - A.a[0][1] * A.a[2][2]) * id;
  R.a[1][1] = (A.a[0][0] * A.a[2][2] - A.a[0][2] * A.a[2][0]) * id;
  R.a[1][2] = (A.a[0][1] * A.a[2][0] - A.a[0][0] * A.a[2][1]) * id;
  R.a[2][0] = (A.a[0][1] * A.a[1][2] - A.a[0][2] * A.a[1][1]) * id;
  R.a[2][1] = (A.a[0][2] * A.a[1][0] - A.a[0][0] * A.a[1][2]) * id;
  R.a[2][2] = (A.a[0][0] * A.a[1][1] - A.a[0][1] * A.a[1][0]) * id;
  return R;
}

// SVD F = U diag(s) V^T with U, V proper rotations, |s0| >= |s1| >= |s2| and
// the sign of det F carried by s2.  Method: cyclic Jacobi eigen-decomposition
// of the symmetric F^T F (V and s^2), eigenvalues sorted descending, then a
// Givens QR factorisation F V = U R whose rotations keep det U = +1; s is the
// diagonal of R (the structure of McAdams et al., "Computing the singular
// value decomposition of 3x3 matrices with minimal branching and elementary
// floating point operations", 2011, here with exact rotations in float64).
__device__ inline void svd3(const M3& F, M3& U, double s[3], M3& V) {
  M3 A = mmul(mtrans(F), F);
  M3 W;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) W.a[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = A.a[0][1] * A.a[0][1] + A.a[0][2] * A.a[0][2] + A.a[1][2] * A.a[1][2];
    double diag = A.a[0][0] * A.a[0][0] + A.a[1][1] * A.a[1][1] + A.a[2][2] * A.a[2][2];
    if (off <= 1.0e-32 * diag || off == 0.0) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      const double apq = A.a[p][q];
      if (apq == 0.0) continue;
      const double theta = (A.a[q][q] - A.a[p][p]) / (2.0 * apq);
      const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0);
      const double sn = t * c;
      // A <- J^T A J, J rotation in the (p, q) plane
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double akp = A.a[k][p], akq = A.a[k][q];
        A.a[k][p] = c * akp - sn * akq;
        A.a[k][q] = sn * akp + c * akq;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double apk = A.a[p][k], aqk = A.a[q][k];
        A.a[p][k] = c * apk - sn * aqk;
        A.a[q][k] = sn * apk + c * aqk;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double wkp = W.a[k][p], wkq = W.a[k][q];
        W.a[k][p] = c * wkp - sn * wkq;
        W.a[k][q] = sn * wkp + c * wkq;
      }
    }
  }
  // sort eigenvalues descending (columns of W follow)
  double lam[3] = {A.a[0][0], A.a[1][1], A.a[2][2]};
  int idx[3] = {0, 1, 2};
  if (lam[idx[0]] < lam[idx[1]]) { int t = idx[0]; idx[0] = idx[1]; idx[1] = t; }
  if (lam[idx[1]] < lam[idx[2]]) { int t = idx[1]; idx[1] = idx[2]; idx[2] = t; }
  if (lam[idx[0]] < lam[idx[1]]) { int t = idx[0]; idx[0] = idx[1]; idx[1] = t; }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V.a[i][j] = W.a[i][idx[j]];
  if (mdet(V) < 0.0) {
    V.a[0][2] = -V.a[0][2];
    V.a[1][2] = -V.a[1][2];
    V.a[2][2] = -V.a[2][2];
  }
  // Givens QR of B = F V: rows (0,1), (0,2), (1,2)
  M3 B = mmul(F, V);
  M3 Q;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Q.a[i][j] = i == j ? 1.0 : 0.0;
  const int ri[3] = {0, 0, 1}, rj[3] = {1, 2, 2}, col[3] = {0, 0, 1};
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const int i = ri[g], j = rj[g], k = col[g];
    const double a = B.a[i][k], b = B.a[j][k];
    const double r = sqrt(a * a + b * b);
    double c = 1.0, sn = 0.0;
    if (r > 0.0) {
      c = a / r;
      sn = b / r;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double bi = B.a[i][m], bj = B.a[j][m];
      B.a[i][m] = c * bi + sn * bj;
      B.a[j][m] = -sn * bi + c * bj;
    }
    // Q <- Q G^T
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double qi = Q.a[m][i], qj = Q.a[m][j];
      Q.a[m][i] = c * qi + sn * qj;
      Q.a[m][j] = -sn * qi + c * qj;
    }
  }
  U = Q;
  s[0] = B.a[0][0];
  s[1] = B.a[1][1];
  s[2] = B.a[2][2];
}

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------

struct MpmArgs {
  // particles (caller-owned)
  float* x;
  float* v;
  float* C;
  float* F;
  const float* mass;
  const float* volume;
  const float* rest_x;
  const int32_t* fixed;
  const int32_t* region;
  const float* f_ext;
  float* normal;
  const float* rest_normal;
  float* area;
  const float* rest_area;
  int32_t count;
  // grid (caller-owned)
  float* gm;
  float* gv;
  float* gf;
  int32_t nx, ny, nz;
  float bmin[3];
  float dx[3];
  float dt, mu, lambda, damping;
  int32_t primary, secondary;
  // reductions
  double* partials;
  unsigned int* counter;
  double* totals;   // per-stage totals
  double* report;   // 34 packed values (device)
};

// Stage totals layout
enum {
  T_PMOM = 0,    // 3
  T_FEXT = 3,    // 3
  T_MASS = 6,
  T_VOL = 7,
  T_PSQ = 8,
  T_MAXJ = 9,    // max
  T_CLAMP = 10,
  T_OOB = 11,
  T_TRANS = 12,  // 3
  T_GMOM = 15,   // 3
  T_ACTIVE = 18,
  T_MAXV = 19,   // max
  T_PDISP = 20,  // 3
  T_PVEL = 23,   // 3
  T_SDISP = 26,  // 3
  T_SVEL = 29,   // 3
  T_PCNT = 32,
  T_SCNT = 33,
  T_CCUR = 34,   // 3
  T_CREST = 37,  // 3
  T_CCNT = 40,
  T_RSUM = 41,
  T_RCNT = 42,
  T_RMAX = 43,   // max
  T_COUNT = 44
};

// max-reduced slots of each stage (relative to the stage's first total)
constexpr uint64_t kP2gMax = 1ull << T_MAXJ;
constexpr uint64_t kGridMax = 1ull << 7;
constexpr uint64_t kRadialMax = 1ull << 2;

__device__ inline void grid_coord(const MpmArgs& a, const float* xp, float X[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) X[d] = (xp[d] - a.bmin[d]) / a.dx[d];
}

// Written as the negation of "inside" so that a NaN coordinate counts as out
// of bounds and never reaches the node-index computation.
__device__ inline bool stencil_oob(const MpmArgs& a, const float X[3]) {
  const bool inside = X[0] >= 0.5f && X[0] < (float)a.nx - 1.5f && X[1] >= 0.5f && X[1] < (float)a.ny - 1.5f &&
                      X[2] >= 0.5f && X[2] < (float)a.nz - 1.5f;
  return !inside;
}

__device__ inline void bspline(float f, float w[3]) {
  w[0] = 0.5f * (1.5f - f) * (1.5f - f);
  w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
  w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}


// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_mpm_clear(uint32_t n, float* gm, float* gv, float* gf) {
  HFSI_CELL_LOOP(n) {
    gm[c] = 0.0f;
    gv[3 * (size_t)c + 0] = 0.0f;
    gv[3 * (size_t)c + 1] = 0.0f;
    gv[3 * (size_t)c + 2] = 0.0f;
    gf[3 * (size_t)c + 0] = 0.0f;
    gf[3 * (size_t)c + 1] = 0.0f;
    gf[3 * (size_t)c + 2] = 0.0f;
  }
}

// P2G (neo_hookean_mpm.py:712-801).  Totals 0..11.
//
// Per particle: the SVD clamp of F (persisted when it changed), the first
// Piola stress of the clamped F and the affine momentum, plus the report
// sums; the 27-node scatter is done by the caller (global atomics, or LDS
// atomics into a block tile in the binned path).
struct P2gPart {
  bool inb;
  int base[3];
  float fx[3], w[3][3];
  float m, mom[3], aff[3][3], fe[3];
};

__device__ inline void p2g_prepare(const MpmArgs& a, uint32_t p, double (&acc)[12], P2gPart& q) {
  float X[3];
  grid_coord(a, a.x + 3 * (size_t)p, X);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    q.base[d] = (int)floorf(X[d] - 0.5f);
    q.fx[d] = X[d] - (float)q.base[d];
    bspline(q.fx[d], q.w[d]);
  }
  M3 F;
  const float* Fp_in = a.F + 9 * (size_t)p;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F.a[i][j] = (double)Fp_in[3 * i + j];
  const double raw_J = (double)(float)mdet(F);
  M3 U, V;
  double sig[3];
  svd3(F, U, sig, V);
  bool clamped = false;
  double sc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    sc[d] = fmin(fmax(fabs(sig[d]), kMinSigma), kMaxSigma);
    if (fabs(sc[d] - sig[d]) > 1.0e-6) clamped = true;
  }
  if (mdet(mmul(U, mtrans(V))) < 0.0) {
    clamped = true;
    U.a[0][2] = -U.a[0][2];
    U.a[1][2] = -U.a[1][2];
    U.a[2][2] = -U.a[2][2];
  }
  M3 Fc;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Fc.a[i][j] = U.a[i][0] * sc[0] * V.a[j][0] + U.a[i][1] * sc[1] * V.a[j][1] + U.a[i][2] * sc[2] * V.a[j][2];
  if (mdet(Fc) <= 0.0) {
    clamped = true;
    Fc.a[0][2] = -Fc.a[0][2];
    Fc.a[1][2] = -Fc.a[1][2];
    Fc.a[2][2] = -Fc.a[2][2];
  }
  if (raw_J <= 0.0) clamped = true;
  // float32 state: the stress uses the stored (rounded) corrected F
  M3 Fs;
  float* Fp_out = a.F + 9 * (size_t)p;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float fv = (float)Fc.a[i][j];
      Fs.a[i][j] = (double)fv;
      if (clamped) Fp_out[3 * i + j] = fv;
    }
  const double J = mdet(Fs);
  const M3 FinvT = minv_t(Fs, J);
  const double lnJ = log(J);
  M3 Pk;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Pk.a[i][j] = (double)a.mu * (Fs.a[i][j] - FinvT.a[i][j]) + (double)a.lambda * lnJ * FinvT.a[i][j];
  const M3 PFt = mmul(Pk, mtrans(Fs));
  const float m_p = a.mass[p];
  const float vol_p = a.volume[p];
  const bool fixed = a.fixed && a.fixed[p] != 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double inv = 4.0 / ((double)a.dx[j] * (double)a.dx[j]);
      const double stress = -(double)a.dt * (double)vol_p * PFt.a[i][j] * inv;
      q.aff[i][j] = fixed ? 0.0f : (float)(stress + (double)m_p * (double)a.C[9 * (size_t)p + 3 * i + j]);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    q.mom[d] = fixed ? 0.0f : m_p * a.v[3 * (size_t)p + d];
    q.fe[d] = (fixed || !a.f_ext) ? 0.0f : a.f_ext[3 * (size_t)p + d];
  }
  q.m = m_p;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    acc[T_PMOM + d] += q.mom[d];
    acc[T_FEXT + d] += q.fe[d];
  }
  acc[T_MASS] += m_p;
  acc[T_VOL] += vol_p;
  acc[T_PSQ] += (double)q.mom[0] * q.mom[0] + (double)q.mom[1] * q.mom[1] + (double)q.mom[2] * q.mom[2];
  acc[T_MAXJ] = fmax(acc[T_MAXJ], fabs(J));
  acc[T_CLAMP] += clamped ? 1.0 : 0.0;
  q.inb = !stencil_oob(a, X);
  if (!q.inb) acc[T_OOB] += 1.0;
}

// The 27 node contributions of a particle: node(ox, oy, oz) gets mass wt m,
// momentum wt (m v + aff dpos) and force wt f_ext.
template <typename Sink>
__device__ inline void p2g_scatter(const MpmArgs& a, const P2gPart& q, Sink sink) {
#pragma unroll
  for (int ox = 0; ox < 3; ++ox)
#pragma unroll
    for (int oy = 0; oy < 3; ++oy)
#pragma unroll
      for (int oz = 0; oz < 3; ++oz) {
        const float wt = q.w[0][ox] * q.w[1][oy] * q.w[2][oz];
        const float dp0 = ((float)ox - q.fx[0]) * a.dx[0];
        const float dp1 = ((float)oy - q.fx[1]) * a.dx[1];
        const float dp2 = ((float)oz - q.fx[2]) * a.dx[2];
        float md[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) md[d] = wt * (q.mom[d] + (q.aff[d][0] * dp0 + q.aff[d][1] * dp1 + q.aff[d][2] * dp2));
        sink(ox, oy, oz, wt * q.m, md, wt);
      }
}

__global__ void __launch_bounds__(kBlock) k_mpm_p2g(MpmArgs a) {
  constexpr int NS = 12;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const uint32_t n = (uint32_t)a.count;
  const uint32_t sx = (uint32_t)a.ny * a.nz, sy = a.nz;
  HFSI_CELL_LOOP(n) {
    P2gPart q;
    p2g_prepare(a, c, acc, q);
    if (!q.inb) continue;
    p2g_scatter(a, q, [&](int ox, int oy, int oz, float mw, const float (&md)[3], float wt) {
      const uint32_t node = (uint32_t)(q.base[0] + ox) * sx + (uint32_t)(q.base[1] + oy) * sy + (uint32_t)(q.base[2] + oz);
      atomicAdd(&a.gm[node], mw);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        atomicAdd(&a.gv[3 * (size_t)node + d], md[d]);
        if (q.fe[d] != 0.0f) atomicAdd(&a.gf[3 * (size_t)node + d], wt * q.fe[d]);
      }
    });
  }
  block_partials<NS>(acc, a.partials, lds, kP2gMax);
}

// ---------------------------------------------------------------------------
// Cell-binned P2G: particles are bucketed by the 8^3-node block of their
// stencil base node; one workgroup accumulates a block's particles into a
// 10^3-node LDS tile with LDS atomics and adds the tile to the grid with one
// global atomic per touched node value (instead of 27 x 7 per particle).
// ---------------------------------------------------------------------------

constexpr int kBinEdge = 8;                 // nodes per block edge
constexpr int kTileEdge = kBinEdge + 2;     // stencil reach of a block
constexpr int kTileNodes = kTileEdge * kTileEdge * kTileEdge;

// Components of a block's P2G tile: mass, momentum (3), external force (3).
constexpr int kTileComps = 7;
constexpr size_t kTileStride = (size_t)kTileComps * kTileNodes;  // floats per block
// Blocks up to this many get a tile buffer (28 KB each); larger grids flush tiles with global atomics.
constexpr size_t kMaxTileBlocks = 65536;

struct BinArgs {
  int32_t nbx, nby, nbz;   // block grid
  int32_t* key;            // per particle: block, or -1 when out of bounds
  int32_t* rank;           // per particle: slot within its block
  int32_t* count;          // per block
  int32_t* offset;         // per block: exclusive prefix sum of count
  int32_t* sorted;         // particles ordered by block
  int32_t* active;         // blocks with particles (unordered)
  int32_t* nactive;
  int32_t* oob;            // out-of-bounds particles (unordered)
  int32_t* noob;
  int32_t* scan_sums;      // per 1024-block chunk
  // tile path: every active block stores its whole tile (plain stores) and the grid update gathers each
  // node from the (up to 8) tiles that cover it, in a fixed order
  float* tiles;            // [block][kTileComps][kTileNodes], or nullptr (atomic flush)
  int32_t* need_stamp;     // per block: stamp of the step that last listed it
  int32_t* need_list;      // blocks with nodes a tile covers: the active blocks and their +1 neighbours
  int32_t* nneed;
  int32_t stamp;           // this step's stamp (never 0: need_stamp starts at 0)
  int32_t has_force;       // particles carry an external force (force components stored)
};

__global__ void k_bin_clear(BinArgs b) {
  const int32_t nb = b.nbx * b.nby * b.nbz;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) b.count[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *b.nactive = 0;
    *b.noob = 0;
    if (b.nneed) *b.nneed = 0;
  }
}

// List every block a tile of an active block reaches (the block and its seven +1 neighbours) once: the
// stamp exchange admits the first marker of a block in this step.
__global__ void __launch_bounds__(kBlock) k_need_mark(BinArgs b) {
  const int32_t nact = *b.nactive;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nact; i += gridDim.x * blockDim.x) {
    const int32_t k = b.active[i];
    const int32_t bx = k / (b.nby * b.nbz), by = (k / b.nbz) % b.nby, bz = k % b.nbz;
    for (int e = 0; e < 8; ++e) {
      const int32_t x = bx + (e >> 2), y = by + ((e >> 1) & 1), z = bz + (e & 1);
      if (x >= b.nbx || y >= b.nby || z >= b.nbz) continue;
      const int32_t k2 = (x * b.nby + y) * b.nbz + z;
      if (atomicExch(&b.need_stamp[k2], b.stamp) != b.stamp) b.need_list[atomicAdd(b.nneed, 1)] = k2;
    }
  }
}

// Key, rank and list pushes of every particle.  The two list pushes (an
// out-of-bounds particle to `oob`, the first particle of a block to `active`)
// are one atomic on per-lane selected pointers: written as two branches, the
// ROCm 7.2 compiler tail-merged them into one block that used the active-list
// pointers for both (out-of-bounds particles then went to the active list).
__global__ void __launch_bounds__(kBlock) k_bin_count(MpmArgs a, BinArgs b) {
  const uint32_t n = (uint32_t)a.count;
  HFSI_CELL_LOOP(n) {
    float X[3];
    grid_coord(a, a.x + 3 * (size_t)c, X);
    const bool out = stencil_oob(a, X);
    int32_t k = -1, r = 0;
    if (!out) {
      const int bx = (int)floorf(X[0] - 0.5f) / kBinEdge, by = (int)floorf(X[1] - 0.5f) / kBinEdge,
                bz = (int)floorf(X[2] - 0.5f) / kBinEdge;
      k = (bx * b.nby + by) * b.nbz + bz;
      r = atomicAdd(&b.count[k], 1);
      b.rank[c] = r;
    }
    b.key[c] = k;
    int32_t* list = out ? b.oob : b.active;
    int32_t* counter = out ? b.noob : b.nactive;
    if (out || r == 0) list[atomicAdd(counter, 1)] = out ? (int32_t)c : k;
  }
}

// Exclusive prefix sum of count over nb blocks: per-chunk (1024) scan, scan
// of the chunk sums (<= 1024 chunks, one workgroup), then add.
constexpr int kScanChunk = 1024;

__global__ void __launch_bounds__(kScanChunk) k_scan_chunks(const int32_t* in, int32_t* out, int32_t n,
                                                            int32_t* sums) {
  __shared__ int32_t t[kScanChunk];
  const int32_t i = blockIdx.x * kScanChunk + threadIdx.x;
  const int32_t v = i < n ? in[i] : 0;
  t[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanChunk; off <<= 1) {
    const int32_t add = threadIdx.x >= (unsigned)off ? t[threadIdx.x - off] : 0;
    __syncthreads();
    t[threadIdx.x] += add;
    __syncthreads();
  }
  if (i < n) out[i] = t[threadIdx.x] - v;  // exclusive
  if (threadIdx.x == kScanChunk - 1) sums[blockIdx.x] = t[threadIdx.x];
}

__global__ void __launch_bounds__(kScanChunk) k_scan_add(int32_t* out, int32_t n, const int32_t* sums,
                                                         int32_t nchunks) {
  __shared__ int32_t t[kScanChunk];
  // every workgroup scans the (<= 1024) chunk sums itself
  const int32_t v = (int32_t)threadIdx.x < nchunks ? sums[threadIdx.x] : 0;
  t[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanChunk; off <<= 1) {
    const int32_t add = threadIdx.x >= (unsigned)off ? t[threadIdx.x - off] : 0;
    __syncthreads();
    t[threadIdx.x] += add;
    __syncthreads();
  }
  const int32_t base = blockIdx.x > 0 ? t[blockIdx.x - 1] : 0;
  const int32_t i = blockIdx.x * kScanChunk + threadIdx.x;
  if (i < n) out[i] += base;
}

__global__ void __launch_bounds__(kBlock) k_bin_scatter(int32_t n, BinArgs b) {
  HFSI_CELL_LOOP((uint32_t)n) {
    const int32_t k = b.key[c];
    if (k >= 0) b.sorted[b.offset[k] + b.rank[c]] = (int32_t)c;
  }
}

__global__ void __launch_bounds__(kBlock) k_mpm_p2g_binned(MpmArgs a, BinArgs b) {
  constexpr int NS = 12;
  __shared__ double lds[NS * (kBlock / 64)];
  __shared__ float tm[kTileNodes];
  __shared__ float tp[3][kTileNodes];
  __shared__ float tf[3][kTileNodes];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const int32_t nact = *b.nactive;
  const uint32_t sx = (uint32_t)a.ny * a.nz, sy = a.nz;
  for (int32_t ab = blockIdx.x; ab < nact; ab += gridDim.x) {
    const int32_t k = b.active[ab];
    const int ox0 = (k / (b.nby * b.nbz)) * kBinEdge, oy0 = ((k / b.nbz) % b.nby) * kBinEdge,
              oz0 = (k % b.nbz) * kBinEdge;
    for (int i = threadIdx.x; i < kTileNodes; i += blockDim.x) {
      tm[i] = 0.0f;
      tp[0][i] = tp[1][i] = tp[2][i] = 0.0f;
      tf[0][i] = tf[1][i] = tf[2][i] = 0.0f;
    }
    __syncthreads();
    const int32_t beg = b.offset[k], cnt = b.count[k];
    for (int32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      P2gPart q;
      p2g_prepare(a, (uint32_t)b.sorted[beg + t], acc, q);
      // q.inb holds: only in-bounds particles are binned
      const int lx = q.base[0] - ox0, ly = q.base[1] - oy0, lz = q.base[2] - oz0;
      p2g_scatter(a, q, [&](int ox, int oy, int oz, float mw, const float (&md)[3], float wt) {
        const int li = ((lx + ox) * kTileEdge + (ly + oy)) * kTileEdge + (lz + oz);
        atomicAdd(&tm[li], mw);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          atomicAdd(&tp[d][li], md[d]);
          if (q.fe[d] != 0.0f) atomicAdd(&tf[d][li], wt * q.fe[d]);
        }
      });
    }
    __syncthreads();
    if (b.tiles) {
      // the whole tile, plain stores (the grid update gathers it)
      float* T = b.tiles + (size_t)k * kTileStride;
      for (int i = threadIdx.x; i < kTileNodes; i += blockDim.x) {
        T[i] = tm[i];
#pragma unroll
        for (int d = 0; d < 3; ++d) T[(1 + d) * kTileNodes + i] = tp[d][i];
        if (b.has_force) {
#pragma unroll
          for (int d = 0; d < 3; ++d) T[(4 + d) * kTileNodes + i] = tf[d][i];
        }
      }
    } else {
      for (int i = threadIdx.x; i < kTileNodes; i += blockDim.x) {
        const int tz = i % kTileEdge, ty = (i / kTileEdge) % kTileEdge, tx = i / (kTileEdge * kTileEdge);
        const int gx = ox0 + tx, gy = oy0 + ty, gz = oz0 + tz;
        if (gx >= a.nx || gy >= a.ny || gz >= a.nz) continue;
        const uint32_t node = (uint32_t)gx * sx + (uint32_t)gy * sy + (uint32_t)gz;
        if (tm[i] != 0.0f) atomicAdd(&a.gm[node], tm[i]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (tp[d][i] != 0.0f) atomicAdd(&a.gv[3 * (size_t)node + d], tp[d][i]);
          if (tf[d][i] != 0.0f) atomicAdd(&a.gf[3 * (size_t)node + d], tf[d][i]);
        }
      }
    }
    __syncthreads();  // the tile is reused by the next block
  }
  block_partials<NS>(acc, a.partials, lds, kP2gMax);
}

// Out-of-bounds particles of the binned path: the clamp and the report sums only.
__global__ void __launch_bounds__(kBlock) k_mpm_p2g_oob(MpmArgs a, BinArgs b, double* partials) {
  constexpr int NS = 12;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const int32_t n = *b.noob;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    P2gPart q;
    p2g_prepare(a, (uint32_t)b.oob[i], acc, q);
  }
  block_partials<NS>(acc, partials, lds, kP2gMax);
}

// Grid update (neo_hookean_mpm.py:803-817).  Totals 12..19.
__global__ void __launch_bounds__(kBlock) k_mpm_grid(MpmArgs a) {
  constexpr int NS = 8;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const uint32_t n = (uint32_t)a.nx * a.ny * a.nz;
  HFSI_CELL_LOOP(n) {
    const float m = a.gm[c];
    if (!(m > 1.0e-20f)) continue;
    float v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = a.gv[3 * (size_t)c + d] / m;
      v[d] += a.dt * a.gf[3 * (size_t)c + d] / m;
      acc[d] += (double)(m * v[d]);
      v[d] *= a.damping;
      a.gv[3 * (size_t)c + d] = v[d];
      acc[3 + d] += (double)(m * v[d]);
    }
    acc[6] += 1.0;
    acc[7] = fmax(acc[7], (double)sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  }
  block_partials<NS>(acc, a.partials, lds, kGridMax);
}

// Grid update of the tile path: every node of a listed block sums the tiles
// that cover it (its own block's and, within 2 nodes of a low face, the lower
// neighbours'), in a fixed order, then updates as k_mpm_grid and writes mass,
// velocity and force.  Nodes of listed blocks without mass get the raw sums
// (zero), as the cleared grid has.  Totals 12..19.
__global__ void __launch_bounds__(kBlock) k_mpm_grid_gather(MpmArgs a, BinArgs b) {
  constexpr int NS = 8;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const int32_t nneed = *b.nneed;
  const uint32_t sx = (uint32_t)a.ny * a.nz, sy = a.nz;
  constexpr int kBlockNodes = kBinEdge * kBinEdge * kBinEdge;
  for (int32_t w = blockIdx.x; w < nneed; w += gridDim.x) {
    const int32_t B = b.need_list[w];
    const int32_t bx = B / (b.nby * b.nbz), by = (B / b.nbz) % b.nby, bz = B % b.nbz;
    for (int t = threadIdx.x; t < kBlockNodes; t += blockDim.x) {
      const int lx = t / (kBinEdge * kBinEdge), ly = (t / kBinEdge) % kBinEdge, lz = t % kBinEdge;
      const int gx = bx * kBinEdge + lx, gy = by * kBinEdge + ly, gz = bz * kBinEdge + lz;
      if (gx >= a.nx || gy >= a.ny || gz >= a.nz) continue;
      float m = 0.0f, p[3] = {0.0f, 0.0f, 0.0f}, f[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int ex = e >> 2, ey = (e >> 1) & 1, ez = e & 1;
        if ((ex && lx >= 2) || (ey && ly >= 2) || (ez && lz >= 2)) continue;
        if (bx - ex < 0 || by - ey < 0 || bz - ez < 0) continue;
        const int32_t k = ((bx - ex) * b.nby + (by - ey)) * b.nbz + (bz - ez);
        if (b.count[k] <= 0) continue;
        const float* T = b.tiles + (size_t)k * kTileStride;
        const int ti = ((lx + ex * kBinEdge) * kTileEdge + (ly + ey * kBinEdge)) * kTileEdge + (lz + ez * kBinEdge);
        m += T[ti];
#pragma unroll
        for (int d = 0; d < 3; ++d) p[d] += T[(1 + d) * kTileNodes + ti];
        if (b.has_force) {
#pragma unroll
          for (int d = 0; d < 3; ++d) f[d] += T[(4 + d) * kTileNodes + ti];
        }
      }
      const uint32_t node = (uint32_t)gx * sx + (uint32_t)gy * sy + (uint32_t)gz;
      a.gm[node] = m;
      float v[3] = {p[0], p[1], p[2]};
      if (m > 1.0e-20f) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          v[d] = p[d] / m;
          v[d] += a.dt * f[d] / m;
          acc[d] += (double)(m * v[d]);
          v[d] *= a.damping;
          acc[3 + d] += (double)(m * v[d]);
        }
        acc[6] += 1.0;
        acc[7] = fmax(acc[7], (double)sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a.gv[3 * (size_t)node + d] = v[d];
        a.gf[3 * (size_t)node + d] = f[d];
      }
    }
  }
  block_partials<NS>(acc, a.partials, lds, kGridMax);
}

__device__ inline void update_surface(const MpmArgs& a, uint32_t p) {
  if (!a.rest_area) return;
  const float ra = a.rest_area[p];
  const float* rn = a.rest_normal + 3 * (size_t)p;
  if (!(ra > 0.0f) || !(sqrtf(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]) > 1.0e-12f)) return;
  M3 F;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F.a[i][j] = (double)a.F[9 * (size_t)p + 3 * i + j];
  const double J = mdet(F);
  if (!(J > 1.0e-12)) return;
  const M3 FiT = minv_t(F, J);
  double av[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) av[i] = J * (FiT.a[i][0] * rn[0] + FiT.a[i][1] * rn[1] + FiT.a[i][2] * rn[2]);
  const double s = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
  if (!(s > 1.0e-12)) return;
#pragma unroll
  for (int i = 0; i < 3; ++i) a.normal[3 * (size_t)p + i] = (float)(av[i] / s);
  a.area[p] = (float)((double)ra * s);
}

// G2P (neo_hookean_mpm.py:819-874).  Totals 20..40.
__global__ void __launch_bounds__(kBlock) k_mpm_g2p(MpmArgs a) {
  constexpr int NS = 21;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = 0.0;
  const uint32_t n = (uint32_t)a.count;
  const uint32_t sx = (uint32_t)a.ny * a.nz, sy = a.nz;
  HFSI_CELL_LOOP(n) {
    const uint32_t p = c;
    float* xp = a.x + 3 * (size_t)p;
    float X[3];
    grid_coord(a, xp, X);
    const bool fixed = a.fixed && a.fixed[p] != 0;
    if (!stencil_oob(a, X)) {
      int base[3];
      float fx[3], w[3][3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        base[d] = (int)floorf(X[d] - 0.5f);
        fx[d] = X[d] - (float)base[d];
        bspline(fx[d], w[d]);
      }
      float nv[3] = {0.0f, 0.0f, 0.0f};
      float nC[3][3] = {};
      const float s0 = 4.0f / (a.dx[0] * a.dx[0]), s1 = 4.0f / (a.dx[1] * a.dx[1]), s2 = 4.0f / (a.dx[2] * a.dx[2]);
#pragma unroll
      for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
          for (int oz = 0; oz < 3; ++oz) {
            const float wt = w[0][ox] * w[1][oy] * w[2][oz];
            const float g0 = s0 * ((float)ox - fx[0]) * a.dx[0];
            const float g1 = s1 * ((float)oy - fx[1]) * a.dx[1];
            const float g2 = s2 * ((float)oz - fx[2]) * a.dx[2];
            const uint32_t node =
                (uint32_t)(base[0] + ox) * sx + (uint32_t)(base[1] + oy) * sy + (uint32_t)(base[2] + oz);
            const float* gvn = a.gv + 3 * (size_t)node;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const float gvd = gvn[d];
              nv[d] += wt * gvd;
              nC[d][0] += wt * gvd * g0;
              nC[d][1] += wt * gvd * g1;
              nC[d][2] += wt * gvd * g2;
            }
          }
      if (!fixed) {
        float* vp = a.v + 3 * (size_t)p;
        float* Cp = a.C + 9 * (size_t)p;
        float* Fp = a.F + 9 * (size_t)p;
        float Fo[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Fo[i] = Fp[i];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          vp[d] = nv[d];
          xp[d] = xp[d] + a.dt * nv[d];
#pragma unroll
          for (int e = 0; e < 3; ++e) Cp[3 * d + e] = nC[d][e];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            float acc_f = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc_f += ((i == k ? 1.0f : 0.0f) + a.dt * nC[i][k]) * Fo[3 * k + j];
            Fp[3 * i + j] = acc_f;
          }
      }
    }
    if (fixed) {
      float* vp = a.v + 3 * (size_t)p;
      float* Cp = a.C + 9 * (size_t)p;
      float* Fp = a.F + 9 * (size_t)p;
#pragma unroll
      for (int d = 0; d < 3; ++d) vp[d] = 0.0f;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        Cp[i] = 0.0f;
        Fp[i] = (i % 4 == 0) ? 1.0f : 0.0f;
      }
    }
    update_surface(a, p);
    float Xn[3];
    grid_coord(a, xp, Xn);
    if (!stencil_oob(a, Xn)) {
      const int r = a.region ? a.region[p] : 0;
      const float* rx = a.rest_x + 3 * (size_t)p;
      const float* vp = a.v + 3 * (size_t)p;
      int slot = -1;
      if (r == a.primary)
        slot = 0;
      else if (r == a.secondary)
        slot = 1;
      if (slot >= 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          acc[slot * 6 + d] += (double)(xp[d] - rx[d]);
          acc[slot * 6 + 3 + d] += (double)vp[d];
        }
        acc[12 + slot] += 1.0;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        acc[14 + d] += (double)xp[d];
        acc[17 + d] += (double)rx[d];
      }
      acc[20] += 1.0;
    }
  }
  block_partials<NS>(acc, a.partials, lds, 0ull);
}

// Radial stretch and the final report snapshot (neo_hookean_mpm.py:876-959).
__global__ void __launch_bounds__(kBlock) k_mpm_radial(MpmArgs a) {
  constexpr int NS = 3;
  __shared__ double lds[NS * (kBlock / 64)];
  double acc[NS] = {0.0, 0.0, 0.0};
  const double cnt = fmax(a.totals[T_CCNT], 1.0);
  const double cc[3] = {a.totals[T_CCUR] / cnt, a.totals[T_CCUR + 1] / cnt, a.totals[T_CCUR + 2] / cnt};
  const double rc[3] = {a.totals[T_CREST] / cnt, a.totals[T_CREST + 1] / cnt, a.totals[T_CREST + 2] / cnt};
  const uint32_t n = (uint32_t)a.count;
  HFSI_CELL_LOOP(n) {
    const float* xp = a.x + 3 * (size_t)c;
    float X[3];
    grid_coord(a, xp, X);
    if (stencil_oob(a, X)) continue;
    const float* rx = a.rest_x + 3 * (size_t)c;
    double rr = 0.0, cr = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rr += ((double)rx[d] - rc[d]) * ((double)rx[d] - rc[d]);
      cr += ((double)xp[d] - cc[d]) * ((double)xp[d] - cc[d]);
    }
    rr = sqrt(rr);
    if (rr > 1.0e-12) {
      const double st = sqrt(cr) / rr;
      acc[0] += st;
      acc[1] += 1.0;
      acc[2] = fmax(acc[2], fabs(st - 1.0));
    }
  }
  block_partials<NS>(acc, a.partials, lds, kRadialMax);

}


struct FinNone {
  __device__ void operator()(const double*) const {}
};

// Report snapshot (neo_hookean_mpm.py:899-959) from the stage totals, run by
// the finishing workgroup of the radial-stretch reduction.
struct FinMpmReport {
  const double* totals;
  double* report;
  float dt;
  __device__ void operator()(const double*) const {
    const double* t = totals;
    double* R = report;
    double e[3], diff2 = 0.0, en = 0.0;
    for (int d = 0; d < 3; ++d) {
      e[d] = t[T_PMOM + d] + (double)dt * t[T_FEXT + d];
      en += e[d] * e[d];
      diff2 += (t[T_TRANS + d] - e[d]) * (t[T_TRANS + d] - e[d]);
    }
    const double den = fmax(fmax(sqrt(en), sqrt(t[T_PSQ])), 1.0e-20);
    R[0] = t[T_MASS];
    R[1] = t[T_VOL];
    for (int d = 0; d < 3; ++d) {
      R[2 + d] = t[T_PMOM + d];
      R[5 + d] = t[T_GMOM + d];
      R[8 + d] = t[T_FEXT + d];
      R[16 + d] = t[T_PDISP + d];
      R[19 + d] = t[T_PVEL + d];
      R[22 + d] = t[T_SDISP + d];
      R[25 + d] = t[T_SVEL + d];
    }
    R[11] = sqrt(diff2) / den;
    R[12] = t[T_MAXV];
    R[13] = t[T_MAXJ];
    R[14] = t[T_RSUM];
    R[15] = t[T_RMAX];
    R[28] = t[T_ACTIVE];
    R[29] = t[T_OOB];
    R[30] = t[T_CLAMP];
    R[31] = t[T_RCNT];
    R[32] = t[T_PCNT];
    R[33] = t[T_SCNT];
  }
};

}  // namespace hfsi

struct hfsi_mpm {
  int32_t capacity = 0;
  int32_t nx = 0, ny = 0, nz = 0;
  float bmin[3] = {0, 0, 0};
  float dx[3] = {0, 0, 0};
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* totals = nullptr;
  double* report = nullptr;
  double* host_report = nullptr;
  int32_t p2g_mode = HFSI_MPM_P2G_BINNED;
  hfsi::BinArgs bins{};
  std::vector<void*> allocs;
};

using namespace hfsi;

static uint32_t mpm_blocks(uint64_t n) { return grid_blocks(n, 2048); }

extern "C" {

int hfsi_mpm_create(hfsi_mpm** out, int32_t capacity, const double* bounds_min3, const double* bounds_max3,
                    const int32_t* grid_nodes3, void* stream) {
  (void)stream;
  if (!out || capacity <= 0 || !bounds_min3 || !bounds_max3 || !grid_nodes3) {
    set_error("hfsi_mpm_create: invalid arguments");
    return HFSI_ERR_ARG;
  }
  for (int d = 0; d < 3; ++d)
    if (grid_nodes3[d] < 4 || !(bounds_max3[d] > bounds_min3[d])) {
      set_error("hfsi_mpm_create: grid must have >= 4 nodes and positive spacing per axis");
      return HFSI_ERR_ARG;
    }
  if ((uint64_t)grid_nodes3[0] * grid_nodes3[1] * grid_nodes3[2] >= (1ull << 31)) {
    set_error("hfsi_mpm_create: grid too large for 32-bit node indices");
    return HFSI_ERR_ARG;
  }
  hfsi_mpm* h = new hfsi_mpm();
  h->capacity = capacity;
  h->nx = grid_nodes3[0];
  h->ny = grid_nodes3[1];
  h->nz = grid_nodes3[2];
  for (int d = 0; d < 3; ++d) {
    h->bmin[d] = (float)bounds_min3[d];
    h->dx[d] = (float)((bounds_max3[d] - bounds_min3[d]) / grid_nodes3[d]);
  }
  {
    BinArgs& b = h->bins;
    b.nbx = (h->nx + kBinEdge - 1) / kBinEdge;
    b.nby = (h->ny + kBinEdge - 1) / kBinEdge;
    b.nbz = (h->nz + kBinEdge - 1) / kBinEdge;
    const size_t nb = (size_t)b.nbx * b.nby * b.nbz;
    const size_t nchunks = (nb + kScanChunk - 1) / kScanChunk;
    if (nchunks > (size_t)kScanChunk) {
      set_error("hfsi_mpm_create: grid too large for the particle binning");
      hfsi_mpm_destroy(h);
      return HFSI_ERR_ARG;
    }
    auto al = [&](int32_t** p, size_t n) {
      void* q = nullptr;
      if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(int32_t)) != hipSuccess) return false;
      h->allocs.push_back(q);
      *p = (int32_t*)q;
      return true;
    };
    const size_t P = (size_t)capacity;
    if (!(al(&b.key, P) && al(&b.rank, P) && al(&b.count, nb) && al(&b.offset, nb) && al(&b.sorted, P) &&
          al(&b.active, nb) && al(&b.nactive, 1) && al(&b.oob, P) && al(&b.noob, 1) && al(&b.scan_sums, nchunks))) {
      set_error("hfsi_mpm_create: device allocation failed");
      hfsi_mpm_destroy(h);
      return HFSI_ERR_HIP;
    }
    if (nb <= kMaxTileBlocks) {
      void* q = nullptr;
      if (!(al(&b.need_stamp, nb) && al(&b.need_list, nb) && al(&b.nneed, 1) &&
            hipMalloc(&q, nb * kTileStride * sizeof(float)) == hipSuccess &&
            hipMemset(b.need_stamp, 0, nb * sizeof(int32_t)) == hipSuccess)) {
        set_error("hfsi_mpm_create: device allocation failed");
        if (q) (void)hipFree(q);
        hfsi_mpm_destroy(h);
        return HFSI_ERR_HIP;
      }
      h->allocs.push_back(q);
      b.tiles = (float*)q;
    }
  }
  bool ok = hipMalloc((void**)&h->partials, sizeof(double) * 2048 * 32) == hipSuccess &&
            hipMalloc((void**)&h->counter, sizeof(unsigned int)) == hipSuccess &&
            hipMalloc((void**)&h->totals, sizeof(double) * T_COUNT) == hipSuccess &&
            hipMalloc((void**)&h->report, sizeof(double) * 34) == hipSuccess &&
            hipHostMalloc((void**)&h->host_report, sizeof(double) * 34, hipHostMallocDefault) == hipSuccess &&
            hipMemset(h->counter, 0, sizeof(unsigned int)) == hipSuccess &&
            hipMemset(h->totals, 0, sizeof(double) * T_COUNT) == hipSuccess &&
            hipMemset(h->report, 0, sizeof(double) * 34) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    set_error("hfsi_mpm_create: device allocation failed");
    hfsi_mpm_destroy(h);
    return HFSI_ERR_HIP;
  }
  *out = h;
  return HFSI_OK;
}

int hfsi_mpm_set_p2g_mode(hfsi_mpm* h, int32_t mode) {
  if (!h || (mode != HFSI_MPM_P2G_BINNED && mode != HFSI_MPM_P2G_ATOMIC)) {
    set_error("hfsi_mpm_set_p2g_mode: invalid arguments");
    return HFSI_ERR_ARG;
  }
  h->p2g_mode = mode;
  return HFSI_OK;
}

int hfsi_mpm_destroy(hfsi_mpm* h) {
  if (!h) return HFSI_OK;
  (void)hipDeviceSynchronize();
  if (h->partials) (void)hipFree(h->partials);
  if (h->counter) (void)hipFree(h->counter);
  if (h->totals) (void)hipFree(h->totals);
  if (h->report) (void)hipFree(h->report);
  if (h->host_report) (void)hipHostFree(h->host_report);
  for (void* p : h->allocs) (void)hipFree(p);
  delete h;
  return HFSI_OK;
}

int hfsi_mpm_step(hfsi_mpm* h, const hfsi_mpm_particles* P, int32_t count, const hfsi_mpm_grid* G, float dt_s,
                  float mu_pa, float lambda_pa, float velocity_damping, int32_t primary_region_id,
                  int32_t secondary_region_id, double* report34, void* stream) {
  if (!h || !P || !G || count <= 0 || count > h->capacity || !P->x || !P->v || !P->C || !P->F || !P->mass ||
      !P->volume || !P->rest_x || !G->mass || !G->velocity || !G->force) {
    set_error("hfsi_mpm_step: invalid arguments");
    return HFSI_ERR_ARG;
  }
  if ((P->rest_area != nullptr) != (P->rest_normal != nullptr) || (P->rest_area && (!P->area || !P->normal))) {
    set_error("hfsi_mpm_step: surface arrays must be given together");
    return HFSI_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  MpmArgs a;
  a.x = P->x;
  a.v = P->v;
  a.C = P->C;
  a.F = P->F;
  a.mass = P->mass;
  a.volume = P->volume;
  a.rest_x = P->rest_x;
  a.fixed = P->fixed;
  a.region = P->region;
  a.f_ext = P->external_force;
  a.normal = P->normal;
  a.rest_normal = P->rest_normal;
  a.area = P->area;
  a.rest_area = P->rest_area;
  a.count = count;
  a.gm = G->mass;
  a.gv = G->velocity;
  a.gf = G->force;
  a.nx = h->nx;
  a.ny = h->ny;
  a.nz = h->nz;
  for (int d = 0; d < 3; ++d) {
    a.bmin[d] = h->bmin[d];
    a.dx[d] = h->dx[d];
  }
  a.dt = dt_s;
  a.mu = mu_pa;
  a.lambda = lambda_pa;
  a.damping = velocity_damping;
  a.primary = primary_region_id;
  a.secondary = secondary_region_id;
  a.partials = h->partials;
  a.counter = h->counter;
  a.totals = h->totals;
  a.report = h->report;
  const uint32_t nodes = (uint32_t)h->nx * h->ny * h->nz;
  hipLaunchKernelGGL(k_mpm_clear, dim3(mpm_blocks(nodes)), dim3(kBlock), 0, st, nodes, a.gm, a.gv, a.gf);
  const uint32_t pb = mpm_blocks(count), nb = mpm_blocks(nodes);
  if (h->p2g_mode == HFSI_MPM_P2G_BINNED) {
    BinArgs& b = h->bins;
    b.stamp = b.stamp == INT32_MAX ? 1 : b.stamp + 1;
    b.has_force = a.f_ext != nullptr;
    const int32_t nb = b.nbx * b.nby * b.nbz;
    const int32_t nchunks = (nb + kScanChunk - 1) / kScanChunk;
    const uint32_t G = (uint32_t)std::min<int32_t>(nb, 2048), G2 = 64;
    hipLaunchKernelGGL(k_bin_clear, dim3(std::min<int32_t>((nb + 255) / 256, 1024)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_bin_count, dim3(pb), dim3(kBlock), 0, st, a, b);
    hipLaunchKernelGGL(k_scan_chunks, dim3(nchunks), dim3(kScanChunk), 0, st, (const int32_t*)b.count, b.offset, nb,
                       b.scan_sums);
    hipLaunchKernelGGL(k_scan_add, dim3(nchunks), dim3(kScanChunk), 0, st, b.offset, nb, (const int32_t*)b.scan_sums,
                       nchunks);
    hipLaunchKernelGGL(k_bin_scatter, dim3(pb), dim3(kBlock), 0, st, count, b);
    if (b.tiles) hipLaunchKernelGGL(k_need_mark, dim3(std::min<int32_t>((nb + 255) / 256, 256)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_mpm_p2g_binned, dim3(G), dim3(kBlock), 0, st, a, b);
    hipLaunchKernelGGL(k_mpm_p2g_oob, dim3(G2), dim3(kBlock), 0, st, a, b, a.partials + (size_t)G * 12);
    hipLaunchKernelGGL((k_finish_sums<12, FinNone>), dim3(1), dim3(kFinishThreads), 0, st, (const double*)a.partials,
                       G + G2, a.totals + 0, FinNone{}, kP2gMax);
  } else {
    hipLaunchKernelGGL(k_mpm_p2g, dim3(pb), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL((k_finish_sums<12, FinNone>), dim3(1), dim3(kFinishThreads), 0, st,
                       (const double*)a.partials, pb, a.totals + 0, FinNone{}, kP2gMax);
  }
  if (h->p2g_mode == HFSI_MPM_P2G_BINNED && h->bins.tiles) {
    const uint32_t GG = 2048;
    hipLaunchKernelGGL(k_mpm_grid_gather, dim3(GG), dim3(kBlock), 0, st, a, h->bins);
    hipLaunchKernelGGL((k_finish_sums<8, FinNone>), dim3(1), dim3(kFinishThreads), 0, st, (const double*)a.partials,
                       GG, a.totals + T_TRANS, FinNone{}, kGridMax);
  } else {
    hipLaunchKernelGGL(k_mpm_grid, dim3(nb), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL((k_finish_sums<8, FinNone>), dim3(1), dim3(kFinishThreads), 0, st, (const double*)a.partials,
                       nb, a.totals + T_TRANS, FinNone{}, kGridMax);
  }
  hipLaunchKernelGGL(k_mpm_g2p, dim3(pb), dim3(kBlock), 0, st, a);
  hipLaunchKernelGGL((k_finish_sums<21, FinNone>), dim3(1), dim3(kFinishThreads), 0, st, (const double*)a.partials,
                     pb, a.totals + T_PDISP, FinNone{}, (uint64_t)0);
  hipLaunchKernelGGL(k_mpm_radial, dim3(pb), dim3(kBlock), 0, st, a);
  hipLaunchKernelGGL((k_finish_sums<3, FinMpmReport>), dim3(1), dim3(kFinishThreads), 0, st,
                     (const double*)a.partials, pb, a.totals + T_RSUM, FinMpmReport{a.totals, a.report, a.dt},
                     kRadialMax);
  HFSI_LAUNCH_CHECK();
  if (report34) {
    HFSI_CHECK(hipMemcpyAsync(h->host_report, h->report, 34 * sizeof(double), hipMemcpyDeviceToHost, st));
    HFSI_CHECK(hipStreamSynchronize(st));
    std::memcpy(report34, h->host_report, 34 * sizeof(double));
  }
  return HFSI_OK;
}

}  // extern "C"
