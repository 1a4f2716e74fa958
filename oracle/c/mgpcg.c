/*
 * C + OpenMP restatement of the reference's plain pressure solve (TEST
 * INFRASTRUCTURE: CPU baseline and fast oracle; never linked into libhfsi).
 *
 * Follows simulation_core/fluid.py:
 *   operator            _fv_pressure_neighbor_sum_denominator   4079-4221
 *   weighted Jacobi     _pressure_fv_jacobi_kernel               4223-4273 (omega = f32(0.8))
 *   residual/restrict   _mg_compute_residual_kernel              5261-5307
 *                       _mg_restrict_residual_kernel             5309-5388
 *   prolongation        _mg_prolongate_add_kernel                5390-5406 (x0.7 graded)
 *   V-cycle             _apply_fv_multigrid_preconditioner       6085-6262 (2 / 24 / 2)
 *   flexible PCG        _solve_pressure_poisson_fv_cg            6535-6811
 * Plain path only: obstacles, z-min outlet and the closed-domain mean
 * projection; no velocity-Dirichlet rows or interface terms.  Level 0 in
 * float64, coarse levels stored as float32 with float64 arithmetic (the
 * reference's dtypes).  Metrics are rounded through float32 like the
 * reference's metric fields.
 *
 * Pinned against the NumPy oracle (tests/test_oracle_c.py), which is pinned
 * by the reference's known-answer tests and scipy.sparse.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 16
#define OBST 1u
#define XM 2u
#define XP 4u
#define YM 8u
#define YP 16u
#define ZM 32u
#define ZP 64u
#define NOTFREE 128u

typedef struct {
  int nx, ny, nz;
  long n, sx, sy;
  double *cl[3], *ch[3], *wd[3];
  float* wf[3];
  unsigned char* flags;
  double outlet_add;
  float *p, *t, *rhs; /* coarse storage */
} Level;

typedef struct {
  int L, uniform;
  Level lv[MAXL];
} Hier;

static void axis_level(const double* w, int n, double origin, Level* m, int a) {
  double* centers = (double*)malloc(sizeof(double) * n);
  double face = origin;
  for (int i = 0; i < n; ++i) {
    centers[i] = face + 0.5 * w[i];
    face += w[i];
  }
  m->cl[a] = (double*)calloc(n, sizeof(double));
  m->ch[a] = (double*)calloc(n, sizeof(double));
  m->wd[a] = (double*)malloc(sizeof(double) * n);
  m->wf[a] = (float*)malloc(sizeof(float) * n);
  float* df = (float*)malloc(sizeof(float) * n);
  for (int i = 0; i < n; ++i) {
    m->wf[a][i] = (float)w[i];
    m->wd[a][i] = (double)m->wf[a][i];
    df[i] = (float)(i == 0 ? w[0] : centers[i] - centers[i - 1]);
  }
  for (int i = 0; i < n; ++i) {
    if (i > 0) m->cl[a][i] = 1.0 / ((double)m->wf[a][i] * (double)df[i]);
    if (i + 1 < n) m->ch[a][i] = 1.0 / ((double)m->wf[a][i] * (double)df[i + 1]);
  }
  if (a == 2) m->outlet_add = 2.0 / ((double)m->wf[2][0] * (double)m->wf[2][0]);
  free(df);
  free(centers);
}

static void open_bits(Level* m) {
  const long n = m->n;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    int k = (int)(c % m->nz), j = (int)((c / m->nz) % m->ny), i = (int)(c / m->sx);
    unsigned char f = m->flags[c] & (OBST | NOTFREE);
    if (i > 0 && !(m->flags[c - m->sx] & NOTFREE)) f |= XM;
    if (i + 1 < m->nx && !(m->flags[c + m->sx] & NOTFREE)) f |= XP;
    if (j > 0 && !(m->flags[c - m->sy] & NOTFREE)) f |= YM;
    if (j + 1 < m->ny && !(m->flags[c + m->sy] & NOTFREE)) f |= YP;
    if (k > 0 && !(m->flags[c - 1] & NOTFREE)) f |= ZM;
    if (k + 1 < m->nz && !(m->flags[c + 1] & NOTFREE)) f |= ZP;
    m->flags[c] = f;
  }
}

static void hier_free(Hier* h) {
  for (int l = 0; l < h->L; ++l) {
    Level* m = &h->lv[l];
    for (int a = 0; a < 3; ++a) {
      free(m->cl[a]);
      free(m->ch[a]);
      free(m->wd[a]);
      free(m->wf[a]);
    }
    free(m->flags);
    free(m->p);
    free(m->t);
    free(m->rhs);
  }
}

static void hier_build(Hier* h, int nx, int ny, int nz, const double* wx, const double* wy, const double* wz,
                       const double* origin, const int32_t* obstacle) {
  memset(h, 0, sizeof(*h));
  int n3[3] = {nx, ny, nz};
  double* cur[3];
  const double* w0[3] = {wx, wy, wz};
  h->uniform = 1;
  for (int a = 0; a < 3; ++a) {
    cur[a] = (double*)malloc(sizeof(double) * n3[a]);
    memcpy(cur[a], w0[a], sizeof(double) * n3[a]);
    double tol = fmax(fabs(cur[a][0]), 1.0) * 1e-12;
    for (int i = 0; i < n3[a]; ++i)
      if (fabs(cur[a][i] - cur[a][0]) > tol) h->uniform = 0;
  }
  int L = 0;
  for (;;) {
    Level* m = &h->lv[L];
    m->nx = n3[0];
    m->ny = n3[1];
    m->nz = n3[2];
    m->n = (long)n3[0] * n3[1] * n3[2];
    m->sx = (long)n3[1] * n3[2];
    m->sy = n3[2];
    for (int a = 0; a < 3; ++a) axis_level(cur[a], n3[a], origin[a], m, a);
    m->flags = (unsigned char*)calloc(m->n, 1);
    if (L > 0) {
      m->p = (float*)calloc(m->n, sizeof(float));
      m->t = (float*)calloc(m->n, sizeof(float));
      m->rhs = (float*)calloc(m->n, sizeof(float));
    }
    ++L;
    int mn = n3[0] < n3[1] ? n3[0] : n3[1];
    mn = mn < n3[2] ? mn : n3[2];
    if (mn <= 3 || L >= MAXL) break;
    for (int a = 0; a < 3; ++a) {
      int cn = (n3[a] + 1) / 2;
      double* nw = (double*)malloc(sizeof(double) * cn);
      for (int i = 0; i < n3[a]; i += 2) nw[i / 2] = (i + 1 < n3[a]) ? cur[a][i] + cur[a][i + 1] : cur[a][i];
      free(cur[a]);
      cur[a] = nw;
      n3[a] = cn;
    }
  }
  for (int a = 0; a < 3; ++a) free(cur[a]);
  h->L = L;
  /* masks */
  Level* m0 = &h->lv[0];
  for (long c = 0; c < m0->n; ++c)
    m0->flags[c] = (unsigned char)((obstacle[c] == 1 ? OBST : 0) | (obstacle[c] != 0 ? NOTFREE : 0));
  open_bits(m0);
  for (int l = 1; l < L; ++l) {
    Level *f = &h->lv[l - 1], *m = &h->lv[l];
    for (long c = 0; c < m->n; ++c) {
      int ck = (int)(c % m->nz), cj = (int)((c / m->nz) % m->ny), ci = (int)(c / m->sx);
      int obst = 0;
      for (int oi = 0; oi < 2; ++oi)
        for (int oj = 0; oj < 2; ++oj)
          for (int ok = 0; ok < 2; ++ok) {
            int fi = 2 * ci + oi, fj = 2 * cj + oj, fk = 2 * ck + ok;
            if (fi < f->nx && fj < f->ny && fk < f->nz && (f->flags[fi * f->sx + fj * f->sy + fk] & OBST)) ++obst;
          }
      m->flags[c] = obst ? (OBST | NOTFREE) : 0;
    }
    open_bits(m);
  }
}

/* neighbour sum / denominator of cell c; P(idx) gives the value */
#define NBDEN(m, c, i, j, k, f, P, nb, den, outlet)          \
  do {                                                      \
    nb = 0.0;                                               \
    den = 0.0;                                              \
    if (f & XM) { double q = m->cl[0][i]; nb += q * (P(c - m->sx)); den += q; } \
    if (f & XP) { double q = m->ch[0][i]; nb += q * (P(c + m->sx)); den += q; } \
    if (f & YM) { double q = m->cl[1][j]; nb += q * (P(c - m->sy)); den += q; } \
    if (f & YP) { double q = m->ch[1][j]; nb += q * (P(c + m->sy)); den += q; } \
    if (f & ZM) { double q = m->cl[2][k]; nb += q * (P(c - 1)); den += q; }     \
    if (f & ZP) { double q = m->ch[2][k]; nb += q * (P(c + 1)); den += q; }     \
    if ((outlet) && k == 0) den += m->outlet_add;          \
  } while (0)

#define DECOMP(m, c, i, j, k) \
  int k = (int)((c) % (m)->nz), j = (int)(((c) / (m)->nz) % (m)->ny), i = (int)((c) / (m)->sx)

static const double OMEGA = (double)0.8f;


static void sweep_d(const Level* m, const double* p, const double* R, double rs, double* out, int outlet,
                    int from_zero) {
  const long n = m->n;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    unsigned char f = m->flags[c];
    if (f & OBST) {
      out[c] = 0.0;
      continue;
    }
    DECOMP(m, c, i, j, k);
    double nb, den;
#define PV(x) (from_zero ? 0.0 : p[x])
    NBDEN(m, c, i, j, k, f, PV, nb, den, outlet);
#undef PV
    if (!(den > 0.0)) {
      out[c] = 0.0;
      continue;
    }
    double pc = from_zero ? 0.0 : p[c];
    double jac = (nb - rs * R[c]) / den;
    out[c] = pc + OMEGA * (jac - pc);
  }
}

static void sweep_f(const Level* m, const float* p, const float* R, float* out, int outlet, int from_zero) {
  const long n = m->n;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    unsigned char f = m->flags[c];
    if (f & OBST) {
      out[c] = 0.0f;
      continue;
    }
    DECOMP(m, c, i, j, k);
    double nb, den;
#define PV(x) (from_zero ? 0.0 : (double)p[x])
    NBDEN(m, c, i, j, k, f, PV, nb, den, outlet);
#undef PV
    if (!(den > 0.0)) {
      out[c] = 0.0f;
      continue;
    }
    double pc = from_zero ? 0.0 : (double)p[c];
    double jac = (nb - (double)R[c]) / den;
    out[c] = (float)(pc + OMEGA * (jac - pc));
  }
}

/* residual of fine level (double or float), restricted into coarse rhs */
static void resid_restrict(const Level* f, const double* pd, const float* pf, const double* Rd, double rs,
                           const float* Rf, int outlet, Level* m) {
  const long n = m->n;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < n; ++c) {
    if (m->flags[c] & OBST) {
      m->rhs[c] = 0.0f;
      continue;
    }
    DECOMP(m, c, ci, cj, ck);
    double rsum = 0.0, vsum = 0.0;
    for (int oi = 0; oi < 2; ++oi)
      for (int oj = 0; oj < 2; ++oj)
        for (int ok = 0; ok < 2; ++ok) {
          int fi = 2 * ci + oi, fj = 2 * cj + oj, fk = 2 * ck + ok;
          if (fi >= f->nx || fj >= f->ny || fk >= f->nz) continue;
          long fc = fi * f->sx + fj * f->sy + fk;
          unsigned char fl = f->flags[fc];
          if (fl & OBST) continue;
          double vol = (double)(f->wf[0][fi] * f->wf[1][fj] * f->wf[2][fk]);
          double nb, den, pc, rhs;
          if (pd) {
#define PV(x) (pd[x])
            NBDEN(f, fc, fi, fj, fk, fl, PV, nb, den, outlet);
#undef PV
            pc = pd[fc];
            rhs = rs * Rd[fc];
          } else {
#define PV(x) ((double)pf[x])
            NBDEN(f, fc, fi, fj, fk, fl, PV, nb, den, outlet);
#undef PV
            pc = pf[fc];
            rhs = Rf[fc];
          }
          double res = rhs - (nb - den * pc);
          if (!pd) res = (double)(float)res;
          rsum += res * vol;
          vsum += vol;
        }
    m->rhs[c] = (float)(rsum / fmax(vsum, 1e-30));
  }
}

static inline long parent(const Level* f, const Level* m, long c) {
  DECOMP(f, c, i, j, k);
  int ci = i / 2 < m->nx - 1 ? i / 2 : m->nx - 1;
  int cj = j / 2 < m->ny - 1 ? j / 2 : m->ny - 1;
  int ck = k / 2 < m->nz - 1 ? k / 2 : m->nz - 1;
  return ci * m->sx + cj * m->sy + ck;
}

/* V-cycle: z (double, level 0) from r; scratch t0 (double) */
static void vcycle(Hier* h, const double* r, double* z, double* t0, int pre, int coarse, int post, int outlet) {
  Level* l0 = &h->lv[0];
  const int L = h->L;
  const double scale = (double)(h->uniform ? 1.0f : 0.7f);
  /* level 0 pre-smoothing: rhs = -r */
  double* a = t0;
  double* b = z;
  double *in = NULL, *out;
  int W = pre + post;
  if (L <= 1) {
    int S = pre + post > coarse ? pre + post : coarse;
    for (int s = 1; s <= S; ++s) {
      out = ((S - s) % 2 == 0) ? b : a;
      sweep_d(l0, in, r, -1.0, out, outlet, s == 1);
      in = out;
    }
    return;
  }
  for (int s = 1; s <= pre; ++s) {
    out = ((W - s) % 2 == 0) ? b : a;
    sweep_d(l0, in, r, -1.0, out, outlet, s == 1);
    in = out;
  }
  double* pre0 = in;
  resid_restrict(l0, pre0, NULL, r, -1.0, NULL, outlet, &h->lv[1]);
  float* cur[MAXL];
  for (int l = 1; l < L - 1; ++l) {
    Level* m = &h->lv[l];
    float* fin = NULL;
    for (int s = 1; s <= pre; ++s) {
      float* fo = ((W - s) % 2 == 0) ? m->p : m->t;
      sweep_f(m, fin, m->rhs, fo, outlet, s == 1);
      fin = fo;
    }
    cur[l] = fin;
    resid_restrict(m, NULL, fin, NULL, 0.0, m->rhs, outlet, &h->lv[l + 1]);
  }
  {
    Level* m = &h->lv[L - 1];
    float* fin = NULL;
    for (int s = 1; s <= coarse; ++s) {
      float* fo = ((coarse - s) % 2 == 0) ? m->p : m->t;
      sweep_f(m, fin, m->rhs, fo, outlet, s == 1);
      fin = fo;
    }
    cur[L - 1] = fin;
  }
  for (int l = L - 2; l >= 1; --l) {
    Level *m = &h->lv[l], *mc = &h->lv[l + 1];
    float* p = cur[l];
#pragma omp parallel for schedule(static)
    for (long c = 0; c < m->n; ++c)
      if (!(m->flags[c] & NOTFREE)) p[c] = (float)((double)p[c] + scale * (double)cur[l + 1][parent(m, mc, c)]);
    float* fin = p;
    for (int s = pre + 1; s <= W; ++s) {
      float* fo = ((W - s) % 2 == 0) ? m->p : m->t;
      sweep_f(m, fin, m->rhs, fo, outlet, 0);
      fin = fo;
    }
    cur[l] = fin;
  }
  {
    Level* mc = &h->lv[1];
#pragma omp parallel for schedule(static)
    for (long c = 0; c < l0->n; ++c)
      if (!(l0->flags[c] & NOTFREE)) pre0[c] = pre0[c] + scale * (double)cur[1][parent(l0, mc, c)];
    in = pre0;
    for (int s = pre + 1; s <= W; ++s) {
      out = ((W - s) % 2 == 0) ? b : a;
      sweep_d(l0, in, r, -1.0, out, outlet, 0);
      in = out;
    }
  }
}

static void apply(const Hier* h, const double* p, double* q, int outlet) {
  const Level* m = &h->lv[0];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < m->n; ++c) {
    unsigned char f = m->flags[c];
    if (f & OBST) {
      q[c] = 0.0;
      continue;
    }
    DECOMP(m, c, i, j, k);
    double nb, den;
#define PV(x) (p[x])
    NBDEN(m, c, i, j, k, f, PV, nb, den, outlet);
#undef PV
    q[c] = den * p[c] - nb;
  }
}

static double wdot(const Hier* h, const double* a, const double* b) {
  const Level* m = &h->lv[0];
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long c = 0; c < m->n; ++c) {
    if (m->flags[c] & NOTFREE) continue;
    DECOMP(m, c, i, j, k);
    s += a[c] * b[c] * (m->wd[0][i] * m->wd[1][j] * m->wd[2][k]);
  }
  return s;
}

static double wsum(const Hier* h, const double* a, double* vol_out) {
  const Level* m = &h->lv[0];
  double s = 0.0, v = 0.0;
#pragma omp parallel for reduction(+ : s, v) schedule(static)
  for (long c = 0; c < m->n; ++c) {
    if (m->flags[c] & NOTFREE) continue;
    DECOMP(m, c, i, j, k);
    double V = m->wd[0][i] * m->wd[1][j] * m->wd[2][k];
    s += a[c] * V;
    v += V;
  }
  if (vol_out) *vol_out = v;
  return s;
}

static void sub_mean(const Hier* h, double* a) {
  double V;
  double s = wsum(h, a, &V);
  double mean = V > 0.0 ? s / V : 0.0;
  const Level* m = &h->lv[0];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < m->n; ++c) a[c] = (m->flags[c] & NOTFREE) ? 0.0 : a[c] - mean;
}

static void jacobi_precond(const Hier* h, const double* r, double* z, int outlet) {
  const Level* m = &h->lv[0];
#pragma omp parallel for schedule(static)
  for (long c = 0; c < m->n; ++c) {
    unsigned char f = m->flags[c];
    if (f & OBST) {
      z[c] = 0.0;
      continue;
    }
    DECOMP(m, c, i, j, k);
    double nb, den;
#define PV(x) (0.0)
    NBDEN(m, c, i, j, k, f, PV, nb, den, outlet);
#undef PV
    (void)nb;
    z[c] = den > 0.0 ? r[c] / den : 0.0;
  }
}

/*
 * Flexible PCG.  rhs0: legacy-sign level-0 right-hand side (float64, 0 on
 * obstacle rows).  x: warm start / result.  precond: 0 Jacobi, 1 MG (2/24/2).
 * Returns the iteration count; *rel_out the relative residual of the last
 * check, *converged_out 1/0.  `iteration_cap` > 0 stops after that many
 * iterations regardless of convergence (bounded CPU-baseline samples).
 */
int oracle_pcg(int nx, int ny, int nz, const double* wx, const double* wy, const double* wz, const double* origin,
               const int32_t* obstacle, const double* rhs0, int outlet, int remove_mean, int precond, int max_iters,
               double tol, double* x, double* rel_out, int* converged_out) {
  Hier h;
  hier_build(&h, nx, ny, nz, wx, wy, wz, origin, obstacle);
  const long n = h.lv[0].n;
  const Level* m = &h.lv[0];
  double* b = (double*)malloc(sizeof(double) * n);
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* d = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* ro = (double*)malloc(sizeof(double) * n);
  double* t0 = (double*)malloc(sizeof(double) * n);
  int it = 0, converged = 0;
  double rel = INFINITY;
  double V;
  double sum_rhs = wsum(&h, rhs0, &V);
  double meanb = (remove_mean && V > 0.0) ? -sum_rhs / V : 0.0; /* cg_weighted_mean, multiplier -1 */
  for (long c = 0; c < n; ++c) b[c] = (m->flags[c] & OBST) ? 0.0 : -rhs0[c] - meanb;
  double bnorm = sqrt(fmax(wdot(&h, b, b), 0.0));
  if (!(bnorm > 1e-30)) {
    memset(x, 0, sizeof(double) * n);
    *rel_out = 0.0;
    *converged_out = 1;
    goto done;
  }
  apply(&h, x, q, outlet);
  for (long c = 0; c < n; ++c) r[c] = (m->flags[c] & OBST) ? 0.0 : b[c] - q[c];
  if (remove_mean) sub_mean(&h, r);
  rel = sqrt(fmax(wdot(&h, r, r), 0.0)) / bnorm;
  if (rel <= tol) {
    converged = 1;
    goto done;
  }
  {
    int use_mg = precond == 1 && h.L > 1;
    if (use_mg)
      vcycle(&h, r, z, t0, 2, 24, 2, outlet);
    else
      jacobi_precond(&h, r, z, outlet);
    if (remove_mean) sub_mean(&h, z);
    double rz = wdot(&h, r, z);
    if (!isfinite(rz) || fabs(rz) <= 1e-300) {
      use_mg = 0;
      jacobi_precond(&h, r, z, outlet);
      rz = wdot(&h, r, z);
    }
    if (!isfinite(rz) || rz <= 0.0) goto done;
    memcpy(d, z, sizeof(double) * n);
    int code = 0;
    while (it < max_iters) {
      ++it;
      apply(&h, d, q, outlet);
      double dAd = wdot(&h, d, q);
      double alpha = 0.0;
      if (dAd <= 0.0)
        code = 1;
      else
        alpha = rz / dAd;
#pragma omp parallel for schedule(static)
      for (long c = 0; c < n; ++c) {
        if (m->flags[c] & OBST) {
          ro[c] = 0.0;
          x[c] = 0.0;
          r[c] = 0.0;
        } else {
          ro[c] = r[c];
          x[c] += alpha * d[c];
          r[c] -= alpha * q[c];
        }
      }
      if (remove_mean && it % 16 == 0) {
        sub_mean(&h, x);
        sub_mean(&h, r);
      }
      if (it % 16 == 0 || it == max_iters) {
        rel = sqrt(fmax(wdot(&h, r, r), 0.0)) / bnorm;
        if (rel <= tol) {
          converged = 1;
          break;
        }
        if (code != 0) break; /* the plain-path baseline never breaks down */
        if (it < max_iters && it % 512 == 0) {
          apply(&h, x, q, outlet);
          for (long c = 0; c < n; ++c) r[c] = (m->flags[c] & OBST) ? 0.0 : b[c] - q[c];
          if (remove_mean) sub_mean(&h, r);
          rel = sqrt(fmax(wdot(&h, r, r), 0.0)) / bnorm;
          if (rel <= tol) {
            converged = 1;
            break;
          }
          if (use_mg)
            vcycle(&h, r, z, t0, 2, 24, 2, outlet);
          else
            jacobi_precond(&h, r, z, outlet);
          if (remove_mean) sub_mean(&h, z);
          rz = wdot(&h, r, z);
          memcpy(d, z, sizeof(double) * n);
          continue;
        }
      }
      if (use_mg)
        vcycle(&h, r, z, t0, 2, 24, 2, outlet);
      else
        jacobi_precond(&h, r, z, outlet);
      if (remove_mean) sub_mean(&h, z);
      double rz_new = wdot(&h, r, z);
      double beta = 0.0;
      if (use_mg) {
#pragma omp parallel for schedule(static)
        for (long c = 0; c < n; ++c) q[c] = (m->flags[c] & OBST) ? 0.0 : r[c] - ro[c];
        double num = wdot(&h, z, q);
        if (rz <= 1e-300 || rz_new <= 0.0)
          code = 2;
        else {
          beta = num / rz;
          if (beta < 0.0) beta = 0.0;
        }
      } else {
        if (rz <= 1e-300 || rz_new <= 0.0)
          code = 2;
        else
          beta = rz_new / rz;
      }
#pragma omp parallel for schedule(static)
      for (long c = 0; c < n; ++c) d[c] = (m->flags[c] & OBST) ? 0.0 : z[c] + beta * d[c];
      rz = rz_new;
    }
  }
done:
  *rel_out = rel;
  *converged_out = converged;
  free(b);
  free(r);
  free(z);
  free(d);
  free(q);
  free(ro);
  free(t0);
  hier_free(&h);
  return it;
}

/* One V-cycle (for pinning against the NumPy oracle). */
int oracle_vcycle(int nx, int ny, int nz, const double* wx, const double* wy, const double* wz, const double* origin,
                  const int32_t* obstacle, const double* r, double* z, int outlet) {
  Hier h;
  hier_build(&h, nx, ny, nz, wx, wy, wz, origin, obstacle);
  double* t0 = (double*)malloc(sizeof(double) * h.lv[0].n);
  vcycle(&h, r, z, t0, 2, 24, 2, outlet);
  free(t0);
  hier_free(&h);
  return 0;
}
