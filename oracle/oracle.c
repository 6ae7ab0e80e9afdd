/*
 * oracle.c -- CPU restatement of heterogpc's PC-stable skeleton path
 * (milanpro/Masterthesis-Heterogeneous-PC, the `--cpu-only` executor plus CPU
 * restatements of the correlation and compaction kernels).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library.  It is the checker the
 * HIP path is compared against; the product never links or calls it.
 *
 * Arithmetic contract: built with -ffp-contract=off; every fused multiply-add
 * below is an explicit fma().  The device code performs the same operations
 * in the same order, so results are bit-identical and so are decisions.
 *
 * Pins (what independent fact checks each part; see tests/test_oracle_*.py):
 *   orc_pval        scipy.special.erfc / scipy.stats.norm.sf golden vectors
 *   orc_r_crit      monotone boundary of orc_pval, checked on neighbours
 *   orc_pcor_l1     numpy.linalg.pinv on the 3x3 submatrix
 *   orc_pcor_*      numpy.linalg.pinv (rcond = Armadillo tolerance) on random
 *                   SPD and rank-deficient matrices; recursive identity
 *   orc_unrank      itertools.combinations, brute force n <= 20, k <= 6
 *   orc_moments /
 *   orc_corr        numpy.corrcoef (fp64) within the float-sd bound, plus a
 *                   numpy emulation of the lane-strided order on small n
 *   orc_level/run   reference known answer data/cooling_house_skeleton.txt;
 *                   brute-force PC-stable over all conditioning sets (p <= 8);
 *                   invariants of assertions.cpp:5-30; order independence
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_CUT_THR 0.9999999        /* cpuInd.cpp:7 */
#define ORC_SQRT2 1.4142135623730951 /* sd * sqrt(2) of the unit normal */
#define ORC_PIV_TOL 1e-12
#define ORC_MMAX 40                  /* largest submatrix the oracle handles */
#define ORC_JACOBI_SWEEPS 60

/* ======================================================================== */
/* scalar statistics                                                         */
/* ======================================================================== */

/* cpuInd.cpp:8-14: r <- NaN ? 0 : min(0.9999999, |r|);
 * z = sqrt(n-3) * 0.5 * log1p(2r/(1-r)); p = 2 * Q(z) = erfc(z / sqrt 2). */
double orc_pval(double r_in, int n) {
  double r = isnan(r_in) ? 0.0 : fmin(ORC_CUT_THR, fabs(r_in));
  double absz = sqrt(n - 3.0) * 0.5 * log1p(2 * r / (1 - r));
  return erfc(absz / ORC_SQRT2);
}

/* gpuRowTests.cu:18 / cpuRowTests.cpp:12: removal iff p >= alpha. */
int orc_removes(double pval, double alpha) { return pval >= alpha; }

/* cpuInd.cpp:16-22 (x1 = r_ij, x2 = r_ik, x3 = r_jk). */
double orc_pcor_l1(double x1, double x2, double x3) {
  return (x1 - x2 * x3) / sqrt((1.0 - x3 * x3) * (1.0 - x2 * x2));
}

/* cpuInd.cpp:16-27: NaN inputs give a NaN p-value (the edge is kept). */
double orc_pval_l1(double x1, double x2, double x3, int n) {
  if (isnan(x1) || isnan(x2) || isnan(x3)) return NAN;
  return orc_pval(orc_pcor_l1(x1, x2, x3), n);
}

/* Symmetric cyclic Jacobi eigen-decomposition of M (m x m), then the
 * pseudoinverse entries P00, P11, P01 with Armadillo's default tolerance
 * tol = max(m,n) * max singular value * eps (singular values of a symmetric
 * matrix are |eigenvalues|).  Returns -P01 / sqrt(P00 * P11)
 * (cpuInd.cpp:35-36). */
double orc_pcor_pinv(int m, const double *M) {
  double A[ORC_MMAX][ORC_MMAX], V[ORC_MMAX][ORC_MMAX];
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      A[a][b] = M[a * m + b];
      V[a][b] = (a == b) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < ORC_JACOBI_SWEEPS; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int a = 0; a < m; ++a) {
      dia = fma(A[a][a], A[a][a], dia);
      for (int b = a + 1; b < m; ++b) off = fma(A[a][b], A[a][b], off);
    }
    if (!(off > 1e-34 * dia)) break;
    for (int pp = 0; pp < m - 1; ++pp) {
      for (int qq = pp + 1; qq < m; ++qq) {
        double apq = A[pp][qq];
        if (apq == 0.0) continue;
        double app = A[pp][pp], aqq = A[qq][qq];
        double theta = (aqq - app) / (2.0 * apq);
        double t = 1.0 / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
        if (theta < 0.0) t = -t;
        double c = 1.0 / sqrt(fma(t, t, 1.0));
        double s = t * c;
        for (int k = 0; k < m; ++k) {
          if (k == pp || k == qq) continue;
          double akp = A[k][pp], akq = A[k][qq];
          double nkp = fma(c, akp, -(s * akq));
          double nkq = fma(s, akp, c * akq);
          A[k][pp] = nkp; A[pp][k] = nkp;
          A[k][qq] = nkq; A[qq][k] = nkq;
        }
        A[pp][pp] = fma(-t, apq, app);
        A[qq][qq] = fma(t, apq, aqq);
        A[pp][qq] = 0.0; A[qq][pp] = 0.0;
        for (int k = 0; k < m; ++k) {
          double vkp = V[k][pp], vkq = V[k][qq];
          V[k][pp] = fma(c, vkp, -(s * vkq));
          V[k][qq] = fma(s, vkp, c * vkq);
        }
      }
    }
  }
  double maxabs = 0.0;
  for (int a = 0; a < m; ++a) maxabs = fmax(maxabs, fabs(A[a][a]));
  double tol = (double)m * maxabs * DBL_EPSILON;
  double p00 = 0.0, p11 = 0.0, p01 = 0.0;
  for (int a = 0; a < m; ++a) {
    double lam = A[a][a];
    if (!(fabs(lam) > tol)) continue;
    double inv = 1.0 / lam;
    p00 = fma(V[0][a] * V[0][a], inv, p00);
    p11 = fma(V[1][a] * V[1][a], inv, p11);
    p01 = fma(V[0][a] * V[1][a], inv, p01);
  }
  return -p01 / sqrt(p00 * p11);
}

/* Canonical partial correlation of 0 and 1 given 2..m-1 (m = l + 2).
 * With C = M[2:,2:] = U D U^T (U unit lower), u = U^-1 M[0,2:],
 * v = U^-1 M[1,2:]:
 *   s00 = 1 - u.D^-1 u, s11 = 1 - v.D^-1 v, s01 = M01 - u.D^-1 v,
 *   r = s01 / sqrt(s00 s11),
 * which equals -P01/sqrt(P00 P11) for P = M^-1 (the 2x2 block of M^-1 is the
 * inverse of the Schur complement [[s00,s01],[s01,s11]]).  Any pivot at or
 * below 1e-12, or a numerically singular Schur block, selects the pinv form
 * so that rank-deficient sets follow Armadillo's truncation. */
double orc_pcor_canon(int m, const double *M, int *used_pinv) {
  const int l = m - 2;
  /* LDL^T of the S block: U unit lower, E = U D, inv = 1/D */
  double U[ORC_MMAX][ORC_MMAX], E[ORC_MMAX][ORC_MMAX], inv[ORC_MMAX], u[ORC_MMAX], v[ORC_MMAX];
  if (used_pinv) *used_pinv = 0;
  for (int r = 0; r < l; ++r) {
    for (int c = 0; c < r; ++c) {
      double e = M[(r + 2) * m + (c + 2)];
      for (int k = 0; k < c; ++k) e = fma(-E[r][k], U[c][k], e);
      E[r][c] = e;
      U[r][c] = e * inv[c];
    }
    double d = 1.0;
    for (int k = 0; k < r; ++k) d = fma(-E[r][k], U[r][k], d);
    if (!(d > ORC_PIV_TOL)) goto singular;
    inv[r] = 1.0 / d;
  }
  for (int c = 0; c < l; ++c) {
    double su = M[0 * m + (c + 2)], sv = M[1 * m + (c + 2)];
    for (int k = 0; k < c; ++k) {
      su = fma(-U[c][k], u[k], su);
      sv = fma(-U[c][k], v[k], sv);
    }
    u[c] = su;
    v[c] = sv;
  }
  {
    double s00 = 1.0, s11 = 1.0, s01 = M[0 * m + 1];
    for (int k = 0; k < l; ++k) {
      const double wu = u[k] * inv[k], wv = v[k] * inv[k];
      s00 = fma(-wu, u[k], s00);
      s11 = fma(-wv, v[k], s11);
      s01 = fma(-wu, v[k], s01);
    }
    if (!(s00 > ORC_PIV_TOL) || !(s11 > ORC_PIV_TOL)) goto singular;
    double q = s00 * s11;
    if (!(fma(-s01, s01, q) > ORC_PIV_TOL * q)) goto singular;
    return s01 / sqrt(q);
  }
singular:
  if (used_pinv) *used_pinv = 1;
  return orc_pcor_pinv(m, M);
}

/* cpuInd.cpp:29-38. */
double orc_pval_ln(int m, const double *M, int n, int mode) {
  for (int a = 0; a < m * m; ++a)
    if (isnan(M[a])) return NAN;
  double r = (mode == ORC_MODE_PINV) ? orc_pcor_pinv(m, M) : orc_pcor_canon(m, M, NULL);
  return orc_pval(r, n);
}

double orc_r_crit(int n, double alpha) {
  double hi = ORC_CUT_THR;
  if (orc_pval(hi, n) >= alpha) return INFINITY;
  if (!(orc_pval(0.0, n) >= alpha)) return -1.0;
  uint64_t a, b;
  double lo = 0.0;
  memcpy(&a, &lo, 8);
  memcpy(&b, &hi, 8);
  while (b - a > 1) {
    uint64_t mid = a + (b - a) / 2;
    double r;
    memcpy(&r, &mid, 8);
    if (orc_pval(r, n) >= alpha) a = mid; else b = mid;
  }
  double out;
  memcpy(&out, &a, 8);
  return out;
}

/* ======================================================================== */
/* combinatorics                                                             */
/* ======================================================================== */

/* cuda_util.cu:20-32, saturating instead of wrapping. */
uint64_t orc_binom(int n, int k) {
  if (k < 0 || n < k) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 res = 1;
  for (int i = 0; i < k; ++i) {
    res = res * (unsigned __int128)(n - i) / (unsigned __int128)(i + 1);
    if (res > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
  }
  return (uint64_t)res;
}

/* cuda_util.cu:34-50 (ithCombination), 0-based: q[0] < ... < q[l-1] is the
 * combination of rank `rank` among l-subsets of {0..nn-1} in lexicographic
 * order. */
void orc_unrank(uint64_t rank, int l, int nn, int *q) {
  int v = 0;
  for (int i = 0; i < l; ++i) {
    for (;;) {
      uint64_t c = orc_binom(nn - 1 - v, l - 1 - i);
      if (rank < c) break;
      rank -= c;
      ++v;
    }
    q[i] = v++;
  }
}

/* lexicographic successor; returns 0 after the last combination */
static int next_comb(int *q, int l, int nn) {
  int i = l - 1;
  while (i >= 0 && q[i] == nn - l + i) --i;
  if (i < 0) return 0;
  ++q[i];
  for (int k = i + 1; k < l; ++k) q[k] = q[k - 1] + 1;
  return 1;
}

/* ======================================================================== */
/* correlation                                                               */
/* ======================================================================== */

/* 64 lanes, lane t owning k = t, t+64, ...; halving tree t += t + stride. */
static double tree64(double *s) {
  for (int stride = 32; stride > 0; stride >>= 1)
    for (int t = 0; t < stride; ++t) s[t] = s[t] + s[t + stride];
  return s[0];
}

/* corHelper.cu:8-42 (means over non-NaN cells) and :44-77
 * (sd = sqrtf(sum((a-mean)^2) / (n - 1.f)), rounded through float). */
void orc_moments(const double *x, int p, int n, double *mean, double *sd) {
  for (int a = 0; a < p; ++a) {
    const double *xa = x + (size_t)a * n;
    double s[64];
    int cnt[64];
    for (int t = 0; t < 64; ++t) { s[t] = 0.0; cnt[t] = 0; }
    for (int k = 0; k < n; ++k) {
      double v = xa[k];
      if (!isnan(v)) { s[k & 63] = s[k & 63] + v; cnt[k & 63] += 1; }
    }
    for (int stride = 32; stride > 0; stride >>= 1)
      for (int t = 0; t < stride; ++t) cnt[t] += cnt[t + stride];
    double mu = tree64(s) / cnt[0];
    mean[a] = mu;
    for (int t = 0; t < 64; ++t) s[t] = 0.0;
    for (int k = 0; k < n; ++k) {
      double v = xa[k];
      if (!isnan(v)) { double d = v - mu; s[k & 63] = fma(d, d, s[k & 63]); }
    }
    double q = tree64(s);
    float nm1 = (float)n - 1.f;
    sd[a] = (double)sqrtf((float)(q / (double)nm1));
  }
}

/* corHelper.cu:79-116: cor[a][b] = sum_k z_a z_b / (n - 1.f), skipping pairs
 * with a NaN cell, z = (x - mean) / sd; both triangles and the diagonal. */
void orc_corr(const double *x, int p, int n, double *cor, int nthreads) {
  double *mean = (double *)malloc(sizeof(double) * (size_t)p);
  double *sd = (double *)malloc(sizeof(double) * (size_t)p);
  orc_moments(x, p, n, mean, sd);
  const int npad = (n + 63) & ~63;
  double *z = (double *)calloc((size_t)p * npad, sizeof(double));
  unsigned char *nanm = (unsigned char *)calloc((size_t)p * npad, 1);
  char *hasnan = (char *)calloc((size_t)p, 1);
  for (int a = 0; a < p; ++a)
    for (int k = 0; k < n; ++k) {
      double v = x[(size_t)a * n + k];
      z[(size_t)a * npad + k] = (v - mean[a]) / sd[a];
      if (isnan(v)) { nanm[(size_t)a * npad + k] = 1; hasnan[a] = 1; }
    }
  const double div = (double)((float)n - 1.f);
  const int TB = 4;
  const int nb = (p + TB - 1) / TB;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t blk = 0; blk < (int64_t)nb * nb; ++blk) {
    int A = (int)(blk / nb), B = (int)(blk % nb);
    if (A > B) continue;
    double acc[4][4][64];
    memset(acc, 0, sizeof(acc));
    int a0 = A * TB, b0 = B * TB;
    int ua = (p - a0 < TB) ? p - a0 : TB, ub = (p - b0 < TB) ? p - b0 : TB;
    int anynan = 0;
    for (int u = 0; u < ua; ++u) anynan |= hasnan[a0 + u];
    for (int w = 0; w < ub; ++w) anynan |= hasnan[b0 + w];
    for (int m0 = 0; m0 < npad; m0 += 64) {
      for (int u = 0; u < ua; ++u) {
        const double *za = z + (size_t)(a0 + u) * npad + m0;
        const unsigned char *na = nanm + (size_t)(a0 + u) * npad + m0;
        for (int w = 0; w < ub; ++w) {
          const double *zb = z + (size_t)(b0 + w) * npad + m0;
          const unsigned char *nb2 = nanm + (size_t)(b0 + w) * npad + m0;
          double *ac = acc[u][w];
          if (!anynan) {
            for (int t = 0; t < 64; ++t) ac[t] = fma(za[t], zb[t], ac[t]);
          } else {
            for (int t = 0; t < 64; ++t)
              if (!(na[t] | nb2[t])) ac[t] = fma(za[t], zb[t], ac[t]);
          }
        }
      }
    }
    for (int u = 0; u < ua; ++u)
      for (int w = 0; w < ub; ++w) {
        double r = tree64(acc[u][w]) / div;
        cor[(size_t)(a0 + u) * p + (b0 + w)] = r;
        cor[(size_t)(b0 + w) * p + (a0 + u)] = r;
      }
  }
  free(mean); free(sd); free(z); free(nanm); free(hasnan);
}

/* ======================================================================== */
/* skeleton                                                                  */
/* ======================================================================== */

struct orc_state {
  int p, n, mc, mode, flags, nthreads;
  double alpha;
  double *cor;       /* p*p copy */
  int32_t *adj;      /* p*p, state.cu:31-38 */
  int32_t *sep;      /* p*p*mc, state.cu:40 */
  int32_t *lock;     /* p*p, used by the early-exit timing mode only */
  int compact_valid;
  int32_t *deg, *rowptr, *nbr;
  int64_t *fp;
  int last_level;
};

orc_state *orc_new(const double *cor, int p, int n, double alpha, int mc,
                   int mode, int flags, int nthreads) {
  orc_state *st = (orc_state *)calloc(1, sizeof(orc_state));
  st->p = p; st->n = n; st->alpha = alpha; st->mc = mc < 1 ? 1 : mc;
  st->mode = mode; st->flags = flags;
  st->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  size_t pp = (size_t)p * p;
  st->cor = (double *)malloc(sizeof(double) * pp);
  memcpy(st->cor, cor, sizeof(double) * pp);
  st->adj = (int32_t *)malloc(sizeof(int32_t) * pp);
  for (size_t a = 0; a < pp; ++a) st->adj[a] = 1;
  for (int i = 0; i < p; ++i) st->adj[(size_t)i * p + i] = 0;
  st->sep = (int32_t *)malloc(sizeof(int32_t) * pp * st->mc);
  for (size_t a = 0; a < pp * st->mc; ++a) st->sep[a] = -1;
  st->lock = (int32_t *)calloc(pp, sizeof(int32_t));
  st->deg = (int32_t *)calloc((size_t)p, sizeof(int32_t));
  st->rowptr = (int32_t *)calloc((size_t)p + 1, sizeof(int32_t));
  st->nbr = (int32_t *)malloc(sizeof(int32_t) * (pp > 0 ? pp : 1));
  st->fp = (int64_t *)malloc(sizeof(int64_t) * (pp > 0 ? pp : 1));
  st->last_level = -1;
  return st;
}

void orc_free(orc_state *st) {
  if (!st) return;
  free(st->cor); free(st->adj); free(st->sep); free(st->lock);
  free(st->deg); free(st->rowptr); free(st->nbr); free(st->fp);
  free(st);
}

int orc_mc(const orc_state *st) { return st->mc; }

static void grow_sepsets(orc_state *st, int mc) {
  size_t pp = (size_t)st->p * st->p;
  int32_t *ns = (int32_t *)malloc(sizeof(int32_t) * pp * mc);
  for (size_t e = 0; e < pp; ++e) {
    for (int k = 0; k < mc; ++k)
      ns[e * mc + k] = k < st->mc ? st->sep[e * st->mc + k] : -1;
  }
  free(st->sep);
  st->sep = ns;
  st->mc = mc;
}

/* compact.cu:4-89: ascending neighbour ids of every row. */
static int compact(orc_state *st) {
  const int p = st->p;
  int maxdeg = 0;
  st->rowptr[0] = 0;
  for (int i = 0; i < p; ++i) {
    int d = 0;
    const int32_t *row = st->adj + (size_t)i * p;
    int32_t *out = st->nbr + st->rowptr[i];
    for (int j = 0; j < p; ++j)
      if (row[j]) out[d++] = j;
    st->deg[i] = d;
    st->rowptr[i + 1] = st->rowptr[i] + d;
    if (d > maxdeg) maxdeg = d;
  }
  st->compact_valid = 1;
  return maxdeg;
}

static int32_t find_pos(const orc_state *st, int row, int node) {
  const int32_t *b = st->nbr + st->rowptr[row];
  int lo = 0, hi = st->deg[row];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (b[mid] < node) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static inline int32_t slot_node(const orc_state *st, int row, int y, int q) {
  return st->nbr[st->rowptr[row] + (q < y ? q : q + 1)];
}

static inline int live_adj(const orc_state *st, int i, int j) {
  return __atomic_load_n(&st->adj[(size_t)i * st->p + j], __ATOMIC_RELAXED);
}

static void delete_edge(orc_state *st, int level, int i, int j, const int32_t *S) {
  const int p = st->p;
  __atomic_store_n(&st->adj[(size_t)i * p + j], 0, __ATOMIC_RELAXED);
  __atomic_store_n(&st->adj[(size_t)j * p + i], 0, __ATOMIC_RELAXED);
  int a = i < j ? i : j, b = i < j ? j : i;
  int32_t *slot = st->sep + ((size_t)a * p + b) * st->mc;
  if (level == 0) { slot[0] = -2; return; }
  for (int k = 0; k < level; ++k) slot[k] = S[k];
}

/* Search of one ordered edge (row i, neighbour position y) at level >= 1,
 * cpuRowTests.cpp:23-56 (level 1) and :58-136 / cpuWorkstealingTests.cpp:
 * 39-112 (any level >= 2): candidates are the l-subsets of the row's
 * level-start neighbours without j, in lexicographic order of positions.
 * Returns the first passing rank or -1; *count = tests performed. */
static int64_t search_edge(orc_state *st, int level, int i, int y, int64_t *count) {
  const int p = st->p, n = st->n;
  const int d = st->deg[i], nn = d - 1;
  const int j = st->nbr[st->rowptr[i] + y];
  const double *cor = st->cor;
  const int early = st->flags & ORC_FLAG_REF_EARLY_EXIT;
  *count = 0;
  if (nn < level) return -1;
  if (level == 1) {
    const double x1 = cor[(size_t)i * p + j];
    for (int q = 0; q < nn; ++q) {
      int k = slot_node(st, i, y, q);
      double pv = orc_pval_l1(x1, cor[(size_t)i * p + k], cor[(size_t)j * p + k], n);
      ++*count;
      if (orc_removes(pv, st->alpha)) return q;
      if (early && !live_adj(st, i, j)) return -1;
    }
    return -1;
  }
  const int m = level + 2;
  int q[ORC_MMAX], S[ORC_MMAX];
  double M[ORC_MMAX * ORC_MMAX];
  for (int k = 0; k < level; ++k) q[k] = k;
  int64_t rank = 0;
  do {
    for (int k = 0; k < level; ++k) S[k] = slot_node(st, i, y, q[k]);
    /* gpuRowTests.cu:143-165 / cpuRowTests.cpp:151-171 */
    for (int a = 0; a < m; ++a) M[a * m + a] = 1.0;
    M[0 * m + 1] = M[1 * m + 0] = cor[(size_t)i * p + j];
    for (int c = 2; c < m; ++c) {
      M[0 * m + c] = M[c * m + 0] = cor[(size_t)i * p + S[c - 2]];
      M[1 * m + c] = M[c * m + 1] = cor[(size_t)j * p + S[c - 2]];
    }
    for (int a = 2; a < m; ++a)
      for (int b = a + 1; b < m; ++b)
        M[a * m + b] = M[b * m + a] = cor[(size_t)S[a - 2] * p + S[b - 2]];
    double pv = orc_pval_ln(m, M, n, st->mode);
    ++*count;
    if (orc_removes(pv, st->alpha)) return rank;
    if (early && !live_adj(st, i, j)) return -1;
    ++rank;
  } while (next_comb(q, level, nn));
  return -1;
}

static void sepset_of(const orc_state *st, int level, int row, int y, int64_t rank, int32_t *S) {
  int q[ORC_MMAX];
  if (level == 1) { q[0] = (int)rank; }
  else orc_unrank((uint64_t)rank, level, st->deg[row] - 1, q);
  for (int k = 0; k < level; ++k) S[k] = slot_node(st, row, y, q[k]);
}

typedef struct { int row, len; } row_len;
static int cmp_len_desc(const void *a, const void *b) {
  const row_len *x = (const row_len *)a, *y = (const row_len *)b;
  if (x->len != y->len) return y->len - x->len;
  return x->row - y->row;
}

int orc_level(orc_state *st, int level, int64_t *stats) {
  const int p = st->p;
  int64_t T = 0, U = 0, E = 0, R = 0;
  int maxdeg = p;
  if (level == 0) {
    /* cpuRowTests.cpp:5-21 over the lower triangle */
    const double *cor = st->cor;
#pragma omp parallel for schedule(dynamic, 10) reduction(+ : T, R) num_threads(st->nthreads)
    for (int i = 0; i < p; ++i) {
      for (int j = 0; j < i; ++j) {
        if (!st->adj[(size_t)i * p + j]) continue;
        double pv = orc_pval(cor[(size_t)i * p + j], st->n);
        ++T;
        if (orc_removes(pv, st->alpha)) { delete_edge(st, 0, j, i, NULL); ++R; }
      }
    }
    U = T; E = T;
    st->last_level = 0;
  } else {
    maxdeg = compact(st);
    if (level > st->mc) grow_sepsets(st, level);
    const int64_t nE = st->rowptr[p];
    for (int64_t e = 0; e < nE; ++e) st->fp[e] = -1;
    /* cpuExecutor.cpp:88-109: rows with length >= level, longest first */
    row_len *rows = (row_len *)malloc(sizeof(row_len) * (size_t)(p > 0 ? p : 1));
    int nrows = 0;
    for (int i = 0; i < p; ++i)
      if (st->deg[i] >= level) { rows[nrows].row = i; rows[nrows].len = st->deg[i]; ++nrows; }
    qsort(rows, (size_t)nrows, sizeof(row_len), cmp_len_desc);
    const int early = st->flags & ORC_FLAG_REF_EARLY_EXIT;
#pragma omp parallel for collapse(2) schedule(dynamic, 10) reduction(+ : T, U, E, R) num_threads(st->nthreads)
    for (int r = 0; r < nrows; ++r) {
      for (int y = 0; y < p; ++y) {
        const int i = rows[r].row;
        if (y >= st->deg[i]) continue;
        const int nn = st->deg[i] - 1;
        if (nn < level) continue;
        int64_t cnt;
        int64_t f = search_edge(st, level, i, y, &cnt);
        st->fp[st->rowptr[i] + y] = f;
        uint64_t tot = (level == 1) ? (uint64_t)nn : orc_binom(nn, level);
        U += (int64_t)tot;
        E += 1;
        if (!early) T += (f >= 0) ? f + 1 : (int64_t)tot;
        else T += cnt;
        if (early && f >= 0) {
          /* gpuRowTests.cu:174-201 style: first CAS on the pair wins */
          const int j = st->nbr[st->rowptr[i] + y];
          int a = i < j ? i : j, b = i < j ? j : i;
          int expect = 0;
          if (__atomic_compare_exchange_n(&st->lock[(size_t)a * p + b], &expect, 1, 0,
                                          __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            int32_t S[ORC_MMAX];
            sepset_of(st, level, i, y, f, S);
            delete_edge(st, level, i, j, S);
            ++R;
          }
        }
      }
    }
    free(rows);
    if (!early) {
      /* deterministic commit: the lower-id row's lowest passing set wins,
       * else the higher-id row's */
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : R) num_threads(st->nthreads)
      for (int i = 0; i < p; ++i) {
        for (int y = 0; y < st->deg[i]; ++y) {
          const int j = st->nbr[st->rowptr[i] + y];
          if (j < i) continue;
          const int yj = find_pos(st, j, i);
          int64_t fi = st->fp[st->rowptr[i] + y], fj = st->fp[st->rowptr[j] + yj];
          int32_t S[ORC_MMAX];
          if (fi >= 0) sepset_of(st, level, i, y, fi, S);
          else if (fj >= 0) sepset_of(st, level, j, yj, fj, S);
          else continue;
          delete_edge(st, level, i, j, S);
          ++R;
        }
      }
    }
    st->last_level = level;
  }
  if (stats) {
    stats[0] = T; stats[1] = U; stats[2] = E; stats[3] = R; stats[4] = maxdeg;
  }
  if (level >= 1 && maxdeg - 1 < level) return 1;
  return 0;
}

int orc_run(orc_state *st, int max_level, int64_t *stats, int stats_rows) {
  int ran = 0;
  for (int l = 0; max_level < 0 || l <= max_level; ++l) {
    int64_t s[5];
    int none = orc_level(st, l, s);
    if (stats && ran < stats_rows) memcpy(stats + 5 * ran, s, sizeof(s));
    ++ran;
    if (none && max_level < 0) break;
  }
  return ran;
}

void orc_get_adj(const orc_state *st, int32_t *adj) {
  memcpy(adj, st->adj, sizeof(int32_t) * (size_t)st->p * st->p);
}

void orc_get_compact(const orc_state *st, int32_t *out) {
  const int p = st->p;
  if (!st->compact_valid) {
    for (size_t a = 0; a < (size_t)p * p; ++a) out[a] = 1;
    for (int i = 0; i < p; ++i) out[(size_t)i * p + i] = 0;
    return;
  }
  for (int i = 0; i < p; ++i) {
    int32_t *row = out + (size_t)i * p;
    for (int k = 0; k < p; ++k) row[k] = 0;
    for (int k = 0; k < st->deg[i]; ++k) row[k] = st->nbr[st->rowptr[i] + k];
    row[p - 1] = st->deg[i];
  }
}

void orc_get_sepsets(const orc_state *st, int32_t *sep) {
  memcpy(sep, st->sep, sizeof(int32_t) * (size_t)st->p * st->p * st->mc);
}

int64_t orc_num_ordered_edges(const orc_state *st) {
  return st->compact_valid ? st->rowptr[st->p] : 0;
}

void orc_get_first_pass(const orc_state *st, int64_t *fp, int32_t *rowptr) {
  if (!st->compact_valid) return;
  memcpy(fp, st->fp, sizeof(int64_t) * (size_t)st->rowptr[st->p]);
  memcpy(rowptr, st->rowptr, sizeof(int32_t) * ((size_t)st->p + 1));
}

double orc_pval_ids(const double *cor, int p, int n, int i, int j,
                    const int32_t *S, int l, int mode) {
  if (l == 0) {
    int a = i > j ? i : j, b = i > j ? j : i;
    return orc_pval(cor[(size_t)a * p + b], n);
  }
  if (l == 1)
    return orc_pval_l1(cor[(size_t)i * p + j], cor[(size_t)i * p + S[0]],
                       cor[(size_t)j * p + S[0]], n);
  const int m = l + 2;
  double M[ORC_MMAX * ORC_MMAX];
  for (int a = 0; a < m; ++a) M[a * m + a] = 1.0;
  M[1] = M[m] = cor[(size_t)i * p + j];
  for (int c = 2; c < m; ++c) {
    M[c] = M[c * m] = cor[(size_t)i * p + S[c - 2]];
    M[m + c] = M[c * m + 1] = cor[(size_t)j * p + S[c - 2]];
  }
  for (int a = 2; a < m; ++a)
    for (int b = a + 1; b < m; ++b)
      M[a * m + b] = M[b * m + a] = cor[(size_t)S[a - 2] * p + S[b - 2]];
  return orc_pval_ln(m, M, n, mode);
}
