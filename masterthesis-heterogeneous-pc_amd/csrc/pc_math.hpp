// Arithmetic of the CI tests, shared by the device kernels and the host share
// of a level (hostshare.hip).
//
// Every function performs a fixed sequence of IEEE-754 double operations
// (+, -, *, /, sqrt, and explicit fma), compiled with -ffp-contract=off, so a
// CPU build of the same sequence yields the same bits.  Removal decisions
// compare |r| against r_crit, the largest double whose Fisher-z p-value is
// still >= alpha (computed once on the host), so no transcendental function
// runs on the device and p >= alpha (gpuRowTests.cu:18) is decided exactly.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpc {

constexpr double kCutThr = 0.9999999;  // gpuInd.cu:3 / cpuInd.cpp:7
constexpr double kPivTol = 1e-12;
constexpr int kJacobiSweeps = 60;

// calcPValue (gpuInd.cu:4-9, cpuInd.cpp:8-14) followed by the p >= alpha
// test: NaN r counts as r = 0 (p = 1, removed); |r| is clamped at 0.9999999.
__host__ __device__ __forceinline__ bool removes(double r, double r_crit) {
  double re = __builtin_isnan(r) ? 0.0 : __builtin_fmin(kCutThr, __builtin_fabs(r));
  return re <= r_crit;
}

// pValL1 (cpuInd.cpp:16-27): x1 = r_ij, x2 = r_ik, x3 = r_jk.
__host__ __device__ __forceinline__ double pcor_l1(double x1, double x2, double x3) {
  return (x1 - x2 * x3) / __builtin_sqrt((1.0 - x3 * x3) * (1.0 - x2 * x2));
}

// Partial correlation -P01 / __builtin_sqrt(P00 P11) of the m x m unit-diagonal matrix
// A (row-major, overwritten) through a symmetric cyclic Jacobi
// eigen-decomposition and the pseudoinverse with Armadillo's default
// tolerance m * max|lambda| * eps (cpuInd.cpp:35-36).  Taken only when the
// Cholesky form below meets a non-positive pivot.  The rotation loops stay
// rolled (the matrices live in private memory): the search kernels call this
// rarely, and an unrolled form would set their register count.
template <int M>
__host__ __device__ __noinline__ double pcor_pinv(double (&A)[M][M]) {
  double V[M][M];
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) V[a][b] = (a == b) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < kJacobiSweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int a = 0; a < M; ++a) {
      dia = __builtin_fma(A[a][a], A[a][a], dia);
      for (int b = a + 1; b < M; ++b) off = __builtin_fma(A[a][b], A[a][b], off);
    }
    if (!(off > 1e-34 * dia)) break;
#pragma unroll 1
    for (int pp = 0; pp < M - 1; ++pp) {
#pragma unroll 1
      for (int qq = pp + 1; qq < M; ++qq) {
        double apq = A[pp][qq];
        if (apq == 0.0) continue;
        double app = A[pp][pp], aqq = A[qq][qq];
        double theta = (aqq - app) / (2.0 * apq);
        double t = 1.0 / (__builtin_fabs(theta) + __builtin_sqrt(__builtin_fma(theta, theta, 1.0)));
        if (theta < 0.0) t = -t;
        double c = 1.0 / __builtin_sqrt(__builtin_fma(t, t, 1.0));
        double s = t * c;
#pragma unroll 1
        for (int k = 0; k < M; ++k) {
          if (k == pp || k == qq) continue;
          double akp = A[k][pp], akq = A[k][qq];
          double nkp = __builtin_fma(c, akp, -(s * akq));
          double nkq = __builtin_fma(s, akp, c * akq);
          A[k][pp] = nkp; A[pp][k] = nkp;
          A[k][qq] = nkq; A[qq][k] = nkq;
        }
        A[pp][pp] = __builtin_fma(-t, apq, app);
        A[qq][qq] = __builtin_fma(t, apq, aqq);
        A[pp][qq] = 0.0; A[qq][pp] = 0.0;
#pragma unroll 1
        for (int k = 0; k < M; ++k) {
          double vkp = V[k][pp], vkq = V[k][qq];
          V[k][pp] = __builtin_fma(c, vkp, -(s * vkq));
          V[k][qq] = __builtin_fma(s, vkp, c * vkq);
        }
      }
    }
  }
  double maxabs = 0.0;
  for (int a = 0; a < M; ++a) maxabs = __builtin_fmax(maxabs, __builtin_fabs(A[a][a]));
  double tol = (double)M * maxabs * 2.220446049250313e-16;
  double p00 = 0.0, p11 = 0.0, p01 = 0.0;
  for (int a = 0; a < M; ++a) {
    double lam = A[a][a];
    if (!(__builtin_fabs(lam) > tol)) continue;
    double inv = 1.0 / lam;
    p00 = __builtin_fma(V[0][a] * V[0][a], inv, p00);
    p11 = __builtin_fma(V[1][a] * V[1][a], inv, p11);
    p01 = __builtin_fma(V[0][a] * V[1][a], inv, p01);
  }
  return -p01 / __builtin_sqrt(p00 * p11);
}

// Partial correlation of i and j given S (|S| = L) from a = r_ij,
// x[c] = r_i,S[c], y[c] = r_j,S[c] and the strictly lower triangle
// C[r][c] = r_S[c],S[r] (r > c).  With C = U D U^T (U unit lower, E = U D),
// u = U^-1 x, v = U^-1 y and w = D^-1 u:
// r = (a - w.v) / sqrt((1 - w.u)(1 - (D^-1 v).v)), the same value as
// -P01/sqrt(P00 P11) for P = inverse of the (L+2)-square submatrix [i, j, S]
// (gpuRowTests.cu:143-170).  One reciprocal per pivot and one square root per
// test.  A pivot at or below 1e-12 or a numerically singular 2x2 Schur block
// switches to the pinv form.
// The part of pcor_canon that depends on S and x only (one row i, one
// conditioning set): the factorisation C = U D U^T, u = U^-1 x, w = D^-1 u
// and s00 = 1 - w.u.  A search that tests one S against many j computes it
// once; every value is formed by the same operations as in a per-test call.
template <int L>
struct PcorPrep {
  double Um[L][L], inv[L], u[L], wu[L];
  double s00;
  bool sing;  // a pivot at or below kPivTol
};

template <int L>
__host__ __device__ __forceinline__ void pcor_prep(const double (&x)[L], const double (&C)[L][L],
                                                   PcorPrep<L> &P) {
  double Em[L][L];
  P.sing = false;
#pragma unroll
  for (int r = 0; r < L; ++r) {
#pragma unroll
    for (int c = 0; c < r; ++c) {
      double e = C[r][c];
#pragma unroll
      for (int k = 0; k < c; ++k) e = __builtin_fma(-Em[r][k], P.Um[c][k], e);
      Em[r][c] = e;
      P.Um[r][c] = e * P.inv[c];
    }
    double d = 1.0;
#pragma unroll
    for (int k = 0; k < r; ++k) d = __builtin_fma(-Em[r][k], P.Um[r][k], d);
    P.sing |= !(d > kPivTol);
    P.inv[r] = 1.0 / d;
  }
#pragma unroll
  for (int c = 0; c < L; ++c) {
    double su = x[c];
#pragma unroll
    for (int k = 0; k < c; ++k) su = __builtin_fma(-P.Um[c][k], P.u[k], su);
    P.u[c] = su;
  }
  double s00 = 1.0;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const double wu = P.u[k] * P.inv[k];
    s00 = __builtin_fma(-wu, P.u[k], s00);
    P.wu[k] = wu;
  }
  P.s00 = s00;
}

// The j part: v = U^-1 y, s11 = 1 - (D^-1 v).v, s01 = a - w.v and
// q = s00 s11 (wv receives D^-1 v).  Returns whether S or the 2 x 2 Schur
// block is numerically singular (then the pinv form decides); else the
// statistic is s01 / sqrt(q).
template <int L>
__host__ __device__ __forceinline__ bool pcor_schur(const PcorPrep<L> &P, double a,
                                                    const double (&y)[L], double &s11o,
                                                    double &s01o, double &qo,
                                                    double (&wv)[L]) {
  double v[L];
#pragma unroll
  for (int c = 0; c < L; ++c) {
    double sv = y[c];
#pragma unroll
    for (int k = 0; k < c; ++k) sv = __builtin_fma(-P.Um[c][k], v[k], sv);
    v[c] = sv;
  }
  double s11 = 1.0, s01 = a;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const double w = v[k] * P.inv[k];
    s11 = __builtin_fma(-w, v[k], s11);
    s01 = __builtin_fma(-P.wu[k], v[k], s01);
    wv[k] = w;
  }
  const double s00 = P.s00;
  const double q = s00 * s11;
  const bool sing = P.sing || !(s00 > kPivTol) || !(s11 > kPivTol) ||
                    !(__builtin_fma(-s01, s01, q) > kPivTol * q);
  s11o = s11;
  s01o = s01;
  qo = q;
  return sing;
}

// pcor_schur, then r = s01 / sqrt(q) unless singular.
template <int L>
__host__ __device__ __forceinline__ bool pcor_finish(const PcorPrep<L> &P, double a,
                                                     const double (&y)[L], double &r,
                                                     double &s11o, double &s01o,
                                                     double (&wv)[L]) {
  double q;
  const bool sing = pcor_schur<L>(P, a, y, s11o, s01o, q, wv);
  if (!sing) r = s01o / __builtin_sqrt(q);
  return sing;
}

// removes(s01 / sqrt(q), r_crit) for finite s01 and q > 0, mostly without the
// square root and the division.  With u = 2^-53, R = r_crit,
// lo = fl(fl(R R) (1 - 64 u)) and hi = fl(fl(R R) (1 + 64 u)) (square_bounds,
// for 1e-100 < R < 0.9999999, where the clamp of removes cannot matter):
// fl(s01 s01) <= fl(lo q) gives s01^2 / q <= R^2 (1 - 64 u)(1 + u)^3 / (1 - u)
// <= R^2 (1 - 59 u), so |s01| / sqrt(q) <= R (1 - 29 u) and the computed
// quotient |s01| / fl(sqrt(q)) <= R (1 - 27 u) < R, which rounds to at most R:
// removed.  fl(s01 s01) >= fl(hi q) gives likewise a quotient of at least
// R (1 + 27 u), more than one ulp above R, which rounds above R: kept.  Only
// squares inside that band take the exact quotient, so the decision is
// always that of the exact statistic.  (A subnormal square only errs towards
// zero, where the quotient is far below R.)
__host__ __device__ __forceinline__ bool removes_sq(double s01, double q, double r_crit, double lo,
                                                    double hi) {
  const double s2 = s01 * s01;
  if (s2 <= lo * q) return true;
  if (s2 >= hi * q) return false;
  return removes(s01 / __builtin_sqrt(q), r_crit);
}

// The bounds of removes_sq; outside 1e-100 < r_crit < 0.9999999 they never
// decide (lo q < 0 <= s01^2 < hi q = inf), so the exact quotient does.
__host__ __device__ inline void square_bounds(double r_crit, double &lo, double &hi) {
  if (r_crit > 1e-100 && r_crit < kCutThr) {
    const double r2 = r_crit * r_crit;
    lo = r2 * (1.0 - 0x1p-47);
    hi = r2 * (1.0 + 0x1p-47);
  } else {
    lo = -1.0;
    hi = __builtin_inf();
  }
}

// Partial correlation of i and j given S (|S| = L) from a = r_ij,
// x[c] = r_i,S[c], y[c] = r_j,S[c] and the strictly lower triangle
// C[r][c] = r_S[c],S[r] (r > c).  With C = U D U^T (U unit lower, E = U D),
// u = U^-1 x, v = U^-1 y and w = D^-1 u:
// r = (a - w.v) / sqrt((1 - w.u)(1 - (D^-1 v).v)), the same value as
// -P01/sqrt(P00 P11) for P = inverse of the (L+2)-square submatrix [i, j, S]
// (gpuRowTests.cu:143-170).  One reciprocal per pivot and one square root per
// test.  A pivot at or below 1e-12 or a numerically singular 2x2 Schur block
// switches to the pinv form.
template <int L, bool SENS>
__host__ __device__ __forceinline__ double pcor_canon_impl(double a, const double (&x)[L],
                                                  const double (&y)[L],
                                                  const double (&C)[L][L], bool &used_pinv,
                                                  double *sens) {
  PcorPrep<L> P;
  pcor_prep<L>(x, C, P);
  double gx[L], gy[L], s11, s01, r = 0.0;
  const bool sing = pcor_finish<L>(P, a, y, r, s11, s01, gy);
  used_pinv = false;
  const double s00 = P.s00;
#pragma unroll
  for (int k = 0; k < L; ++k) gx[k] = P.wu[k];
  if (!sing) {
    if (SENS) {
      const double q = s00 * s11;
      const auto &Um = P.Um;
      // rows 0 and 1 of P = M^-1: [S^-1, -S^-1 G^T] with S the 2x2 Schur
      // block and G = C^-1 [x y] = U^-T D^-1 [u v] (back substitution)
#pragma unroll
      for (int c = L - 1; c >= 0; --c) {
#pragma unroll
        for (int k = c + 1; k < L; ++k) {
          gx[c] = __builtin_fma(-Um[k][c], gx[k], gx[c]);
          gy[c] = __builtin_fma(-Um[k][c], gy[k], gy[c]);
        }
      }
      const double id = 1.0 / __builtin_fma(-s01, s01, q);
      const double p00 = s11 * id, p11 = s00 * id, p01 = -s01 * id;
      double rho0 = __builtin_fabs(p00) + __builtin_fabs(p01);
      double rho1 = __builtin_fabs(p01) + __builtin_fabs(p11);
#pragma unroll
      for (int k = 0; k < L; ++k) {
        rho0 += __builtin_fabs(p00 * gx[k] + p01 * gy[k]);
        rho1 += __builtin_fabs(p01 * gx[k] + p11 * gy[k]);
      }
      *sens = rho0 * rho1 / __builtin_sqrt(p00 * p11) +
              __builtin_fabs(r) * 0.5 * (rho0 * rho0 / p00 + rho1 * rho1 / p11);
    }
    return r;
  }
  used_pinv = true;
  if (SENS) *sens = __builtin_inf();
  double A[L + 2][L + 2];
  for (int r = 0; r < L + 2; ++r) A[r][r] = 1.0;
  A[0][1] = A[1][0] = a;
  for (int c = 0; c < L; ++c) {
    A[0][c + 2] = A[c + 2][0] = x[c];
    A[1][c + 2] = A[c + 2][1] = y[c];
    for (int r = c + 1; r < L; ++r) A[r + 2][c + 2] = A[c + 2][r + 2] = C[r][c];
  }
  return pcor_pinv<L + 2>(A);
}

template <int L>
__host__ __device__ __forceinline__ double pcor_canon(double a, const double (&x)[L],
                                             const double (&y)[L],
                                             const double (&C)[L][L],
                                             bool &used_pinv) {
  return pcor_canon_impl<L, false>(a, x, y, C, used_pinv, nullptr);
}

// pcor_canon plus the first-order sensitivity factor of pcor_sensitivity
// (+inf when the pinv form is taken), sharing the factorisation.
template <int L>
__host__ __device__ __forceinline__ double pcor_canon_sens(double a, const double (&x)[L],
                                                  const double (&y)[L],
                                                  const double (&C)[L][L], bool &used_pinv,
                                                  double &sens) {
  return pcor_canon_impl<L, true>(a, x, y, C, used_pinv, &sens);
}

// First-order sensitivity of the partial correlation to its inputs: with
// P = M^-1 for the (L+2)-square matrix M = [i, j, S] (unit diagonal) and
// rho_a the absolute sum of row a of P,
//   |dr| <= eps * (rho_0 rho_1 / sqrt(P00 P11)
//                  + |r| (rho_0^2 / P00 + rho_1^2 / P11) / 2)
// when every off-diagonal entry of M moves by at most eps (dP = -P dM P).
// Returns the factor in parentheses, or +inf when M is numerically singular.
template <int L>
__host__ __device__ __forceinline__ double pcor_sensitivity(double a, const double (&x)[L],
                                                   const double (&y)[L],
                                                   const double (&C)[L][L]) {
  double Um[L][L], Em[L][L], inv[L], u[L], v[L], gx[L], gy[L];
#pragma unroll
  for (int r = 0; r < L; ++r) {
#pragma unroll
    for (int c = 0; c < r; ++c) {
      double e = C[r][c];
#pragma unroll
      for (int k = 0; k < c; ++k) e = __builtin_fma(-Em[r][k], Um[c][k], e);
      Em[r][c] = e;
      Um[r][c] = e * inv[c];
    }
    double d = 1.0;
#pragma unroll
    for (int k = 0; k < r; ++k) d = __builtin_fma(-Em[r][k], Um[r][k], d);
    if (!(d > kPivTol)) return __builtin_inf();
    inv[r] = 1.0 / d;
  }
#pragma unroll
  for (int c = 0; c < L; ++c) {
    double su = x[c], sv = y[c];
#pragma unroll
    for (int k = 0; k < c; ++k) {
      su = __builtin_fma(-Um[c][k], u[k], su);
      sv = __builtin_fma(-Um[c][k], v[k], sv);
    }
    u[c] = su;
    v[c] = sv;
  }
  double s00 = 1.0, s11 = 1.0, s01 = a;
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const double wu = u[k] * inv[k], wv = v[k] * inv[k];
    s00 = __builtin_fma(-wu, u[k], s00);
    s11 = __builtin_fma(-wv, v[k], s11);
    s01 = __builtin_fma(-wu, v[k], s01);
    gx[k] = wu;
    gy[k] = wv;
  }
  // C^-1 x = U^-T D^-1 u (back substitution, in place)
#pragma unroll
  for (int c = L - 1; c >= 0; --c) {
#pragma unroll
    for (int k = c + 1; k < L; ++k) {
      gx[c] = __builtin_fma(-Um[k][c], gx[k], gx[c]);
      gy[c] = __builtin_fma(-Um[k][c], gy[k], gy[c]);
    }
  }
  const double det = __builtin_fma(-s01, s01, s00 * s11);
  if (!(s00 > kPivTol) || !(s11 > kPivTol) || !(det > kPivTol * s00 * s11)) return __builtin_inf();
  const double id = 1.0 / det;
  const double p00 = s11 * id, p11 = s00 * id, p01 = -s01 * id;
  double rho0 = __builtin_fabs(p00) + __builtin_fabs(p01);
  double rho1 = __builtin_fabs(p01) + __builtin_fabs(p11);
#pragma unroll
  for (int k = 0; k < L; ++k) {
    rho0 += __builtin_fabs(p00 * gx[k] + p01 * gy[k]);
    rho1 += __builtin_fabs(p01 * gx[k] + p11 * gy[k]);
  }
  const double r = s01 / __builtin_sqrt(s00 * s11);
  return rho0 * rho1 / __builtin_sqrt(p00 * p11) +
         __builtin_fabs(r) * 0.5 * (rho0 * rho0 / p00 + rho1 * rho1 / p11);
}

// Whether a test on fast correlations (error <= eps per entry) decides the
// same as on exact ones: the statistic is farther from the threshold than
// twice its first-order error bound, and that bound is small enough for the
// first order to hold.
__host__ __device__ __forceinline__ bool fast_certain(double r, double r_crit, double sens, double eps) {
  const double bound = 2.0 * sens * eps;
  if (!(bound < 0.05)) return false;
  const double re = __builtin_isnan(r) ? 0.0 : __builtin_fmin(kCutThr, __builtin_fabs(r));
  return __builtin_fabs(re - r_crit) > bound;
}

// Number of l-subsets of an nn-set, saturating (cuda_util.cu:20-32).
__host__ __device__ __forceinline__ unsigned long long binom(const unsigned long long *tab,
                                                    int kcols, int nn, int l) {
  if (nn < l || l < 0) return 0ull;
  return tab[(size_t)nn * kcols + l];
}

// Largest task count of one row.  A row's tasks, and their prefix sums over
// up to 46340 rows, then stay far below 2^63; a row beyond it would hold more
// than 2^46 tasks of at least 64 tests, more than any run completes.
constexpr long long kMaxRowTasks = 1ll << 46;

// Search tasks of a row of length deg at a level with `total` = C(deg - 1,
// level) conditioning sets per ordered edge (saturated at 2^64 - 1 by the
// binomial table) and rpt ranks per task: deg * ceil(total / rpt).  -1 when
// total saturated or the count exceeds kMaxRowTasks: the level cannot be
// planned.  Monotone in deg, so the longest row decides for all.
__host__ __device__ __forceinline__ long long row_task_count(unsigned long long total, int deg,
                                                             int rpt) {
  if (total == 0ull || deg <= 0) return 0;
  if (total == ~0ull) return -1;
  const unsigned long long per = (total - 1ull) / (unsigned long long)rpt + 1ull;
  if (per > (unsigned long long)kMaxRowTasks / (unsigned long long)deg) return -1;
  return (long long)deg * (long long)per;
}

// ithCombination (cuda_util.cu:34-50), 0-based: positions q[0] < ... < q[L-1]
// of the rank-th L-subset of {0..nn-1} in lexicographic order.
//
// With v0 the smallest free position and k = L - i members still to place,
// the sets whose member i lies in [v0, v] number C(nn-v0, k) - C(nn-1-v, k)
// (hockey-stick identity), so member i is nn-1-m for the largest m with
// C(m, k) < C(nn-v0, k) - rank: a binary search of log2(nn) table reads per
// member instead of a walk over every candidate position.
template <int L>
__host__ __device__ __forceinline__ void unrank(const unsigned long long *tab, int kcols,
                                       unsigned long long rank, int nn, int (&q)[L]) {
  int v0 = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const int k = L - i;
    const unsigned long long top = tab[(size_t)(nn - v0) * kcols + k];
    const unsigned long long X = top - rank;
    int lo = k - 1, hi = nn - 1 - v0;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[(size_t)mid * kcols + k] < X) lo = mid; else hi = mid - 1;
    }
    const int v = nn - 1 - lo;
    q[i] = v;
    rank = rank + tab[(size_t)(lo + 1) * kcols + k] - top;
    v0 = v + 1;
  }
}

// Lexicographic successor of the L-subset q of {0..nn-1}: the rightmost
// member that can grow grows by one and the members after it follow it
// consecutively, so next_comb(unrank(r)) == unrank(r + 1).  The last subset
// is left unchanged.
template <int L>
__host__ __device__ __forceinline__ void next_comb(int (&q)[L], int nn) {
  int piv = -1;
#pragma unroll
  for (int i = L - 1; i >= 0; --i)
    if (piv < 0 && q[i] < nn - L + i) piv = i;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    if (i == piv) q[i] += 1;
    else if (piv >= 0 && i > piv) q[i] = q[i - 1] + 1;
  }
}

}  // namespace hpc
