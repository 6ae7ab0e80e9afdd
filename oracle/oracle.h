/*
 * oracle.h -- CPU restatement of heterogpc's PC-stable skeleton path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load liboracle.so, and only as the checker
 * (or as the timed CPU baseline).  The HIP library never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Parity pins (what independent fact each function is checked against) are
 * listed in oracle.c next to each function.
 */
#ifndef HETEROGPC_ORACLE_H
#define HETEROGPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar statistics ------------------------------------------------- */
double orc_pval(double r, int n);                          /* cpuInd.cpp:8-14   */
double orc_pcor_l1(double x1, double x2, double x3);       /* cpuInd.cpp:16-22  */
double orc_pval_l1(double x1, double x2, double x3, int n);/* cpuInd.cpp:16-27  */
/* Partial correlation of variables 0 and 1 of the m x m unit-diagonal matrix
 * M (row-major) given the rest.  canon: Cholesky/Schur order shared with the
 * device (falls back to the pinv form on a non-positive pivot); pinv: the
 * -P01/sqrt(P00 P11) form of cpuInd.cpp:35-36 with P = pinv(M) and
 * Armadillo's default tolerance. */
double orc_pcor_canon(int m, const double *M, int *used_pinv);
double orc_pcor_pinv(int m, const double *M);
/* cpuInd.cpp:29-38: NaN anywhere -> NaN p-value (edge kept).  mode 0 canon,
 * mode 1 pinv. */
double orc_pval_ln(int m, const double *M, int n, int mode);
/* Decision p >= alpha (gpuRowTests.cu:18, cpuRowTests.cpp:12); NaN -> keep. */
int orc_removes(double pval, double alpha);

/* ---- combinatorics (cuda_util.cu:20-50) -------------------------------- */
uint64_t orc_binom(int n, int k);                 /* saturates at UINT64_MAX */
void orc_unrank(uint64_t rank, int l, int nn, int *q);   /* 0-based, lexicographic */

/* ---- correlation (corHelper.cu:8-116) ---------------------------------- */
/* x: variable-major [p][n].  Writes mean[p], sd[p] (sd rounded through float,
 * corHelper.cu:75) and the full p x p correlation, both triangles, in the
 * reference's summation order (64 lane-strided FMA partial sums, then a
 * halving tree). */
void orc_moments(const double *x, int p, int n, double *mean, double *sd);
void orc_corr(const double *x, int p, int n, double *cor, int nthreads);

/* ---- skeleton ----------------------------------------------------------- */
typedef struct orc_state orc_state;

enum { ORC_MODE_CANON = 0, ORC_MODE_PINV = 1 };
/* flags */
enum {
  ORC_FLAG_REF_EARLY_EXIT = 1, /* delete immediately and stop a search once
                                  the live adjacency lost the edge, as the
                                  reference CPU executor does (timing mode;
                                  sepset winner then depends on scheduling) */
};

orc_state *orc_new(const double *cor, int p, int n, double alpha, int mc,
                   int mode, int flags, int nthreads);
void orc_free(orc_state *st);
/* One level as SkeletonCalculator::calcLevel does it (skeleton.cpp:104-150):
 * compaction for level >= 1, then every ordered edge's search.  stats[0] =
 * canonical test count T_l, stats[1] = U_l (all sets), stats[2] = ordered
 * edges searched, stats[3] = edges removed, stats[4] = max row length.
 * Returns 0, or 1 when no row has enough neighbours (nothing tested). */
int orc_level(orc_state *st, int level, int64_t *stats);
/* Runs levels 0..max_level (max_level < 0: until no row has enough
 * neighbours).  Returns the number of levels that ran; stats is
 * [levels][5]. */
int orc_run(orc_state *st, int max_level, int64_t *stats, int stats_rows);
void orc_get_adj(const orc_state *st, int32_t *adj);          /* p*p */
void orc_get_compact(const orc_state *st, int32_t *compact);  /* p*p, compact.cu layout */
void orc_get_sepsets(const orc_state *st, int32_t *sep);      /* p*p*mc, state.cu layout */
/* First passing rank per ordered edge of the last level, CSR order
 * (rows ascending, neighbours ascending); -1 when none passed. */
int64_t orc_num_ordered_edges(const orc_state *st);
void orc_get_first_pass(const orc_state *st, int64_t *fp, int32_t *rowptr);
int orc_mc(const orc_state *st);

/* p-value of the test "i _||_ j | S" in the submatrix order [i, j, S]
 * (gpuRowTests.cu:143-170); |S| = 0 uses the level-0 form, |S| = 1 the
 * closed form. */
double orc_pval_ids(const double *cor, int p, int n, int i, int j,
                    const int32_t *S, int l, int mode);

/* ---- exact decision boundary of orc_pval ------------------------------- */
/* Largest double r in [0, 0.9999999] with orc_pval(r, n) >= alpha, or +inf
 * if even r = 0.9999999 is removed. */
double orc_r_crit(int n, double alpha);

#ifdef __cplusplus
}
#endif
#endif
