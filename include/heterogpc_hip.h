/*
 * heterogpc_hip.h -- C ABI of libheterogpc_hip.so, the MI355X (gfx950)
 * PC-stable skeleton path of heterogpc.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative HPC_E* code; hpc_last_error() then describes the failure (per
 * calling thread).  One hpc_state owns one skeleton instance: its device
 * buffers and one HIP stream.  Calls on one state must not overlap; different
 * states may be driven from different threads.
 *
 * Host arrays use the reference's layouts (src/util/state.cu:3-44):
 *   observations  variable-major [p][n] doubles (an arma::mat n x p,
 *                 column-major, as gpuPMCC receives it, corOwn.cu:10)
 *   cor, adj      row-major p x p
 *   adj_compact   row-major p x p: ascending neighbour ids, zero padding,
 *                 row length at [i][p-1] (compact.cu:57-88)
 *   sepSets       p x p x mc ints at ((min*p + max)*mc + k); -1 empty or end,
 *                 -2 removed at level 0 (cpuUtil.cpp:3-48)
 */
#ifndef HETEROGPC_HIP_H
#define HETEROGPC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HPC_OK = 0,
  HPC_EINVAL = -1,   /* bad argument */
  HPC_EHIP = -2,     /* HIP runtime error (message has the HIP string) */
  HPC_ENOMEM = -3,   /* device or host allocation failed */
  HPC_ESTATE = -4,   /* call out of order (e.g. no correlation yet) */
  HPC_ELEVEL = -5    /* level beyond what the device kernels support */
};

/* Correlation modes for hpc_corr / hpc_pmcc. */
enum {
  HPC_CORR_CANON = 0 /* FP64, the reference's summation order (bit-exact) */
};

typedef struct hpc_state hpc_state;

/* Per-level record (replaces TestResult, src/util/indep_util.hpp:4-16, whose
 * test count the reference never filled: "Test count not implemented yet"). */
typedef struct hpc_level_stats {
  int32_t level;
  int32_t max_deg;       /* largest row length at level start (max_adj) */
  int64_t tests;         /* canonical CI tests: per ordered edge, rank of the
                            first passing set + 1, else all its sets */
  int64_t tests_max;     /* all conditioning sets of the level */
  int64_t edges;         /* ordered edges searched */
  int64_t removed;       /* undirected edges removed by this level */
  int64_t edges_left;    /* undirected edges after this level */
  int64_t pinv_fallbacks;/* tests resolved through the pseudoinverse form */
  double ms_compact;     /* device time of the compaction kernels */
  double ms_test;        /* device time of the CI-test kernel(s) */
  double ms_commit;      /* device time of the removal kernel */
  double ms_wall;        /* host wall time of the whole level */
} hpc_level_stats;

/* ---- runtime ----------------------------------------------------------- */
const char *hpc_last_error(void);
/* getDeviceCount (src/util/cuda_util.cuh:58-62). */
int hpc_device_count(int *count);
int hpc_device_synchronize(int device);

/* Largest |r| (in [0, 0.9999999]) for which the Fisher-z p-value of
 * calcPValue (src/executor/independence/cpuInd.cpp:8-14) is still >= alpha;
 * +inf when every test removes, -1 when none does.  Host only. */
double hpc_fisher_r_crit(int n, double alpha);

/* Linear-Gaussian SEM sample (host only; the reference ships no generator):
 * random DAG on 0..p-1 in topological order, edge i->j (i<j) with
 * probability degree/(p-1), weights +-U[wmin,wmax], N(0,1) noise, n samples,
 * written variable-major to x[p][n]; fp32 != 0 rounds every value to float.
 * dag (optional, p*p bytes) receives the true skeleton (symmetric 0/1). */
int hpc_sem_generate(int p, int n, double degree, double wmin, double wmax,
                     uint64_t seed, int fp32, double *x, uint8_t *dag);

/* ---- state (MMState, src/util/state.cuh:29-55, state.cu:3-44) ----------- */
/* maxCondSize = max(max_level, 1) (state.cu:7); max_level < 0 means levels
 * run until no row has enough neighbours and the sepset width grows as
 * levels are reached. */
int hpc_state_create(uint64_t p, int observations, double alpha, int max_level,
                     int device, hpc_state **out);
/* MMState::destroy (state.cu:96-117). */
int hpc_state_destroy(hpc_state *st);
/* Back to the initial adjacency (all pairs) with no sepsets; keeps the
 * correlation matrix. */
int hpc_reset(hpc_state *st);
int hpc_sync(hpc_state *st);
/* Current maxCondSize (sepset width). */
int hpc_max_cond_size(const hpc_state *st);

/* Launch-shape knobs (set by the autotuner; results do not depend on them). */
enum {
  HPC_TUNE_RANKS_PER_TASK = 1 /* conditioning sets per search task at levels
                                 >= 2: a multiple of 64, default 512 */
};
int hpc_set_tuning(hpc_state *st, int param, int value);

/* ---- correlation (src/correlation/corOwn.cu:10-48) ---------------------- */
/* gpuPMCC(h_mat, p, n, state.cor, device): uploads the variable-major
 * observations and computes the correlation matrix on the device. */
int hpc_pmcc(hpc_state *st, const double *h_mat, int mode);
/* The two halves of hpc_pmcc: upload once, recompute from device-resident
 * observations. */
int hpc_set_observations(hpc_state *st, const double *h_mat);
int hpc_corr(hpc_state *st, int mode);
/* Device time of the last hpc_corr: moments (mean, sd, z-scores) and the
 * pairwise sums (waits for them to finish). */
int hpc_corr_ms(hpc_state *st, double *ms_moments, double *ms_corr);
/* add_correlation_matrix (src/independence/skeleton.cpp:60-65): a given
 * p x p correlation matrix. */
int hpc_set_corr(hpc_state *st, const double *h_cor);

/* ---- levels ------------------------------------------------------------- */
/* callCompact (src/independence/compact.cuh:7-8, compact.cu:91-99): builds
 * the level-start neighbour lists; *max_deg receives max_adj. */
int hpc_compact(hpc_state *st, int *max_deg);
/* The kernel switch of GPUExecutor::executeLevel (src/executor/
 * gpuExecutor.cu:51-106): searches every ordered edge of rows[0..nrows) (all
 * rows when rows == NULL), then removes edges and records sepsets.  Level 0
 * tests all adjacent pairs.  Requires hpc_compact first for level >= 1. */
int hpc_test_level(hpc_state *st, int level, const int32_t *rows, int nrows,
                   hpc_level_stats *stats);
/* The two halves of hpc_test_level, for a level split between executors
 * (Balancer::execute, src/loadbalance/balancer.cpp:138-184):
 * hpc_search_level runs the device search of rows[0..nrows) (all rows when
 * rows == NULL) and records each ordered edge's first passing rank; level 0
 * also removes its edges here.  hpc_commit_level then removes the edges and
 * records the sepsets of every ordered edge searched since hpc_compact. */
int hpc_search_level(hpc_state *st, int level, const int32_t *rows, int nrows,
                     hpc_level_stats *stats);
int hpc_commit_level(hpc_state *st, int level, hpc_level_stats *stats);
/* CPUExecutor::executeLevel (src/executor/cpuExecutor.cpp:80-119): the
 * host's share of a level >= 1 -- rows[0..nrows) searched on `threads` host
 * threads (all cores when <= 0) with the device's arithmetic, ranks written
 * to the device.  May run concurrently with hpc_search_level on the same
 * state when the row sets are disjoint; hpc_commit_level after both. */
int hpc_search_level_host(hpc_state *st, int level, const int32_t *rows,
                          int nrows, int threads, hpc_level_stats *stats);
/* SkeletonCalculator::calcLevel (src/independence/skeleton.cpp:104-150):
 * compaction (level >= 1) followed by hpc_test_level on all rows. */
int hpc_level(hpc_state *st, int level, hpc_level_stats *stats);
/* SkeletonCalculator::run's level loop (skeleton.cpp:172-176): levels
 * 0..max_level (the state's max_level when max_level == -2; until no row
 * has enough neighbours when max_level == -1).  stats[0..cap) receives one
 * record per level; *levels_run the number of levels that ran. */
int hpc_run(hpc_state *st, int max_level, hpc_level_stats *stats, int cap,
            int *levels_run);

/* ---- results in the reference layouts ----------------------------------- */
int hpc_get_cor(hpc_state *st, double *cor);               /* p*p */
int hpc_get_adj(hpc_state *st, int32_t *adj);              /* p*p */
int hpc_get_compact(hpc_state *st, int32_t *adj_compact);  /* p*p */
/* sepSets with width mc (pass hpc_max_cond_size(st) for the full width). */
int hpc_get_sepsets(hpc_state *st, int32_t *sep, int mc);
/* First passing rank per ordered edge of the last level (CSR order, -1 =
 * none), rowptr[p+1]; *n_edges receives the CSR size (fp may be NULL to
 * query it). */
int hpc_get_first_pass(hpc_state *st, int64_t *fp, int32_t *rowptr,
                       int64_t *n_edges);

#ifdef __cplusplus
}
#endif
#endif
