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
 *
 * Sepset list (the compact form of sepSets, no counterpart in the
 * reference): one record per edge removed at a level l >= 1, l + 3 int32
 * words: i, j (i < j), l, then the l members in ascending order; records
 * sorted by (i, j).  Absent pairs without a record were removed at level 0.
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
  HPC_CORR_CANON = 0, /* FP64, the reference's summation order (bit-exact) */
  HPC_CORR_F32 = 1    /* z-scores rounded to float, products on MFMA with FP32
                         accumulation over 64-sample chunks and FP64 across
                         chunks; |cor - exact| <= 74 * 2^-24 * 1.01.  Every test
                         whose statistic lies within its first-order error bound
                         of the threshold (every level-0 pair within that bound)
                         is re-run on exact FP64 correlations, so the adjacency,
                         sepsets and test counts equal HPC_CORR_CANON's.
                         Instances with NaN cells use the exact kernels. */
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
  int64_t rechecks;      /* tests re-run on exact correlations (HPC_CORR_F32) */
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
  HPC_TUNE_RANKS_PER_TASK = 1, /* conditioning sets per search task at levels
                                  >= 2: a multiple of 64, default 512 */
  HPC_TUNE_RECHECK_QUEUE = 2,  /* HPC_CORR_F32: length of the exact-recheck
                                  queues (0 = 1 Mi); a level whose queue
                                  overflows is redone on exact correlations */
  HPC_TUNE_KERNEL_TIMING = 3,  /* 1 (default): hpc_run records device-time
                                  events around each level's kernels (the
                                  ms_compact / ms_test / ms_commit fields);
                                  0: no events (each marker between two
                                  kernels delays the second by a few
                                  microseconds) and those fields stay 0 */
  HPC_TUNE_LDS_LIMIT = 4,      /* bytes of LDS per workgroup for staging a
                                  row's neighbourhood in the CI-test kernels,
                                  1..65536 (0 = 65536, the default).  A row
                                  whose staging needs more is searched by a
                                  gather kernel that reads the correlation
                                  matrix; results do not depend on it */
  HPC_TUNE_WAIT_LIMIT = 5,     /* seconds the host waits for the device's
                                  counts of a level before the call fails with
                                  HPC_EHIP (0 = no limit; default 900) */
  HPC_TUNE_QUEUE = 6,          /* 1 (default): the level >= 2 CI-test kernel
                                  is persistent and its workgroups claim
                                  blocks of tasks from a device queue; 0:
                                  blocks go to workgroups by id.  Results do
                                  not depend on it */
  HPC_TUNE_GRAPH = 7           /* 1 (default): hpc_run on a freshly reset state
                                  (hpc_reset) with kernel timing off, with a
                                  level cap or none, records its launches and
                                  captures them in a hipGraph; later such runs
                                  replay it with one host wait, checking every
                                  level's sizes against the recorded run (a
                                  difference recomputes the run without the
                                  graph).  0: never */
};
int hpc_set_tuning(hpc_state *st, int param, int value);
/* Level-loop graph of a state (HPC_TUNE_GRAPH): runs replayed, graphs
 * captured, and replays whose sizes differed from the recorded run (and
 * were recomputed). */
int hpc_graph_stats(hpc_state *st, int64_t *replays, int64_t *captures, int64_t *mismatches);

/* Autotuner (no counterpart in the reference): runs the state's levels
 * 1..max_level once, timing each level's CI-test kernel on that level's
 * snapshot for every candidate launch shape -- level 1: ordered edges per
 * workgroup in {8..256}; levels >= 2: conditioning sets per task in
 * {256..4096} x tasks per workgroup in {4..128} x task order (device queue
 * or workgroup id), and with each conditioning-set size one contiguous
 * equal share of the tasks per resident workgroup -- and keeps the fastest per
 * level for this state's later runs.  Results do not depend on the shape.
 * The state is left reset (adjacency and sepsets back to the start). */
typedef struct hpc_tune_entry {
  int32_t level;
  int32_t ranks_per_task;    /* levels >= 2 */
  int32_t tasks_per_block;   /* levels >= 2 */
  int32_t edges_per_block;   /* level 1 */
  int32_t queue;             /* levels >= 2: 1 = blocks claimed from the device
                                task queue, 2 = blocks by workgroup id, 3 = one
                                contiguous share of the tasks per resident
                                workgroup (exact instances; the others take
                                blocks by workgroup id), 0 = the state's
                                HPC_TUNE_QUEUE */
  int32_t pad_;
  double ms_best;            /* CI-test kernel time with the chosen shape */
  double ms_default;         /* with the default shape */
} hpc_tune_entry;
int hpc_autotune(hpc_state *st, int max_level, hpc_tune_entry *out, int cap,
                 int *levels);
/* Launch shapes chosen by hpc_autotune (its entries, any subset of levels)
 * for this state's later runs; no counterpart in the reference.  Every entry
 * is checked before any is applied (level 1: edges_per_block in 1..4096;
 * levels >= 2: ranks_per_task a multiple of 64 in [64, 2^24] and
 * tasks_per_block in 1..1024); results do not depend on the shapes. */
int hpc_apply_tuning(hpc_state *st, const hpc_tune_entry *entries, int count);

/* ---- correlation (src/correlation/corOwn.cu:10-48) ---------------------- */
/* gpuPMCC(h_mat, p, n, state.cor, device): uploads the variable-major
 * observations and computes the correlation matrix on the device. */
int hpc_pmcc(hpc_state *st, const double *h_mat, int mode);
/* The two halves of hpc_pmcc: upload once, recompute from device-resident
 * observations.  hpc_set_observations returns once h_mat has been read (the
 * caller may reuse it); the copy to the device ends, in stream order, before
 * any later work of the state.  Values that FP32 holds exactly cross the bus
 * at half width; the device receives the same doubles either way. */
int hpc_set_observations(hpc_state *st, const double *h_mat);
int hpc_corr(hpc_state *st, int mode);
/* Device time of the last hpc_corr: moments (mean, sd, z-scores) and the
 * pairwise sums (waits for them to finish). */
int hpc_corr_ms(hpc_state *st, double *ms_moments, double *ms_corr);
/* hpc_corr's exact (HPC_CORR_CANON) correlation without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  x: device observations [p][n] (as hpc_set_observations
 * takes them on the host); cor: device output, p x p row-major, both
 * triangles, bit-identical to hpc_corr's.  The kernels' working memory (means,
 * standard deviations, z-scores, NaN mask, instance table) is one static
 * 64 MiB device array per device, so 9 p ceil(n/64) 64 + 20 p bytes plus a
 * few KiB must fit it (p = 1000: n <= 7400); larger shapes fail with
 * HPC_EINVAL.  Queues its kernels on `stream` (a hipStream_t; NULL = the
 * default stream) and returns without waiting.  Calls that overlap on one
 * device share the array and race; calls in one stream's order do not. */
int hpc_corr_noalloc(int device, const double *x, uint64_t p, int n, double *cor,
                     void *stream);
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
/* The work-stealing level of Balancer::executeWorkstealing
 * (src/loadbalance/balancer.cpp:186-234) on one GPU: the level's search tasks
 * (ordered edge, range of conditioning-set ranks) form one queue.  A
 * persistent device kernel claims blocks of tasks from its front while
 * `threads` host threads (all cores when <= 0) take blocks from its back, as
 * the CPU executor claims rows longest first beside the GPU
 * (src/executor/cpuExecutor.cpp:15-78, gpuWorkstealingTests.cu:21-39); each
 * side skips blocks the other has finished and the host stops when the
 * device has claimed everything.  The host's first passing ranks are merged
 * with the device's (lowest wins), so the result equals a device-only
 * search.  Levels 1..16, exact correlation, after hpc_compact;
 * hpc_commit_level then removes the edges.  host_stats receives the host's
 * share (edges and tests it searched; `removed` holds the blocks it
 * finished), device_stats the level's search time. */
int hpc_search_level_shared(hpc_state *st, int level, int threads,
                            hpc_level_stats *device_stats, hpc_level_stats *host_stats);
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
 * record per level; *levels_run the number of levels that ran.  With no
 * cap, the level found to have no testable row ends the loop after its
 * compaction and gets no record. */
int hpc_run(hpc_state *st, int max_level, hpc_level_stats *stats, int cap,
            int *levels_run);

/* ---- results in the reference layouts ----------------------------------- */
int hpc_get_cor(hpc_state *st, double *cor);               /* p*p */
int hpc_get_adj(hpc_state *st, int32_t *adj);              /* p*p */
int hpc_get_compact(hpc_state *st, int32_t *adj_compact);  /* p*p */
/* sepSets with width mc (pass hpc_max_cond_size(st) for the full width). */
int hpc_get_sepsets(hpc_state *st, int32_t *sep, int mc);
/* Sepset list; *words receives its length (out may be NULL to query it). */
int hpc_get_sepset_list(hpc_state *st, int32_t *out, int64_t cap, int64_t *words);
/* First passing rank per ordered edge of the last level (CSR order, -1 =
 * none), rowptr[p+1]; *n_edges receives the CSR size (fp may be NULL to
 * query it). */
int hpc_get_first_pass(hpc_state *st, int64_t *fp, int32_t *rowptr,
                       int64_t *n_edges);

/* ---- batched skeletons --------------------------------------------------
 * No counterpart in the reference, which runs one dataset per process
 * (src/main.cpp:57-151).  A batch holds many independent instances on one
 * device; each step runs the next level of every unfinished instance in one
 * set of launches (level 0, compaction, the CI-test kernels of every level
 * present and the removal kernel), whatever their p, n, alpha and level.
 * Per instance the results equal a single hpc_state run on the same input:
 * the arithmetic of a test and the sepset choice do not depend on the other
 * instances.  Instance ids are small non-negative ints, reused after
 * hpc_batch_remove. */
typedef struct hpc_batch hpc_batch;

/* Device bytes one instance of (p, n, max_level) needs at most. */
int hpc_footprint(uint64_t p, int n, int max_level, uint64_t *bytes);
/* Whether instances with these p can share one batch or server step (host
 * only): HPC_EINVAL for a p outside [2, 46340], HPC_ENOMEM when their ordered
 * edges, the sum of p (p - 1), exceed 2^31 - 1 (one step indexes them with
 * int32).  hpc_batch_add refuses an instance that breaks it; the server
 * admits a request only while it holds. */
int hpc_pack_check(const uint64_t *p, int count);
int hpc_batch_create(int device, hpc_batch **out);
int hpc_batch_destroy(hpc_batch *b);
/* New instance on observations x ([p][n], as hpc_set_observations); max_level
 * as the state's (-1 = until no row has enough neighbours). */
int hpc_batch_add(hpc_batch *b, uint64_t p, int n, double alpha, int max_level,
                  const double *x, int *id);
int hpc_batch_remove(hpc_batch *b, int id);
/* One level of every unfinished instance (correlation first for those that
 * have none); *unfinished receives how many remain. */
int hpc_batch_step(hpc_batch *b, int *unfinished);
/* Steps until every instance has finished; *steps receives the count. */
int hpc_batch_run(hpc_batch *b, int *steps);
int hpc_batch_status(hpc_batch *b, int id, int *finished, int *levels, int *error);
int hpc_batch_get_stats(hpc_batch *b, int id, hpc_level_stats *stats, int cap,
                        int *levels);
int hpc_batch_max_cond_size(hpc_batch *b, int id);
int hpc_batch_get_cor(hpc_batch *b, int id, double *cor);
int hpc_batch_get_adj(hpc_batch *b, int id, int32_t *adj);
int hpc_batch_get_sepsets(hpc_batch *b, int id, int32_t *sep, int mc);
int hpc_batch_get_sepset_list(hpc_batch *b, int id, int32_t *out, int64_t cap,
                              int64_t *words);
/* HPC_TUNE_RANKS_PER_TASK, HPC_TUNE_LDS_LIMIT or HPC_TUNE_WAIT_LIMIT for
 * every instance. */
int hpc_batch_set_tuning(hpc_batch *b, int param, int value);
/* hpc_apply_tuning for every instance of the batch. */
int hpc_batch_apply_tuning(hpc_batch *b, const hpc_tune_entry *entries, int count);
/* Autotuner of a batch (no counterpart in the reference): runs the batch's
 * unfinished instances once from their start, timing each level's CI-test
 * kernel with every candidate shape (those of hpc_autotune) on the first step
 * that meets the level -- the step's own mix of instances -- and keeps the
 * fastest per level for the batch's later steps; out[0..cap) receives them
 * and *levels their number.  Every instance is left at its start with its
 * correlation.  Exact instances only; results do not depend on the shapes. */
int hpc_batch_autotune(hpc_batch *b, hpc_tune_entry *out, int cap, int *levels);
/* Correlation mode (HPC_CORR_*) of an instance, before its first step. */
int hpc_batch_set_corr_mode(hpc_batch *b, int id, int mode);

/* ---- continuous-batching server -------------------------------------------
 * No counterpart in the reference.  A worker thread admits submitted requests
 * in arrival order while the sum of their hpc_footprint stays within the HBM
 * budget (and at most max_active run), steps every admitted request together
 * and retires each as soon as it finishes; results are copied to host memory
 * and the request's device blocks freed.  A request's result is the same as
 * a single hpc_state run on its input. */
typedef struct hpc_server hpc_server;

typedef struct hpc_request_info {
  int32_t p;
  int32_t levels;         /* levels run */
  int32_t max_cond_size;  /* sepset width held */
  int32_t pad_;
  int64_t tests;          /* canonical CI tests over all levels */
  double queue_ms;        /* submit -> admission */
  double latency_ms;      /* submit -> result ready */
} hpc_request_info;

typedef struct hpc_server_info {
  int64_t steps;          /* batch steps run */
  int64_t peak_active;    /* most requests admitted at once */
  uint64_t peak_bytes;    /* largest admitted footprint */
  uint64_t budget;        /* HBM budget in bytes */
  double busy_ms;         /* wall time spent in steps */
  int32_t queued, running;
} hpc_server_info;

/* hbm_budget 0 = 80% of the device's free memory; max_active <= 0 = 1024. */
int hpc_server_create(int device, uint64_t hbm_budget, int max_active,
                      hpc_server **out);
/* Stops the worker; queued requests are dropped, running ones finish first. */
int hpc_server_destroy(hpc_server *s);
/* Copies x; *req receives the request id.  HPC_ENOMEM when the request alone
 * exceeds the budget. */
int hpc_server_submit(hpc_server *s, uint64_t p, int n, double alpha,
                      int max_level, const double *x, int64_t *req);
/* As hpc_server_submit without the copy: x is read in place, so it must stay
 * valid and unchanged until hpc_server_wait reports the request done. */
int hpc_server_submit_borrowed(hpc_server *s, uint64_t p, int n, double alpha,
                               int max_level, const double *x, int64_t *req);
/* Waits up to timeout_s seconds (< 0: no limit); *done = 1 when finished. */
int hpc_server_wait(hpc_server *s, int64_t req, double timeout_s, int *done);
/* Adjacency (p*p) and sepsets (width mc) of a finished request; either may be
 * NULL.  Returns the request's error when it failed. */
int hpc_server_result(hpc_server *s, int64_t req, int32_t *adj, int32_t *sep,
                      int mc, hpc_request_info *info);
/* Sepset list of a finished request (out may be NULL to query *words). */
int hpc_server_sepset_list(hpc_server *s, int64_t req, int32_t *out, int64_t cap,
                           int64_t *words);
/* Correlation mode (HPC_CORR_*) of the requests submitted from now on. */
int hpc_server_set_corr_mode(hpc_server *s, int mode);
/* hpc_apply_tuning for the server's later steps (taken up by the worker
 * before its next step). */
int hpc_server_apply_tuning(hpc_server *s, const hpc_tune_entry *entries, int count);
/* HPC_TUNE_RANKS_PER_TASK, HPC_TUNE_LDS_LIMIT or HPC_TUNE_WAIT_LIMIT for the
 * server's later steps (taken up by the worker before its next step). */
int hpc_server_set_tuning(hpc_server *s, int param, int value);
/* Drops a finished request's results. */
int hpc_server_release(hpc_server *s, int64_t req);
int hpc_server_stats(hpc_server *s, hpc_server_info *info);

/* ---- fault injection (tests of failure isolation) -------------------------
 * No counterpart in the reference.  The host side of the instance's (or the
 * request's) given level fails with HPC_ENOMEM, as an allocation failure
 * would, before any of its kernels are queued; device work is untouched.
 * The instance stops with that error (hpc_batch_status, hpc_server_result)
 * while the others run on.  level < 0 clears it. */
int hpc_batch_inject_failure(hpc_batch *b, int id, int level);
/* For a request not yet admitted, or one not yet submitted: request ids
 * count up from 1 in submission order. */
int hpc_server_inject_failure(hpc_server *s, int64_t req, int level);
/* The search plan of one row (host only): *tasks receives the tasks a row of
 * deg neighbours gets at `level` with ranks_per_task conditioning sets per
 * task, deg * ceil(C(deg - 1, level) / ranks_per_task), from the same
 * saturating binomial table and overflow-checked arithmetic the device
 * plans with; HPC_ELEVEL when C(deg - 1, level) does not fit 64 bits or the
 * count exceeds 2^46 (such a level fails with HPC_ELEVEL). */
int hpc_plan_tasks(uint64_t deg, int level, int ranks_per_task, int64_t *tasks);
/* The host wait for device counts on its own (host only): spins until *word
 * equals value, failing with HPC_EHIP after limit_s seconds (<= 0: no
 * limit), as a level's wait does under HPC_TUNE_WAIT_LIMIT. */
int hpc_wait_word(const uint64_t *word, uint64_t value, double limit_s);

#ifdef __cplusplus
}
#endif
#endif
