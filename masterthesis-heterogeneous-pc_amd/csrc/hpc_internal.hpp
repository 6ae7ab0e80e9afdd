// Internal declarations shared by the HIP translation units of
// libheterogpc_hip.so.  Not part of the C ABI (include/heterogpc_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/heterogpc_hip.h"

namespace hpc {

// Device-wide counters one level writes; read back by the host once per level.
struct Counters {
  unsigned long long tests;      // canonical test count T_l
  unsigned long long tests_max;  // U_l
  unsigned long long edges;      // ordered edges searched
  unsigned long long removed;    // undirected edges removed
  unsigned long long n_edges;    // ordered edges at level start (CSR size)
  int max_deg;                   // largest row length at level start
  int pad_;
  unsigned long long pinv_fallbacks;  // tests that took the pinv form
  unsigned long long n_tasks;         // search tasks of a level >= 2
};

// Largest level the device kernels accept (conditioning-set size).
constexpr int kMaxLevel = 16;
// Binomial table columns: C(n, k) for k <= kMaxLevel.
constexpr int kBinomK = kMaxLevel + 1;

struct Views {
  // read-only inputs of the test kernels
  const double *cor;        // p x p
  const int32_t *rowptr;    // p + 1
  const int32_t *nbr;       // CSR neighbour ids (level-start snapshot)
  const int32_t *erow;      // row of every ordered edge
  const unsigned long long *binom;  // (p + 1) x kBinomK, saturating
  int64_t *fp;              // first passing rank per ordered edge, -1 = none
  Counters *ctr;
  int p;
  int max_deg;              // largest row length of the snapshot
  double r_crit;
};

// ---- launchers (each defined in the .hip file that owns the kernel) --------
void launch_moments(hipStream_t s, const double *x, int p, int n, int npad,
                    double *mean, double *sd, double *z, uint8_t *nanm,
                    int32_t *hasnan);
void launch_corr_canon(hipStream_t s, const double *z, const uint8_t *nanm,
                       const int32_t *hasnan, int p, int n, int npad,
                       double *cor);
void launch_corr_tiled(hipStream_t s, const double *z, int p, int n, int npad,
                       double *cor);
void launch_adj_init(hipStream_t s, unsigned long long *adj, int p, int W);
void launch_level0(hipStream_t s, const double *cor, unsigned long long *adj,
                   int p, int W, double r_crit, Counters *ctr);
void launch_compact(hipStream_t s, const unsigned long long *adj, int p, int W,
                    int32_t *deg, int32_t *rowptr, int32_t *nbr, int32_t *erow,
                    Counters *ctr);
void launch_search_l1(hipStream_t s, const Views &v, int64_t n_edges,
                      const int32_t *row_mask);
// Levels >= 2: per-row task counts and their scan (total to ctr->n_tasks),
// then the task list, then the search over it.
void launch_task_plan(hipStream_t s, int level, const Views &v, int rpt,
                      const int32_t *row_mask, long long *row_tasks, long long *task_off);
void launch_task_fill(hipStream_t s, int level, const Views &v, int rpt,
                      const long long *row_tasks, const long long *task_off,
                      unsigned long long *tasks);
void launch_search_ln(hipStream_t s, int level, const Views &v,
                      const unsigned long long *tasks, long long n_tasks, int rpt);
void launch_commit(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   unsigned long long *adj, int W, int32_t *sep, int8_t *rmlev);

}  // namespace hpc

// One skeleton instance (include/heterogpc_hip.h).
struct hpc_state {
  int device = 0;
  hipStream_t stream = nullptr;
  int p = 0, n = 0, npad = 0, W = 0;
  int max_level = 0, mc = 1;
  double alpha = 0.0, r_crit = 0.0;
  // observations and correlation
  double *d_x = nullptr, *d_z = nullptr, *d_mean = nullptr, *d_sd = nullptr;
  uint8_t *d_nanm = nullptr;
  int32_t *d_hasnan = nullptr;
  double *d_cor = nullptr;
  bool have_x = false, have_cor = false, x_has_nan = false;
  // graph
  unsigned long long *d_adj = nullptr;
  int32_t *d_deg = nullptr, *d_rowptr = nullptr, *d_nbr = nullptr, *d_erow = nullptr;
  int32_t *d_rowmask = nullptr;
  int64_t *d_fp = nullptr;
  int32_t *d_sep = nullptr;  // [mc][p][p]
  int8_t *d_rmlev = nullptr;
  unsigned long long *d_binom = nullptr;
  hpc::Counters *d_ctr = nullptr;
  hpc::Counters *h_ctr = nullptr;  // pinned
  // snapshot bookkeeping
  bool compacted = false;     // neighbour lists match the adjacency
  bool ever_compacted = false;
  int64_t n_edges = 0;        // ordered edges in the snapshot
  int max_deg = 0;
  int64_t undirected = 0;     // edges currently present
  hipEvent_t ev[7] = {};
  bool corr_timed = false;
  // host share of a level (CPU executor): a second stream for its copies and
  // host copies of the snapshot and correlation
  hipStream_t copy_stream = nullptr;
  std::vector<double> h_cor;
  bool h_cor_valid = false;
  std::vector<int32_t> h_rowptr, h_nbr;
  std::vector<unsigned long long> h_binom;
  // level >= 2 search tasks
  long long *d_row_tasks = nullptr, *d_task_off = nullptr;
  unsigned long long *d_tasks = nullptr;
  size_t task_cap = 0;
  int ranks_per_task = 512;
  int64_t *d_scatter_idx = nullptr;
  int64_t *d_scatter_val = nullptr;
  size_t scatter_cap = 0;
};

// Error reporting shared by the ABI translation units.
int hpc_fail(int code, const char *fmt, ...);
