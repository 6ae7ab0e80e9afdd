// Internal declarations shared by the HIP translation units of
// libheterogpc_hip.so.  Not part of the C ABI (include/heterogpc_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

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
void launch_search(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   const int32_t *row_mask);
void launch_commit(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   unsigned long long *adj, int W, int32_t *sep, int mc,
                   int8_t *rmlev);

}  // namespace hpc
