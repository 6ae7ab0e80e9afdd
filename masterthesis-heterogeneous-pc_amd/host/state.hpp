// Host-side view of one skeleton instance: the reference's MMState
// (src/util/state.cuh:29-55) with the same fields, backed by a device-resident
// hpc_state (include/heterogpc_hip.h).  The device holds the authoritative
// state; the host arrays are reference-layout mirrors refreshed by pull()
// and callCompact().
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/heterogpc_hip.h"

// src/util/indep_util.hpp:4-16
struct SplitTask {
  int row;
  int rowCount;
};

struct TestResult {
  unsigned long long duration;  // milliseconds
  int64_t tests;                // canonical CI tests of this executor's rows
};

// Every ABI status is checked; a failure prints and exits like the
// reference's checkCudaErrors (src/util/cuda_util.cuh:28-36).
void hpcCheck(int rc, const char *what);

struct MMState {
  bool ats = false;             // accepted for --power9-ats; no effect
  double *pMax = nullptr;       // p*p, p-value of the removing test
  int *adj = nullptr;           // p*p
  int *adj_compact = nullptr;   // p*p, compact.cu layout
  double *cor = nullptr;        // p*p
  int *sepSets = nullptr;       // p*p*maxCondSize
  uint64_t p = 0;
  int observations = 0;
  double alpha = 0.0;
  int maxCondSize = 1;
  int *max_adj = nullptr;       // 1
  int *lock = nullptr;          // p*p, unused: removal is deterministic
  int maxLevel = 0;
  bool *node_status = nullptr;  // p*p, unused: see lock
  bool gpu_done = false;
  hpc_state *dev = nullptr;     // device-resident instance

  MMState() {}
  MMState(uint64_t p, int observations, double alpha, int maxLevel, int mainDeviceId,
          bool ats = false);

  // Memory advice and prefetch of the unified-memory design
  // (state.cu:46-94): the state already lives in HBM, so these are no-ops.
  void adviceReadonlyCor(std::vector<int> gpuList);
  void memAdvise(std::vector<int> gpuList);
  void prefetchRows(int startRow, int rowCount, int deviceId);

  // Copies adj, adj_compact, sepSets (and cor when asked) from the device.
  void pull(bool with_cor = false);
  void destroy();
};
