// Work split between the device and the host threads for one level
// (src/loadbalance/balancer.{hpp,cpp}).  One GPU: the reference's per-GPU
// row shards collapse to a single device share.
#pragma once

#include <memory>
#include <tuple>
#include <vector>

#include "executors.hpp"

enum Heterogeneity { All = 0, GPUOnly = 1, CPUOnly = 2 };  // balancer.hpp:8-12

struct Balancer {
  bool verbose = false;
  Heterogeneity heterogeneity = All;
  std::vector<int> gpuList;
  MMState *state = nullptr;
  int ompThreadCount = 1;
  float rows_multiplier = 0.2f, rows_multiplier_l2 = 0.2f, rows_multiplier_l3 = 0.1f;
  std::shared_ptr<CPUExecutor> cpuExecutor;
  std::shared_ptr<GPUExecutor> gpuExecutor;

  // balancer.cpp:31-136: CPU-only puts every row on the host; level 0 and
  // GPU-only put every row on the device; otherwise the rows longer than the
  // floor(mult * threads)-th longest row go to the host.
  int64_t balance(int level);
  // balancer.cpp:138-184: both executors concurrently, then the level's
  // removals (one commit over both shares).
  std::tuple<TestResult, TestResult> execute(int level, int numThreads);
  // balancer.cpp:186-234: every level is tested.  With both sides (All) the
  // level's tasks form one queue that the device takes from the front and
  // host threads from the back (hpc_search_level_shared); GPU-only and
  // CPU-only keep their one-sided shares.
  std::tuple<TestResult, TestResult> executeWorkstealing(int level, int numThreads);

  Balancer() {}
  Balancer(std::vector<int> gpuList, MMState *state, std::tuple<float, float, float> row_multipliers,
           Heterogeneity heterogeneity = All, bool verbose = false, bool allLevels = false);

 private:
  void split_longest(int level, float mult);
  void commit(int level, std::tuple<TestResult, TestResult> &res);
};

// omp_get_max_threads() equivalent: OMP_NUM_THREADS when set, else the
// hardware thread count.
int defaultThreadCount();
