// Data preparation and executors above the C ABI:
//   gpuPMCC      src/correlation/corOwn.cuh:3
//   callCompact  src/independence/compact.cuh:7-8
//   Executor, GPUExecutor, CPUExecutor  src/executor/*.hpp
#pragma once

#include <stdint.h>

#include <tuple>
#include <vector>

#include "state.hpp"

// corOwn.cu:10-48: correlation of the variable-major [n][dim] observations
// h_mat into cors (n x n).  Standalone form: a temporary device instance.
void gpuPMCC(const double *h_mat, uint64_t n, int dim, double *cors, int deviceId,
             bool verbose);
// The same into a state's device-resident correlation (and its host mirror).
void gpuPMCC(MMState *state, const double *h_mat, bool verbose);

// compact.cu:91-99: level-start neighbour lists; refreshes the host
// adj_compact mirror and max_adj.  idx / numberOfGPUs / d_rowCount are the
// reference's multi-GPU row shard, which is always the whole matrix here.
void callCompact(MMState *state, int deviceId, int idx, int numberOfGPUs, int d_rowCount);

struct Executor {
  std::vector<SplitTask> tasks;  // pre-balanced row ranges (executor.hpp:14-21)
  void enqueueSplitTask(SplitTask task) { tasks.push_back(task); }
  void cleanupSplitTasks() { tasks.clear(); }
};

struct GPUExecutor : Executor {
  int maxEdgeCount;
  std::vector<int> gpuList;
  MMState *state;
  // Tests every ordered edge of the executor's rows with length >= level
  // (gpuExecutor.cu:9-122).  With allLevels == false levels >= 4 launch
  // nothing, as the reference's kernel switch has no case for them
  // (gpuExecutor.cu:69-101); workstealing or allLevels test every level.
  TestResult executeLevel(int level, bool workstealing, bool verbose = false);
  bool allLevels = false;

  GPUExecutor(MMState *state, int maxEdgeCount, std::vector<int> gpuList)
      : maxEdgeCount(maxEdgeCount), gpuList(gpuList), state(state) {}
};

struct CPUExecutor : Executor {
  MMState *state;
  std::vector<std::tuple<int, int>> rowLengthMap;
  bool allLevels = false;

  // cpuExecutor.cpp:80-119: the executor's rows searched on host threads
  // with the device's arithmetic (levels 0..3 unless allLevels; level 0 is
  // tested on the device, which holds the adjacency).
  TestResult executeLevel(int level, int numThreads, bool verbose = false);
  // cpuExecutor.cpp:15-78: the host end of a work-stealing level -- the
  // longest rows, searched concurrently with the device's share.
  TestResult workstealingExecuteLevel(int level, int numThreads, bool verbose);
  // cpuExecutor.cpp:142-156: (row, length) of rows with length >= level,
  // longest first.
  void calculateRowLengthMap(int level);

  explicit CPUExecutor(MMState *state) : state(state) {}
};
