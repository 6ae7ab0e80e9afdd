#include "balancer.hpp"

#include <stdlib.h>

#include <chrono>
#include <cmath>
#include <future>
#include <iostream>
#include <thread>

int defaultThreadCount() {
  if (const char *e = getenv("OMP_NUM_THREADS")) {
    const int v = atoi(e);
    if (v > 0) return v;
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

Balancer::Balancer(std::vector<int> gpuList_, MMState *state_,
                   std::tuple<float, float, float> row_multipliers, Heterogeneity het,
                   bool verbose_, bool allLevels)
    : verbose(verbose_), heterogeneity(het), gpuList(gpuList_), state(state_) {
  int maxGPUCount = 0;
  hpcCheck(hpc_device_count(&maxGPUCount), "device count");
  if ((int)gpuList.size() > maxGPUCount) {
    std::cout << "Only " << maxGPUCount << " GPUs available for kernel execution" << std::endl;
    exit(-1);
  }
  if (gpuList.size() > 1) {
    std::cout << "One GPU per skeleton instance: using device " << gpuList[0] << std::endl;
    gpuList.resize(1);
  }
  ompThreadCount = defaultThreadCount();
  const int maxEdgeCount = (int)(state->p * (state->p - 1L) / 2);
  gpuExecutor = std::make_shared<GPUExecutor>(state, maxEdgeCount, gpuList);
  cpuExecutor = std::make_shared<CPUExecutor>(state);
  gpuExecutor->allLevels = allLevels;
  cpuExecutor->allLevels = allLevels;
  rows_multiplier = std::get<0>(row_multipliers);
  rows_multiplier_l2 = std::get<1>(row_multipliers);
  rows_multiplier_l3 = std::get<2>(row_multipliers);
}

// The rows longer than the floor(mult * threads)-th longest row go to the
// host, the runs in between to the device (balancer.cpp:74-126).
void Balancer::split_longest(int level, float mult) {
  const int variableCount = (int)state->p;
  const int max_rows_on_cpu = (int)(mult * ompThreadCount);
  cpuExecutor->calculateRowLengthMap(level);
  if (cpuExecutor->rowLengthMap.empty()) {
    gpuExecutor->enqueueSplitTask(SplitTask{0, variableCount});
    return;
  }
  int idx = max_rows_on_cpu < variableCount ? max_rows_on_cpu
                                            : (int)std::ceil(variableCount * mult);
  // the reference indexes past the map on sparse graphs (balancer.cpp:88-90)
  idx = std::min(idx, (int)cpuExecutor->rowLengthMap.size() - 1);
  const int max_row_length = std::get<1>(cpuExecutor->rowLengthMap[idx]);
  int balancedRows = -1;
  for (int row = 0; row < variableCount; row++) {
    const int len = state->adj_compact[(size_t)row * variableCount + variableCount - 1];
    if (len < level) {
      if (balancedRows + 1 == row) balancedRows = row;
      continue;
    }
    if (len > max_row_length) {
      if (balancedRows < row - 1)
        gpuExecutor->enqueueSplitTask(SplitTask{balancedRows + 1, row - balancedRows - 1});
      cpuExecutor->enqueueSplitTask(SplitTask{row, 1});
      balancedRows = row;
    }
  }
  if (balancedRows < variableCount - 1)
    gpuExecutor->enqueueSplitTask(SplitTask{balancedRows + 1, variableCount - balancedRows - 1});
}

int64_t Balancer::balance(int level) {
  if (verbose) std::cout << "Start balancing rows..." << std::endl;
  auto start = std::chrono::system_clock::now();
  cpuExecutor->cleanupSplitTasks();
  gpuExecutor->cleanupSplitTasks();
  const int variableCount = (int)state->p;
  if (heterogeneity == CPUOnly) {
    for (int row = 0; row < variableCount; row++) cpuExecutor->enqueueSplitTask(SplitTask{row, 1});
  } else if (level == 0 || heterogeneity == GPUOnly) {
    gpuExecutor->enqueueSplitTask(SplitTask{0, variableCount});
  } else {
    float mult = rows_multiplier;
    if (level == 2) mult = rows_multiplier_l2;
    else if (level == 3) mult = rows_multiplier_l3;
    split_longest(level, mult);
  }
  const int64_t duration = std::chrono::duration_cast<std::chrono::milliseconds>(
                               std::chrono::system_clock::now() - start)
                               .count();
  if (verbose) {
    int onGPU = 0;
    for (auto &t : gpuExecutor->tasks) onGPU += t.rowCount;
    std::cout << "Balanced " << cpuExecutor->tasks.size() << " rows on the CPU and " << onGPU
              << " rows on the GPU in " << duration << " ms." << std::endl;
  }
  return duration;
}

void Balancer::commit(int level, std::tuple<TestResult, TestResult> &res) {
  hpc_level_stats s{};
  hpcCheck(hpc_commit_level(state->dev, level, &s), "level commit");
  const auto &cpu = std::get<0>(res);
  const auto &gpu = std::get<1>(res);
  // the canonical count of the level, known once both shares are in
  const int64_t tests = level == 0 ? cpu.tests + gpu.tests : s.tests;
  if (verbose)
    std::cout << "Order " << level << " finished with " << tests << " tests in "
              << std::max(cpu.duration, gpu.duration) << " ms.\n"
              << std::endl;
}

std::tuple<TestResult, TestResult> Balancer::execute(int level, int numThreads) {
  if (verbose) std::cout << "Start execution..." << std::endl;
  auto gpu = gpuExecutor;
  auto cpu = cpuExecutor;
  const bool v = verbose;
  auto resGPU = std::async(std::launch::async, [gpu, level, v] { return gpu->executeLevel(level, false, v); });
  auto resCPU = std::async(std::launch::async,
                           [cpu, level, numThreads, v] { return cpu->executeLevel(level, numThreads, v); });
  std::tuple<TestResult, TestResult> res{resCPU.get(), resGPU.get()};
  commit(level, res);
  return res;
}

std::tuple<TestResult, TestResult> Balancer::executeWorkstealing(int level, int numThreads) {
  if (verbose) std::cout << "Start workstealing execution..." << std::endl;
  if (level >= 1 && heterogeneity == All) {
    // one task queue: the device claims from the front, host threads from
    // the back, until they meet (balancer.cpp:186-234 on one GPU)
    hpc_level_stats dev{}, host{};
    hpcCheck(hpc_search_level_shared(state->dev, level, numThreads, &dev, &host),
             "work-stealing level");
    if (verbose)
      std::cout << "Shared queue: the host finished " << host.removed << " blocks ("
                << host.tests << " tests) beside the device." << std::endl;
    std::tuple<TestResult, TestResult> res{
        TestResult{(unsigned long long)host.ms_wall, host.tests},
        TestResult{(unsigned long long)dev.ms_wall, dev.tests}};
    commit(level, res);
    return res;
  }
  cpuExecutor->cleanupSplitTasks();
  gpuExecutor->cleanupSplitTasks();
  if (level == 0 || heterogeneity == GPUOnly)
    gpuExecutor->enqueueSplitTask(SplitTask{0, (int)state->p});
  else if (heterogeneity == CPUOnly)
    for (int row = 0; row < (int)state->p; row++) cpuExecutor->enqueueSplitTask(SplitTask{row, 1});
  else
    split_longest(level, rows_multiplier);
  auto gpu = gpuExecutor;
  auto cpu = cpuExecutor;
  const bool v = verbose;
  auto resGPU = std::async(std::launch::async, [gpu, level, v] { return gpu->executeLevel(level, true, v); });
  auto resCPU = std::async(std::launch::async, [cpu, level, numThreads, v] {
    return cpu->workstealingExecuteLevel(level, numThreads, v);
  });
  std::tuple<TestResult, TestResult> res{resCPU.get(), resGPU.get()};
  commit(level, res);
  return res;
}
