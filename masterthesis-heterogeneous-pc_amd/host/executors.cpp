#include "executors.hpp"

#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <iostream>

namespace {

using Clock = std::chrono::system_clock;

unsigned long long ms_since(Clock::time_point t0) {
  return (unsigned long long)std::chrono::duration_cast<std::chrono::milliseconds>(
             Clock::now() - t0)
      .count();
}

// rows of the task list whose level-start length is >= level
// (gpuExecutor.cu:22-33, cpuExecutor.cpp:88-100)
std::vector<int32_t> task_rows(const MMState *state, const std::vector<SplitTask> &tasks,
                               int level) {
  std::vector<int32_t> rows;
  const uint64_t p = state->p;
  for (const auto &t : tasks)
    for (int i = t.row; i < t.row + t.rowCount; ++i)
      if (state->adj_compact[(size_t)i * p + p - 1] >= level) rows.push_back(i);
  return rows;
}

}  // namespace

void gpuPMCC(const double *h_mat, uint64_t n, int dim, double *cors, int deviceId,
             bool verbose) {
  if (verbose) printf("Cor started with N=%lu, dim=%i\n", (unsigned long)n, dim);
  hpc_state *st = nullptr;
  hpcCheck(hpc_state_create(n, dim, 0.05, 1, deviceId, &st), "gpuPMCC");
  hpcCheck(hpc_pmcc(st, h_mat, HPC_CORR_CANON), "gpuPMCC");
  hpcCheck(hpc_get_cor(st, cors), "gpuPMCC");
  hpc_state_destroy(st);
  if (verbose) printf("PMCC successful \n");
}

void gpuPMCC(MMState *state, const double *h_mat, bool verbose) {
  if (verbose)
    printf("Cor started with N=%lu, dim=%i\n", (unsigned long)state->p, state->observations);
  hpcCheck(hpc_pmcc(state->dev, h_mat, HPC_CORR_CANON), "gpuPMCC");
  hpcCheck(hpc_get_cor(state->dev, state->cor), "gpuPMCC");
  if (verbose) printf("PMCC successful \n");
}

void callCompact(MMState *state, int, int, int, int) {
  int max_deg = 0;
  hpcCheck(hpc_compact(state->dev, &max_deg), "Compacting failed");
  hpcCheck(hpc_get_compact(state->dev, state->adj_compact), "Compacting failed");
  state->max_adj[0] = max_deg;
}

TestResult GPUExecutor::executeLevel(int level, bool workstealing, bool verbose) {
  if (tasks.empty()) return {0, 0};
  std::vector<int32_t> rows = task_rows(state, tasks, level);
  if (rows.empty()) return {0, 0};
  auto start = Clock::now();
  if (verbose) {
    int maxRowLength = 0;
    for (int i : rows)
      maxRowLength = std::max(maxRowLength, state->adj_compact[(size_t)i * state->p + state->p - 1]);
    std::cout << "\tRows per GPU: " << rows.size() << " Maximum GPU row length: " << maxRowLength
              << std::endl;
  }
  hpc_level_stats s{};
  if (level <= 3 || workstealing || allLevels) {
    const bool all = rows.size() == state->p;
    hpcCheck(hpc_search_level(state->dev, level, all ? nullptr : rows.data(),
                              all ? 0 : (int)rows.size(), &s),
             ("L " + std::to_string(level) + " Kernel execution failed").c_str());
  }
  if (workstealing) state->gpu_done = level % 2 == 1;
  const unsigned long long duration = ms_since(start);
  if (verbose) std::cout << "\tGPU is done. Time: " << duration << " ms." << std::endl;
  return TestResult{duration, s.tests};
}

TestResult CPUExecutor::executeLevel(int level, int numThreads, bool verbose) {
  if (tasks.empty()) return {0, 0};
  auto start = Clock::now();
  std::vector<int32_t> rows = task_rows(state, tasks, level);
  hpc_level_stats s{};
  if (!rows.empty() && (level <= 3 || allLevels)) {
    if (level == 0)  // the adjacency lives on the device
      hpcCheck(hpc_search_level(state->dev, 0, nullptr, 0, &s), "CPU level 0");
    else
      hpcCheck(hpc_search_level_host(state->dev, level, rows.data(), (int)rows.size(),
                                     numThreads, &s),
               "CPU level");
  }
  const unsigned long long duration = ms_since(start);
  if (verbose) std::cout << "\tCPU is done. Time: " << duration << " ms." << std::endl;
  return TestResult{duration, s.tests};
}

TestResult CPUExecutor::workstealingExecuteLevel(int level, int numThreads, bool verbose) {
  if (level == 0 || tasks.empty()) return TestResult{0, 0};
  auto start = Clock::now();
  std::vector<int32_t> rows = task_rows(state, tasks, level);
  hpc_level_stats s{};
  if (!rows.empty())
    hpcCheck(hpc_search_level_host(state->dev, level, rows.data(), (int)rows.size(),
                                   numThreads, &s),
             "CPU workstealing level");
  const unsigned long long duration = ms_since(start);
  if (verbose)
    std::cout << "\tCPU is done. Time: " << (int)duration << " ms. Edges stolen: " << s.edges
              << std::endl;
  return TestResult{duration, s.tests};
}

void CPUExecutor::calculateRowLengthMap(int level) {
  rowLengthMap.clear();
  const uint64_t p = state->p;
  for (int row = 0; row < (int)p; row++) {
    const int row_length = state->adj_compact[(size_t)row * p + p - 1];
    if (row_length >= level) rowLengthMap.push_back({row, row_length});
  }
  std::stable_sort(rowLengthMap.begin(), rowLengthMap.end(),
                   [](const std::tuple<int, int> &a, const std::tuple<int, int> &b) {
                     return std::get<1>(a) > std::get<1>(b);
                   });
}
