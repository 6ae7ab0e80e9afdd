#include "skeleton.hpp"

#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <iostream>

using namespace std;

SkeletonCalculator::SkeletonCalculator(int maxLevel_, double alpha_, bool workstealing_,
                                       string csvExportFile_, int numThreads_, bool verbose_)
    : maxLevel(maxLevel_), numThreads(numThreads_), alpha(alpha_), verbose(verbose_),
      workstealing(workstealing_), csvExportFile(csvExportFile_) {
  gpuList = {0};
  heterogeneity = All;
}

SkeletonCalculator::~SkeletonCalculator() { state.destroy(); }

void SkeletonCalculator::set_heterogeneity(Heterogeneity h) { heterogeneity = h; }
void SkeletonCalculator::set_gpu_list(vector<int> list) { gpuList = list; }
void SkeletonCalculator::set_all_levels(bool all) { allLevels = all; }

shared_ptr<Mat> SkeletonCalculator::read_csv_file(string input_file) {
  if (verbose) cout << "Reading file: " << input_file << endl;
  vector<string> column_names(0);
  if (input_file.find(".csv") == string::npos) {
    cout << "Cannot process file '" << input_file << "'." << endl;
    cout << "Has to be .csv format." << endl;
    exit(-1);
  }
  return CSVParser::read_csv_to_mat(input_file, column_names);
}

void SkeletonCalculator::add_observations(string input_file, bool use_p9_ats) {
  auto data = read_csv_file(input_file);
  state = MMState(data->n_cols, (int)data->n_rows, alpha, maxLevel, gpuList[0], use_p9_ats);
  gpuPMCC(&state, data->begin(), verbose);
}

void SkeletonCalculator::add_correlation_matrix(string input_file, int observation_count,
                                                bool use_p9_ats) {
  auto data = read_csv_file(input_file);
  state = MMState(data->n_cols, observation_count, alpha, maxLevel, gpuList[0], use_p9_ats);
  // the file's rows are the matrix rows; column-major storage transposes, and
  // the reference copies that storage as is (skeleton.cpp:64)
  memcpy(state.cor, data->begin(), state.p * state.p * sizeof(double));
  hpcCheck(hpc_set_corr(state.dev, state.cor), "add_correlation_matrix");
}

void SkeletonCalculator::initialize_balancer(tuple<float, float, float> thresholds) {
  balancer = Balancer(gpuList, &state, thresholds, heterogeneity, verbose, allLevels);
}

void SkeletonCalculator::export_metrics(vector<LevelMetrics> levelMetrics,
                                        int64_t executionDuration, int nrEdges) {
  if (csvExportFile.empty()) return;
  if (verbose) cout << "Export metrics to CSV file: " << csvExportFile << endl;
  ofstream csvFile(csvExportFile, ios::app | ios::out);
  const int numGPUs = heterogeneity == CPUOnly ? 0 : (int)gpuList.size();
  csvFile << numGPUs << "," << defaultThreadCount() << "," << nrEdges << ",";
  for (auto &metric : levelMetrics) {
    auto [levelDur, balanceDur, execRes] = metric;
    auto [cpuRes, gpuRes] = execRes;
    csvFile << levelDur << "," << balanceDur << "," << cpuRes.duration << "," << gpuRes.duration
            << ",";
  }
  csvFile << executionDuration << endl;
}

LevelMetrics SkeletonCalculator::calcLevel(int level) {
  auto start = chrono::system_clock::now();
  if (level >= 1) callCompact(&state, gpuList[0], 0, 1, (int)state.p);
  if (verbose) cout << "Max row length: " << state.max_adj[0] << endl;
  tuple<TestResult, TestResult> execRes;
  int64_t balanceDur = 0;
  if (workstealing) {
    execRes = balancer.executeWorkstealing(level, numThreads);
  } else {
    balanceDur = balancer.balance(level);
    execRes = balancer.execute(level, numThreads);
  }
  const int64_t levelDur =
      chrono::duration_cast<chrono::milliseconds>(chrono::system_clock::now() - start).count();
  return {levelDur, balanceDur, execRes};
}

void SkeletonCalculator::run(bool print_sepsets) {
  if (verbose) {
    cout << "Using following GPUs:" << endl;
    for (auto deviceId : gpuList) cout << "\t" << deviceId << endl;
    cout << "maxCondSize: " << state.maxCondSize << "  observations: " << state.observations
         << "  p: " << state.p << " number of GPUS: " << gpuList.size() << endl;
  }
  auto start = chrono::system_clock::now();
  state.adviceReadonlyCor(gpuList);
  state.memAdvise(gpuList);
  vector<LevelMetrics> levelMetrics;
  for (int lvl = 0; lvl <= state.maxLevel; lvl++) levelMetrics.push_back(calcLevel(lvl));
  const int64_t executionDuration =
      chrono::duration_cast<chrono::milliseconds>(chrono::system_clock::now() - start).count();
  if (verbose) cout << "Execution duration: " << executionDuration << " ms." << endl;
  state.pull();
  nrEdges = printSepsets(&state, print_sepsets, verbose);
  export_metrics(levelMetrics, executionDuration, nrEdges);
}
