// Level loop of one skeleton (src/independence/skeleton.{hpp,cpp}).
#pragma once

#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "balancer.hpp"
#include "io.hpp"

// (level ms, balance ms, (cpu, gpu)) -- skeleton.hpp:8
typedef std::tuple<int64_t, int64_t, std::tuple<TestResult, TestResult>> LevelMetrics;

class SkeletonCalculator {
  int maxLevel;
  int numThreads;
  double alpha;
  bool verbose;
  bool workstealing;
  bool allLevels = false;
  std::string csvExportFile;
  Heterogeneity heterogeneity;
  std::vector<int> gpuList;
  MMState state;
  Balancer balancer;
  int nrEdges = 0;

 public:
  SkeletonCalculator(int maxLevel, double alpha, bool workstealing, std::string csvExportFile,
                     int numThreads, bool verbose = false);
  ~SkeletonCalculator();

  void set_heterogeneity(Heterogeneity heterogeneity);
  void set_gpu_list(std::vector<int> gpuList);
  // Test levels >= 4 in the pre-balanced path too (the reference's row
  // executors only test levels 0..3).
  void set_all_levels(bool all);
  std::shared_ptr<Mat> read_csv_file(std::string input_file);
  // skeleton.cpp:53-58: observations -> state -> correlation on the device
  void add_observations(std::string input_file, bool use_p9_ats = false);
  // skeleton.cpp:60-65: a correlation matrix file plus the sample count
  void add_correlation_matrix(std::string input_file, int observation_count,
                              bool use_p9_ats = false);
  void initialize_balancer(std::tuple<float, float, float> balancer_thresholds);
  // skeleton.cpp:152-190: levels 0..maxLevel, sepset printout, metrics CSV
  void run(bool print_sepsets);
  // edges left after the last run()
  int edges_left() const { return nrEdges; }
  LevelMetrics calcLevel(int level);
  void export_metrics(std::vector<LevelMetrics> levelMetrics, int64_t executionDuration,
                      int nrEdges);
  MMState &get_state() { return state; }
};
