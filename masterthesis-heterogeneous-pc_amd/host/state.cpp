#include "state.hpp"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

void hpcCheck(int rc, const char *what) {
  if (rc != HPC_OK) {
    printf("HIP path error in %s: %s\n", what, hpc_last_error());
    exit(EXIT_FAILURE);
  }
}

// state.cu:3-44: maxCondSize = max(maxLevel, 1); adj all ones off the
// diagonal; adj_compact the same; pMax 0; sepSets -1; max_adj = p.
MMState::MMState(uint64_t p_, int observations_, double alpha_, int maxLevel_,
                 int mainDeviceId, bool ats_)
    : ats(ats_), p(p_), observations(observations_), alpha(alpha_), maxLevel(maxLevel_) {
  maxCondSize = std::max(maxLevel, 1);
  hpcCheck(hpc_state_create(p, observations, alpha, maxLevel, mainDeviceId, &dev),
           "MMState");
  const size_t pp = (size_t)p * p;
  cor = new double[pp]();
  adj = new int[pp];
  adj_compact = new int[pp];
  pMax = new double[pp]();
  sepSets = new int[pp * maxCondSize];
  lock = new int[pp]();
  node_status = new bool[pp]();
  max_adj = new int[1];
  std::fill_n(adj, pp, 1);
  std::fill_n(adj_compact, pp, 1);
  for (uint64_t i = 0; i < p; ++i) {
    adj[i * p + i] = 0;
    adj_compact[i * p + i] = 0;
  }
  std::fill_n(sepSets, pp * maxCondSize, -1);
  max_adj[0] = (int)p;
  gpu_done = false;
}

void MMState::adviceReadonlyCor(std::vector<int>) {}
void MMState::memAdvise(std::vector<int>) {}
void MMState::prefetchRows(int, int, int) {}

void MMState::pull(bool with_cor) {
  hpcCheck(hpc_get_adj(dev, adj), "pull adj");
  hpcCheck(hpc_get_compact(dev, adj_compact), "pull adj_compact");
  const int mc = hpc_max_cond_size(dev);
  if (mc == maxCondSize) {
    hpcCheck(hpc_get_sepsets(dev, sepSets, maxCondSize), "pull sepSets");
  } else {  // the device grew the width (levels beyond maxLevel)
    delete[] sepSets;
    maxCondSize = mc;
    sepSets = new int[(size_t)p * p * maxCondSize];
    hpcCheck(hpc_get_sepsets(dev, sepSets, maxCondSize), "pull sepSets");
  }
  if (with_cor) hpcCheck(hpc_get_cor(dev, cor), "pull cor");
}

void MMState::destroy() {
  if (dev) hpc_state_destroy(dev);
  dev = nullptr;
  delete[] cor; delete[] adj; delete[] adj_compact; delete[] pMax;
  delete[] sepSets; delete[] lock; delete[] node_status; delete[] max_adj;
  cor = pMax = nullptr;
  adj = adj_compact = sepSets = lock = max_adj = nullptr;
  node_status = nullptr;
}
