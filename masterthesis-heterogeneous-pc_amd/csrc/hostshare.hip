// Host share of a level: the CPU executor's rows (CPUExecutor::executeLevel,
// cpuExecutor.cpp:80-119) searched on the host beside the device.
//
// The search runs the arithmetic of pc_math.hpp compiled for the host, so a
// row gives the same first passing rank on either side; the ranks are then
// scattered into the device's per-edge array and removed by the same commit.
// Threads take (row, neighbour) work items longest row first, ten at a time,
// as the reference's `collapse(2) schedule(dynamic, 10)` loop does.
#pragma clang fp contract(off)

#include <string.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "hpc_internal.hpp"
#include "pc_math.hpp"

using namespace hpc;
using std::isnan;

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return hpc_fail(HPC_EHIP, "%s failed: %s (%s:%d)", #call,             \
                      hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

namespace {

struct HostSnap {
  const double *cor;
  const int32_t *rowptr, *nbr;
  const unsigned long long *binom;
  int p;
  double r_crit;
};

inline int slot_node(const HostSnap &s, int base, int y, int q) {
  return s.nbr[base + (q < y ? q : q + 1)];
}

// lexicographic successor of q over {0..nn-1}; false after the last set
template <int L>
inline bool next_comb(int (&q)[L], int nn) {
  int i = L - 1;
  while (i >= 0 && q[i] == nn - L + i) --i;
  if (i < 0) return false;
  ++q[i];
  for (int k = i + 1; k < L; ++k) q[k] = q[k - 1] + 1;
  return true;
}

// First passing rank of ordered edge (i, position y), ranks in order.
template <int L>
long long search_edge(const HostSnap &s, int i, int y, unsigned long long &tests,
                      unsigned long long &total_out, unsigned long long &fallbacks) {
  const int base = s.rowptr[i];
  const int nn = s.rowptr[i + 1] - base - 1;
  const int j = s.nbr[base + y];
  const double *cor = s.cor;
  const size_t p = s.p;
  const double a = cor[(size_t)i * p + j];
  if (L == 1) {
    total_out = nn > 0 ? (unsigned long long)nn : 0ull;
    for (int q = 0; q < nn; ++q) {
      const int k = slot_node(s, base, y, q);
      const double x2 = cor[(size_t)i * p + k], x3 = cor[(size_t)j * p + k];
      ++tests;
      if (!(isnan(a) || isnan(x2) || isnan(x3)) && removes(pcor_l1(a, x2, x3), s.r_crit))
        return q;
    }
    return -1;
  }
  const unsigned long long total = binom(s.binom, kBinomK, nn, L);
  total_out = total;
  if (total == 0) return -1;
  int q[L];
  for (int c = 0; c < L; ++c) q[c] = c;
  long long rank = 0;
  do {
    int S[L];
    for (int c = 0; c < L; ++c) S[c] = slot_node(s, base, y, q[c]);
    double x[L], yv[L], C[L][L];
    bool anynan = isnan(a);
    for (int c = 0; c < L; ++c) {
      x[c] = cor[(size_t)i * p + S[c]];
      yv[c] = cor[(size_t)j * p + S[c]];
      anynan |= isnan(x[c]) || isnan(yv[c]);
    }
    for (int c = 0; c < L; ++c)
      for (int r = c + 1; r < L; ++r) {
        C[r][c] = cor[(size_t)S[c] * p + S[r]];
        anynan |= isnan(C[r][c]);
      }
    ++tests;
    if (!anynan) {
      bool fb;
      const double r = pcor_canon<L>(a, x, yv, C, fb);
      fallbacks += fb ? 1 : 0;
      if (removes(r, s.r_crit)) return rank;
    }
    ++rank;
  } while (next_comb<L>(q, nn));
  return -1;
}

// Ranks [r_begin, r_end) of ordered edge (i, N(i)[y]) in lexicographic
// order, as one task of the device's plan: the lowest passing rank or -1;
// `tests` counts the sets tested.
template <int L>
long long search_range(const HostSnap &s, int i, int y, unsigned long long r_begin,
                       unsigned long long r_end, unsigned long long &tests,
                       unsigned long long &fallbacks) {
  const int base = s.rowptr[i];
  const int nn = s.rowptr[i + 1] - base - 1;
  const int j = s.nbr[base + y];
  const double *cor = s.cor;
  const size_t p = s.p;
  const double a = cor[(size_t)i * p + j];
  if (r_begin >= r_end) return -1;
  if (L == 1) {
    for (unsigned long long q = r_begin; q < r_end; ++q) {
      const int k = slot_node(s, base, y, (int)q);
      const double x2 = cor[(size_t)i * p + k], x3 = cor[(size_t)j * p + k];
      ++tests;
      if (!(isnan(a) || isnan(x2) || isnan(x3)) && removes(pcor_l1(a, x2, x3), s.r_crit))
        return (long long)q;
    }
    return -1;
  }
  int q[L];
  unrank<L>(s.binom, kBinomK, r_begin, nn, q);
  for (unsigned long long rank = r_begin; rank < r_end; ++rank) {
    int S[L];
    for (int c = 0; c < L; ++c) S[c] = slot_node(s, base, y, q[c]);
    double x[L], yv[L], C[L][L];
    bool anynan = isnan(a);
    for (int c = 0; c < L; ++c) {
      x[c] = cor[(size_t)i * p + S[c]];
      yv[c] = cor[(size_t)j * p + S[c]];
      anynan |= isnan(x[c]) || isnan(yv[c]);
    }
    for (int c = 0; c < L; ++c)
      for (int r = c + 1; r < L; ++r) {
        C[r][c] = cor[(size_t)S[c] * p + S[r]];
        anynan |= isnan(C[r][c]);
      }
    ++tests;
    if (!anynan) {
      bool fb;
      const double r = pcor_canon<L>(a, x, yv, C, fb);
      fallbacks += fb ? 1 : 0;
      if (removes(r, s.r_crit)) return (long long)rank;
    }
    next_comb<L>(q, nn);
  }
  return -1;
}

using RangeFn = long long (*)(const HostSnap &, int, int, unsigned long long,
                              unsigned long long, unsigned long long &, unsigned long long &);

template <int L>
RangeFn pick_range() { return &search_range<L>; }

RangeFn range_for(int level) {
  switch (level) {
    case 1: return pick_range<1>();   case 2: return pick_range<2>();   case 3: return pick_range<3>();
    case 4: return pick_range<4>();   case 5: return pick_range<5>();   case 6: return pick_range<6>();
    case 7: return pick_range<7>();   case 8: return pick_range<8>();   case 9: return pick_range<9>();
    case 10: return pick_range<10>(); case 11: return pick_range<11>(); case 12: return pick_range<12>();
    case 13: return pick_range<13>(); case 14: return pick_range<14>(); case 15: return pick_range<15>();
    case 16: return pick_range<16>();
    default: return nullptr;
  }
}

using SearchFn = long long (*)(const HostSnap &, int, int, unsigned long long &,
                               unsigned long long &, unsigned long long &);

template <int L>
SearchFn pick() { return &search_edge<L>; }

SearchFn search_for(int level) {
  switch (level) {
    case 1: return pick<1>();   case 2: return pick<2>();   case 3: return pick<3>();
    case 4: return pick<4>();   case 5: return pick<5>();   case 6: return pick<6>();
    case 7: return pick<7>();   case 8: return pick<8>();   case 9: return pick<9>();
    case 10: return pick<10>(); case 11: return pick<11>(); case 12: return pick<12>();
    case 13: return pick<13>(); case 14: return pick<14>(); case 15: return pick<15>();
    case 16: return pick<16>();
    default: return nullptr;
  }
}


}  // namespace

extern "C" int hpc_search_level_host(hpc_state *st, int level, const int32_t *rows,
                                     int nrows, int threads, hpc_level_stats *stats) {
  if (!st) return hpc_fail(HPC_EINVAL, "state is NULL");
  Engine &e = st->eng;
  InstHost &h = e.inst[st->id];
  if (level < 1 || level > kMaxLevel)
    return hpc_fail(HPC_ELEVEL, "host share supports levels 1..%d, not %d", kMaxLevel, level);
  if (!h.compacted) return hpc_fail(HPC_ESTATE, "level %d needs hpc_compact first", level);
  if (h.fast)
    return hpc_fail(HPC_ESTATE, "the host share needs an exact correlation (HPC_CORR_CANON)");
  if (!rows && nrows > 0) return hpc_fail(HPC_EINVAL, "rows is NULL");
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(e.device));
  const int p = h.p;
  // host copies of the snapshot (and the correlation, once per matrix); the
  // snapshot of a one-instance engine uses local ids
  e.h_rowptr.resize(p + 1);
  e.h_nbr.resize(e.n_edges > 0 ? e.n_edges : 1);
  HIPCHK(hipMemcpyAsync(e.h_rowptr.data(), e.d_rowptr, (p + 1) * sizeof(int32_t),
                        hipMemcpyDeviceToHost, e.copy_stream));
  if (e.n_edges > 0)
    HIPCHK(hipMemcpyAsync(e.h_nbr.data(), e.d_nbr, e.n_edges * sizeof(int32_t),
                          hipMemcpyDeviceToHost, e.copy_stream));
  if (!e.h_cor_valid) {
    e.h_cor.resize((size_t)p * p);
    HIPCHK(hipMemcpyAsync(e.h_cor.data(), h.cor, (size_t)p * p * sizeof(double),
                          hipMemcpyDeviceToHost, e.copy_stream));
  }
  HIPCHK(hipStreamSynchronize(e.copy_stream));
  e.h_cor_valid = true;
  HostSnap snap{e.h_cor.data(), e.h_rowptr.data(), e.h_nbr.data(), e.h_binom.data(), p, h.r_crit};
  // work items: (row, neighbour position), rows longest first
  std::vector<std::pair<int, int>> order;
  for (int k = 0; k < nrows; ++k) {
    const int i = rows[k];
    if (i < 0 || i >= p) return hpc_fail(HPC_EINVAL, "row %d out of range", i);
    order.push_back({snap.rowptr[i + 1] - snap.rowptr[i], i});
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first > b.first; });
  // the longest row's conditioning sets per edge (saturated: beyond any run)
  if (!order.empty() && level >= 2 && order[0].first - 1 >= level &&
      binom(snap.binom, kBinomK, order[0].first - 1, level) == ~0ull)
    return hpc_fail(HPC_ELEVEL, "level %d: C(%d, %d) conditioning sets per edge saturates",
                    level, order[0].first - 1, level);
  std::vector<std::pair<int, int>> items;
  for (auto &o : order)
    for (int y = 0; y < o.first; ++y) items.push_back({o.second, y});
  std::vector<int64_t> idx(items.size()), val(items.size());
  SearchFn fn = search_for(level);
  std::atomic<size_t> next{0};
  std::atomic<unsigned long long> T{0}, U{0}, E{0}, F{0};
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  auto worker = [&]() {
    unsigned long long t = 0, u = 0, ed = 0, f = 0;
    for (;;) {
      const size_t b = next.fetch_add(10);
      if (b >= items.size()) break;
      const size_t end = std::min(items.size(), b + 10);
      for (size_t w = b; w < end; ++w) {
        const int i = items[w].first, y = items[w].second;
        unsigned long long tests = 0, total = 0;
        const long long f0 = fn(snap, i, y, tests, total, f);
        idx[w] = snap.rowptr[i] + y;
        val[w] = f0;
        if (total > 0) {
          t += f0 >= 0 ? (unsigned long long)f0 + 1 : total;
          u += total;
          ed += 1;
        }
      }
    }
    T += t; U += u; E += ed; F += f;
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < threads; ++k) pool.emplace_back(worker);
  worker();
  for (auto &th : pool) th.join();
  // scatter the ranks into the device array
  const size_t cnt = items.size();
  if (cnt > 0) {
    if (cnt > e.scatter_cap) {
      if (e.d_scatter_idx) HIPCHK(hipFree(e.d_scatter_idx));
      if (e.d_scatter_val) HIPCHK(hipFree(e.d_scatter_val));
      e.d_scatter_idx = e.d_scatter_val = nullptr;
      HIPCHK(hipMalloc(&e.d_scatter_idx, cnt * sizeof(int64_t)));
      HIPCHK(hipMalloc(&e.d_scatter_val, cnt * sizeof(int64_t)));
      e.scatter_cap = cnt;
    }
    // after the compaction's reset of the array (the device share may still
    // be searching on the main stream: its edges are disjoint from these)
    HIPCHK(hipStreamWaitEvent(e.copy_stream, e.fp_ready, 0));
    HIPCHK(hipMemcpyAsync(e.d_scatter_idx, idx.data(), cnt * sizeof(int64_t),
                          hipMemcpyHostToDevice, e.copy_stream));
    HIPCHK(hipMemcpyAsync(e.d_scatter_val, val.data(), cnt * sizeof(int64_t),
                          hipMemcpyHostToDevice, e.copy_stream));
    launch_scatter_fp(e.copy_stream, e.d_scatter_idx, e.d_scatter_val, (long long)cnt, e.d_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e.copy_stream));
  }
  if (stats) {
    hpc_level_stats s;
    memset(&s, 0, sizeof(s));
    s.level = level;
    s.max_deg = h.max_deg;
    s.tests = (int64_t)T.load();
    s.tests_max = (int64_t)U.load();
    s.edges = (int64_t)E.load();
    s.pinv_fallbacks = (int64_t)F.load();
    s.edges_left = h.undirected;
    s.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *stats = s;
  }
  return HPC_OK;
}

extern "C" int hpc_search_level_shared(hpc_state *st, int level, int threads,
                                       hpc_level_stats *device_stats,
                                       hpc_level_stats *host_stats) {
  if (!st) return hpc_fail(HPC_EINVAL, "state is NULL");
  Engine &e = st->eng;
  InstHost &h = e.inst[st->id];
  if (level < 1 || level > kMaxLevel)
    return hpc_fail(HPC_ELEVEL, "the shared queue supports levels 1..%d, not %d", kMaxLevel, level);
  if (!h.compacted) return hpc_fail(HPC_ESTATE, "level %d needs hpc_compact first", level);
  if (h.fast)
    return hpc_fail(HPC_ESTATE, "the shared queue needs an exact correlation (HPC_CORR_CANON)");
  auto t0 = std::chrono::steady_clock::now();
  engine_note_level(e, st->id, level);
  int rc = engine_pack(e, {st->id}, {level});
  SharedPlan plan;
  if (!rc) rc = engine_search_shared_launch(e, level, plan);
  if (rc) return rc;
  // host threads: blocks from the back, until they reach one the device has
  // finished or the device's kernels are done (every block is then claimed
  // there, and each claimed block that the host had not finished was run)
  std::atomic<long long> tail{plan.n_blocks};
  std::atomic<bool> device_over{plan.n_blocks == 0};
  std::atomic<unsigned long long> T{0}, E{0}, F{0}, B{0};
  std::mutex out_mu;
  std::vector<int64_t> idx, val;
  HostSnap snap{e.h_cor.data(), e.h_rowptr.data(), e.h_nbr.data(), e.h_binom.data(), h.p, h.r_crit};
  RangeFn fn = range_for(level);
  const int p = e.p_total;
  auto device_finished = [&]() -> bool {
    if (device_over.load(std::memory_order_relaxed)) return true;
    if (hipEventQuery(plan.device_done) == hipSuccess) device_over.store(true);
    return device_over.load(std::memory_order_relaxed);
  };
  auto worker = [&]() {
    unsigned long long t = 0, ed = 0, f = 0, blocks = 0;
    std::vector<int64_t> li, lv;
    for (;;) {
      const long long c = tail.fetch_sub(1) - 1;
      if (c < 0 || device_finished()) break;
      if (__atomic_load_n(&plan.done[c], __ATOMIC_ACQUIRE)) break;  // the device's side
      const long long b0 = c * plan.tpb, b1 = std::min(plan.n_tasks, b0 + plan.tpb);
      bool whole = true;
      for (long long task = b0; task < b1; ++task) {
        if (device_finished()) { whole = false; break; }
        // the task's row: the last row whose offset is <= task
        const long long *off = e.h_task_off.data();
        int lo = 0, hi = p - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (off[mid] <= task) lo = mid; else hi = mid - 1;
        }
        const int i = lo;
        const int deg = snap.rowptr[i + 1] - snap.rowptr[i];
        const long long nch = e.h_row_tasks[i] / deg;
        const long long local = task - off[i];
        const int y = (int)(local / nch);
        const unsigned long long total = binom(snap.binom, kBinomK, deg - 1, level);
        const unsigned long long r_begin = (unsigned long long)(local % nch) * plan.rpt;
        const unsigned long long r_end = std::min(total, r_begin + (unsigned long long)plan.rpt);
        const long long r = fn(snap, i, y, r_begin, r_end, t, f);
        if (r_begin == 0) ++ed;
        if (r >= 0) {
          li.push_back(snap.rowptr[i] + y);
          lv.push_back(r);
        }
      }
      if (!whole) break;
      __atomic_store_n(&plan.done[c], 1, __ATOMIC_RELEASE);
      ++blocks;
    }
    T += t; E += ed; F += f; B += blocks;
    std::lock_guard<std::mutex> lk(out_mu);
    idx.insert(idx.end(), li.begin(), li.end());
    val.insert(val.end(), lv.begin(), lv.end());
  };
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::thread> pool;
  if (plan.n_blocks > 0) {
    for (int k = 1; k < threads; ++k) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
  }
  float ms = 0.f;
  rc = engine_search_shared_finish(e, idx, val, plan.n_blocks > 0 ? &ms : nullptr);
  if (rc) return rc;
  const double wall =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (device_stats) {
    hpc_level_stats s;
    memset(&s, 0, sizeof(s));
    s.level = level;
    s.max_deg = h.max_deg;
    s.edges_left = h.undirected;
    s.ms_test = ms;
    s.ms_wall = wall;
    *device_stats = s;
  }
  if (host_stats) {
    hpc_level_stats s;
    memset(&s, 0, sizeof(s));
    s.level = level;
    s.max_deg = h.max_deg;
    s.tests = (int64_t)T.load();
    s.edges = (int64_t)E.load();
    s.removed = (int64_t)B.load();
    s.pinv_fallbacks = (int64_t)F.load();
    s.edges_left = h.undirected;
    s.ms_wall = wall;
    *host_stats = s;
  }
  return HPC_OK;
}
