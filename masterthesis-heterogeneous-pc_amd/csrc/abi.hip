// Single-instance C ABI of libheterogpc_hip.so (include/heterogpc_hip.h): one
// hpc_state is an engine holding one instance.  Runtime helpers, the
// synthetic SEM generator and result export live here too.
#pragma clang fp contract(off)

#include <atomic>
#include <mutex>
#include <thread>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "hpc_internal.hpp"
#include "pc_math.hpp"

using namespace hpc;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int hpc_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define fail hpc_fail
#define HIPCHK HPC_HIPCHK

extern "C" const char *hpc_last_error(void) { return g_err.c_str(); }

extern "C" int hpc_device_count(int *count) {
  if (!count) return fail(HPC_EINVAL, "count is NULL");
  HIPCHK(hipGetDeviceCount(count));
  return HPC_OK;
}

extern "C" int hpc_device_synchronize(int device) {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipDeviceSynchronize());
  return HPC_OK;
}

extern "C" double hpc_fisher_r_crit(int n, double alpha) { return fisher_r_crit(n, alpha); }

extern "C" int hpc_plan_tasks(uint64_t deg, int level, int ranks_per_task, int64_t *tasks) {
  if (!tasks) return fail(HPC_EINVAL, "tasks is NULL");
  if (deg < 1 || deg > (uint64_t)kMaxP) return fail(HPC_EINVAL, "row length %llu out of range",
                                                    (unsigned long long)deg);
  if (level < 2 || level > kMaxLevel) return fail(HPC_ELEVEL, "level %d outside 2..%d", level, kMaxLevel);
  if (ranks_per_task < 64 || ranks_per_task % 64 != 0 || ranks_per_task > (1 << 24))
    return fail(HPC_EINVAL, "ranks per task must be a multiple of 64 in [64, 2^24]");
  const std::vector<unsigned long long> tab = binom_table((size_t)deg);
  const long long t = row_task_count(binom(tab.data(), kBinomK, (int)deg - 1, level), (int)deg,
                                     ranks_per_task);
  if (t < 0)
    return fail(HPC_ELEVEL, "level %d: a row of %llu neighbours is beyond what one search plans",
                level, (unsigned long long)deg);
  *tasks = (int64_t)t;
  return HPC_OK;
}

extern "C" int hpc_wait_word(const uint64_t *word, uint64_t value, double limit_s) {
  if (!word) return fail(HPC_EINVAL, "word is NULL");
  return spin_wait_word(reinterpret_cast<const unsigned long long *>(word),
                        (unsigned long long)value, limit_s);
}

// ---------------------------------------------------------------------------
// synthetic SEM
// ---------------------------------------------------------------------------
namespace {
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    for (int k = 0; k < 4; ++k) s[k] = splitmix(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256** (Blackman & Vigna 2018)
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  void normal2(double &z0, double &z1) {  // Box-Muller
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    z0 = r * cos(th);
    z1 = r * sin(th);
  }
};
}  // namespace

extern "C" int hpc_sem_generate(int p, int n, double degree, double wmin,
                                double wmax, uint64_t seed, int fp32, double *x,
                                uint8_t *dag) {
  if (p < 2 || n < 1 || !x) return fail(HPC_EINVAL, "bad SEM arguments");
  Rng rng(seed);
  const double prob = degree / (double)(p - 1);
  std::vector<std::vector<std::pair<int, double>>> parents(p);
  if (dag) memset(dag, 0, (size_t)p * p);
  for (int i = 0; i < p; ++i)
    for (int j = i + 1; j < p; ++j) {
      if (rng.uniform() < prob) {
        double sign = rng.uniform() < 0.5 ? -1.0 : 1.0;
        double w = sign * (wmin + (wmax - wmin) * rng.uniform());
        parents[j].push_back({i, w});
        if (dag) { dag[(size_t)i * p + j] = 1; dag[(size_t)j * p + i] = 1; }
      }
    }
  for (int j = 0; j < p; ++j) {
    double *xj = x + (size_t)j * n;
    for (int k = 0; k < n; k += 2) {
      double z0, z1;
      rng.normal2(z0, z1);
      xj[k] = z0;
      if (k + 1 < n) xj[k + 1] = z1;
    }
    for (auto &pw : parents[j]) {
      const double *xi = x + (size_t)pw.first * n;
      for (int k = 0; k < n; ++k) xj[k] = __builtin_fma(pw.second, xi[k], xj[k]);
    }
  }
  if (fp32)
    for (size_t k = 0; k < (size_t)p * n; ++k) x[k] = (double)(float)x[k];
  return HPC_OK;
}

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
static InstHost &H(hpc_state *st) { return st->eng.inst[st->id]; }

extern "C" int hpc_state_destroy(hpc_state *st) {
  if (!st) return HPC_OK;
  engine_free(st->eng);
  delete st;
  return HPC_OK;
}

extern "C" int hpc_state_create(uint64_t p64, int observations, double alpha, int max_level,
                                int device, hpc_state **out) {
  if (!out) return fail(HPC_EINVAL, "out is NULL");
  *out = nullptr;
  if (p64 < 2 || p64 > (uint64_t)kMaxP)
    return fail(HPC_EINVAL, "p = %llu outside [2, %d]", (unsigned long long)p64, kMaxP);
  if (observations < 1) return fail(HPC_EINVAL, "observations must be >= 1");
  hpc_state *st = new hpc_state();
  int rc = engine_init(st->eng, device);
  if (!rc) rc = engine_add(st->eng, (int)p64, observations, alpha, max_level, &st->id);
  if (!rc && hipStreamSynchronize(st->eng.stream) != hipSuccess)
    rc = fail(HPC_EHIP, "state initialisation failed");
  if (rc) { hpc_state_destroy(st); return rc; }
  *out = st;
  return HPC_OK;
}

extern "C" int hpc_reset(hpc_state *st) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  return engine_reset(st->eng, st->id);
}

extern "C" int hpc_sync(hpc_state *st) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  HIPCHK(hipSetDevice(st->eng.device));
  HIPCHK(hipStreamSynchronize(st->eng.stream));
  return HPC_OK;
}

extern "C" int hpc_max_cond_size(const hpc_state *st) {
  return st ? st->eng.inst[st->id].mc : 0;
}

extern "C" int hpc_set_tuning(hpc_state *st, int param, int value) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  switch (param) {
    case HPC_TUNE_RANKS_PER_TASK:
      if (value < 64 || value % 64 != 0 || value > (1 << 24))
        return fail(HPC_EINVAL, "ranks per task must be a multiple of 64 in [64, 2^24]");
      st->eng.ranks_per_task = value;
      return HPC_OK;
    case HPC_TUNE_RECHECK_QUEUE:
      if (value < 0) return fail(HPC_EINVAL, "recheck queue length must be >= 0");
      st->eng.rchk_limit = (size_t)value;
      return HPC_OK;
    case HPC_TUNE_KERNEL_TIMING:
      if (value != 0 && value != 1) return fail(HPC_EINVAL, "kernel timing must be 0 or 1");
      st->eng.kernel_timing = value != 0;
      return HPC_OK;
    case HPC_TUNE_LDS_LIMIT:
      if (value < 0 || (size_t)value > kSearchLdsLimit)
        return fail(HPC_EINVAL, "LDS limit must be in [0, %zu] bytes", kSearchLdsLimit);
      st->eng.lds_limit = (size_t)value;
      return HPC_OK;
    case HPC_TUNE_WAIT_LIMIT:
      if (value < 0) return fail(HPC_EINVAL, "wait limit must be >= 0 seconds");
      st->eng.wait_limit_s = (double)value;
      return HPC_OK;
    case HPC_TUNE_QUEUE:
      if (value != 0 && value != 1) return fail(HPC_EINVAL, "queue must be 0 or 1");
      st->eng.use_queue = value != 0;
      return HPC_OK;
    case HPC_TUNE_GRAPH:
      if (value != 0 && value != 1) return fail(HPC_EINVAL, "graph must be 0 or 1");
      st->eng.use_graph = value != 0;
      if (!value) graph_drop(st->eng);
      return HPC_OK;
    default:
      return fail(HPC_EINVAL, "unknown tuning parameter %d", param);
  }
}

extern "C" int hpc_graph_stats(hpc_state *st, int64_t *replays, int64_t *captures,
                               int64_t *mismatches) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (replays) *replays = st->eng.graph_replays;
  if (captures) *captures = st->eng.graph_captures;
  if (mismatches) *mismatches = st->eng.graph_mismatches;
  return HPC_OK;
}

extern "C" int hpc_apply_tuning(hpc_state *st, const hpc_tune_entry *entries, int count) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  return engine_apply_tuning(st->eng, entries, count);
}

extern "C" int hpc_autotune(hpc_state *st, int max_level, hpc_tune_entry *out, int cap,
                            int *levels) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (H(st).fast) return fail(HPC_ESTATE, "the autotuner needs an exact correlation");
  std::vector<hpc_tune_entry> res;
  int rc = engine_autotune(st->eng, st->id, max_level, res);
  if (rc) return rc;
  for (int k = 0; out && k < (int)res.size() && k < cap; ++k) out[k] = res[k];
  if (levels) *levels = (int)res.size();
  return HPC_OK;
}

// packs the state's one instance at `level`
static int pack_at(hpc_state *st, int level) {
  return engine_pack(st->eng, {st->id}, {level});
}

// ---------------------------------------------------------------------------
// correlation
// ---------------------------------------------------------------------------
extern "C" int hpc_set_observations(hpc_state *st, const double *h_mat) {
  if (!st || !h_mat) return fail(HPC_EINVAL, "NULL argument");
  return engine_set_observations(st->eng, st->id, h_mat);
}

extern "C" int hpc_corr(hpc_state *st, int mode) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!H(st).have_x) return fail(HPC_ESTATE, "no observations uploaded");
  if (mode != HPC_CORR_CANON && mode != HPC_CORR_F32)
    return fail(HPC_EINVAL, "unknown correlation mode %d", mode);
  H(st).corr_mode = mode;
  int rc = engine_prepare_corr(st->eng, st->id);
  if (!rc) rc = pack_at(st, 0);
  if (!rc) rc = engine_corr(st->eng, &st->ms_moments, &st->ms_corr);
  if (rc) return rc;
  st->corr_timed = true;
  return HPC_OK;
}

extern "C" int hpc_corr_ms(hpc_state *st, double *ms_moments, double *ms_corr) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!st->corr_timed) return fail(HPC_ESTATE, "no correlation has run");
  if (ms_moments) *ms_moments = st->ms_moments;
  if (ms_corr) *ms_corr = st->ms_corr;
  return HPC_OK;
}

extern "C" int hpc_pmcc(hpc_state *st, const double *h_mat, int mode) {
  int rc = hpc_set_observations(st, h_mat);
  if (rc) return rc;
  rc = hpc_corr(st, mode);
  if (rc) return rc;
  return hpc_sync(st);
}

extern "C" int hpc_set_corr(hpc_state *st, const double *h_cor) {
  if (!st || !h_cor) return fail(HPC_EINVAL, "NULL argument");
  return engine_set_corr(st->eng, st->id, h_cor);
}

// ---------------------------------------------------------------------------
// levels
// ---------------------------------------------------------------------------
static int check_level(hpc_state *st, int level) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!H(st).have_cor) return fail(HPC_ESTATE, "no correlation matrix");
  if (level < 0) return fail(HPC_EINVAL, "level %d < 0", level);
  if (level > kMaxLevel)
    return fail(HPC_ELEVEL, "level %d exceeds the device maximum %d", level, kMaxLevel);
  if (level >= 1 && !H(st).compacted)
    return fail(HPC_ESTATE, "level %d needs hpc_compact first", level);
  return HPC_OK;
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" int hpc_compact(hpc_state *st, int *max_deg) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  int rc = pack_at(st, 1);
  float ms = 0.f;
  if (!rc) rc = engine_compact(st->eng, &ms);
  if (rc) return rc;
  if (max_deg) *max_deg = H(st).max_deg;
  return HPC_OK;
}

static void fill_stats(hpc_state *st, int level, hpc_level_stats &s) {
  const InstCounters &c = st->eng.h_ictr[0];
  s.level = level;
  s.max_deg = level == 0 ? H(st).p : H(st).max_deg;
  s.tests = (int64_t)c.tests;
  s.tests_max = (int64_t)c.tests_max;
  s.edges = (int64_t)c.edges;
  s.removed = (int64_t)c.removed;
  s.edges_left = H(st).undirected;
}

extern "C" int hpc_search_level(hpc_state *st, int level, const int32_t *rows, int nrows,
                                hpc_level_stats *stats) {
  int rc = check_level(st, level);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  hpc_level_stats s;
  memset(&s, 0, sizeof(s));
  float ms = 0.f;
  if (level == 0) {
    // level 0 tests and removes in one pass (testRowL0, gpuRowTests.cu:4-27)
    rc = pack_at(st, 0);
    if (!rc) rc = engine_level0(st->eng, &ms);
    if (rc) return rc;
    fill_stats(st, 0, s);
  } else {
    engine_note_level(st->eng, st->id, level);
    rc = pack_at(st, level);  // counters of the compaction stay
    if (!rc && H(st).max_deg - 1 >= level) rc = engine_search(st->eng, level, rows, nrows, &ms);
    if (rc) return rc;
    s.level = level;
    s.max_deg = H(st).max_deg;
    s.edges_left = H(st).undirected;
  }
  HIPCHK(hipMemcpyAsync(st->eng.h_ctr, st->eng.d_ctr, sizeof(Counters), hipMemcpyDeviceToHost,
                        st->eng.stream));
  HIPCHK(hipStreamSynchronize(st->eng.stream));
  s.pinv_fallbacks = (int64_t)st->eng.h_ctr->pinv_fallbacks;
  s.rechecks = (int64_t)st->eng.rechecks;
  s.ms_test = ms;
  s.ms_wall = ms_since(t0);
  if (stats) *stats = s;
  return HPC_OK;
}

extern "C" int hpc_commit_level(hpc_state *st, int level, hpc_level_stats *stats) {
  int rc = check_level(st, level);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  hpc_level_stats s;
  memset(&s, 0, sizeof(s));
  s.level = level;
  s.max_deg = level == 0 ? H(st).p : H(st).max_deg;
  if (level >= 1) {
    // the host share searches without the device table: the commit kernel
    // takes the level from a table packed here
    engine_note_level(st->eng, st->id, level);
    rc = pack_at(st, level);
    if (rc) return rc;
    float ms = 0.f;
    rc = engine_commit(st->eng, &ms);
    if (rc) return rc;
    fill_stats(st, level, s);
    s.ms_commit = ms;
  }
  s.edges_left = H(st).undirected;
  s.ms_wall = ms_since(t0);
  if (stats) *stats = s;
  return HPC_OK;
}

extern "C" int hpc_test_level(hpc_state *st, int level, const int32_t *rows, int nrows,
                              hpc_level_stats *stats) {
  hpc_level_stats a, b;
  int rc = hpc_search_level(st, level, rows, nrows, &a);
  if (rc) return rc;
  rc = hpc_commit_level(st, level, &b);
  if (rc) return rc;
  if (level >= 1) {
    a.tests = b.tests;
    a.tests_max = b.tests_max;
    a.edges = b.edges;
    a.removed = b.removed;
    a.edges_left = b.edges_left;
    a.ms_commit = b.ms_commit;
  }
  a.ms_wall += b.ms_wall;
  if (stats) *stats = a;
  return HPC_OK;
}

extern "C" int hpc_level(hpc_state *st, int level, hpc_level_stats *stats) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  auto t0 = std::chrono::steady_clock::now();
  float ms_c = 0.f;
  if (level >= 1) {
    int rc = pack_at(st, level);
    if (!rc) rc = engine_compact(st->eng, &ms_c);
    if (rc) return rc;
  }
  hpc_level_stats s;
  int rc = hpc_test_level(st, level, nullptr, 0, &s);
  if (rc) return rc;
  s.ms_compact = ms_c;
  s.ms_wall = ms_since(t0);
  if (stats) *stats = s;
  return HPC_OK;
}

extern "C" int hpc_run(hpc_state *st, int max_level, hpc_level_stats *stats, int cap,
                       int *levels_run) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  InstHost &h = H(st);
  if (!h.have_cor) return fail(HPC_ESTATE, "no correlation matrix");
  // levels 0..maxLevel (skeleton.cpp:172) from the current adjacency; with no
  // cap, until no row has enough neighbours.  The one-instance loop queues
  // each level's compaction behind the previous removal and waits once per
  // level; a level voided by a recheck-queue overflow and levels beyond the
  // device maximum continue in the engine's step, which waits twice.
  const int saved = h.max_level;
  if (max_level != -2) h.max_level = max_level;
  h.next_level = 0;
  h.finished = false;
  h.error = 0;
  const size_t first = h.hist.size();
  int rc = HPC_OK;
  rc = engine_run_single(st->eng, st->id);
  while (!h.finished && !rc) rc = engine_step(st->eng, {st->id});
  if (!rc && h.error) rc = fail(h.error, "level %d exceeds the device maximum %d", h.next_level - 1,
                               kMaxLevel);
  h.max_level = saved;
  const int ran = (int)(h.hist.size() - first);
  for (int k = 0; stats && k < ran && k < cap; ++k) stats[k] = h.hist[first + k];
  if (levels_run) *levels_run = ran;
  return rc;
}

// ---------------------------------------------------------------------------
// export (shared with the batched ABI through the engine instance)
// ---------------------------------------------------------------------------
namespace hpc {

int export_cor(Engine &e, int id, double *cor) {
  InstHost &h = e.inst[id];
  if (!h.have_cor) return fail(HPC_ESTATE, "no correlation matrix");
  HIPCHK(hipSetDevice(e.device));
  HIPCHK(hipMemcpyAsync(cor, h.cor, (size_t)h.p * h.p * sizeof(double), hipMemcpyDeviceToHost,
                        e.stream));
  HIPCHK(hipStreamSynchronize(e.stream));
  return HPC_OK;
}

int fetch_adj_bits(Engine &e, int id, std::vector<unsigned long long> &bits) {
  InstHost &h = e.inst[id];
  if (h.res_bits_valid) {
    bits = h.res_bits;
    return HPC_OK;
  }
  HIPCHK(hipSetDevice(e.device));
  bits.resize((size_t)h.p * h.W);
  HIPCHK(hipMemcpyAsync(bits.data(), h.adj, bits.size() * sizeof(bits[0]), hipMemcpyDeviceToHost,
                        e.stream));
  HIPCHK(hipStreamSynchronize(e.stream));
  if (h.finished) {  // final until the instance is packed or reset again
    h.res_bits = bits;
    h.res_bits_valid = true;
  }
  return HPC_OK;
}

// Records are appended in scheduling order: sorted by (i, j) here, by a
// counting sort on i (each pair has at most one record) and a sort on j
// within each i (insertion sort for short buckets).
static int sort_records(const InstHost &h, const int32_t *raw, size_t words,
                        std::vector<int32_t> &rec) {
  const int p = h.p;
  std::vector<uint32_t> first(p + 1, 0);
  size_t n = 0;
  for (size_t a = 0; a + 3 <= words; ++n) {
    const int i = raw[a], j = raw[a + 1], l = raw[a + 2];
    if (l < 1 || a + 3 + (size_t)l > words || i < 0 || i >= p || j < 0 || j >= p)
      return fail(HPC_EHIP, "corrupt sepset record at word %zu", a);
    ++first[i + 1];
    a += 3 + (size_t)l;
  }
  for (int i = 0; i < p; ++i) first[i + 1] += first[i];
  std::vector<size_t> at(n);  // record start words, bucketed by i
  {
    std::vector<uint32_t> fill(first.begin(), first.end() - 1);
    for (size_t a = 0; a + 3 <= words; a += 3 + (size_t)raw[a + 2]) at[fill[raw[a]]++] = a;
  }
  for (int i = 0; i < p; ++i) {
    if (first[i + 1] - first[i] > 32) {  // a long bucket: O(k log k)
      std::sort(at.begin() + first[i], at.begin() + first[i + 1],
                [&](size_t u, size_t v) { return raw[u + 1] < raw[v + 1]; });
      continue;
    }
    for (uint32_t x = first[i] + 1; x < first[i + 1]; ++x) {
      const size_t v = at[x];
      uint32_t y = x;
      while (y > first[i] && raw[at[y - 1] + 1] > raw[v + 1]) {
        at[y] = at[y - 1];
        --y;
      }
      at[y] = v;
    }
  }
  rec.clear();
  rec.reserve(words);
  for (size_t a : at) rec.insert(rec.end(), raw + a, raw + a + 3 + raw[a + 2]);
  return HPC_OK;
}

// Pinned host memory that batch results land in, one per device, grown on
// demand and kept for the process (shared by every engine under a mutex): a
// copy into pageable memory, above all freshly allocated memory, took
// milliseconds for a few MB; a pinned buffer per engine costs its pinning on
// every new batch.
namespace {
struct ResultStage {
  std::mutex mu;
  char *buf = nullptr;
  size_t cap = 0;
};
constexpr int kResultDevices = 64;
ResultStage g_result_stage[kResultDevices];
}  // namespace

int engine_prefetch_results(Engine &e, const std::vector<int> &ids) {
  HIPCHK(hipSetDevice(e.device));
  struct Item {
    int id;
    size_t bits_at, rec_at;
  };
  std::vector<Item> items;
  std::vector<CopyOut> copies;
  size_t bytes = 0;
  for (int id : ids) {
    if (id < 0 || id >= (int)e.inst.size()) continue;
    const InstHost &h = e.inst[id];
    if (!h.used || !h.finished || h.error || !h.adj) continue;
    if (h.res_bits_valid && h.res_rec_valid) continue;
    Item it{id, bytes, 0};
    copies.push_back(CopyOut{h.adj, (unsigned long long)h.p * h.W * 8, (unsigned long long)bytes});
    bytes += (size_t)h.p * h.W * 8;
    it.rec_at = bytes;
    if (h.rec_used) copies.push_back(CopyOut{h.rec, (unsigned long long)h.rec_used * 4, (unsigned long long)bytes});
    bytes += (h.rec_used * 4 + 7) & ~(size_t)7;
    items.push_back(it);
  }
  if (items.empty()) return HPC_OK;
  if (bytes > e.d_stage_cap) {
    if (e.d_stage) HIPCHK(hipFree(e.d_stage));
    e.d_stage = nullptr;
    e.d_stage_cap = 0;
    const size_t cap = bytes + bytes / 4;
    HIPCHK(hipMalloc((void **)&e.d_stage, cap));
    e.d_stage_cap = cap;
  }
  if (copies.size() > e.cap_copies) {
    if (e.d_copies) HIPCHK(hipFree(e.d_copies));
    e.d_copies = nullptr;
    e.cap_copies = 0;
    const size_t cap = copies.size() + copies.size() / 4 + 16;
    HIPCHK(hipMalloc((void **)&e.d_copies, cap * sizeof(CopyOut)));
    e.cap_copies = cap;
  }
  // one kernel gathers every block into one device buffer, one copy brings
  // it to host memory: one upload of the table, one launch, one copy, one
  // wait (a device-to-host copy per block measured about 80 us each)
  HIPCHK(hipMemcpyAsync(e.d_copies, copies.data(), copies.size() * sizeof(CopyOut),
                        hipMemcpyHostToDevice, e.stream));
  launch_copy_out(e.stream, e.d_copies, (int)copies.size(), e.d_stage);
  HIPCHK(hipGetLastError());
  std::unique_lock<std::mutex> pin_lock;
  char *host = nullptr;
  if (e.device >= 0 && e.device < kResultDevices) {
    ResultStage &rs = g_result_stage[e.device];
    pin_lock = std::unique_lock<std::mutex>(rs.mu);
    if (bytes > rs.cap) {
      if (rs.buf) (void)hipHostFree(rs.buf);
      rs.buf = nullptr;
      rs.cap = 0;
      const size_t cap = ((bytes + bytes / 4) + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
      if (hipHostMalloc((void **)&rs.buf, cap, hipHostMallocDefault) == hipSuccess) rs.cap = cap;
      else rs.buf = nullptr;  // (then ordinary host memory below)
    }
    host = rs.buf;
  }
  if (!host) {
    pin_lock = std::unique_lock<std::mutex>();
    e.h_stage.resize(bytes);
    host = e.h_stage.data();
  }
  HIPCHK(hipMemcpyAsync(host, e.d_stage, bytes, hipMemcpyDeviceToHost, e.stream));
  HIPCHK(hipStreamSynchronize(e.stream));
  const char *stage = host;
  // the instances' unpacking (the record sort is most of it) on a few host
  // threads; each touches only its own instances
  std::atomic<size_t> next{0};
  std::atomic<int> bad{HPC_OK};
  auto unpack = [&]() {
    for (size_t k; (k = next.fetch_add(1)) < items.size();) {
      const Item &it = items[k];
      InstHost &h = e.inst[it.id];
      const unsigned long long *bw = reinterpret_cast<const unsigned long long *>(stage + it.bits_at);
      h.res_bits.assign(bw, bw + (size_t)h.p * h.W);
      h.res_bits_valid = true;
      std::vector<int32_t> rec;
      if (sort_records(h, reinterpret_cast<const int32_t *>(stage + it.rec_at), h.rec_used, rec)) {
        bad = HPC_EHIP;
        continue;
      }
      h.res_rec.swap(rec);
      h.res_rec_valid = true;
    }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nthr = std::min<size_t>(std::min<size_t>(8, hw), (items.size() + 15) / 16);
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nthr; ++t) pool.emplace_back(unpack);
  unpack();
  for (auto &t : pool) t.join();
  if (bad != HPC_OK) return fail(HPC_EHIP, "corrupt sepset records in a batch result");
  return HPC_OK;
}

int fetch_records(Engine &e, int id, std::vector<int32_t> &rec) {
  InstHost &h = e.inst[id];
  if (h.res_rec_valid) {
    rec = h.res_rec;
    return HPC_OK;
  }
  HIPCHK(hipSetDevice(e.device));
  std::vector<int32_t> raw(h.rec_used);
  if (h.rec_used) {
    HIPCHK(hipMemcpyAsync(raw.data(), h.rec, h.rec_used * sizeof(int32_t), hipMemcpyDeviceToHost,
                          e.stream));
    HIPCHK(hipStreamSynchronize(e.stream));
  }
  const int rc = sort_records(h, raw.data(), raw.size(), rec);
  if (rc) return rc;
  if (h.finished) {
    h.res_rec = rec;
    h.res_rec_valid = true;
  }
  return HPC_OK;
}


void dense_adj(const std::vector<unsigned long long> &bits, int p, int32_t *adj) {
  const int W = (p + 63) / 64;
  for (int i = 0; i < p; ++i) {
    const unsigned long long *row = bits.data() + (size_t)i * W;
    int32_t *out = adj + (size_t)i * p;
    for (int j = 0; j < p; ++j) out[j] = (int32_t)((row[j >> 6] >> (j & 63)) & 1ull);
  }
}

void dense_sepsets(const std::vector<unsigned long long> &bits, const std::vector<int32_t> &rec,
                   int p, int32_t *sep, int mc) {
  const int W = (p + 63) / 64;
  const size_t plane = (size_t)p * p;
  std::fill(sep, sep + plane * mc, -1);
  // every absent pair was removed at level 0 unless a record says otherwise
  for (int i = 0; i < p; ++i)
    for (int j = i + 1; j < p; ++j)
      if (!((bits[(size_t)i * W + (j >> 6)] >> (j & 63)) & 1ull)) sep[((size_t)i * p + j) * mc] = -2;
  for (size_t a = 0; a + 3 <= rec.size(); a += 3 + (size_t)rec[a + 2]) {
    const int i = rec[a], j = rec[a + 1], l = rec[a + 2];
    int32_t *slot = sep + ((size_t)i * p + j) * mc;
    slot[0] = -1;
    for (int m = 0; m < l && m < mc; ++m) slot[m] = rec[a + 3 + m];
  }
}

int export_adj(Engine &e, int id, int32_t *adj) {
  const InstHost &h = e.inst[id];
  if (h.res_bits_valid) {
    dense_adj(h.res_bits, h.p, adj);
    return HPC_OK;
  }
  std::vector<unsigned long long> bits;
  int rc = fetch_adj_bits(e, id, bits);
  if (rc) return rc;
  dense_adj(bits, e.inst[id].p, adj);
  return HPC_OK;
}

int export_sepsets(Engine &e, int id, int32_t *sep, int mc) {
  std::vector<unsigned long long> bits;
  std::vector<int32_t> rec;
  int rc = fetch_adj_bits(e, id, bits);
  if (!rc) rc = fetch_records(e, id, rec);
  if (rc) return rc;
  dense_sepsets(bits, rec, e.inst[id].p, sep, mc);
  return HPC_OK;
}

int export_sepset_list(Engine &e, int id, int32_t *out, int64_t cap, int64_t *words) {
  if (!out) {
    // the size query: the sorted list has the records' word count
    if (words) *words = (int64_t)e.inst[id].rec_used;
    return HPC_OK;
  }
  const InstHost &h = e.inst[id];
  if (h.res_rec_valid) {
    if (words) *words = (int64_t)h.res_rec.size();
    if ((int64_t)h.res_rec.size() > cap)
      return fail(HPC_EINVAL, "sepset list needs %zu words, cap is %lld", h.res_rec.size(),
                  (long long)cap);
    std::copy(h.res_rec.begin(), h.res_rec.end(), out);
    return HPC_OK;
  }
  std::vector<int32_t> rec;
  int rc = fetch_records(e, id, rec);
  if (rc) return rc;
  if (words) *words = (int64_t)rec.size();
  if (out) {
    if ((int64_t)rec.size() > cap)
      return fail(HPC_EINVAL, "sepset list needs %zu words, cap is %lld", rec.size(), (long long)cap);
    std::copy(rec.begin(), rec.end(), out);
  }
  return HPC_OK;
}

}  // namespace hpc

extern "C" int hpc_get_cor(hpc_state *st, double *cor) {
  if (!st || !cor) return fail(HPC_EINVAL, "NULL argument");
  return export_cor(st->eng, st->id, cor);
}

extern "C" int hpc_get_adj(hpc_state *st, int32_t *adj) {
  if (!st || !adj) return fail(HPC_EINVAL, "NULL argument");
  return export_adj(st->eng, st->id, adj);
}

extern "C" int hpc_get_sepsets(hpc_state *st, int32_t *sep, int mc) {
  if (!st || !sep || mc < 1) return fail(HPC_EINVAL, "bad argument");
  return export_sepsets(st->eng, st->id, sep, mc);
}

extern "C" int hpc_get_sepset_list(hpc_state *st, int32_t *out, int64_t cap, int64_t *words) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  return export_sepset_list(st->eng, st->id, out, cap, words);
}

extern "C" int hpc_get_compact(hpc_state *st, int32_t *out) {
  if (!st || !out) return fail(HPC_EINVAL, "NULL argument");
  Engine &e = st->eng;
  InstHost &h = H(st);
  const int p = h.p;
  if (!h.ever_compacted) {  // state.cu:32-38: all ones off the diagonal
    for (size_t a = 0; a < (size_t)p * p; ++a) out[a] = 1;
    for (int i = 0; i < p; ++i) out[(size_t)i * p + i] = 0;
    return HPC_OK;
  }
  HIPCHK(hipSetDevice(e.device));
  std::vector<int32_t> rp(p + 1), nb(e.n_edges > 0 ? e.n_edges : 1);
  HIPCHK(hipMemcpyAsync(rp.data(), e.d_rowptr, (p + 1) * sizeof(int32_t), hipMemcpyDeviceToHost,
                        e.stream));
  if (e.n_edges > 0)
    HIPCHK(hipMemcpyAsync(nb.data(), e.d_nbr, e.n_edges * sizeof(int32_t), hipMemcpyDeviceToHost,
                          e.stream));
  HIPCHK(hipStreamSynchronize(e.stream));
  for (int i = 0; i < p; ++i) {
    int32_t *row = out + (size_t)i * p;
    memset(row, 0, sizeof(int32_t) * p);
    const int d = rp[i + 1] - rp[i];
    for (int k = 0; k < d; ++k) row[k] = nb[rp[i] + k];
    row[p - 1] = d;
  }
  return HPC_OK;
}

extern "C" int hpc_get_first_pass(hpc_state *st, int64_t *fp, int32_t *rowptr, int64_t *n_edges) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  Engine &e = st->eng;
  const bool have = H(st).ever_compacted;
  if (n_edges) *n_edges = have ? e.n_edges : 0;
  if (!fp || !have) return HPC_OK;
  HIPCHK(hipSetDevice(e.device));
  if (e.n_edges > 0)
    HIPCHK(hipMemcpyAsync(fp, e.d_fp, e.n_edges * sizeof(int64_t), hipMemcpyDeviceToHost, e.stream));
  if (rowptr)
    HIPCHK(hipMemcpyAsync(rowptr, e.d_rowptr, (H(st).p + 1) * sizeof(int32_t),
                          hipMemcpyDeviceToHost, e.stream));
  HIPCHK(hipStreamSynchronize(e.stream));
  return HPC_OK;
}
