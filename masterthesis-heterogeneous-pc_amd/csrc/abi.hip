// C ABI of libheterogpc_hip.so (include/heterogpc_hip.h): state management,
// the level loop and result export.  Host code only; kernels live in
// corr.hip, graph.hip and search.hip.
#pragma clang fp contract(off)

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "hpc_internal.hpp"

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

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(HPC_EHIP, "%s failed: %s (%s:%d)", #call,                 \
                  hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

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

// ---------------------------------------------------------------------------
// Fisher-z decision boundary
// ---------------------------------------------------------------------------
// calcPValue (cpuInd.cpp:8-14): 2 * Q(z) = erfc(z / sqrt 2).
static double fisher_p(double r_in, int n) {
  double r = isnan(r_in) ? 0.0 : fmin(0.9999999, fabs(r_in));
  double absz = sqrt(n - 3.0) * 0.5 * log1p(2 * r / (1 - r));
  return erfc(absz / 1.4142135623730951);
}

extern "C" double hpc_fisher_r_crit(int n, double alpha) {
  double hi = 0.9999999, lo = 0.0;
  if (fisher_p(hi, n) >= alpha) return INFINITY;
  if (!(fisher_p(lo, n) >= alpha)) return -1.0;
  // positive doubles are ordered like their bit patterns: bisect on bits
  uint64_t a, b;
  memcpy(&a, &lo, 8);
  memcpy(&b, &hi, 8);
  while (b - a > 1) {
    uint64_t m = a + (b - a) / 2;
    double r;
    memcpy(&r, &m, 8);
    if (fisher_p(r, n) >= alpha) a = m; else b = m;
  }
  double out;
  memcpy(&out, &a, 8);
  return out;
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

static int alloc(void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess)
    return fail(HPC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return HPC_OK;
}

#define ALLOC(ptr, bytes)                                     \
  do {                                                        \
    int rc_ = alloc((void **)&(ptr), (bytes));                \
    if (rc_) { hpc_state_destroy(st); return rc_; }           \
  } while (0)

extern "C" int hpc_state_destroy(hpc_state *st) {
  if (!st) return HPC_OK;
  hipSetDevice(st->device);
  if (st->stream) hipStreamSynchronize(st->stream);
  void *bufs[] = {st->d_x, st->d_z, st->d_mean, st->d_sd, st->d_nanm,
                  st->d_hasnan, st->d_cor, st->d_adj, st->d_deg, st->d_rowptr,
                  st->d_nbr, st->d_erow, st->d_rowmask, st->d_fp, st->d_sep,
                  st->d_rmlev, st->d_binom, st->d_ctr, st->d_row_tasks,
                  st->d_task_off, st->d_tasks};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (st->h_ctr) hipHostFree(st->h_ctr);
  if (st->d_scatter_idx) hipFree(st->d_scatter_idx);
  if (st->d_scatter_val) hipFree(st->d_scatter_val);
  if (st->copy_stream) {
    hipStreamSynchronize(st->copy_stream);
    hipStreamDestroy(st->copy_stream);
  }
  for (auto &e : st->ev)
    if (e) hipEventDestroy(e);
  if (st->stream) hipStreamDestroy(st->stream);
  delete st;
  return HPC_OK;
}

static int upload_binom(hpc_state *st) {
  const int rows = st->p + 1;
  std::vector<unsigned long long> tab((size_t)rows * kBinomK, 0ull);
  for (int nn = 0; nn < rows; ++nn) {
    tab[(size_t)nn * kBinomK] = 1ull;
    for (int k = 1; k < kBinomK && k <= nn; ++k) {
      unsigned long long a = tab[(size_t)(nn - 1) * kBinomK + k - 1];
      unsigned long long b = (k <= nn - 1) ? tab[(size_t)(nn - 1) * kBinomK + k] : 0ull;
      unsigned long long s = a + b;
      tab[(size_t)nn * kBinomK + k] = (s < a) ? ~0ull : s;  // saturate
    }
  }
  HIPCHK(hipMemcpy(st->d_binom, tab.data(), tab.size() * sizeof(tab[0]),
                   hipMemcpyHostToDevice));
  st->h_binom = std::move(tab);
  return HPC_OK;
}

extern "C" int hpc_reset(hpc_state *st) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  HIPCHK(hipSetDevice(st->device));
  launch_adj_init(st->stream, st->d_adj, st->p, st->W);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(st->d_rmlev, 0xFF, (size_t)st->p * st->p, st->stream));
  st->compacted = false;
  st->ever_compacted = false;
  st->n_edges = 0;
  st->max_deg = st->p;
  st->undirected = (int64_t)st->p * (st->p - 1) / 2;
  return HPC_OK;
}

extern "C" int hpc_state_create(uint64_t p64, int observations, double alpha,
                                int max_level, int device, hpc_state **out) {
  if (!out) return fail(HPC_EINVAL, "out is NULL");
  *out = nullptr;
  if (p64 < 2 || p64 > 46340)
    return fail(HPC_EINVAL, "p = %llu outside [2, 46340]", (unsigned long long)p64);
  if (observations < 1) return fail(HPC_EINVAL, "observations must be >= 1");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(HPC_EINVAL, "Only %d GPUs available for kernel execution", ndev);
  HIPCHK(hipSetDevice(device));
  hpc_state *st = new hpc_state();
  st->device = device;
  st->p = (int)p64;
  st->n = observations;
  st->npad = (observations + 63) & ~63;
  st->W = (st->p + 63) / 64;
  st->alpha = alpha;
  st->max_level = max_level;
  st->mc = max_level > 1 ? max_level : (max_level < 0 ? 4 : 1);
  st->r_crit = hpc_fisher_r_crit(observations, alpha);
  if (hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess) {
    delete st;
    return fail(HPC_EHIP, "hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&st->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    hpc_state_destroy(st);
    return fail(HPC_EHIP, "hipStreamCreate failed");
  }
  for (auto &e : st->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      hpc_state_destroy(st);
      return fail(HPC_EHIP, "hipEventCreate failed");
    }
  }
  const size_t p = st->p, pp = p * p;
  ALLOC(st->d_cor, pp * sizeof(double));
  ALLOC(st->d_adj, p * st->W * sizeof(unsigned long long));
  ALLOC(st->d_deg, p * sizeof(int32_t));
  ALLOC(st->d_rowptr, (p + 1) * sizeof(int32_t));
  ALLOC(st->d_nbr, pp * sizeof(int32_t));
  ALLOC(st->d_erow, pp * sizeof(int32_t));
  ALLOC(st->d_rowmask, p * sizeof(int32_t));
  ALLOC(st->d_fp, pp * sizeof(int64_t));
  ALLOC(st->d_sep, (size_t)st->mc * pp * sizeof(int32_t));
  ALLOC(st->d_rmlev, pp);
  ALLOC(st->d_row_tasks, p * sizeof(long long));
  ALLOC(st->d_task_off, p * sizeof(long long));
  ALLOC(st->d_binom, (p + 1) * kBinomK * sizeof(unsigned long long));
  ALLOC(st->d_ctr, sizeof(Counters));
  if (hipHostMalloc((void **)&st->h_ctr, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
    hpc_state_destroy(st);
    return fail(HPC_ENOMEM, "hipHostMalloc failed");
  }
  int rc = upload_binom(st);
  if (!rc) rc = hpc_reset(st);
  if (!rc && hipStreamSynchronize(st->stream) != hipSuccess)
    rc = fail(HPC_EHIP, "state initialisation failed");
  if (rc) { hpc_state_destroy(st); return rc; }
  *out = st;
  return HPC_OK;
}

extern "C" int hpc_sync(hpc_state *st) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  HIPCHK(hipStreamSynchronize(st->stream));
  return HPC_OK;
}

extern "C" int hpc_max_cond_size(const hpc_state *st) { return st ? st->mc : 0; }

extern "C" int hpc_set_tuning(hpc_state *st, int param, int value) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  switch (param) {
    case HPC_TUNE_RANKS_PER_TASK:
      if (value < 64 || value % 64 != 0 || value > (1 << 24))
        return fail(HPC_EINVAL, "ranks per task must be a multiple of 64 in [64, 2^24]");
      st->ranks_per_task = value;
      return HPC_OK;
    default:
      return fail(HPC_EINVAL, "unknown tuning parameter %d", param);
  }
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// ---------------------------------------------------------------------------
// correlation
// ---------------------------------------------------------------------------
extern "C" int hpc_set_observations(hpc_state *st, const double *h_mat) {
  if (!st || !h_mat) return fail(HPC_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(st->device));
  const size_t p = st->p;
  if (!st->d_x) {
    hpc_state *s2 = st;
    int rc;
    if ((rc = alloc((void **)&s2->d_x, p * st->n * sizeof(double)))) return rc;
    if ((rc = alloc((void **)&s2->d_z, p * st->npad * sizeof(double)))) return rc;
    if ((rc = alloc((void **)&s2->d_nanm, p * st->npad))) return rc;
    if ((rc = alloc((void **)&s2->d_mean, p * sizeof(double)))) return rc;
    if ((rc = alloc((void **)&s2->d_sd, p * sizeof(double)))) return rc;
    if ((rc = alloc((void **)&s2->d_hasnan, p * sizeof(int32_t)))) return rc;
  }
  HIPCHK(hipMemcpyAsync(st->d_x, h_mat, p * st->n * sizeof(double),
                        hipMemcpyHostToDevice, st->stream));
  // NaN cells need the masked pairwise kernel (corHelper.cu:101 skips them)
  bool any_nan = false;
  for (size_t k = 0; k < p * st->n; ++k)
    if (h_mat[k] != h_mat[k]) { any_nan = true; break; }
  st->x_has_nan = any_nan;
  HIPCHK(hipStreamSynchronize(st->stream));
  st->have_x = true;
  return HPC_OK;
}

extern "C" int hpc_corr(hpc_state *st, int mode) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!st->have_x) return fail(HPC_ESTATE, "no observations uploaded");
  if (mode != HPC_CORR_CANON) return fail(HPC_EINVAL, "unknown correlation mode %d", mode);
  HIPCHK(hipSetDevice(st->device));
  HIPCHK(hipEventRecord(st->ev[4], st->stream));
  launch_moments(st->stream, st->d_x, st->p, st->n, st->npad, st->d_mean,
                 st->d_sd, st->d_z, st->d_nanm, st->d_hasnan);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(st->ev[5], st->stream));
  if (st->x_has_nan)
    launch_corr_canon(st->stream, st->d_z, st->d_nanm, st->d_hasnan, st->p, st->n,
                      st->npad, st->d_cor);
  else
    launch_corr_tiled(st->stream, st->d_z, st->p, st->n, st->npad, st->d_cor);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(st->ev[6], st->stream));
  st->have_cor = true;
  st->h_cor_valid = false;
  st->corr_timed = true;
  return HPC_OK;
}

extern "C" int hpc_corr_ms(hpc_state *st, double *ms_moments, double *ms_corr) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!st->corr_timed) return fail(HPC_ESTATE, "no correlation has run");
  HIPCHK(hipEventSynchronize(st->ev[6]));
  if (ms_moments) *ms_moments = elapsed(st->ev[4], st->ev[5]);
  if (ms_corr) *ms_corr = elapsed(st->ev[5], st->ev[6]);
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
  HIPCHK(hipSetDevice(st->device));
  HIPCHK(hipMemcpyAsync(st->d_cor, h_cor, (size_t)st->p * st->p * sizeof(double),
                        hipMemcpyHostToDevice, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  st->have_cor = true;
  st->h_cor_valid = false;
  return HPC_OK;
}

// ---------------------------------------------------------------------------
// levels
// ---------------------------------------------------------------------------
static int read_counters(hpc_state *st) {
  HIPCHK(hipMemcpyAsync(st->h_ctr, st->d_ctr, sizeof(Counters),
                        hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return HPC_OK;
}


static Views views(hpc_state *st) {
  Views v;
  v.cor = st->d_cor;
  v.rowptr = st->d_rowptr;
  v.nbr = st->d_nbr;
  v.erow = st->d_erow;
  v.binom = st->d_binom;
  v.fp = st->d_fp;
  v.ctr = st->d_ctr;
  v.p = st->p;
  v.max_deg = st->max_deg;
  v.r_crit = st->r_crit;
  return v;
}

static int compact_impl(hpc_state *st, float *ms) {
  HIPCHK(hipMemsetAsync(st->d_ctr, 0, sizeof(Counters), st->stream));
  HIPCHK(hipEventRecord(st->ev[0], st->stream));
  launch_compact(st->stream, st->d_adj, st->p, st->W, st->d_deg, st->d_rowptr,
                 st->d_nbr, st->d_erow, st->d_ctr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(st->ev[1], st->stream));
  int rc = read_counters(st);
  if (rc) return rc;
  st->n_edges = (int64_t)st->h_ctr->n_edges;
  st->max_deg = st->h_ctr->max_deg;
  // every ordered edge of the new snapshot starts with no passing set
  if (st->n_edges > 0)
    HIPCHK(hipMemsetAsync(st->d_fp, 0xFF, (size_t)st->n_edges * sizeof(int64_t), st->stream));
  st->compacted = true;
  st->ever_compacted = true;
  if (ms) *ms = elapsed(st->ev[0], st->ev[1]);
  return HPC_OK;
}

extern "C" int hpc_compact(hpc_state *st, int *max_deg) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  HIPCHK(hipSetDevice(st->device));
  int rc = compact_impl(st, nullptr);
  if (rc) return rc;
  if (max_deg) *max_deg = st->max_deg;
  return HPC_OK;
}

static int grow_sepsets(hpc_state *st, int mc) {
  const size_t plane = (size_t)st->p * st->p;
  int32_t *ns = nullptr;
  int rc = alloc((void **)&ns, (size_t)mc * plane * sizeof(int32_t));
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(ns, st->d_sep, (size_t)st->mc * plane * sizeof(int32_t),
                        hipMemcpyDeviceToDevice, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  HIPCHK(hipFree(st->d_sep));
  st->d_sep = ns;
  st->mc = mc;
  return HPC_OK;
}

static int check_level(hpc_state *st, int level) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (!st->have_cor) return fail(HPC_ESTATE, "no correlation matrix");
  if (level < 0) return fail(HPC_EINVAL, "level %d < 0", level);
  if (level > kMaxLevel)
    return fail(HPC_ELEVEL, "level %d exceeds the device maximum %d", level, kMaxLevel);
  if (level >= 1 && !st->compacted)
    return fail(HPC_ESTATE, "level %d needs hpc_compact first", level);
  return HPC_OK;
}

static void counters_to_stats(const hpc_state *st, hpc_level_stats &s) {
  const Counters &c = *st->h_ctr;
  s.tests = (int64_t)c.tests;
  s.tests_max = (int64_t)c.tests_max;
  s.edges = (int64_t)c.edges;
  s.pinv_fallbacks = (int64_t)c.pinv_fallbacks;
}

extern "C" int hpc_search_level(hpc_state *st, int level, const int32_t *rows,
                                int nrows, hpc_level_stats *stats) {
  int rc = check_level(st, level);
  if (rc) return rc;
  HIPCHK(hipSetDevice(st->device));
  auto t0 = std::chrono::steady_clock::now();
  hpc_level_stats s;
  memset(&s, 0, sizeof(s));
  s.level = level;
  s.max_deg = level == 0 ? st->p : st->max_deg;
  HIPCHK(hipMemsetAsync(st->d_ctr, 0, sizeof(Counters), st->stream));
  HIPCHK(hipEventRecord(st->ev[2], st->stream));
  if (level == 0) {
    // level 0 tests and removes in one pass (testRowL0, gpuRowTests.cu:4-27)
    launch_level0(st->stream, st->d_cor, st->d_adj, st->p, st->W, st->r_crit, st->d_ctr);
  } else {
    if (level > st->mc) {
      rc = grow_sepsets(st, level);
      if (rc) return rc;
    }
    const int32_t *mask = nullptr;
    if (rows) {
      std::vector<int32_t> m(st->p, 0);
      for (int k = 0; k < nrows; ++k) {
        if (rows[k] < 0 || rows[k] >= st->p) return fail(HPC_EINVAL, "row %d out of range", rows[k]);
        m[rows[k]] = 1;
      }
      HIPCHK(hipMemcpyAsync(st->d_rowmask, m.data(), st->p * sizeof(int32_t),
                            hipMemcpyHostToDevice, st->stream));
      mask = st->d_rowmask;
    }
    if (st->max_deg - 1 >= level && (!rows || nrows > 0)) {
      const Views v = views(st);
      if (level == 1) {
        launch_search_l1(st->stream, v, st->n_edges, mask);
      } else {
        const int rpt = st->ranks_per_task;
        launch_task_plan(st->stream, level, v, rpt, mask, st->d_row_tasks, st->d_task_off);
        HIPCHK(hipGetLastError());
        rc = read_counters(st);
        if (rc) return rc;
        const size_t n_tasks = (size_t)st->h_ctr->n_tasks;
        if (n_tasks > st->task_cap) {
          if (st->d_tasks) HIPCHK(hipFree(st->d_tasks));
          st->d_tasks = nullptr;
          st->task_cap = 0;
          const size_t cap = n_tasks + n_tasks / 4 + 1024;
          rc = alloc((void **)&st->d_tasks, cap * sizeof(unsigned long long));
          if (rc) return rc;
          st->task_cap = cap;
        }
        launch_task_fill(st->stream, level, v, rpt, st->d_row_tasks, st->d_task_off, st->d_tasks);
        HIPCHK(hipGetLastError());
        launch_search_ln(st->stream, level, v, st->d_tasks, (long long)n_tasks, rpt);
      }
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(st->ev[3], st->stream));
  rc = read_counters(st);
  if (rc) return rc;
  s.ms_test = elapsed(st->ev[2], st->ev[3]);
  counters_to_stats(st, s);
  if (level == 0) {
    s.removed = (int64_t)st->h_ctr->removed;
    st->undirected -= s.removed;
    if (s.removed) st->compacted = false;
  }
  s.edges_left = st->undirected;
  s.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
  s.max_deg = level == 0 ? st->p : st->max_deg;
  if (level >= 1 && st->max_deg - 1 >= level) {
    HIPCHK(hipSetDevice(st->device));
    HIPCHK(hipMemsetAsync(st->d_ctr, 0, sizeof(Counters), st->stream));
    HIPCHK(hipEventRecord(st->ev[3], st->stream));
    launch_commit(st->stream, level, views(st), st->n_edges, st->d_adj, st->W, st->d_sep,
                  st->d_rmlev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(st->ev[4], st->stream));
    rc = read_counters(st);
    if (rc) return rc;
    s.ms_commit = elapsed(st->ev[3], st->ev[4]);
    counters_to_stats(st, s);
    s.pinv_fallbacks = 0;
    s.removed = (int64_t)st->h_ctr->removed;
    st->undirected -= s.removed;
    if (s.removed) st->compacted = false;
  }
  s.edges_left = st->undirected;
  s.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = s;
  return HPC_OK;
}

extern "C" int hpc_test_level(hpc_state *st, int level, const int32_t *rows,
                              int nrows, hpc_level_stats *stats) {
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
    HIPCHK(hipSetDevice(st->device));
    int rc = compact_impl(st, &ms_c);
    if (rc) return rc;
  }
  hpc_level_stats s;
  int rc = hpc_test_level(st, level, nullptr, 0, &s);
  if (rc) return rc;
  s.ms_compact = ms_c;
  s.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = s;
  return HPC_OK;
}

extern "C" int hpc_run(hpc_state *st, int max_level, hpc_level_stats *stats,
                       int cap, int *levels_run) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (max_level == -2) max_level = st->max_level;
  int ran = 0;
  for (int l = 0; max_level < 0 || l <= max_level; ++l) {
    if (l >= 1) {
      // stop once no row has enough neighbours (skeleton.cpp:172 runs all
      // levels up to maxLevel; with no cap there is nothing left to test)
      HIPCHK(hipSetDevice(st->device));
      auto t0 = std::chrono::steady_clock::now();
      float ms_c = 0.f;
      int rc = compact_impl(st, &ms_c);
      if (rc) return rc;
      if (max_level < 0 && st->max_deg - 1 < l) break;
      if (l > kMaxLevel && st->max_deg - 1 >= l)
        return fail(HPC_ELEVEL, "level %d exceeds the device maximum %d", l, kMaxLevel);
      hpc_level_stats s;
      if (l <= kMaxLevel) {
        rc = hpc_test_level(st, l, nullptr, 0, &s);
        if (rc) return rc;
      } else {
        memset(&s, 0, sizeof(s));
        s.level = l;
        s.max_deg = st->max_deg;
        s.edges_left = st->undirected;
      }
      s.ms_compact = ms_c;
      s.ms_wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (stats && ran < cap) stats[ran] = s;
    } else {
      hpc_level_stats s;
      int rc = hpc_test_level(st, 0, nullptr, 0, &s);
      if (rc) return rc;
      if (stats && ran < cap) stats[ran] = s;
    }
    ++ran;
  }
  if (levels_run) *levels_run = ran;
  return HPC_OK;
}

// ---------------------------------------------------------------------------
// export
// ---------------------------------------------------------------------------
extern "C" int hpc_get_cor(hpc_state *st, double *cor) {
  if (!st || !cor) return fail(HPC_EINVAL, "NULL argument");
  if (!st->have_cor) return fail(HPC_ESTATE, "no correlation matrix");
  HIPCHK(hipSetDevice(st->device));
  HIPCHK(hipMemcpyAsync(cor, st->d_cor, (size_t)st->p * st->p * sizeof(double),
                        hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return HPC_OK;
}

static int fetch_adj(hpc_state *st, std::vector<unsigned long long> &bits) {
  bits.resize((size_t)st->p * st->W);
  HIPCHK(hipMemcpyAsync(bits.data(), st->d_adj, bits.size() * sizeof(bits[0]),
                        hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return HPC_OK;
}

extern "C" int hpc_get_adj(hpc_state *st, int32_t *adj) {
  if (!st || !adj) return fail(HPC_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(st->device));
  std::vector<unsigned long long> bits;
  int rc = fetch_adj(st, bits);
  if (rc) return rc;
  const int p = st->p, W = st->W;
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < p; ++j)
      adj[(size_t)i * p + j] = (int32_t)((bits[(size_t)i * W + (j >> 6)] >> (j & 63)) & 1ull);
  return HPC_OK;
}

extern "C" int hpc_get_compact(hpc_state *st, int32_t *out) {
  if (!st || !out) return fail(HPC_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(st->device));
  const int p = st->p;
  if (!st->ever_compacted) {  // state.cu:32-38: all ones off the diagonal
    for (size_t a = 0; a < (size_t)p * p; ++a) out[a] = 1;
    for (int i = 0; i < p; ++i) out[(size_t)i * p + i] = 0;
    return HPC_OK;
  }
  std::vector<int32_t> rp(p + 1), nb(st->n_edges > 0 ? st->n_edges : 1);
  HIPCHK(hipMemcpyAsync(rp.data(), st->d_rowptr, (p + 1) * sizeof(int32_t),
                        hipMemcpyDeviceToHost, st->stream));
  if (st->n_edges > 0)
    HIPCHK(hipMemcpyAsync(nb.data(), st->d_nbr, st->n_edges * sizeof(int32_t),
                          hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  for (int i = 0; i < p; ++i) {
    int32_t *row = out + (size_t)i * p;
    memset(row, 0, sizeof(int32_t) * p);
    const int d = rp[i + 1] - rp[i];
    for (int k = 0; k < d; ++k) row[k] = nb[rp[i] + k];
    row[p - 1] = d;
  }
  return HPC_OK;
}

extern "C" int hpc_get_sepsets(hpc_state *st, int32_t *sep, int mc) {
  if (!st || !sep || mc < 1) return fail(HPC_EINVAL, "bad argument");
  HIPCHK(hipSetDevice(st->device));
  const int p = st->p, W = st->W;
  const size_t plane = (size_t)p * p;
  std::vector<unsigned long long> bits;
  int rc = fetch_adj(st, bits);
  if (rc) return rc;
  std::vector<int8_t> lev(plane);
  std::vector<int32_t> dev((size_t)st->mc * plane);
  HIPCHK(hipMemcpyAsync(lev.data(), st->d_rmlev, plane, hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipMemcpyAsync(dev.data(), st->d_sep, dev.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  for (size_t a = 0; a < plane * mc; ++a) sep[a] = -1;
  for (int i = 0; i < p; ++i)
    for (int j = i + 1; j < p; ++j) {
      if ((bits[(size_t)i * W + (j >> 6)] >> (j & 63)) & 1ull) continue;
      int32_t *slot = sep + ((size_t)i * p + j) * mc;
      const int l = lev[(size_t)i * p + j];
      if (l <= 0) { slot[0] = -2; continue; }  // removed at level 0
      for (int k = 0; k < l && k < mc; ++k) slot[k] = dev[(size_t)k * plane + (size_t)i * p + j];
    }
  return HPC_OK;
}

extern "C" int hpc_get_first_pass(hpc_state *st, int64_t *fp, int32_t *rowptr,
                                  int64_t *n_edges) {
  if (!st) return fail(HPC_EINVAL, "state is NULL");
  if (n_edges) *n_edges = st->ever_compacted ? st->n_edges : 0;
  if (!fp || !st->ever_compacted) return HPC_OK;
  HIPCHK(hipSetDevice(st->device));
  if (st->n_edges > 0)
    HIPCHK(hipMemcpyAsync(fp, st->d_fp, st->n_edges * sizeof(int64_t),
                          hipMemcpyDeviceToHost, st->stream));
  if (rowptr)
    HIPCHK(hipMemcpyAsync(rowptr, st->d_rowptr, (st->p + 1) * sizeof(int32_t),
                          hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return HPC_OK;
}
