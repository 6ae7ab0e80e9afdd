// Host side of the device state: instances, the packed layout of a phase and
// the phases themselves (correlation, level 0, compaction, search, commit).
// The single-instance ABI (abi.hip), the batched ABI and the server
// (batch.hip) all drive these functions.
#pragma clang fp contract(off)

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {

// calcPValue (cpuInd.cpp:8-14): 2 * Q(z) = erfc(z / sqrt 2).
static double fisher_p(double r_in, int n) {
  double r = isnan(r_in) ? 0.0 : fmin(0.9999999, fabs(r_in));
  double absz = sqrt(n - 3.0) * 0.5 * log1p(2 * r / (1 - r));
  return erfc(absz / 1.4142135623730951);
}

double fisher_r_crit(int n, double alpha) {
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

int spin_wait_word(const unsigned long long *word, unsigned long long value, double limit_s) {
  return spin_wait(word, value, limit_s, []() -> int { return 0; });
}

// ---- slab cache ------------------------------------------------------------
// Arena slabs of engines that have ended, kept per device for the next engine
// (batches and servers often come one after another): a hipMalloc / hipFree
// pair of a 512 MiB slab costs about a millisecond.  At most kSlabCacheBytes
// per device are kept; an allocation that fails empties the device's cache
// and tries once more.
namespace {
constexpr size_t kSlabCacheBytes = (size_t)8 << 30;
constexpr int kSlabCacheDevices = 64;
struct SlabCache {
  std::mutex mu;
  std::multimap<size_t, void *> free;  // size -> base
  size_t bytes = 0;
};
SlabCache g_slabs[kSlabCacheDevices];

// a cached slab of at least size bytes (and at most twice that); size
// becomes its own
void *slab_take(int device, size_t &size) {
  if (device < 0 || device >= kSlabCacheDevices) return nullptr;
  SlabCache &c = g_slabs[device];
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.free.lower_bound(size);
  if (it == c.free.end() || it->first > 2 * size) return nullptr;
  void *p = it->second;
  size = it->first;
  c.bytes -= it->first;
  c.free.erase(it);
  return p;
}

// true when the cache keeps the slab (its last user has finished with it)
bool slab_put(int device, void *base, size_t size) {
  if (device < 0 || device >= kSlabCacheDevices) return false;
  SlabCache &c = g_slabs[device];
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.bytes + size > kSlabCacheBytes) return false;
  c.free.emplace(size, base);
  c.bytes += size;
  return true;
}

// frees the current device's cached slabs
void slab_flush() {
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kSlabCacheDevices) return;
  SlabCache &c = g_slabs[device];
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto &kv : c.free) (void)hipFree(kv.second);
  c.free.clear();
  c.bytes = 0;
}

// hipMalloc on the current device; after a failure, once more with the
// device's slab cache emptied (the failure's error is cleared, so a later
// hipGetLastError does not report it)
hipError_t cached_malloc(void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e == hipSuccess) return e;
  (void)hipGetLastError();
  slab_flush();
  e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}
}  // namespace

static int dalloc(void **ptr, size_t bytes) {
  *ptr = nullptr;
  hipError_t e = cached_malloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return hpc_fail(HPC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return HPC_OK;
}

template <class T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// ---- arena ----------------------------------------------------------------
constexpr size_t kArenaAlign = 256;

static int arena_alloc(Engine &e, size_t bytes, void **out) {
  Arena &A = e.arena;
  const size_t need = std::max<size_t>((bytes + kArenaAlign - 1) & ~(kArenaAlign - 1), kArenaAlign);
  for (size_t si = 0; si < A.slabs.size(); ++si) {
    auto &fr = A.slabs[si].free;
    for (auto it = fr.begin(); it != fr.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, len = it->second;
      fr.erase(it);
      if (len > need) fr.emplace(off + need, len - need);
      *out = A.slabs[si].base + off;
      A.live[*out] = {(int)si, need};
      A.used += need;
      return HPC_OK;
    }
  }
  Arena::Slab sl;
  sl.size = std::max(need, A.slab_min);
  void *base = slab_take(e.device, sl.size);
  if (!base) {
    hipError_t err = cached_malloc(&base, sl.size);
    if (err != hipSuccess)
      return hpc_fail(HPC_ENOMEM, "hipMalloc(%zu) failed: %s", sl.size, hipGetErrorString(err));
  }
  sl.base = (char *)base;
  if (sl.size > need) sl.free.emplace(need, sl.size - need);
  A.slabs.push_back(std::move(sl));
  A.reserved += A.slabs.back().size;
  *out = base;
  A.live[base] = {(int)A.slabs.size() - 1, need};
  A.used += need;
  return HPC_OK;
}

static void arena_free(Engine &e, const void *ptr) {
  if (!ptr) return;
  Arena &A = e.arena;
  auto it = A.live.find(ptr);
  if (it == A.live.end()) return;
  const int si = it->second.first;
  const size_t len = it->second.second;
  A.live.erase(it);
  A.used -= len;
  auto &fr = A.slabs[si].free;
  size_t off = (size_t)((const char *)ptr - A.slabs[si].base), l = len;
  auto nx = fr.lower_bound(off);
  if (nx != fr.end() && nx->first == off + l) { l += nx->second; nx = fr.erase(nx); }
  if (nx != fr.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) { off = pv->first; l += pv->second; fr.erase(pv); }
  }
  fr.emplace(off, l);
}

template <class T>
static int ialloc(Engine &e, T *&p, size_t count) {
  void *v = nullptr;
  int rc = arena_alloc(e, count * sizeof(T), &v);
  p = (T *)v;
  return rc;
}

template <class T>
static void ifree(Engine &e, T *&p) {
  arena_free(e, (const void *)p);
  p = nullptr;
}

// grows a device array to hold n elements (contents dropped)
template <class T>
static int grow(T *&p, size_t &cap, size_t n) {
  if (n <= cap && p) return HPC_OK;
  dfree(p);
  cap = 0;
  const size_t want = std::max<size_t>(n + n / 4, 64);
  int rc = dalloc((void **)&p, want * sizeof(T));
  if (rc) return rc;
  cap = want;
  return HPC_OK;
}

static int ensure_inst_cap(Engine &e, size_t K);

int engine_init(Engine &e, int device) {
  int ndev = 0;
  HPC_HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return hpc_fail(HPC_EINVAL, "Only %d GPUs available for kernel execution", ndev);
  HPC_HIPCHK(hipSetDevice(device));
  e.device = device;
  HPC_HIPCHK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
  HPC_HIPCHK(hipStreamCreateWithFlags(&e.copy_stream, hipStreamNonBlocking));
  // timing events only: without the system-scope fence a recorded event
  // costs no cache writeback between the kernels it separates (counters reach
  // the host through the kernels' own system-scope releases)
  for (auto &ev : e.ev) HPC_HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
  for (auto &ev : e.sev) HPC_HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
  HPC_HIPCHK(hipEventCreateWithFlags(&e.fp_ready, hipEventDisableTiming));
  HPC_HIPCHK(hipHostMalloc((void **)&e.h_seq, 64, hipHostMallocMapped | hipHostMallocCoherent));
  HPC_HIPCHK(hipHostGetDevicePointer((void **)&e.d_seq, e.h_seq, 0));
  __atomic_store_n(e.h_seq, 0ull, __ATOMIC_RELEASE);
  e.seq = 0;
  // the task-queue head starts at 0; every queue launch leaves it at 0
  int rc = dalloc((void **)&e.d_qhead, 64);
  if (rc) return rc;
  HPC_HIPCHK(hipMemset(e.d_qhead, 0, 64));
  return ensure_inst_cap(e, 1);
}

// Bytes of the z-score buffers the instance holds.
static size_t zbytes(const InstHost &h) {
  size_t b = 0;
  if (h.z) b += (size_t)h.p * h.npad * 8;
  if (h.z32) b += (size_t)h.p * h.npad * 4;
  if (h.zc) b += zc_size(h.p, h.npad) * 8;
  return b;
}

static void free_obs(Engine &e, InstHost &h) {
  h.bytes -= std::min(h.bytes, zbytes(h) + (h.x ? (size_t)h.p * h.n * 8 : 0) +
                                   (h.nanm ? (size_t)h.p * h.npad : 0));
  ifree(e, h.x); ifree(e, h.z); ifree(e, h.nanm); ifree(e, h.z32); ifree(e, h.zc);
}

static void free_blocks(Engine &e, InstHost &h) {
  free_obs(e, h);
  ifree(e, h.mean); ifree(e, h.sd); ifree(e, h.cor); ifree(e, h.adj); ifree(e, h.rec);
  h.rec_cap = h.rec_used = 0;
}

void engine_free(Engine &e) {
  (void)hipSetDevice(e.device);
  if (e.stream) (void)hipStreamSynchronize(e.stream);
  if (e.copy_stream) (void)hipStreamSynchronize(e.copy_stream);
  e.inst.clear();
  // (the streams have drained: the slabs are idle)
  for (auto &sl : e.arena.slabs)
    if (sl.base && !slab_put(e.device, sl.base, sl.size)) (void)hipFree(sl.base);
  e.arena = Arena();
  dfree(e.d_inst); dfree(e.d_row_inst); dfree(e.d_deg); dfree(e.d_rowptr);
  dfree(e.d_nbr); dfree(e.d_erow); dfree(e.d_rowmask); dfree(e.d_fp); dfree(e.d_ccor);
  dfree(e.d_erev);
  dfree(e.d_row_tasks);
  dfree(e.d_task_off); dfree(e.d_binom); dfree(e.d_prefix); dfree(e.d_ctr);
  e.d_ictr = nullptr;
  graph_drop(e);
  if (e.h_slots) (void)hipHostFree(e.h_slots);
  e.h_slots = nullptr;
  e.d_slots = nullptr;
  if (e.d_slots_local) (void)hipFree(e.d_slots_local);
  e.d_slots_local = nullptr;
  if (e.d_stage) (void)hipFree(e.d_stage);
  e.d_stage = nullptr;
  e.d_stage_cap = 0;
  e.h_stage = std::vector<char>();
  if (e.d_copies) (void)hipFree(e.d_copies);
  e.d_copies = nullptr;
  e.cap_copies = 0;
  dfree(e.d_scatter_idx); dfree(e.d_scatter_val); dfree(e.d_rchk); dfree(e.d_qhead);
  if (e.h_qdone) (void)hipHostFree(e.h_qdone);
  e.h_qdone = e.d_qdone = nullptr;
  e.cap_qdone = 0;
  dfree(e.d_blk); dfree(e.d_l0slot);
  if (e.h_ctr) (void)hipHostFree(e.h_ctr);
  if (e.h_ctr2) (void)hipHostFree(e.h_ctr2);
  e.h_ctr = e.h_ctr_dev = e.h_ctr2 = e.h_ctr2_dev = nullptr;
  e.h_ictr = nullptr;
  if (e.lev_ev_ready)
    for (auto &row : e.lev_ev)
      for (auto &ev : row)
        if (ev) (void)hipEventDestroy(ev);
  e.lev_ev_ready = false;
  if (e.h_seq) (void)hipHostFree(e.h_seq);
  e.h_seq = e.d_seq = nullptr;
  for (auto &ev : e.ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : e.sev)
    if (ev) (void)hipEventDestroy(ev);
  if (e.fp_ready) (void)hipEventDestroy(e.fp_ready);
  e.fp_ready = nullptr;
  if (e.stream) (void)hipStreamDestroy(e.stream);
  if (e.copy_stream) (void)hipStreamDestroy(e.copy_stream);
  e.stream = e.copy_stream = nullptr;
}

std::vector<unsigned long long> binom_table(size_t rows) {
  std::vector<unsigned long long> tab(rows * kBinomK, 0ull);
  for (size_t nn = 0; nn < rows; ++nn) {
    tab[nn * kBinomK] = 1ull;
    for (int k = 1; k < kBinomK && (size_t)k <= nn; ++k) {
      unsigned long long a = tab[(nn - 1) * kBinomK + k - 1];
      unsigned long long b = ((size_t)k <= nn - 1) ? tab[(nn - 1) * kBinomK + k] : 0ull;
      unsigned long long s = a + b;
      // Pascal's rule; a saturated operand or a carry saturates the sum
      tab[nn * kBinomK + k] = (s < a || a == ~0ull || b == ~0ull) ? ~0ull : s;
    }
  }
  return tab;
}

// saturating C(m, k), k <= kMaxLevel, for m < rows
static int ensure_binom(Engine &e, int rows) {
  if ((size_t)rows <= e.cap_binom) return HPC_OK;
  const size_t R = std::max<size_t>(rows, 64);
  std::vector<unsigned long long> tab = binom_table(R);
  dfree(e.d_binom);
  int rc = dalloc((void **)&e.d_binom, tab.size() * sizeof(tab[0]));
  if (rc) return rc;
  HPC_HIPCHK(hipMemcpy(e.d_binom, tab.data(), tab.size() * sizeof(tab[0]), hipMemcpyHostToDevice));
  e.h_binom = std::move(tab);
  e.cap_binom = R;
  return HPC_OK;
}

// Sepset width before any level has run: the level cap (at least 1), or 4
// with no cap (the reference's default maxCondSize).
static int initial_mc(int max_level) {
  return max_level > 1 ? std::min(max_level, kMaxLevel) : (max_level < 0 ? 4 : 1);
}

int engine_add(Engine &e, int p, int n, double alpha, int max_level, int *id) {
  if (p < 2 || p > kMaxP) return hpc_fail(HPC_EINVAL, "p = %d outside [2, %d]", p, kMaxP);
  if (n < 1) return hpc_fail(HPC_EINVAL, "observations must be >= 1");
  HPC_HIPCHK(hipSetDevice(e.device));
  int slot = -1;
  for (size_t k = 0; k < e.inst.size(); ++k)
    if (!e.inst[k].used) { slot = (int)k; break; }
  if (slot < 0) { slot = (int)e.inst.size(); e.inst.emplace_back(); }
  InstHost &h = e.inst[slot];
  h = InstHost();
  h.used = true;
  h.p = p;
  h.n = n;
  h.npad = (n + 63) & ~63;
  h.W = (p + 63) / 64;
  h.alpha = alpha;
  h.max_level = max_level;
  h.mc = initial_mc(max_level);
  h.r_crit = fisher_r_crit(n, alpha);
  const size_t pp = (size_t)p * p;
  int rc = 0;
  if (!rc) rc = ialloc(e, h.cor, pp);
  if (!rc) rc = ialloc(e, h.adj, (size_t)p * h.W);
  if (!rc) rc = ialloc(e, h.mean, (size_t)p);
  if (!rc) rc = ialloc(e, h.sd, (size_t)p);
  if (!rc) rc = ensure_binom(e, p + 1);
  if (rc) { free_blocks(e, h); h.used = false; return rc; }
  h.bytes = pp * 8 + (size_t)p * h.W * 8 + 16 * (size_t)p;
  rc = engine_reset(e, slot);
  if (rc) { free_blocks(e, h); h.used = false; return rc; }
  *id = slot;
  return HPC_OK;
}

void engine_remove(Engine &e, int id) {
  if (id < 0 || id >= (int)e.inst.size() || !e.inst[id].used) return;
  (void)hipSetDevice(e.device);
  free_blocks(e, e.inst[id]);
  e.inst[id] = InstHost();
}

void engine_drop_results(InstHost &h) {
  h.res_bits_valid = h.res_rec_valid = false;
  h.res_bits.clear();
  h.res_rec.clear();
}

int engine_reset(Engine &e, int id) {
  InstHost &h = e.inst[id];
  engine_drop_results(h);
  HPC_HIPCHK(hipSetDevice(e.device));
  launch_adj_init(e.stream, h.adj, h.p, h.W);
  HPC_HIPCHK(hipGetLastError());
  h.fresh = true;
  h.rec_used = 0;
  h.mc = initial_mc(h.max_level);
  h.next_level = 0;
  h.finished = false;
  h.compacted = false;
  h.ever_compacted = false;
  h.max_deg = h.p;
  h.undirected = (int64_t)h.p * (h.p - 1) / 2;
  return HPC_OK;
}

int engine_alloc_observations(Engine &e, int id, double **x) {
  InstHost &h = e.inst[id];
  HPC_HIPCHK(hipSetDevice(e.device));
  int rc = 0;
  const bool fresh = !h.x;
  if (!h.x) rc = ialloc(e, h.x, (size_t)h.p * h.n);
  if (!rc && !h.nanm) rc = ialloc(e, h.nanm, (size_t)h.p * h.npad);
  if (rc) return rc;
  if (fresh) h.bytes += (size_t)h.p * h.n * 8 + (size_t)h.p * h.npad;
  *x = h.x;
  return HPC_OK;
}

void engine_observations_ready(Engine &e, int id) {
  InstHost &h = e.inst[id];
  h.have_x = true;
  h.have_cor = false;
}

// Host-to-device copies of observations.  A copy from pageable memory goes
// through the runtime's own staging at about 27 GB/s on the MI355X host; here
// a large one goes through a ring of pinned buffers per device, each filled by
// kStageThreads host threads (the caller and a pool of helpers) while the
// others' DMAs run.  The ring carries over from one call to the next: a
// buffer is refilled once its last use has ended, whichever call and stream
// issued it, so a caller that does not wait for its copy (a later kernel on
// the same stream orders itself after it) fills the next while the DMA runs.
// Arrays of doubles that FP32 holds exactly (float data widened to double,
// as sensor and benchmark files often are) cross the bus at half width: the
// fill narrows each slice and checks every value round-trips, the DMA moves
// the floats into a device buffer and widen_kernel writes the doubles, bit
// for bit the caller's (a slice with any other value, NaN included, is sent
// as doubles and so is the rest of that copy).  Buffers and helpers are
// created on first use and kept for the process (one set per device, shared
// by every engine under a mutex; never destroyed, so no helper outlives its
// synchronisation objects at exit).
namespace {
constexpr size_t kStageChunk = (size_t)4 << 20;
constexpr size_t kStageMin = (size_t)1 << 20;
constexpr size_t kPoolMin = (size_t)256 << 10;  // smaller slices: the caller alone
constexpr int kStageBufs = 4;
constexpr int kStageThreads = 8;
constexpr int kStageDevices = 64;

// narrows n doubles to floats; true when every one round-trips exactly
// ((double)(float)v == v keeps the bits too: the one pair of equal doubles
// with different bits, +0 and -0, narrows to +0f and -0f)
bool narrow_exact(float *to, const double *from, size_t n) {
  bool ok = true;
  for (size_t i = 0; i < n; ++i) {
    const float f = (float)from[i];
    to[i] = f;
    ok &= (double)f == from[i];
  }
  return ok;
}

// kStageThreads - 1 helper threads that each take one part of a slice.
struct CopyPool {
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  unsigned long long gen = 0;
  int pending = 0;
  const void *src = nullptr;
  void *dst = nullptr;
  size_t count = 0, per = 0;  // bytes, or doubles when narrowing
  bool narrow = false, exact = true;
  CopyPool() {
    for (int t = 1; t < kStageThreads; ++t) std::thread([this, t] { run(t); }).detach();
  }
  bool part(int t, const void *s, void *d, size_t n, size_t each, bool nar) {
    const size_t a = std::min(n, (size_t)t * each), b = std::min(n, a + each);
    if (a >= b) return true;
    if (nar) return narrow_exact(static_cast<float *>(d) + a, static_cast<const double *>(s) + a, b - a);
    memcpy(static_cast<char *>(d) + a, static_cast<const char *>(s) + a, b - a);
    return true;
  }
  void run(int t) {
    unsigned long long seen = 0;
    for (;;) {
      const void *s;
      void *d;
      size_t n, each;
      bool nar;
      {
        std::unique_lock<std::mutex> lk(m);
        cv_work.wait(lk, [&] { return gen != seen; });
        seen = gen;
        s = src; d = dst; n = count; each = per; nar = narrow;
      }
      const bool ok = part(t, s, d, n, each, nar);
      std::lock_guard<std::mutex> lk(m);
      exact = exact && ok;
      if (--pending == 0) cv_done.notify_one();
    }
  }
  // one slice with every helper; returns when all parts are done (and, when
  // narrowing, whether every value round-tripped)
  bool job(void *to, const void *from, size_t n, bool nar) {
    const size_t each = (n + kStageThreads - 1) / kStageThreads;
    {
      std::lock_guard<std::mutex> lk(m);
      src = from; dst = to; count = n; per = each; narrow = nar; exact = true;
      pending = kStageThreads - 1;
      ++gen;
    }
    cv_work.notify_all();
    const bool ok = part(0, from, to, n, each, nar);
    std::unique_lock<std::mutex> lk(m);
    cv_done.wait(lk, [&] { return pending == 0; });
    return ok && exact;
  }
};

struct Staging {
  std::mutex mu;
  char *buf[kStageBufs] = {};    // pinned host slices
  float *wide[kStageBufs] = {};  // device copies of narrowed slices
  hipEvent_t ev[kStageBufs] = {};
  bool busy[kStageBufs] = {};  // the buffer's last use was issued (ev marks its end)
  unsigned next = 0;           // the ring's next buffer
  CopyPool *pool = nullptr;
  bool ready = false, failed = false;
};
Staging g_staging[kStageDevices];

__global__ void widen_kernel(const float *__restrict__ src, double *__restrict__ dst, long long n) {
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n) {
    const float4 v = *reinterpret_cast<const float4 *>(src + i);
    double2 *d = reinterpret_cast<double2 *>(dst + i);
    d[0] = make_double2((double)v.x, (double)v.y);
    d[1] = make_double2((double)v.z, (double)v.w);
  } else {
    for (long long k = i; k < n; ++k) dst[k] = (double)src[k];
  }
}

bool staging_init(Staging &g, int device) {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(device) != hipSuccess) return false;
  bool ok = true;
  for (int k = 0; k < kStageBufs && ok; ++k)
    ok = hipHostMalloc((void **)&g.buf[k], kStageChunk, hipHostMallocDefault) == hipSuccess &&
         hipMalloc((void **)&g.wide[k], kStageChunk) == hipSuccess &&
         hipEventCreateWithFlags(&g.ev[k], hipEventDisableTiming) == hipSuccess;
  if (ok) g.pool = new CopyPool();
  (void)hipSetDevice(cur);
  return ok;
}
}  // namespace

static hipError_t upload_staged(int device, void *dst, const void *src, size_t bytes,
                                hipStream_t s, bool complete, bool doubles) {
  auto direct = [&]() {
    hipError_t err = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    return err == hipSuccess ? hipStreamSynchronize(s) : err;
  };
  if (bytes == 0) return hipSuccess;
  if ((complete && bytes < kStageMin) || device < 0 || device >= kStageDevices) return direct();
  Staging &g = g_staging[device];
  std::lock_guard<std::mutex> lk(g.mu);
  if (!g.ready && !g.failed) {
    g.ready = staging_init(g, device);
    g.failed = !g.ready;  // (no pinned memory: the runtime's own path from now on)
  }
  if (!g.ready) return direct();
  const char *from = static_cast<const char *>(src);
  char *to = static_cast<char *>(dst);
  // narrowing needs whole doubles and 16-byte aligned stores of them
  bool narrow = doubles && bytes % 8 == 0 && ((uintptr_t)dst & 15) == 0;
  const size_t slice_dbl = kStageChunk / sizeof(float);
  for (size_t off = 0; off < bytes;) {
    const unsigned b = g.next;
    g.next = (g.next + 1) % kStageBufs;
    // the buffer's (and its device copy's) last use has ended
    if (g.busy[b]) {
      hipError_t err = hipEventSynchronize(g.ev[b]);
      if (err != hipSuccess) return err;
      g.busy[b] = false;
    }
    hipError_t err = hipSuccess;
    if (narrow) {
      const size_t cnt = std::min(slice_dbl, (bytes - off) / 8);
      const double *sd = reinterpret_cast<const double *>(from + off);
      float *hf = reinterpret_cast<float *>(g.buf[b]);
      const bool exact = cnt * 8 >= kPoolMin ? g.pool->job(hf, sd, cnt, true)
                                             : narrow_exact(hf, sd, cnt);
      if (exact) {
        err = hipMemcpyAsync(g.wide[b], hf, cnt * sizeof(float), hipMemcpyHostToDevice, s);
        if (err == hipSuccess) {
          const long long quads = ((long long)cnt + 3) / 4;
          hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((quads + 255) / 256)), dim3(256), 0,
                             s, g.wide[b], reinterpret_cast<double *>(to + off), (long long)cnt);
          err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(g.ev[b], s);
        if (err != hipSuccess) return err;
        g.busy[b] = true;
        off += cnt * 8;
        continue;
      }
      narrow = false;  // this slice and the rest as doubles (the buffer is idle)
    }
    const size_t len = std::min(kStageChunk, bytes - off);
    if (len >= kPoolMin)
      (void)g.pool->job(g.buf[b], from + off, len, false);
    else
      memcpy(g.buf[b], from + off, len);
    err = hipMemcpyAsync(to + off, g.buf[b], len, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipEventRecord(g.ev[b], s);
    if (err != hipSuccess) return err;
    g.busy[b] = true;
    off += len;
  }
  return complete ? hipStreamSynchronize(s) : hipSuccess;
}

hipError_t upload_h2d(int device, void *dst, const void *src, size_t bytes, hipStream_t s,
                      bool complete) {
  return upload_staged(device, dst, src, bytes, s, complete, false);
}

hipError_t upload_f64_h2d(int device, double *dst, const double *src, size_t count,
                          hipStream_t s, bool complete) {
  return upload_staged(device, dst, src, count * sizeof(double), s, complete, true);
}

int engine_set_observations(Engine &e, int id, const double *h_x) {
  InstHost &h = e.inst[id];
  double *dst = nullptr;
  int rc = engine_alloc_observations(e, id, &dst);
  if (rc) return rc;
  const size_t cnt = (size_t)h.p * h.n;
  // only kernels on e.stream read the observations: no wait for the copy
  HPC_HIPCHK(upload_f64_h2d(e.device, dst, h_x, cnt, e.stream, false));
  engine_observations_ready(e, id);
  return HPC_OK;
}

int engine_set_corr(Engine &e, int id, const double *h_cor) {
  InstHost &h = e.inst[id];
  HPC_HIPCHK(hipSetDevice(e.device));
  HPC_HIPCHK(
      upload_h2d(e.device, h.cor, h_cor, (size_t)h.p * h.p * sizeof(double), e.stream, true));
  h.have_cor = true;
  h.fast = false;
  e.h_cor_valid = false;
  h.cor_nan = 0;
  const size_t pp = (size_t)h.p * h.p;
  for (size_t t = 0; t < pp; ++t)
    if (h_cor[t] != h_cor[t]) {
      h.cor_nan = 1;
      break;
    }
  return HPC_OK;
}

void engine_note_level(Engine &e, int id, int level) {
  InstHost &h = e.inst[id];
  if (level > h.mc) h.mc = std::min(level, kMaxLevel);
}

// Host-side preparation of an instance's level: room for its sepset records
// (ensure_rec), and the failure a test asked for (hpc_batch_inject_failure).
static int ensure_rec(Engine &e, InstHost &h, int level);
static int prepare_level(Engine &e, InstHost &h, int level) {
  if (h.inject_level >= 0 && h.inject_level == level)
    return hpc_fail(HPC_ENOMEM, "injected host failure at level %d", level);
  return level >= 1 ? ensure_rec(e, h, level) : HPC_OK;
}

// Room for the worst case of every level the run can still reach at once:
// each remaining edge removed, its record at most 3 + deepest words (an edge
// is removed once); contents kept.  The deepest level is the cap, or without
// one the last level whose tests an endpoint's degree still allows (an edge
// is tested at level l only when an endpoint keeps l neighbours besides the
// other; degrees only fall, so the last counted maximum bounds them).  The
// buffer then grows once per run, not between levels: a growth is a device
// copy per instance.
static int ensure_rec(Engine &e, InstHost &h, int level) {
  const int reach = std::min(h.max_deg - 1, kMaxLevel);
  if (h.rec && level > reach) return HPC_OK;  // no test, so no record, at this level
  const int deepest = h.max_level >= 0 ? std::min(std::max(h.max_level, level), kMaxLevel)
                                       : std::max(level, reach);
  const size_t need = h.rec_used + (size_t)std::max<int64_t>(h.undirected, 0) * (3 + deepest);
  if (need <= h.rec_cap && h.rec) return HPC_OK;
  const size_t cap = std::max(need, 2 * h.rec_cap);
  int32_t *nr = nullptr;
  int rc = ialloc(e, nr, cap);
  if (rc) return rc;
  if (h.rec_used)
    HPC_HIPCHK(hipMemcpyAsync(nr, h.rec, h.rec_used * sizeof(int32_t), hipMemcpyDeviceToDevice,
                              e.stream));
  ifree(e, h.rec);
  h.bytes += (cap - h.rec_cap) * sizeof(int32_t);
  h.rec = nr;
  h.rec_cap = cap;
  return HPC_OK;
}

// Capacity groups: arrays of one group are reallocated together (contents
// dropped; every phase rewrites them).
static int ensure_inst_cap(Engine &e, size_t K) {
  if (K <= e.cap_inst && e.d_inst) return HPC_OK;
  const size_t c = std::max<size_t>(K + K / 4, 8);
  dfree(e.d_inst); dfree(e.d_ctr); dfree(e.d_prefix); dfree(e.d_l0slot);
  if (e.h_ctr) (void)hipHostFree(e.h_ctr);
  if (e.h_ctr2) (void)hipHostFree(e.h_ctr2);
  e.h_ctr = e.h_ctr_dev = e.h_ctr2 = e.h_ctr2_dev = nullptr;
  e.d_ictr = e.h_ictr = nullptr;
  e.cap_inst = 0;
  const size_t block = kIctrOffset + c * sizeof(InstCounters);
  int rc = dalloc((void **)&e.d_inst, c * sizeof(Inst));
  if (!rc) rc = dalloc((void **)&e.d_ctr, block);
  if (!rc) rc = dalloc((void **)&e.d_prefix, 4 * (c + 1) * sizeof(long long));
  const size_t slot_bytes = c * kL0Slots * kL0SlotWords * sizeof(unsigned long long);
  if (!rc) rc = dalloc((void **)&e.d_l0slot, slot_bytes);
  // fine-grained, so the publication kernel's stores reach host memory
  if (!rc && hipHostMalloc((void **)&e.h_ctr, block, hipHostMallocMapped | hipHostMallocCoherent) !=
                 hipSuccess) {
    e.h_ctr = nullptr;
    rc = hpc_fail(HPC_ENOMEM, "hipHostMalloc failed");
  }
  if (!rc && hipHostMalloc((void **)&e.h_ctr2, block, hipHostMallocMapped | hipHostMallocCoherent) !=
                 hipSuccess) {
    e.h_ctr2 = nullptr;
    rc = hpc_fail(HPC_ENOMEM, "hipHostMalloc failed");
  }
  if (!rc && (hipHostGetDevicePointer((void **)&e.h_ctr_dev, e.h_ctr, 0) != hipSuccess ||
              hipHostGetDevicePointer((void **)&e.h_ctr2_dev, e.h_ctr2, 0) != hipSuccess))
    rc = hpc_fail(HPC_EHIP, "hipHostGetDevicePointer failed");
  if (rc) return rc;
  e.d_ictr = reinterpret_cast<InstCounters *>(reinterpret_cast<char *>(e.d_ctr) + kIctrOffset);
  e.h_ictr = reinterpret_cast<InstCounters *>(reinterpret_cast<char *>(e.h_ctr) + kIctrOffset);
  // a fresh block holds no counts
  HPC_HIPCHK(hipMemset(e.d_ctr, 0, block));
  HPC_HIPCHK(hipMemset(e.d_l0slot, 0, slot_bytes));
  e.cap_inst = c;
  return HPC_OK;
}

// The packed instance table to the device: a kernel argument when it is
// small, else a copy of e.h_inst (the next pack waits for it).
static size_t counter_words(size_t K) {
  return (K ? kIctrOffset + K * sizeof(InstCounters) : sizeof(Counters)) / 8;
}

// zero: the counter block of the K packed instances is cleared too
static int upload_table(Engine &e, bool zero = false) {
  const int K = (int)e.h_inst.size();
  if (K == 0) return HPC_OK;
  if (K <= kTableArg) {
    TableArg a;
    memset(&a, 0, sizeof(a));
    for (int k = 0; k < K; ++k) a.inst[k] = e.h_inst[k];
    a.K = K;
    if (zero) {
      a.zero = reinterpret_cast<unsigned long long *>(e.d_ctr);
      a.zero_words = (int)counter_words(K);
    }
    launch_table(e.stream, e.d_inst, a);
    HPC_HIPCHK(hipGetLastError());
    return HPC_OK;
  }
  HPC_HIPCHK(hipMemcpyAsync(e.d_inst, e.h_inst.data(), K * sizeof(Inst), hipMemcpyHostToDevice,
                            e.stream));
  e.host_upload_pending = true;
  if (zero) HPC_HIPCHK(hipMemsetAsync(e.d_ctr, 0, counter_words(K) * 8, e.stream));
  return HPC_OK;
}

static int ensure_row_cap(Engine &e, size_t rows) {
  if (rows <= e.cap_rows && e.d_row_inst) return HPC_OK;
  const size_t c = std::max<size_t>(rows + rows / 4, 64);
  dfree(e.d_row_inst); dfree(e.d_deg); dfree(e.d_rowptr); dfree(e.d_rowmask);
  dfree(e.d_row_tasks); dfree(e.d_task_off); dfree(e.d_blk);
  e.cap_rows = 0;
  e.row_inst_valid = false;
  int rc = dalloc((void **)&e.d_row_inst, c * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_deg, c * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_rowptr, (c + 1) * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_rowmask, c * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_row_tasks, c * sizeof(long long));
  if (!rc) rc = dalloc((void **)&e.d_task_off, c * sizeof(long long));
  const size_t nb = (size_t)compact_fused_blocks((int)c);
  if (!rc) rc = dalloc((void **)&e.d_blk, 3 * nb * sizeof(unsigned long long));
  if (rc) return rc;
  e.cap_rows = c;
  return HPC_OK;
}

static int ensure_edge_cap(Engine &e, size_t edges) {
  if (edges <= e.cap_edges && e.d_nbr) return HPC_OK;
  const size_t c = std::max<size_t>(edges + edges / 4, 1024);
  dfree(e.d_nbr); dfree(e.d_erow); dfree(e.d_fp); dfree(e.d_ccor); dfree(e.d_erev);
  e.cap_edges = 0;
  int rc = dalloc((void **)&e.d_nbr, c * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_erow, c * sizeof(int32_t));
  if (!rc) rc = dalloc((void **)&e.d_fp, c * sizeof(int64_t));
  if (!rc) rc = dalloc((void **)&e.d_ccor, c * sizeof(double));
  if (!rc) rc = dalloc((void **)&e.d_erev, c * sizeof(int32_t));
  if (!rc) e.cap_edges = c;
  return rc;
}

// Recheck queues of fast instances (1 Mi entries each; an overflow voids the
// phase and the host redoes it with exact correlations).
static int ensure_recheck_cap(Engine &e) {
  const size_t cap = (size_t)1 << 20;
  if (e.d_rchk) return HPC_OK;
  int rc = dalloc((void **)&e.d_rchk, cap * sizeof(Recheck));
  if (rc) return rc;
  e.cap_rchk = cap;
  return HPC_OK;
}

int engine_pack(Engine &e, const std::vector<int> &ids, const std::vector<int> &levels,
                bool zero_ctr) {
  HPC_HIPCHK(hipSetDevice(e.device));
  // a previous phase's upload may still read these host vectors
  if (e.host_upload_pending) {
    HPC_HIPCHK(hipStreamSynchronize(e.stream));
    e.host_upload_pending = false;
  }
  const int K = (int)ids.size();
  int p_total = 0, pmax = 0;
  size_t edges = 0;
  for (int id : ids) {
    p_total += e.inst[id].p;
    pmax = std::max(pmax, e.inst[id].p);
    edges += (size_t)e.inst[id].p * (e.inst[id].p - 1);
  }
  if (edges > (size_t)kMaxPackEdges)
    return hpc_fail(HPC_ENOMEM, "%zu ordered edges in one step, more than int32 row pointers index",
                    edges);
  int rc = ensure_inst_cap(e, (size_t)K);
  if (!rc) rc = ensure_row_cap(e, (size_t)p_total + 1);
  if (!rc) rc = ensure_edge_cap(e, edges + 1);
  if (!rc) rc = ensure_binom(e, pmax + 1);
  bool any_fast = false;
  for (int id : ids) any_fast |= e.inst[id].fast;
  if (!rc && any_fast) rc = ensure_recheck_cap(e);
  if (rc) return rc;
  e.h_inst.assign(K, Inst());
  e.h_row_inst.resize(p_total);
  int start = 0, grp0 = 0;
  for (int k = 0; k < K; ++k) {
    InstHost &h = e.inst[ids[k]];
    engine_drop_results(h);  // a packed instance's device state may change
    Inst &I = e.h_inst[k];
    I.x = h.x; I.z = h.z; I.nanm = h.nanm; I.mean = h.mean; I.sd = h.sd; I.cor = h.cor;
    I.adj = h.adj; I.rec = h.rec; I.rec_base = h.rec_used; I.rec_cap = h.rec_cap;
    I.z32 = h.z32;
    I.zc = h.zc;
    I.fast = h.fast ? 1 : 0;
    I.eps = kFastEps;
    I.r_crit = h.r_crit;
    square_bounds(h.r_crit, I.r2_lo, I.r2_hi);
    I.start = start;
    I.grp0 = grp0;
    grp0 += (h.p + 7) / 8;
    I.p = h.p; I.n = h.n; I.npad = h.npad; I.W = h.W;
    I.level = levels[k];
    I.flags = kActive | (h.cor_nan != 0 ? kCorNaN : 0);
    h.start = start;
    std::fill(e.h_row_inst.begin() + start, e.h_row_inst.begin() + start + h.p, k);
    start += h.p;
  }
  e.packed = ids;
  e.p_total = p_total;
  e.pmax = pmax;
  if (K > 0) {
    rc = upload_table(e, zero_ctr);
    if (rc) return rc;
    // the row map changes only with the set of instances
    if (!e.row_inst_valid || e.h_row_inst != e.h_row_inst_dev) {
      HPC_HIPCHK(hipMemcpyAsync(e.d_row_inst, e.h_row_inst.data(), p_total * sizeof(int32_t),
                                hipMemcpyHostToDevice, e.stream));
      e.h_row_inst_dev = e.h_row_inst;
      e.row_inst_valid = true;
      e.host_upload_pending = true;
    }
  }
  return HPC_OK;
}

static bool any_fast_at(const Engine &e, int level) {
  for (const Inst &I : e.h_inst)
    if (I.fast && (level < 0 || I.level == level)) return true;
  return false;
}

static Layout layout(Engine &e) {
  Layout L;
  L.inst = e.d_inst;
  L.ictr = e.d_ictr;
  L.row_inst = e.d_row_inst;
  L.K = (int)e.packed.size();
  L.p_total = e.p_total;
  L.max_deg = e.max_deg;
  L.pmax = e.pmax;
  L.deg = e.d_deg;
  L.rowptr = e.d_rowptr;
  L.nbr = e.d_nbr;
  L.erow = e.d_erow;
  L.fp = e.d_fp;
  L.ccor = e.d_ccor;
  L.erev = e.d_erev;
  L.binom = e.d_binom;
  L.ctr = e.d_ctr;
  L.rchk = e.d_rchk;
  L.rchk_cap = e.rchk_limit ? std::min(e.rchk_limit, e.cap_rchk) : e.cap_rchk;
  L.l0slot = e.d_l0slot;
  return L;
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  hipError_t r = hipEventElapsedTime(&ms, a, b);
  if (r == hipErrorNotReady && hipEventSynchronize(b) == hipSuccess)
    r = hipEventElapsedTime(&ms, a, b);
  return r == hipSuccess ? ms : 0.f;
}

// Device counters and the packed instances' counters into the host mirror,
// and a wait until all work so far on the stream is done.  A kernel stores
// them into fine-grained host memory and then a sequence number; the host
// spins on that number, which comes back sooner than a copy plus a stream
// synchronisation.  The spin asks the stream every few hundred polls, so a
// failed launch or a stream error ends it.
// Waits until the publication numbered `seq` has arrived in host memory.
// A publication that has not arrived after the engine's wait limit ends the
// wait with HPC_EHIP (a kernel that never completes would otherwise hold the
// caller, and a server's worker, for good).
static int wait_published(Engine &e, unsigned long long seq) {
  const hipStream_t stream = e.stream;
  const int rc = spin_wait(e.h_seq, seq, e.wait_limit_s, [stream, seq]() -> int {
    const hipError_t q = hipStreamQuery(stream);
    if (q == hipSuccess) return 1;  // idle: the word is final
    if (q != hipErrorNotReady)
      return hpc_fail(HPC_EHIP, "HIP error while waiting for publication %llu: %s", seq,
                      hipGetErrorString(q));
    return 0;
  });
  if (rc) return rc;
  e.host_upload_pending = false;  // everything before the publication is done
  return HPC_OK;
}

// Publication arguments for a kernel that publishes the counters itself.
static Publish next_publish(Engine &e) {
  Publish pub;
  pub.dst = reinterpret_cast<unsigned long long *>(e.h_ctr_dev);
  pub.flag = e.d_seq;
  pub.seq = ++e.seq;
  pub.words = (int)counter_words(e.packed.size());
  pub.advance = 0;
  pub.sys = 1;
  return pub;
}

static int read_counters(Engine &e) {
  const Publish pub = next_publish(e);
  launch_publish(e.stream, reinterpret_cast<const unsigned long long *>(e.d_ctr), pub.words,
                 pub.dst, pub.flag, pub.seq);
  HPC_HIPCHK(hipGetLastError());
  return wait_published(e, pub.seq);
}

int engine_read_inst_counters(Engine &e) { return read_counters(e); }

static int zero_counters(Engine &e) {
  const size_t K = e.packed.size();
  const size_t bytes = K ? kIctrOffset + K * sizeof(InstCounters) : sizeof(Counters);
  HPC_HIPCHK(hipMemsetAsync(e.d_ctr, 0, bytes, e.stream));
  return HPC_OK;
}

int engine_corr(Engine &e, float *ms_moments, float *ms_corr) {
  HPC_HIPCHK(hipSetDevice(e.device));
  for (int id : e.packed)
    if (!e.inst[id].have_x) return hpc_fail(HPC_ESTATE, "no observations uploaded");
  // the moments kernel raises the per-instance NaN flags
  int rc = zero_counters(e);
  if (rc) return rc;
  Layout L = layout(e);
  HPC_HIPCHK(hipEventRecord(e.ev[0], e.stream));
  launch_moments(e.stream, L, e.h_inst);
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[1], e.stream));
  launch_corr(e.stream, L, e.h_inst, e.d_prefix, e.h_prefix);
  e.host_upload_pending = true;  // the block prefixes
  HPC_HIPCHK(hipGetLastError());
  if (any_fast_at(e, -1)) {
    launch_corr_fast(e.stream, L, e.h_inst, e.d_prefix + 3 * (L.K + 1), e.h_prefix_fast);
    e.host_upload_pending = true;
    HPC_HIPCHK(hipGetLastError());
  }
  HPC_HIPCHK(hipEventRecord(e.ev[2], e.stream));
  for (int id : e.packed) {
    e.inst[id].have_cor = true;
    e.inst[id].cor_nan = -1;
  }
  e.h_cor_valid = false;
  if (any_fast_at(e, -1)) {
    // instances with NaN cells took the exact masked kernel; the float
    // z-scores are no longer needed
    rc = engine_read_inst_counters(e);
    if (rc) return rc;
    for (size_t k = 0; k < e.packed.size(); ++k) {
      InstHost &h = e.inst[e.packed[k]];
      if (!h.fast) continue;
      if (e.h_ictr[k].has_nan) h.fast = false;
      ifree(e, h.z32);
    }
  }
  if (ms_moments || ms_corr) {
    HPC_HIPCHK(hipEventSynchronize(e.ev[2]));
    if (ms_moments) *ms_moments = elapsed(e.ev[0], e.ev[1]);
    if (ms_corr) *ms_corr = elapsed(e.ev[1], e.ev[2]);
  }
  return HPC_OK;
}

namespace {
constexpr int kZeroSpans = 32;
struct ZeroSpanArg {
  double *ptr[kZeroSpans];
  unsigned long long words[kZeroSpans];
};
// blockIdx.y picks the span; the x blocks stride over it
__global__ void zero_spans_kernel(ZeroSpanArg a) {
  double *p = a.ptr[blockIdx.y];
  const unsigned long long w = a.words[blockIdx.y];
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < w;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
}  // namespace

int engine_zero_spans(Engine &e, const std::vector<ZeroSpan> &spans) {
  for (size_t a = 0; a < spans.size(); a += kZeroSpans) {
    const int n = (int)std::min<size_t>(kZeroSpans, spans.size() - a);
    ZeroSpanArg arg{};
    for (int k = 0; k < n; ++k) {
      arg.ptr[k] = spans[a + k].ptr;
      arg.words[k] = spans[a + k].words;
    }
    hipLaunchKernelGGL(zero_spans_kernel, dim3(64, (unsigned)n), dim3(256), 0, e.stream, arg);
    HPC_HIPCHK(hipGetLastError());
  }
  return HPC_OK;
}

// The z-score buffers of the instance's correlation mode: the exact form's
// class-major copy, or the fast form's float copy with the canonical FP64
// one (its exact rechecks read it).
int engine_prepare_corr(Engine &e, int id, std::vector<ZeroSpan> *clears) {
  InstHost &h = e.inst[id];
  h.fast = h.corr_mode == HPC_CORR_F32;
  const size_t before = zbytes(h);
  int rc = HPC_OK;
  if (h.fast) {
    if (!h.z) rc = ialloc(e, h.z, (size_t)h.p * h.npad);
    if (!rc && !h.z32) rc = ialloc(e, h.z32, (size_t)h.p * h.npad);
  } else if (!h.zc) {
    rc = ialloc(e, h.zc, zc_size(h.p, h.npad));
    // the rows past p (read, never stored, by the matrix-core kernel) are
    // zeros: the last pair of row blocks, cleared once
    const size_t tail = 2 * zc_block(h.npad);
    if (!rc) {
      const ZeroSpan z{h.zc + zc_size(h.p, h.npad) - tail, tail};
      if (clears) clears->push_back(z);
      else rc = engine_zero_spans(e, {z});
    }
  }
  h.bytes += zbytes(h) - before;
  return rc;
}

int engine_exact_from_z(Engine &e, const std::vector<int> &ids) {
  std::vector<int> redo;
  for (int id : ids)
    if (e.inst[id].fast) {
      e.inst[id].fast = false;
      redo.push_back(id);
    }
  if (redo.empty()) return HPC_OK;
  int rc = engine_pack(e, redo, std::vector<int>(redo.size(), 0));
  if (!rc) rc = zero_counters(e);
  if (rc) return rc;
  Layout L = layout(e);
  launch_corr(e.stream, L, e.h_inst, e.d_prefix, e.h_prefix);
  e.host_upload_pending = true;  // the block prefixes
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipStreamSynchronize(e.stream));
  return HPC_OK;
}

int engine_level0(Engine &e, float *ms) {
  HPC_HIPCHK(hipSetDevice(e.device));
  for (int id : e.packed) {
    e.inst[id].fresh = false;
    engine_drop_results(e.inst[id]);
  }
  int rc = zero_counters(e);
  if (rc) return rc;
  Layout L = layout(e);
  HPC_HIPCHK(hipEventRecord(e.ev[3], e.stream));
  launch_level0(e.stream, L, e.h_inst, e.d_prefix + 2 * (L.K + 1), e.h_prefix);
  if (L.K > 1) e.host_upload_pending = true;  // the tile offsets
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[4], e.stream));
  rc = read_counters(e);
  if (rc) return rc;
  if (ms) *ms = elapsed(e.ev[3], e.ev[4]);
  for (size_t k = 0; k < e.packed.size(); ++k) {
    InstHost &h = e.inst[e.packed[k]];
    if (e.h_inst[k].level != 0) continue;
    h.undirected -= (int64_t)e.h_ictr[k].removed;
    if (e.h_ictr[k].removed) h.compacted = false;
    h.cor_nan = e.h_ictr[k].has_nan ? 1 : 0;
  }
  return HPC_OK;
}

int engine_compact(Engine &e, float *ms) {
  HPC_HIPCHK(hipSetDevice(e.device));
  int rc = zero_counters(e);
  if (rc) return rc;
  Layout L = layout(e);
  HPC_HIPCHK(hipEventRecord(e.ev[3], e.stream));
  launch_compact(e.stream, L);
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[4], e.stream));
  rc = read_counters(e);
  if (rc) return rc;
  e.n_edges = (int64_t)e.h_ctr->n_edges;
  e.max_deg = e.h_ctr->max_deg;
  // every ordered edge of the new snapshot starts with no passing set (set
  // by the compaction); the host share scatters its ranks from the copy
  // stream after this point
  HPC_HIPCHK(hipEventRecord(e.fp_ready, e.stream));
  for (size_t k = 0; k < e.packed.size(); ++k) {
    if (e.h_inst[k].level < 1) continue;
    InstHost &h = e.inst[e.packed[k]];
    h.max_deg = e.h_ictr[k].max_deg;
    h.compacted = true;
    h.ever_compacted = true;
  }
  if (ms) *ms = elapsed(e.ev[3], e.ev[4]);
  return HPC_OK;
}

int engine_search(Engine &e, int level, const int32_t *rows, int nrows, float *ms) {
  HPC_HIPCHK(hipSetDevice(e.device));
  e.rechecks = 0;
  for (size_t k = 0; k < e.packed.size(); ++k) {
    const InstHost &h = e.inst[e.packed[k]];
    if (e.h_inst[k].level == level && !engine_plannable(e, h.max_deg, level))
      return hpc_fail(HPC_ELEVEL,
                      "level %d: C(%d, %d) conditioning sets per edge is beyond what one search "
                      "plans", level, h.max_deg - 1, level);
  }
  Layout L = layout(e);
  const int32_t *mask = nullptr;
  if (rows) {
    std::vector<int32_t> m(e.p_total, 0);
    for (int k = 0; k < nrows; ++k) {
      if (rows[k] < 0 || rows[k] >= e.p_total) return hpc_fail(HPC_EINVAL, "row %d out of range", rows[k]);
      m[rows[k]] = 1;
    }
    HPC_HIPCHK(hipMemcpyAsync(e.d_rowmask, m.data(), e.p_total * sizeof(int32_t),
                              hipMemcpyHostToDevice, e.stream));
    HPC_HIPCHK(hipStreamSynchronize(e.stream));
    mask = e.d_rowmask;
  }
  // the event pair brackets the CI-test kernel alone (the task plan of a
  // level >= 2 runs before it)
  bool any = e.n_edges > 0 && (!rows || nrows > 0);
  if (any && level >= 2) {
    launch_task_plan(e.stream, level, L, engine_rpt(e, level), mask, e.d_row_tasks, e.d_task_off);
    HPC_HIPCHK(hipGetLastError());
  }
  HPC_HIPCHK(hipEventRecord(e.ev[5], e.stream));
  const bool fast = any && any_fast_at(e, level);
  if (any) {
    if (level == 1) launch_search_l1(e.stream, L, e.n_edges, mask, fast, engine_epb(e),
                                     search_lds_limit(e.lds_limit));
    else
      launch_search_ln(e.stream, level, L, engine_rpt(e, level), e.d_row_tasks, e.d_task_off, fast,
                       engine_tpb(e, level), search_lds_limit(e.lds_limit), device_queue(e, level));
    if (fast) launch_recheck(e.stream, level, L);
  }
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[6], e.stream));
  if (fast) {
    int rc = read_counters(e);
    if (rc) return rc;
    e.rechecks = e.h_ctr->rechecks;
    if (e.h_ctr->overflow) {
      // too many tests near the threshold: exact correlations, then the
      // search again from scratch
      const std::vector<int> ids = e.packed;
      std::vector<int> lv;
      for (const Inst &I : e.h_inst) lv.push_back(I.level);
      rc = engine_exact_from_z(e, ids);
      if (!rc) rc = engine_pack(e, ids, lv);
      if (rc) return rc;
      HPC_HIPCHK(hipMemsetAsync(e.d_ctr, 0, sizeof(Counters), e.stream));
      if (e.n_edges > 0)
        HPC_HIPCHK(hipMemsetAsync(e.d_fp, 0xFF, (size_t)e.n_edges * sizeof(int64_t), e.stream));
      return engine_search(e, level, rows, nrows, ms);
    }
  }
  if (ms) {
    HPC_HIPCHK(hipEventSynchronize(e.ev[6]));
    *ms = elapsed(e.ev[5], e.ev[6]);
  }
  return HPC_OK;
}

int engine_search_shared_launch(Engine &e, int level, SharedPlan &plan) {
  HPC_HIPCHK(hipSetDevice(e.device));
  e.rechecks = 0;
  plan = SharedPlan{};
  if (e.packed.size() != 1) return hpc_fail(HPC_ESTATE, "the shared queue takes one instance");
  const InstHost &h = e.inst[e.packed[0]];
  if (h.fast) return hpc_fail(HPC_ESTATE, "the shared queue needs an exact correlation");
  if (e.n_edges <= 0 || h.max_deg - 1 < level) return HPC_OK;
  if (!engine_plannable(e, h.max_deg, level))
    return hpc_fail(HPC_ELEVEL,
                    "level %d: C(%d, %d) conditioning sets per edge is beyond what one search "
                    "plans", level, h.max_deg - 1, level);
  Layout L = layout(e);
  plan.rpt = engine_rpt(e, level);
  plan.tpb = engine_tpb(e, level);
  launch_task_plan(e.stream, level, L, plan.rpt, nullptr, e.d_row_tasks, e.d_task_off);
  HPC_HIPCHK(hipGetLastError());
  int rc = read_counters(e);
  if (rc) return rc;
  plan.n_tasks = (long long)e.h_ctr->n_tasks;
  plan.n_blocks = (plan.n_tasks + plan.tpb - 1) / plan.tpb;
  if (plan.n_blocks == 0) return HPC_OK;
  // host copies of the plan and the snapshot (the stream is idle here)
  const int p = e.p_total;
  e.h_rowptr.resize(p + 1);
  e.h_nbr.resize(e.n_edges);
  e.h_task_off.resize(p);
  e.h_row_tasks.resize(p);
  HPC_HIPCHK(hipMemcpy(e.h_rowptr.data(), e.d_rowptr, (p + 1) * sizeof(int32_t),
                       hipMemcpyDeviceToHost));
  HPC_HIPCHK(hipMemcpy(e.h_nbr.data(), e.d_nbr, e.n_edges * sizeof(int32_t), hipMemcpyDeviceToHost));
  HPC_HIPCHK(hipMemcpy(e.h_task_off.data(), e.d_task_off, p * sizeof(long long),
                       hipMemcpyDeviceToHost));
  HPC_HIPCHK(hipMemcpy(e.h_row_tasks.data(), e.d_row_tasks, p * sizeof(long long),
                       hipMemcpyDeviceToHost));
  if (!e.h_cor_valid) {
    e.h_cor.resize((size_t)h.p * h.p);
    HPC_HIPCHK(hipMemcpy(e.h_cor.data(), h.cor, (size_t)h.p * h.p * sizeof(double),
                         hipMemcpyDeviceToHost));
    e.h_cor_valid = true;
  }
  // the done flags, one per block, in host-visible memory
  if ((size_t)plan.n_blocks > e.cap_qdone) {
    if (e.h_qdone) HPC_HIPCHK(hipHostFree(e.h_qdone));
    e.h_qdone = nullptr;
    e.cap_qdone = 0;
    const size_t cap = std::max<size_t>((size_t)plan.n_blocks, 1024);
    HPC_HIPCHK(hipHostMalloc((void **)&e.h_qdone, cap * sizeof(int),
                             hipHostMallocMapped | hipHostMallocCoherent));
    HPC_HIPCHK(hipHostGetDevicePointer((void **)&e.d_qdone, e.h_qdone, 0));
    e.cap_qdone = cap;
  }
  for (long long b = 0; b < plan.n_blocks; ++b) __atomic_store_n(&e.h_qdone[b], 0, __ATOMIC_RELAXED);
  __atomic_thread_fence(__ATOMIC_SEQ_CST);
  plan.done = e.h_qdone;
  L.max_deg = e.max_deg;
  HPC_HIPCHK(hipEventRecord(e.ev[5], e.stream));
  launch_search_ln(e.stream, level, L, plan.rpt, e.d_row_tasks, e.d_task_off, false, plan.tpb,
                   search_lds_limit(e.lds_limit), TaskQueue{e.d_qhead, e.d_qdone});
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[6], e.stream));
  plan.device_done = e.ev[6];
  return HPC_OK;
}

int engine_search_shared_finish(Engine &e, const std::vector<int64_t> &idx,
                                const std::vector<int64_t> &val, float *ms) {
  HPC_HIPCHK(hipSetDevice(e.device));
  const size_t cnt = idx.size();
  if (cnt > 0) {
    if (cnt > e.scatter_cap) {
      dfree(e.d_scatter_idx);
      dfree(e.d_scatter_val);
      e.scatter_cap = 0;
      int rc = dalloc((void **)&e.d_scatter_idx, cnt * sizeof(int64_t));
      if (!rc) rc = dalloc((void **)&e.d_scatter_val, cnt * sizeof(int64_t));
      if (rc) return rc;
      e.scatter_cap = cnt;
    }
    HPC_HIPCHK(hipMemcpyAsync(e.d_scatter_idx, idx.data(), cnt * sizeof(int64_t),
                              hipMemcpyHostToDevice, e.stream));
    HPC_HIPCHK(hipMemcpyAsync(e.d_scatter_val, val.data(), cnt * sizeof(int64_t),
                              hipMemcpyHostToDevice, e.stream));
    launch_scatter_min_fp(e.stream, e.d_scatter_idx, e.d_scatter_val, (long long)cnt, e.d_fp);
    HPC_HIPCHK(hipGetLastError());
  }
  HPC_HIPCHK(hipStreamSynchronize(e.stream));
  if (ms) *ms = elapsed(e.ev[5], e.ev[6]);
  return HPC_OK;
}

int engine_commit(Engine &e, float *ms) {
  HPC_HIPCHK(hipSetDevice(e.device));
  for (int id : e.packed) {
    e.inst[id].fresh = false;
    engine_drop_results(e.inst[id]);
  }
  // room for the level's sepset records; the table is rewritten, so wait for
  // earlier uploads of it
  if (e.host_upload_pending) {
    HPC_HIPCHK(hipStreamSynchronize(e.stream));
    e.host_upload_pending = false;
  }
  for (size_t k = 0; k < e.packed.size(); ++k) {
    const int l = e.h_inst[k].level;
    if (l < 1 || l > kMaxLevel) continue;
    InstHost &h = e.inst[e.packed[k]];
    int rc = ensure_rec(e, h, l);
    if (rc) return rc;
    e.h_inst[k].rec = h.rec;
    e.h_inst[k].rec_base = h.rec_used;
    e.h_inst[k].rec_cap = h.rec_cap;
  }
  if (!e.packed.empty()) {
    int rc = upload_table(e);
    if (rc) return rc;
    // the removal kernel's counts (and record offsets) start from zero
    HPC_HIPCHK(hipMemsetAsync(e.d_ictr, 0, e.packed.size() * sizeof(InstCounters), e.stream));
  }
  Layout L = layout(e);
  HPC_HIPCHK(hipEventRecord(e.ev[6], e.stream));
  Publish none;
  memset(&none, 0, sizeof(none));
  launch_commit(e.stream, L, e.n_edges, none);
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[7], e.stream));
  int rc = engine_read_inst_counters(e);
  if (rc) return rc;
  if (ms) *ms = elapsed(e.ev[6], e.ev[7]);
  for (size_t k = 0; k < e.packed.size(); ++k) {
    if (e.h_inst[k].level < 1) continue;
    InstHost &h = e.inst[e.packed[k]];
    h.undirected -= (int64_t)e.h_ictr[k].removed;
    h.rec_used += (size_t)e.h_ictr[k].rec_words;
    if (e.h_ictr[k].removed) h.compacted = false;
  }
  return HPC_OK;
}

}  // namespace hpc

namespace hpc {

size_t engine_footprint(int p, int n, int max_level) {
  const size_t pp = (size_t)p * p;
  const int mc = max_level > 1 ? std::min(max_level, kMaxLevel) : (max_level < 0 ? 4 : 1);
  const size_t npad = ((size_t)n + 63) & ~(size_t)63;
  const size_t W = ((size_t)p + 63) / 64;
  (void)mc;
  size_t b = pp * 8                 // correlation
             + pp * 8               // sepset records (level 1, every edge removed)
             + (size_t)p * W * 8    // bit adjacency
             + (size_t)p * n * 8    // observations
             // z-scores (class-major FP64, or canonical FP64 + FP32 for
             // HPC_CORR_F32: this bounds both) and the NaN mask
             + std::max(zc_size(p, (int)npad) * 8, (size_t)p * npad * 12) + (size_t)p * npad
             + (size_t)p * 16;      // mean, sd
  b += pp * (4 + 4 + 8 + 8 + 4);    // worst-case share of nbr / erow / first-pass / ccor / erev arrays
  b += (size_t)p * 40;              // row arrays
  return b;
}

// One instance, levels h.next_level .. its cap, with one host wait per
// level.  Level 0's fold kernel and every removal publish their counts into
// h_ctr2, clear the counter block and advance the device table (level,
// record base), so no table kernel runs between levels.  Per level l >= 1:
//  - l = 1 (or after a level with no edges): compaction, publishing
//    n_edges / max_deg into h_ctr; [host wait]; CI tests; removal, with the
//    next compaction queued right behind it (its counts are read at the next
//    wait: publications arrive in stream order);
//  - l >= 2: CI tests; a removal that also counts level l + 1 and publishes
//    those sizes with its counts; the fill kernel; in an exact capped run
//    also level l + 1's search; [host wait].
// Results and per-level records equal engine_step's.  Kernel times are read
// after the run.  With fast correlations a level whose recheck queue
// overflowed did nothing (the removal skips itself); the next wait finds it,
// the correlations are made exact and the loop returns with h.next_level at
// that level for engine_step to continue.
// Publication slot of a level in a captured run: kind 0 = its compaction,
// 1 = its removal (level 0: the fold kernel), device and host addresses.
// A captured level writes its slot in device memory (no system-scope fence
// per level); the capture's last publication copies all of them to the host
// slots.
constexpr size_t kSlotsBytes = (size_t)2 * (kMaxLevel + 2) * kSlotBytes;
static unsigned long long *slot_dev(Engine &e, int level, int kind) {
  return reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(e.d_slots_local) +
                                                (size_t)(2 * level + kind) * kSlotBytes);
}
static const Counters *slot_host(const Engine &e, int level, int kind) {
  return reinterpret_cast<const Counters *>(reinterpret_cast<const char *>(e.h_slots) +
                                            (size_t)(2 * level + kind) * kSlotBytes);
}
static const InstCounters *slot_inst(const Engine &e, int level, int kind) {
  return reinterpret_cast<const InstCounters *>(reinterpret_cast<const char *>(slot_host(e, level, kind)) +
                                                kIctrOffset);
}

// The one-instance level loop.  Eager (cap == nullptr): it waits for each
// level's published counts and, with `rec`, records the sizes it read.  In
// capture mode it enqueues the same launches without waiting, taking every
// size from `cap` (a recorded run of the same state) and sending each
// publication to its own per-level slot; the host-side bookkeeping is left to
// the replay.
static int run_single_impl(Engine &e, int id, const RunShape *cap, RunShape *rec) {
  HPC_HIPCHK(hipSetDevice(e.device));
  InstHost &h = e.inst[id];
  if (!h.have_cor) return hpc_fail(HPC_ESTATE, "no correlation matrix");
  const bool capmode = cap != nullptr;
  h.fresh = false;
  if (!e.lev_ev_ready) {
    for (auto &row : e.lev_ev)
      for (auto &ev : row) HPC_HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    e.lev_ev_ready = true;
  }
  using Clock = std::chrono::steady_clock;
  const std::vector<int> ids{id};
  // Each event marker between two queued kernels delays the second by a few
  // microseconds, so one marker serves both neighbours: the search's start
  // ends the compaction, the search's end starts the removal, and a
  // removal's end starts the next compaction.
  hipEvent_t comp_start[kMaxLevel + 1] = {};
  // the level-loop events, unless kernel timing is off
  auto mark = [&](hipEvent_t ev) -> hipError_t {
    return (e.kernel_timing && !capmode) ? hipEventRecord(ev, e.stream) : hipSuccess;
  };
  hipEvent_t prev_commit_end = nullptr;
  const InstCounters *i2 =
      reinterpret_cast<const InstCounters *>(reinterpret_cast<const char *>(e.h_ctr2) + kIctrOffset);
  auto publish2 = [&](int lv) {
    Publish pb = next_publish(e);
    pb.dst = capmode ? slot_dev(e, lv, 1) : reinterpret_cast<unsigned long long *>(e.h_ctr2_dev);
    if (capmode) pb.sys = 0;
    return pb;
  };
  // the level whose counts sit in h_ctr2 (or will, once the next wait ends)
  int pend = -1;
  unsigned long long pend_seq = 0;
  Clock::time_point pend_t0 = Clock::now();
  std::vector<std::pair<size_t, int>> timed;  // (history entry, level)
  auto finalize = [&](int lv, int max_deg, const Clock::time_point &t_end) {
    const InstCounters &c = i2[0];
    hpc_level_stats st;
    memset(&st, 0, sizeof(st));
    st.level = lv;
    st.max_deg = lv == 0 ? h.p : max_deg;
    st.tests = (int64_t)c.tests;
    st.tests_max = (int64_t)c.tests_max;
    st.edges = (int64_t)c.edges;
    st.removed = (int64_t)c.removed;
    h.undirected -= st.removed;
    st.edges_left = h.undirected;
    st.pinv_fallbacks = (int64_t)e.h_ctr2->pinv_fallbacks;
    st.rechecks = (int64_t)e.h_ctr2->rechecks;
    e.rechecks = e.h_ctr2->rechecks;
    if (lv == 0) {
      h.cor_nan = c.has_nan ? 1 : 0;
    } else {
      h.rec_used += (size_t)c.rec_words;
      engine_note_level(e, id, lv);
    }
    if (st.removed) h.compacted = false;
    st.ms_wall = std::chrono::duration<double, std::milli>(t_end - pend_t0).count();
    // kernel times once the run is over: a level's counts arrive from inside
    // its removal kernel, before the events after it have completed
    timed.push_back({h.hist.size(), lv});
    h.hist.push_back(st);
  };
  int pend_max_deg = 0;
  // a voided level (recheck queue overflow): exact correlations, and the step
  // loop takes over at that level
  auto voided = [&](int lv) -> int {
    int r = engine_exact_from_z(e, ids);
    if (r) return r;
    h.next_level = lv;
    h.finished = false;
    return HPC_OK;
  };
  // the level loop proper; its kernel times are filled in once it returns
  auto levels = [&]() -> int {
    int l = h.next_level;
    int rc = HPC_OK;
    if (l == 0) {
      pend_t0 = Clock::now();
      rc = engine_pack(e, ids, {0}, /*zero_ctr=*/true);
      if (rc) return rc;
      Layout L = layout(e);
      HPC_HIPCHK(mark(e.lev_ev[0][4]));
      // the fold kernel publishes level 0's counts, clears the counters and
      // sets the device table to level 1
      Publish pb = publish2(0);
      pb.advance = 1;
      launch_level0(e.stream, L, e.h_inst, e.d_prefix + 2 * (L.K + 1), e.h_prefix, &pb);
      HPC_HIPCHK(hipGetLastError());
      HPC_HIPCHK(mark(e.lev_ev[0][5]));
      prev_commit_end = e.lev_ev[0][5];  // level 1's compaction starts here
      pend = 0;
      pend_seq = pb.seq;
      pend_max_deg = h.p;
      l = 1;
      h.next_level = 1;
      if (h.max_level == 0) {
        rc = wait_published(e, pend_seq);
        if (rc) return rc;
        finalize(0, pend_max_deg, Clock::now());
        h.finished = true;
        return HPC_OK;
      }
      // the device table is at level 1 on clear counters (level 0's are in
      // h_ctr2): the fold kernel saw to both
      e.h_inst[0].level = 1;
    } else {
      // the device table starts at level l
      rc = engine_pack(e, ids, {l}, /*zero_ctr=*/true);
      if (rc) return rc;
    }
    // the previous removal counted this level (its publication, already waited
    // for, holds the level's sizes) and its fill kernel is queued
    bool counted = false;
    // this level's search is already in the stream (queued before the wait)
    bool searched = false;
    for (;;) {
      if (l > kMaxLevel) break;  // the step loop takes levels the kernels do not reach
      const Clock::time_point t_lv = Clock::now();
      Layout L = layout(e);
      if (counted) {
        if (capmode) {
          e.n_edges = cap->lv[l].n_edges;
          e.max_deg = cap->lv[l].max_deg;
          h.max_deg = cap->lv[l].inst_max_deg;
        } else {
          e.n_edges = (int64_t)e.h_ctr2->n_edges;
          e.max_deg = e.h_ctr2->max_deg;
          h.max_deg = i2[0].max_deg;
        }
        if (rec) rec->lv[l] = LevelShape{e.n_edges, e.max_deg, h.max_deg, 1};
        counted = false;
      } else {
        if (prev_commit_end) {
          comp_start[l] = prev_commit_end;
        } else {
          HPC_HIPCHK(mark(e.lev_ev[l][0]));
          comp_start[l] = e.lev_ev[l][0];
        }
        prev_commit_end = nullptr;
        Publish pa = next_publish(e);
        if (capmode) {
          pa.dst = slot_dev(e, l, 0);
          pa.sys = 0;
        }
        launch_compact_fused(e.stream, L, l, engine_rpt(e, l), e.d_blk, e.d_row_tasks,
                             e.d_task_off, e.cap_edges, pa);
        HPC_HIPCHK(hipGetLastError());
        if (capmode) {
          if (pend == 0) h.cor_nan = cap->cor_nan;  // what level 0's counts would say
          pend = -1;
          e.n_edges = cap->lv[l].n_edges;
          e.max_deg = cap->lv[l].max_deg;
          h.max_deg = cap->lv[l].inst_max_deg;
        } else {
          rc = wait_published(e, pa.seq);
          if (rc) return rc;
          if (pend >= 0) {
            if (e.h_ctr2->overflow) return voided(pend);
            finalize(pend, pend_max_deg, t_lv);
            pend = -1;
          }
          e.n_edges = (int64_t)e.h_ctr->n_edges;
          e.max_deg = e.h_ctr->max_deg;
          h.max_deg = e.h_ictr[0].max_deg;
        }
        if (rec) rec->lv[l] = LevelShape{e.n_edges, e.max_deg, h.max_deg, 0};
      }
      pend_t0 = t_lv;
      h.compacted = true;
      h.ever_compacted = true;
      const bool testable = h.max_deg - 1 >= l;
      if (!testable && h.max_level < 0) {
        h.finished = true;  // nothing left to test: no record (as engine_step)
        return HPC_OK;
      }
      if (testable && !engine_plannable(e, h.max_deg, l))
        return hpc_fail(HPC_ELEVEL,
                        "level %d: C(%d, %d) conditioning sets per edge is beyond what one search "
                        "plans", l, h.max_deg - 1, l);
      // record space, the NaN flag of level 1 on: the table when it changed
      // (a capture reuses the recorded run's record buffer, sized for every
      // level at once)
      if (!capmode) {
        rc = prepare_level(e, h, l);
        if (rc) return rc;
      }
      Inst &I = e.h_inst[0];
      const int flags = kActive | (h.cor_nan != 0 ? kCorNaN : 0);
      const bool retable = I.rec != h.rec || I.flags != flags || I.rec_cap != h.rec_cap;
      I.level = l;
      I.rec = h.rec;
      I.rec_cap = h.rec_cap;
      I.rec_base = h.rec_used;
      I.flags = flags;
      if (retable) {
        rc = upload_table(e, /*zero=*/false);
        if (rc) return rc;
      }
      L = layout(e);
      L.max_deg = e.max_deg;
      if (!searched) HPC_HIPCHK(mark(e.lev_ev[l][4]));
      if (e.n_edges > 0 && !searched) {
        if (l == 1)
          launch_search_l1(e.stream, L, e.n_edges, nullptr, h.fast, engine_epb(e),
                           search_lds_limit(e.lds_limit));
        else
          launch_search_ln(e.stream, l, L, engine_rpt(e, l), e.d_row_tasks, e.d_task_off, h.fast,
                           engine_tpb(e, l), search_lds_limit(e.lds_limit), device_queue(e, l));
        // exact re-runs of the tests near the threshold (fast correlations);
        // the removal below clears the queue with the other counters
        if (h.fast) launch_recheck(e.stream, l, L, /*clear_queue=*/false);
        HPC_HIPCHK(hipGetLastError());
      }
      if (!searched) HPC_HIPCHK(mark(e.lev_ev[l][5]));
      searched = false;
      const bool last = h.max_level >= 0 && l + 1 > h.max_level;
      // the removal also counts the next level when there is one to run.  Not
      // after level 1: it removes most edges, and their row-length updates
      // (device atomics on a few cache lines) cost more than the separate
      // count kernel
      const bool count_next = e.n_edges > 0 && !last && l + 1 <= kMaxLevel && l >= 2;
      Publish pb = publish2(l);
      if (e.n_edges > 0) {
        // publishes, clears the counters, advances level and record base
        pb.advance = 1;
        const NextCount nc{e.d_blk, e.d_row_tasks, l + 1, count_next ? engine_rpt(e, l + 1) : 0,
                           count_next ? 1 : 0};
        launch_commit(e.stream, L, e.n_edges, pb, nc);
        HPC_HIPCHK(hipGetLastError());
      } else {
        launch_publish(e.stream, reinterpret_cast<const unsigned long long *>(e.d_ctr), pb.words,
                       pb.dst, pb.flag, pb.seq);
        HPC_HIPCHK(hipGetLastError());
        if (!last) {
          I.level = l + 1;
          rc = upload_table(e, /*zero=*/true);
          if (rc) return rc;
        }
      }
      HPC_HIPCHK(mark(e.lev_ev[l][3]));
      prev_commit_end = e.lev_ev[l][3];
      if (count_next) {
        // the next level's neighbour lists while the host takes the counts
        comp_start[l + 1] = e.lev_ev[l][3];
        prev_commit_end = nullptr;
        launch_compact_fill(e.stream, L, e.d_blk, e.d_row_tasks, e.d_task_off, e.cap_edges);
        HPC_HIPCHK(hipGetLastError());
        // Exact runs with a level cap (the next level runs whatever its
        // sizes): the next level's search goes into the stream before the
        // wait.  It reads its task count on the device, and rows only shrink,
        // so this level's longest row bounds its LDS block.
        if (!h.fast && h.max_level >= 0) {
          HPC_HIPCHK(mark(e.lev_ev[l + 1][4]));
          launch_search_ln(e.stream, l + 1, L, engine_rpt(e, l + 1), e.d_row_tasks, e.d_task_off,
                           false, engine_tpb(e, l + 1), search_lds_limit(e.lds_limit),
                           device_queue(e, l + 1));
          HPC_HIPCHK(hipGetLastError());
          HPC_HIPCHK(mark(e.lev_ev[l + 1][5]));
          searched = true;
        }
        if (!capmode) {
          rc = wait_published(e, pb.seq);
          if (rc) return rc;
          if (e.h_ctr2->overflow) return voided(l);
          finalize(l, h.max_deg, Clock::now());
        }
        counted = true;
      } else {
        pend = l;
        pend_seq = pb.seq;
        pend_max_deg = h.max_deg;
      }
      ++l;
      h.next_level = l;
      e.h_inst[0].level = l;  // the device table's level after the removal
      if (last) break;
    }
    if (pend >= 0 && !capmode) {
      rc = wait_published(e, pend_seq);
      if (rc) return rc;
      if (e.h_ctr2->overflow) return voided(pend);
      finalize(pend, pend_max_deg, Clock::now());
    }
    if (h.max_level >= 0 && h.next_level > h.max_level) h.finished = true;
    return HPC_OK;
  };
  int rc_run = levels();
  if (capmode && rc_run == HPC_OK) {
    // every level's slot to host memory, then the run's last flag
    const Publish fin = next_publish(e);
    launch_publish(e.stream, e.d_slots_local, kSlotsBytes / 8,
                   reinterpret_cast<unsigned long long *>(e.d_slots), fin.flag, fin.seq);
    if (hipGetLastError() != hipSuccess) rc_run = hpc_fail(HPC_EHIP, "slot publication launch failed");
  }
  if (!e.kernel_timing) timed.clear();
  for (const auto &t : timed) {
    hpc_level_stats &st = h.hist[t.first];
    hipEvent_t *ev = e.lev_ev[t.second];
    st.ms_test = elapsed(ev[4], ev[5]);
    if (t.second > 0) {
      // the compaction runs up to the search's start event
      st.ms_compact = elapsed(comp_start[t.second], ev[4]);
      st.ms_commit = elapsed(ev[5], ev[3]);
    }
  }
  return rc_run;
}

// The launch shapes a captured graph bakes in (a change of any voids it).
static std::vector<long long> tuning_key(const Engine &e) {
  std::vector<long long> k{e.ranks_per_task, (long long)e.lds_limit, e.use_queue ? 1 : 0,
                           e.epb_level1};
  for (int l = 0; l <= kMaxLevel; ++l) {
    k.push_back(e.rpt_level[l]);
    k.push_back(e.tpb_level[l]);
    k.push_back(e.queue_level[l]);
  }
  return k;
}

void graph_drop(Engine &e) {
  if (e.graph.exec) (void)hipGraphExecDestroy(e.graph.exec);
  if (e.graph.graph) (void)hipGraphDestroy(e.graph.graph);
  e.graph.exec = nullptr;
  e.graph.graph = nullptr;
  e.graph.id = -1;
}

// Captures the level loop of a fresh run of instance id with the sizes of
// `shape` into a graph (the instance's host state is left as it was).
static int graph_capture(Engine &e, int id, const RunShape &shape) {
  graph_drop(e);
  InstHost &h = e.inst[id];
  if (!e.d_slots_local) HPC_HIPCHK(hipMalloc((void **)&e.d_slots_local, kSlotsBytes));
  if (!e.h_slots) {
    HPC_HIPCHK(hipHostMalloc((void **)&e.h_slots, kSlotsBytes,
                             hipHostMallocMapped | hipHostMallocCoherent));
    HPC_HIPCHK(hipHostGetDevicePointer((void **)&e.d_slots, e.h_slots, 0));
  }
  HPC_HIPCHK(hipStreamSynchronize(e.stream));
  e.host_upload_pending = false;
  const InstHost saved = h;
  const int64_t n_edges = e.n_edges;
  const int max_deg = e.max_deg;
  // a fresh run from level 0, as the recorded one
  h.next_level = 0;
  h.finished = false;
  // the recorded NaN state from the start (only the searches read the flag,
  // and a replay whose level 0 finds otherwise is void), so the table the
  // levels start from needs no rewrite before level 1
  h.cor_nan = shape.cor_nan;
  h.rec_used = 0;
  h.undirected = (int64_t)h.p * (h.p - 1) / 2;
  h.max_deg = h.p;
  const unsigned long long seq0 = e.seq;
  hipGraph_t graph = nullptr;
  HPC_HIPCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeRelaxed));
  const int rc = run_single_impl(e, id, &shape, nullptr);
  const hipError_t ec = hipStreamEndCapture(e.stream, &graph);
  h = saved;
  e.n_edges = n_edges;
  e.max_deg = max_deg;
  e.packed.clear();  // the device table is whatever the graph leaves
  if (rc || ec != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return rc ? rc : hpc_fail(HPC_EHIP, "level-loop capture failed: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hpc_fail(HPC_EHIP, "level-loop graph instantiation failed: %s", hipGetErrorString(ei));
  }
  e.graph.graph = graph;
  e.graph.exec = exec;
  e.graph.id = id;
  e.graph.shape = shape;
  e.graph.first_seq = seq0 + 1;
  e.graph.final_seq = e.seq;
  e.graph.tuning = tuning_key(e);
  ++e.graph_captures;
  return HPC_OK;
}

// Replays the captured level loop on instance id (fresh, same tuning).
// Returns 1 when the device's sizes differ from the recorded ones: the run's
// results are then void and the caller recomputes them.
static int graph_replay(Engine &e, int id) {
  using Clock = std::chrono::steady_clock;
  InstHost &h = e.inst[id];
  const RunShape &sh = e.graph.shape;
  const auto t0 = Clock::now();
  // every earlier publication has arrived (each run waits for its last)
  __atomic_store_n(e.h_seq, 0ull, __ATOMIC_RELEASE);
  HPC_HIPCHK(hipGraphLaunch(e.graph.exec, e.stream));
  int rc = wait_published(e, e.graph.final_seq);
  if (rc) return rc;
  h.fresh = false;
  e.packed.clear();
  // the sizes each level ran with, as the device computed them this time
  if ((slot_inst(e, 0, 1)->has_nan ? 1 : 0) != sh.cor_nan) return 1;
  const int checked = sh.stop >= 0 ? sh.stop : sh.last;
  for (int l = 1; l <= checked; ++l) {
    const LevelShape &x = sh.lv[l];
    const int kind = x.src ? 1 : 0, at = x.src ? l - 1 : l;
    const Counters *c = slot_host(e, at, kind);
    if ((int64_t)c->n_edges != x.n_edges || c->max_deg != x.max_deg ||
        slot_inst(e, at, kind)->max_deg != x.inst_max_deg)
      return 1;
  }
  const double wall = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  h.cor_nan = sh.cor_nan;
  for (int l = 0; l <= sh.last; ++l) {
    const Counters *c = slot_host(e, l, 1);
    const InstCounters &ic = *slot_inst(e, l, 1);
    hpc_level_stats st;
    memset(&st, 0, sizeof(st));
    st.level = l;
    st.max_deg = l == 0 ? h.p : sh.lv[l].inst_max_deg;
    st.tests = (int64_t)ic.tests;
    st.tests_max = (int64_t)ic.tests_max;
    st.edges = (int64_t)ic.edges;
    st.removed = (int64_t)ic.removed;
    h.undirected -= st.removed;
    st.edges_left = h.undirected;
    st.pinv_fallbacks = (int64_t)c->pinv_fallbacks;
    st.rechecks = (int64_t)c->rechecks;
    if (l >= 1) {
      h.rec_used += (size_t)ic.rec_words;
      engine_note_level(e, id, l);
    }
    st.ms_wall = wall / (sh.last + 1);  // the replay's wall time, split evenly
    h.hist.push_back(st);
  }
  h.max_deg = sh.lv[checked].inst_max_deg;
  e.n_edges = sh.lv[checked].n_edges;
  e.max_deg = sh.lv[checked].max_deg;
  // (a run with no cap ends on a compaction, a capped one on a removal)
  h.compacted = sh.stop >= 0;
  h.ever_compacted = true;
  h.next_level = sh.last + 1;
  h.finished = true;
  return HPC_OK;
}

int engine_run_single(Engine &e, int id) {
  InstHost &h = e.inst[id];
  engine_drop_results(h);
  // graph replay: a fresh run of an exact instance (a level cap of 1..16, or
  // none) without per-level kernel timing, with the launch shapes and record
  // buffer of the capture
  const bool graphable = e.use_graph && !e.kernel_timing && !h.fast && h.fresh &&
                         h.next_level == 0 &&
                         (h.max_level < 0 || (h.max_level >= 1 && h.max_level <= kMaxLevel)) &&
                         h.have_cor && h.inject_level < 0;
  if (graphable && e.graph.exec && e.graph.id == id && e.graph.shape.max_level == h.max_level &&
      e.graph.shape.rec == h.rec && e.graph.tuning == tuning_key(e)) {
    const size_t first = h.hist.size();
    const int r = graph_replay(e, id);
    if (r == 0) ++e.graph_replays;
    if (r <= 0) return r;
    // the device's sizes left the recorded ones: recompute eagerly
    h.hist.resize(first);
    graph_drop(e);
    int rc = engine_reset(e, id);
    if (rc) return rc;
    ++e.graph_mismatches;
  }
  RunShape rec;
  rec.max_level = h.max_level;
  const bool record = graphable && h.fresh;
  int rc = run_single_impl(e, id, nullptr, record ? &rec : nullptr);
  if (rc || !record || !h.finished) return rc;
  if (h.max_level >= 0) {
    if (h.next_level != h.max_level + 1) return rc;
    rec.last = h.max_level;
    rec.stop = -1;
  } else {
    // no cap: levels 0 .. next_level - 1 ran, and next_level's compaction
    // found no row with enough neighbours
    if (h.next_level < 1 || h.next_level > kMaxLevel) return rc;
    rec.last = h.next_level - 1;
    rec.stop = h.next_level;
  }
  rec.cor_nan = h.cor_nan;
  rec.rec = h.rec;
  rec.valid = true;
  // the next fresh run of this instance replays these launches; a failed
  // capture only means it runs eagerly again
  if (graph_capture(e, id, rec)) graph_drop(e);
  return HPC_OK;
}

// Candidate launch shapes of a level's CI-test kernel: (ranks per task,
// tasks per workgroup, task order 1 = device queue / 2 = workgroup order),
// or for level 1 edges per workgroup in the second field.
struct TuneCand {
  int rpt, tpb, queue;
};
static std::vector<TuneCand> tune_candidates(int level) {
  std::vector<TuneCand> cand;
  if (level == 1) {
    for (int epb : {8, 16, 32, 64, 128, 256}) cand.push_back({0, epb, 0});
  } else {
    for (int rpt : {256, 512, 1024, 2048, 4096}) {
      for (int tpb : {4, 8, 16, 32, 64, 128})
        for (int q : {1, 2}) cand.push_back({rpt, tpb, q});
      cand.push_back({rpt, 64, 3});  // one share per workgroup (tpb unused)
    }
  }
  return cand;
}

static void set_shape(Engine &e, int level, const TuneCand &c) {
  if (level == 1) {
    e.epb_level1 = c.tpb;
  } else {
    e.rpt_level[level] = c.rpt;
    e.tpb_level[level] = c.tpb;
    e.queue_level[level] = c.queue;
  }
}

// The batch tuning pass's timing of one level on the current step: every
// candidate twice (its own task plan first), best of the two; the fastest
// shape stays set for the level and is recorded in e.tune_out.
static int tune_step_level(Engine &e, Layout L, int level) {
  L.max_deg = e.max_deg;
  const std::vector<TuneCand> cand = tune_candidates(level);
  hpc_tune_entry t;
  memset(&t, 0, sizeof(t));
  t.level = level;
  const int q_default = e.use_queue ? 1 : 2;
  double best = 1e300;
  TuneCand pick = cand[0];
  for (const TuneCand &c : cand) {
    set_shape(e, level, c);
    double ms_min = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      HPC_HIPCHK(hipMemsetAsync(e.d_fp, 0xFF, (size_t)e.n_edges * sizeof(int64_t), e.stream));
      if (level >= 2)
        launch_task_plan(e.stream, level, L, engine_rpt(e, level), nullptr, e.d_row_tasks,
                         e.d_task_off);
      HPC_HIPCHK(hipEventRecord(e.ev[6], e.stream));
      if (level == 1)
        launch_search_l1(e.stream, L, e.n_edges, nullptr, false, engine_epb(e),
                         search_lds_limit(e.lds_limit));
      else
        launch_search_ln(e.stream, level, L, engine_rpt(e, level), e.d_row_tasks, e.d_task_off,
                         false, engine_tpb(e, level), search_lds_limit(e.lds_limit),
                         device_queue(e, level));
      HPC_HIPCHK(hipGetLastError());
      HPC_HIPCHK(hipEventRecord(e.ev[7], e.stream));
      HPC_HIPCHK(hipEventSynchronize(e.ev[7]));
      ms_min = std::min(ms_min, (double)elapsed(e.ev[6], e.ev[7]));
    }
    const bool dflt = level == 1 ? c.tpb == 32 : (c.rpt == 512 && c.tpb == 16 && c.queue == q_default);
    if (dflt) t.ms_default = ms_min;
    if (ms_min < best) {
      best = ms_min;
      pick = c;
    }
  }
  set_shape(e, level, pick);
  if (level == 1) {
    t.edges_per_block = pick.tpb;
  } else {
    t.ranks_per_task = pick.rpt;
    t.tasks_per_block = pick.tpb;
    t.queue = pick.queue;
  }
  t.ms_best = best;
  e.tune_out.push_back(t);
  e.tuned[level] = true;
  return HPC_OK;
}

int engine_batch_autotune(Engine &e, std::vector<hpc_tune_entry> &out) {
  std::vector<int> ids;
  for (size_t k = 0; k < e.inst.size(); ++k)
    if (e.inst[k].used && !e.inst[k].finished) ids.push_back((int)k);
  out.clear();
  if (ids.empty()) return HPC_OK;
  for (int id : ids)
    if (e.inst[id].fast) return hpc_fail(HPC_ESTATE, "the batch tuner needs exact instances");
  // from every instance's start (instances with a correlation keep it)
  for (int id : ids) {
    InstHost &h = e.inst[id];
    if (h.have_cor) {
      int rc = engine_reset(e, id);
      if (rc) return rc;
      h.hist.clear();
    }
  }
  e.tune_pass = true;
  e.tune_out.clear();
  for (auto &v : e.tuned) v = false;
  int rc = HPC_OK;
  for (int guard = 0; guard < 4 * kMaxLevel + 8; ++guard) {
    bool left = false;
    for (int id : ids) left |= !e.inst[id].finished;
    if (!left) break;
    rc = engine_step(e, ids);
    if (rc) break;
  }
  e.tune_pass = false;
  if (rc) return rc;
  out = e.tune_out;
  std::sort(out.begin(), out.end(),
            [](const hpc_tune_entry &a, const hpc_tune_entry &b) { return a.level < b.level; });
  // the pass's results go: every instance back to its start
  for (int id : ids) {
    InstHost &h = e.inst[id];
    rc = engine_reset(e, id);
    if (rc) return rc;
    h.hist.clear();
    h.error = 0;
    h.message.clear();
  }
  return HPC_OK;
}

int engine_step(Engine &e, const std::vector<int> &ids) {
  HPC_HIPCHK(hipSetDevice(e.device));
  int rc = 0;
  // correlation of instances that have none yet
  std::vector<int> need, zero;
  for (int id : ids) {
    InstHost &h = e.inst[id];
    if (h.finished || h.have_cor) continue;
    if (!h.have_x) return hpc_fail(HPC_ESTATE, "instance %d has no observations", id);
    need.push_back(id);
    zero.push_back(0);
  }
  // an instance whose host-side preparation fails (an allocation) stops
  // alone with its error; the others go on
  auto drop = [&](int id, int code) {
    InstHost &h = e.inst[id];
    h.error = code;
    h.message = hpc_last_error();
    h.finished = true;
  };
  if (!need.empty()) {
    std::vector<int> ok;
    std::vector<ZeroSpan> clears;
    for (int id : need) {
      rc = engine_prepare_corr(e, id, &clears);
      if (rc) drop(id, rc);
      else ok.push_back(id);
    }
    rc = engine_zero_spans(e, clears);
    if (rc) return rc;
    need.swap(ok);
    zero.assign(need.size(), 0);
  }
  if (!need.empty()) {
    rc = engine_pack(e, need, zero);
    if (!rc) rc = engine_corr(e, nullptr, nullptr);
    if (rc) return rc;
  }
  std::vector<int> act, lev;
  for (int id : ids) {
    InstHost &h = e.inst[id];
    if (h.finished) continue;
    act.push_back(id);
    lev.push_back(h.next_level);
  }
  if (act.empty()) return HPC_OK;
  auto t0 = std::chrono::steady_clock::now();
  // room for each level's sepset records before the table is packed (the
  // worst case depends only on the edges left, known from the last step)
  for (size_t k = 0; k < act.size();) {
    if (lev[k] >= 0 && lev[k] <= kMaxLevel) {
      rc = prepare_level(e, e.inst[act[k]], lev[k]);
      if (rc) {
        drop(act[k], rc);
        act.erase(act.begin() + (long)k);
        lev.erase(lev.begin() + (long)k);
        continue;
      }
    }
    ++k;
  }
  if (act.empty()) return HPC_OK;
  rc = engine_pack(e, act, lev, /*zero_ctr=*/true);
  if (rc) return rc;
  for (int id : act) e.inst[id].fresh = false;
  // observations and z-scores are only read by the correlation, which the
  // pack above has waited for; blocks freed here may go to an upload on
  // another stream
  for (int id : need)
    if (e.inst[id].drop_obs) {
      InstHost &h = e.inst[id];
      if (h.fast) {
        // exact rechecks read the z-scores until the instance finishes
        h.bytes -= std::min(h.bytes, (size_t)h.p * h.n * 8 + (size_t)h.p * h.npad);
        ifree(e, h.x); ifree(e, h.nanm);
      } else {
        free_obs(e, h);
      }
      h.have_x = false;
    }
  const int K = (int)act.size();
  Layout L = layout(e);
  bool any0 = false, any1 = false;
  for (int l : lev) { any0 |= l == 0; any1 |= l >= 1; }
  // every instance at one level >= 1: one compaction kernel that also plans
  // the level's tasks and publishes the counters
  int lvl_u = lev[0];
  for (int l : lev) if (l != lvl_u) lvl_u = -1;
  const bool fused = lvl_u >= 1 && lvl_u <= kMaxLevel;
  HPC_HIPCHK(hipEventRecord(e.ev[0], e.stream));
  if (any0) {
    launch_level0(e.stream, L, e.h_inst, e.d_prefix + 2 * (K + 1), e.h_prefix);
    if (K > 1) e.host_upload_pending = true;  // the tile offsets
  }
  HPC_HIPCHK(hipGetLastError());
  HPC_HIPCHK(hipEventRecord(e.ev[1], e.stream));
  if (fused) {
    const Publish pub = next_publish(e);
    launch_compact_fused(e.stream, L, lvl_u, engine_rpt(e, lvl_u), e.d_blk, e.d_row_tasks,
                         e.d_task_off, e.cap_edges, pub);
    HPC_HIPCHK(hipGetLastError());
    HPC_HIPCHK(hipEventRecord(e.ev[2], e.stream));
    // the totals arrive while the fill kernel runs
    rc = wait_published(e, pub.seq);
    if (rc) return rc;
  } else {
    if (any1) launch_compact(e.stream, L);
    HPC_HIPCHK(hipGetLastError());
    HPC_HIPCHK(hipEventRecord(e.ev[2], e.stream));
    rc = read_counters(e);
    if (rc) return rc;
  }
  e.n_edges = any1 ? (int64_t)e.h_ctr->n_edges : 0;
  e.max_deg = any1 ? e.h_ctr->max_deg : 0;
  L.max_deg = e.max_deg;
  // (the compaction set every first passing rank to -1)
  // instances whose level has nothing to test and no cap stop here; a level
  // beyond the kernels' reach is an error for its instance when it has tests,
  // else an empty record.  skip: 1 = no record, 2 = empty record.
  std::vector<char> skip(K, 0);
  bool retable = false;
  for (int k = 0; k < K; ++k) {
    if (lev[k] < 1) continue;
    InstHost &h = e.inst[act[k]];
    h.max_deg = e.h_ictr[k].max_deg;
    h.compacted = true;
    h.ever_compacted = true;
    const bool testable = h.max_deg - 1 >= lev[k];
    // beyond the kernels' levels, or more conditioning sets than a plan holds
    const bool beyond = lev[k] > kMaxLevel || (testable && !engine_plannable(e, h.max_deg, lev[k]));
    if ((!testable && h.max_level < 0) || beyond) {
      if (testable) {
        h.error = HPC_ELEVEL;
        h.message = "level " + std::to_string(lev[k]) + ": C(" + std::to_string(h.max_deg - 1) +
                    ", level) conditioning sets per edge is beyond what one search plans";
      }
      skip[k] = (testable || h.max_level < 0) ? 1 : 2;
      e.h_inst[k].level = -1;
      retable = true;
      continue;
    }
  }
  if (retable) {
    rc = upload_table(e);
    if (rc) return rc;
  }
  std::vector<int> levels_ge2;
  bool run1 = false;
  for (int k = 0; k < K; ++k) {
    const int l = e.h_inst[k].level;
    if (l == 1) run1 = true;
    if (l >= 2 && l <= kMaxLevel &&
        std::find(levels_ge2.begin(), levels_ge2.end(), l) == levels_ge2.end())
      levels_ge2.push_back(l);
  }
  std::sort(levels_ge2.begin(), levels_ge2.end());
  // a tuning pass (engine_batch_autotune): each exact search level met for
  // the first time is timed on this step's snapshot with every candidate
  // shape and the fastest kept; the searches below then run with it (and
  // plan their tasks again: the candidates' plans replaced the compaction's)
  bool replan = false;
  if (e.tune_pass && e.n_edges > 0) {
    std::vector<int> lv;
    if (run1 && !any_fast_at(e, 1)) lv.push_back(1);
    for (int l : levels_ge2)
      if (!any_fast_at(e, l)) lv.push_back(l);
    for (int l : lv) {
      if (e.tuned[l]) continue;
      rc = tune_step_level(e, L, l);
      if (rc) return rc;
      replan = true;
    }
    if (replan)
      HPC_HIPCHK(hipMemsetAsync(e.d_fp, 0xFF, (size_t)e.n_edges * sizeof(int64_t), e.stream));
  }
  // search kernels, each bracketed by its own event pair (task plans of
  // levels >= 2 run outside the pairs)
  int npairs = 0;
  if (e.n_edges > 0) {
    if (run1) {
      HPC_HIPCHK(hipEventRecord(e.sev[2 * npairs], e.stream));
      const bool f1 = any_fast_at(e, 1);
      launch_search_l1(e.stream, L, e.n_edges, nullptr, f1, engine_epb(e),
                       search_lds_limit(e.lds_limit));
      if (f1) launch_recheck(e.stream, 1, L);
      HPC_HIPCHK(hipEventRecord(e.sev[2 * npairs + 1], e.stream));
      ++npairs;
    }
    for (int l : levels_ge2) {
      // (the fused compaction planned the tasks of its level)
      if (!fused || replan)
        launch_task_plan(e.stream, l, L, engine_rpt(e, l), nullptr, e.d_row_tasks, e.d_task_off);
      HPC_HIPCHK(hipEventRecord(e.sev[2 * npairs], e.stream));
      const bool fl = any_fast_at(e, l);
      launch_search_ln(e.stream, l, L, engine_rpt(e, l), e.d_row_tasks, e.d_task_off, fl,
                       engine_tpb(e, l), search_lds_limit(e.lds_limit), device_queue(e, l));
      if (fl) launch_recheck(e.stream, l, L);
      HPC_HIPCHK(hipEventRecord(e.sev[2 * npairs + 1], e.stream));
      ++npairs;
    }
    HPC_HIPCHK(hipGetLastError());
  }
  if (any1) {
    HPC_HIPCHK(hipEventRecord(e.ev[4], e.stream));
    if (e.n_edges > 0) {
      // the removal kernel publishes the counters itself
      const Publish pub = next_publish(e);
      launch_commit(e.stream, L, e.n_edges, pub);
      HPC_HIPCHK(hipGetLastError());
      HPC_HIPCHK(hipEventRecord(e.ev[5], e.stream));
      rc = wait_published(e, pub.seq);
    } else {
      HPC_HIPCHK(hipEventRecord(e.ev[5], e.stream));
      rc = read_counters(e);
    }
    if (rc) return rc;
  }
  if (e.h_ctr->overflow) {
    // a recheck queue overflowed and the removal skipped itself.  Level 0 of
    // the step's level-0 instances did run: their adjacency goes back to the
    // complete graph level 0 starts from.  Then exact correlations for the
    // fast instances and the same step again.
    for (int k = 0; k < K; ++k)
      if (lev[k] == 0) {
        const InstHost &h = e.inst[act[k]];
        launch_adj_init(e.stream, h.adj, h.p, h.W);
      }
    HPC_HIPCHK(hipGetLastError());
    rc = engine_exact_from_z(e, act);
    if (rc) return rc;
    return engine_step(e, ids);
  }
  e.rechecks = e.h_ctr->rechecks;
  const double wall =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const double ms_l0 = elapsed(e.ev[0], e.ev[1]), ms_c = elapsed(e.ev[1], e.ev[2]);
  const double ms_m = any1 ? elapsed(e.ev[4], e.ev[5]) : 0.0;
  double ms_s = 0.0;
  for (int k = 0; k < npairs; ++k) ms_s += elapsed(e.sev[2 * k], e.sev[2 * k + 1]);
  for (int k = 0; k < K; ++k) {
    InstHost &h = e.inst[act[k]];
    const InstCounters &c = e.h_ictr[k];
    if (skip[k] == 2) {
      hpc_level_stats s;
      memset(&s, 0, sizeof(s));
      s.level = lev[k];
      s.max_deg = h.max_deg;
      s.edges_left = h.undirected;
      s.ms_compact = ms_c;
      s.ms_wall = wall;
      h.hist.push_back(s);
    } else if (!skip[k]) {
      hpc_level_stats s;
      memset(&s, 0, sizeof(s));
      s.level = lev[k];
      s.max_deg = lev[k] == 0 ? h.p : h.max_deg;
      s.tests = (int64_t)c.tests;
      s.tests_max = (int64_t)c.tests_max;
      s.edges = (int64_t)c.edges;
      s.removed = (int64_t)c.removed;
      h.undirected -= s.removed;
      if (lev[k] == 0) h.cor_nan = c.has_nan ? 1 : 0;
      s.edges_left = h.undirected;
      if (K == 1) {
        s.pinv_fallbacks = (int64_t)e.h_ctr->pinv_fallbacks;
        s.rechecks = (int64_t)e.h_ctr->rechecks;
      }
      if (lev[k] == 0) {
        s.ms_test = ms_l0;
      } else {
        s.ms_compact = ms_c;
        s.ms_test = ms_s;
        s.ms_commit = ms_m;
      }
      s.ms_wall = wall;  // of the whole step (shared by the instances in it)
      h.hist.push_back(s);
      h.rec_used += (size_t)c.rec_words;
      if (lev[k] >= 1) engine_note_level(e, act[k], lev[k]);
      if (s.removed) h.compacted = false;
    }
    h.next_level += 1;
    if (skip[k] == 1 || h.error || (h.max_level >= 0 && h.next_level > h.max_level))
      h.finished = true;
  }
  return HPC_OK;
}

TaskQueue device_queue(const Engine &e, int level) {
  const int q = (level >= 0 && level <= kMaxLevel) ? e.queue_level[level] : 0;
  const bool on = q == 1 || (q == 0 && e.use_queue);
  return TaskQueue{on ? e.d_qhead : nullptr, nullptr, q == 3 ? 1 : 0};
}

bool engine_plannable(const Engine &e, int max_deg, int level) {
  if (level < 2 || max_deg - 1 < level) return true;
  const size_t row = (size_t)(max_deg - 1);
  if (row >= e.cap_binom) return false;
  const unsigned long long total = e.h_binom[row * kBinomK + level];
  return row_task_count(total, max_deg, engine_rpt(e, level)) >= 0;
}

int engine_rpt(const Engine &e, int level) {
  const int v = (level >= 0 && level <= kMaxLevel) ? e.rpt_level[level] : 0;
  return v > 0 ? v : e.ranks_per_task;
}

int engine_tpb(const Engine &e, int level) {
  const int v = (level >= 0 && level <= kMaxLevel) ? e.tpb_level[level] : 0;
  return v > 0 ? v : 16;
}

int engine_epb(const Engine &e) { return e.epb_level1 > 0 ? e.epb_level1 : 64; }

int engine_check_tuning(const hpc_tune_entry *t, int count) {
  if (count < 0 || (count > 0 && !t)) return hpc_fail(HPC_EINVAL, "no tuning entries");
  for (int k = 0; k < count; ++k) {
    const hpc_tune_entry &x = t[k];
    if (x.level < 1 || x.level > kMaxLevel)
      return hpc_fail(HPC_EINVAL, "tuning entry %d: level %d outside 1..%d", k, x.level, kMaxLevel);
    if (x.level == 1) {
      if (x.edges_per_block < 1 || x.edges_per_block > 4096)
        return hpc_fail(HPC_EINVAL, "tuning entry %d: edges per workgroup %d outside 1..4096", k,
                        x.edges_per_block);
    } else {
      if (x.ranks_per_task < 64 || x.ranks_per_task % 64 != 0 || x.ranks_per_task > (1 << 24))
        return hpc_fail(HPC_EINVAL,
                        "tuning entry %d: ranks per task %d is not a multiple of 64 in [64, 2^24]",
                        k, x.ranks_per_task);
      if (x.tasks_per_block < 1 || x.tasks_per_block > 1024)
        return hpc_fail(HPC_EINVAL, "tuning entry %d: tasks per workgroup %d outside 1..1024", k,
                        x.tasks_per_block);
      if (x.queue < 0 || x.queue > 3)
        return hpc_fail(HPC_EINVAL, "tuning entry %d: queue %d outside 0..3", k, x.queue);
    }
  }
  return HPC_OK;
}

int engine_apply_tuning(Engine &e, const hpc_tune_entry *t, int count) {
  const int rc = engine_check_tuning(t, count);
  if (rc) return rc;
  for (int k = 0; k < count; ++k) {
    const hpc_tune_entry &x = t[k];
    if (x.level == 1) {
      e.epb_level1 = x.edges_per_block;
    } else {
      e.rpt_level[x.level] = x.ranks_per_task;
      e.tpb_level[x.level] = x.tasks_per_block;
      e.queue_level[x.level] = x.queue;
    }
  }
  return HPC_OK;
}

int engine_autotune(Engine &e, int id, int max_level, std::vector<hpc_tune_entry> &out) {
  InstHost &h = e.inst[id];
  if (!h.have_cor) return hpc_fail(HPC_ESTATE, "no correlation matrix");
  if (max_level < 0 || max_level > kMaxLevel) max_level = kMaxLevel;
  out.clear();
  int rc = engine_reset(e, id);
  if (!rc) rc = engine_pack(e, {id}, {0});
  if (!rc) rc = engine_level0(e, nullptr);
  if (rc) return rc;
  for (int l = 1; l <= max_level; ++l) {
    rc = engine_pack(e, {id}, {l});
    if (!rc) rc = engine_compact(e, nullptr);
    if (rc) return rc;
    if (h.max_deg - 1 < l || e.n_edges == 0) break;
    hpc_tune_entry t;
    memset(&t, 0, sizeof(t));
    t.level = l;
    const std::vector<TuneCand> cand = tune_candidates(l);
    double best = 1e300;
    TuneCand pick = cand[0];
    const int q_default = e.use_queue ? 1 : 2;
    for (const auto &c : cand) {
      set_shape(e, l, c);
      double ms_min = 1e300;
      for (int rep = 0; rep < 2; ++rep) {
        HPC_HIPCHK(hipMemsetAsync(e.d_fp, 0xFF, (size_t)e.n_edges * sizeof(int64_t), e.stream));
        float ms = 0.f;
        rc = engine_search(e, l, nullptr, 0, &ms);
        if (rc) return rc;
        ms_min = std::min(ms_min, (double)ms);
      }
      const bool dflt = l == 1 ? c.tpb == 32 : (c.rpt == 512 && c.tpb == 16 && c.queue == q_default);
      if (dflt) t.ms_default = ms_min;
      if (ms_min < best) { best = ms_min; pick = c; }
    }
    if (l == 1) {
      e.epb_level1 = pick.tpb;
      t.edges_per_block = pick.tpb;
    } else {
      e.rpt_level[l] = pick.rpt;
      e.tpb_level[l] = pick.tpb;
      e.queue_level[l] = pick.queue;
      t.ranks_per_task = pick.rpt;
      t.tasks_per_block = pick.tpb;
      t.queue = pick.queue;
    }
    t.ms_best = best;
    out.push_back(t);
    // the level itself, with the chosen shape
    HPC_HIPCHK(hipMemsetAsync(e.d_fp, 0xFF, (size_t)e.n_edges * sizeof(int64_t), e.stream));
    rc = engine_search(e, l, nullptr, 0, nullptr);
    if (!rc) rc = engine_commit(e, nullptr);
    if (rc) return rc;
  }
  return engine_reset(e, id);
}

}  // namespace hpc
