// Batched and served forms of the skeleton (no counterpart in the reference):
//
//   hpc_batch  -- many independent instances advanced together: every step
//                 runs one level of every unfinished instance through one set
//                 of launches (engine_step), whatever their p, n, alpha and
//                 level.
//   hpc_server -- continuous batching: a worker thread admits submitted
//                 requests while their device footprint fits a budget, steps
//                 every admitted instance, retires finished ones and copies
//                 their results out.  A request's result is the same alone and
//                 in any mix: the per-test arithmetic and the sepset choice do
//                 not depend on what else is in a step.
#pragma clang fp contract(off)

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "hpc_internal.hpp"

using namespace hpc;

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------
struct hpc_batch {
  Engine eng;
  int steps = 0;
};

// Instance blocks of batched and served engines come from 512 MiB slabs.
static constexpr size_t kSlabBytes = (size_t)512 << 20;

extern "C" int hpc_footprint(uint64_t p, int n, int max_level, uint64_t *bytes) {
  if (!bytes || p < 2 || n < 1) return hpc_fail(HPC_EINVAL, "bad argument");
  if (p > (uint64_t)kMaxP) return hpc_fail(HPC_EINVAL, "p = %llu outside [2, %d]",
                                           (unsigned long long)p, kMaxP);
  *bytes = engine_footprint((int)p, n, max_level);
  return HPC_OK;
}

extern "C" int hpc_pack_check(const uint64_t *p, int count) {
  if (count < 0 || (count > 0 && !p)) return hpc_fail(HPC_EINVAL, "bad argument");
  int64_t edges = 0;
  for (int k = 0; k < count; ++k) {
    if (p[k] < 2 || p[k] > (uint64_t)kMaxP)
      return hpc_fail(HPC_EINVAL, "p = %llu outside [2, %d]", (unsigned long long)p[k], kMaxP);
    edges += (int64_t)p[k] * (int64_t)(p[k] - 1);
    if (edges > kMaxPackEdges)
      return hpc_fail(HPC_ENOMEM,
                      "%d instances hold more than %lld ordered edges, what one step indexes",
                      k + 1, (long long)kMaxPackEdges);
  }
  return HPC_OK;
}

extern "C" int hpc_batch_create(int device, hpc_batch **out) {
  if (!out) return hpc_fail(HPC_EINVAL, "out is NULL");
  *out = nullptr;
  hpc_batch *b = new hpc_batch();
  b->eng.arena.slab_min = kSlabBytes;
  int rc = engine_init(b->eng, device);
  if (rc) { engine_free(b->eng); delete b; return rc; }
  *out = b;
  return HPC_OK;
}

extern "C" int hpc_batch_destroy(hpc_batch *b) {
  if (!b) return HPC_OK;
  engine_free(b->eng);
  delete b;
  return HPC_OK;
}

static int check_id(hpc_batch *b, int id) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  if (id < 0 || id >= (int)b->eng.inst.size() || !b->eng.inst[id].used)
    return hpc_fail(HPC_EINVAL, "no instance %d", id);
  return HPC_OK;
}

extern "C" int hpc_batch_add(hpc_batch *b, uint64_t p, int n, double alpha, int max_level,
                             const double *x, int *id) {
  if (!b || !x || !id) return hpc_fail(HPC_EINVAL, "NULL argument");
  if (p < 2 || p > (uint64_t)kMaxP)
    return hpc_fail(HPC_EINVAL, "p = %llu outside [2, %d]", (unsigned long long)p, kMaxP);
  // every unfinished instance is packed into one step, whose neighbour lists
  // are indexed with int32 row pointers
  std::vector<uint64_t> ps{p};
  for (const InstHost &h : b->eng.inst)
    if (h.used && !h.finished) ps.push_back((uint64_t)h.p);
  int rc = hpc_pack_check(ps.data(), (int)ps.size());
  if (rc) return rc;
  rc = engine_add(b->eng, (int)p, n, alpha, max_level, id);
  if (rc) return rc;
  rc = engine_set_observations(b->eng, *id, x);
  if (rc) { engine_remove(b->eng, *id); return rc; }
  b->eng.inst[*id].drop_obs = true;
  return HPC_OK;
}

extern "C" int hpc_batch_remove(hpc_batch *b, int id) {
  int rc = check_id(b, id);
  if (rc) return rc;
  engine_remove(b->eng, id);
  return HPC_OK;
}

static std::vector<int> live(Engine &e) {
  std::vector<int> ids;
  for (size_t k = 0; k < e.inst.size(); ++k)
    if (e.inst[k].used && !e.inst[k].finished) ids.push_back((int)k);
  return ids;
}

extern "C" int hpc_batch_step(hpc_batch *b, int *unfinished) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  std::vector<int> ids = live(b->eng);
  int rc = engine_step(b->eng, ids);
  if (rc) return rc;
  ++b->steps;
  if (unfinished) *unfinished = (int)live(b->eng).size();
  return HPC_OK;
}

extern "C" int hpc_batch_run(hpc_batch *b, int *steps) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  int n = 0, left = 0;
  do {
    int rc = hpc_batch_step(b, &left);
    if (rc) return rc;
    ++n;
  } while (left > 0);
  if (steps) *steps = n;
  // every instance's results to host memory in one round (the exports then
  // read them without a device wait each)
  std::vector<int> all;
  for (size_t k = 0; k < b->eng.inst.size(); ++k)
    if (b->eng.inst[k].used) all.push_back((int)k);
  return engine_prefetch_results(b->eng, all);
}

extern "C" int hpc_batch_status(hpc_batch *b, int id, int *finished, int *levels, int *error) {
  int rc = check_id(b, id);
  if (rc) return rc;
  const InstHost &h = b->eng.inst[id];
  if (finished) *finished = h.finished ? 1 : 0;
  if (levels) *levels = (int)h.hist.size();
  if (error) *error = h.error;
  return HPC_OK;
}

extern "C" int hpc_batch_get_stats(hpc_batch *b, int id, hpc_level_stats *stats, int cap,
                                   int *levels) {
  int rc = check_id(b, id);
  if (rc) return rc;
  const InstHost &h = b->eng.inst[id];
  for (int k = 0; k < (int)h.hist.size() && k < cap; ++k) stats[k] = h.hist[k];
  if (levels) *levels = (int)h.hist.size();
  return HPC_OK;
}

extern "C" int hpc_batch_max_cond_size(hpc_batch *b, int id) {
  return check_id(b, id) ? 0 : b->eng.inst[id].mc;
}

extern "C" int hpc_batch_get_adj(hpc_batch *b, int id, int32_t *adj) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (!adj) return hpc_fail(HPC_EINVAL, "adj is NULL");
  return export_adj(b->eng, id, adj);
}

extern "C" int hpc_batch_get_sepsets(hpc_batch *b, int id, int32_t *sep, int mc) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (!sep || mc < 1) return hpc_fail(HPC_EINVAL, "bad argument");
  return export_sepsets(b->eng, id, sep, mc);
}

extern "C" int hpc_batch_get_sepset_list(hpc_batch *b, int id, int32_t *out, int64_t cap,
                                         int64_t *words) {
  int rc = check_id(b, id);
  if (rc) return rc;
  return export_sepset_list(b->eng, id, out, cap, words);
}

extern "C" int hpc_batch_get_cor(hpc_batch *b, int id, double *cor) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (!cor) return hpc_fail(HPC_EINVAL, "cor is NULL");
  return export_cor(b->eng, id, cor);
}

extern "C" int hpc_batch_set_corr_mode(hpc_batch *b, int id, int mode) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (mode != HPC_CORR_CANON && mode != HPC_CORR_F32)
    return hpc_fail(HPC_EINVAL, "unknown correlation mode %d", mode);
  InstHost &h = b->eng.inst[id];
  if (h.have_cor) return hpc_fail(HPC_ESTATE, "instance %d already has its correlation", id);
  h.corr_mode = mode;
  return HPC_OK;
}

// The tuning knobs a batch or a server takes: checked, then applied.
static int check_engine_tuning(int param, int value) {
  switch (param) {
    case HPC_TUNE_RANKS_PER_TASK:
      if (value < 64 || value % 64 != 0 || value > (1 << 24))
        return hpc_fail(HPC_EINVAL, "ranks per task must be a multiple of 64 in [64, 2^24]");
      return HPC_OK;
    case HPC_TUNE_LDS_LIMIT:
      if (value < 0 || (size_t)value > kSearchLdsLimit)
        return hpc_fail(HPC_EINVAL, "LDS limit must be in [0, %zu] bytes", kSearchLdsLimit);
      return HPC_OK;
    case HPC_TUNE_WAIT_LIMIT:
      if (value < 0) return hpc_fail(HPC_EINVAL, "wait limit must be >= 0 seconds");
      return HPC_OK;
    case HPC_TUNE_QUEUE:
      if (value != 0 && value != 1) return hpc_fail(HPC_EINVAL, "queue must be 0 or 1");
      return HPC_OK;
    default:
      return hpc_fail(HPC_EINVAL, "tuning parameter %d does not apply here", param);
  }
}

static int set_engine_tuning(Engine &e, int param, int value) {
  const int rc = check_engine_tuning(param, value);
  if (rc) return rc;
  if (param == HPC_TUNE_RANKS_PER_TASK) e.ranks_per_task = value;
  else if (param == HPC_TUNE_LDS_LIMIT) e.lds_limit = (size_t)value;
  else if (param == HPC_TUNE_QUEUE) e.use_queue = value != 0;
  else e.wait_limit_s = (double)value;
  return HPC_OK;
}

extern "C" int hpc_batch_set_tuning(hpc_batch *b, int param, int value) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  const int rc = set_engine_tuning(b->eng, param, value);
  return rc;
}

extern "C" int hpc_batch_autotune(hpc_batch *b, hpc_tune_entry *out, int cap, int *levels) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  if (cap < 0 || (cap > 0 && !out)) return hpc_fail(HPC_EINVAL, "bad output buffer");
  std::vector<hpc_tune_entry> res;
  const int rc = engine_batch_autotune(b->eng, res);
  if (rc) return rc;
  for (int k = 0; k < (int)res.size() && k < cap; ++k) out[k] = res[k];
  if (levels) *levels = (int)res.size();
  return HPC_OK;
}

extern "C" int hpc_batch_inject_failure(hpc_batch *b, int id, int level) {
  int rc = check_id(b, id);
  if (rc) return rc;
  b->eng.inst[id].inject_level = level < 0 ? -1 : level;
  return HPC_OK;
}

extern "C" int hpc_batch_apply_tuning(hpc_batch *b, const hpc_tune_entry *entries, int count) {
  if (!b) return hpc_fail(HPC_EINVAL, "batch is NULL");
  return engine_apply_tuning(b->eng, entries, count);
}

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int p = 0, n = 0, max_level = 0;
  double alpha = 0.0;
  std::vector<double> x;        // copy of the observations (unless borrowed)
  const double *src = nullptr;  // observations to upload: x, or the caller's
  size_t count = 0;
  double *dst = nullptr;        // device block of the observations
  int corr_mode = HPC_CORR_CANON;
  hipError_t up_err = hipSuccess;
  size_t bytes = 0;
  int slot = -1;       // engine instance while admitted
  int state = 0;       // 0 queued, 1 running, 2 done, 3 failed
  int error = 0;
  std::string message;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<unsigned long long> bits;  // adjacency, [p][W] words
  std::vector<int32_t> rec;              // sepset records sorted by (i, j)
  int mc = 0;
  int levels = 0;
  int64_t tests = 0;
  int inject_level = -1;  // hpc_server_inject_failure
};

}  // namespace

struct hpc_server {
  Engine eng;
  size_t budget = 0, in_use = 0;
  // ordered edges p(p - 1) of the admitted requests: one step packs them all
  // and indexes its neighbour lists with int32 row pointers
  int64_t edges_in_use = 0;
  int max_active = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::vector<std::shared_ptr<Request>> running;
  std::map<int64_t, std::shared_ptr<Request>> all;
  int64_t next_id = 1;
  int64_t submits = 0;
  int corr_mode = HPC_CORR_CANON;  // of requests admitted from now on
  // launch shapes and knobs for the worker to apply before its next step
  std::vector<hpc_tune_entry> tune;
  std::vector<std::pair<int, int>> knobs;
  bool tune_pending = false;
  std::map<int64_t, int> inject;  // failures for requests not yet submitted
  bool stop = false;
  std::thread worker;
  // observation uploads: the uploader thread copies admitted requests' data
  // on its own stream while the worker keeps stepping the running ones
  hipStream_t up_stream = nullptr;
  std::thread uploader;
  std::condition_variable cv_up;
  std::deque<std::shared_ptr<Request>> up_queue;
  std::vector<std::shared_ptr<Request>> up_done;
  int uploading = 0;
  bool stop_up = false;
  // statistics
  int64_t steps = 0;
  int peak_active = 0;
  size_t peak_bytes = 0;
  double busy_ms = 0.0;
};

// Copies a finished request's results to the host and frees its instance;
// returns its error code.  Runs on the worker without the lock: the request
// is published (state set) afterwards, under the lock.
static int server_collect(hpc_server *s, Request &r) {
  InstHost &h = s->eng.inst[r.slot];
  r.levels = (int)h.hist.size();
  for (auto &l : h.hist) r.tests += l.tests;
  r.mc = h.mc;
  int rc = h.error;
  if (rc && r.message.empty() && !h.message.empty()) r.message = h.message;
  if (!rc) rc = fetch_adj_bits(s->eng, r.slot, r.bits);
  if (!rc) rc = fetch_records(s->eng, r.slot, r.rec);
  if (rc && r.message.empty()) r.message = hpc_last_error();
  engine_remove(s->eng, r.slot);
  r.slot = -1;
  return rc;
}

static int64_t request_edges(const Request &r) { return (int64_t)r.p * (r.p - 1); }

static void fail_request(hpc_server *s, Request &r, int rc, const std::string &msg) {
  r.error = rc;
  r.message = msg;
  r.state = 3;
  r.t_done = Clock::now();
  s->in_use -= r.bytes;
  s->edges_in_use -= request_edges(r);
}

static void uploader_loop(hpc_server *s) {
  (void)hipSetDevice(s->eng.device);
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_up.wait(lk, [&] { return s->stop_up || !s->up_queue.empty(); });
      if (s->up_queue.empty()) return;
      r = s->up_queue.front();
      s->up_queue.pop_front();
    }
    const hipError_t err =
        upload_f64_h2d(s->eng.device, r->dst, r->src, r->count, s->up_stream, true);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      r->up_err = err;
      s->up_done.push_back(r);
    }
    s->cv_work.notify_all();
  }
}

static void server_loop(hpc_server *s) {
  (void)hipSetDevice(s->eng.device);
  int64_t seen = -1;
  for (;;) {
    // admission: queued requests in arrival order while they fit; uploads
    // that have completed
    std::vector<std::shared_ptr<Request>> admit, ready;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        return s->stop || !s->queue.empty() || !s->running.empty() || !s->up_done.empty() ||
               s->uploading > 0;
      });
      if (s->stop && s->running.empty() && s->uploading == 0) return;
      seen = s->submits;
      size_t used = s->in_use;
      int active = (int)s->running.size() + s->uploading;
      while (!s->queue.empty()) {
        auto &r = s->queue.front();
        if (active >= s->max_active) break;
        if (used + r->bytes > s->budget && active > 0) break;
        if (s->edges_in_use + request_edges(*r) > kMaxPackEdges && active > 0) break;
        used += r->bytes;
        ++active;
        s->in_use += r->bytes;
        s->edges_in_use += request_edges(*r);
        admit.push_back(r);
        s->queue.pop_front();
      }
      ready.swap(s->up_done);
    }
    for (auto &r : admit) {
      int slot = -1;
      double *dst = nullptr;
      int rc = engine_add(s->eng, r->p, r->n, r->alpha, r->max_level, &slot);
      if (!rc) {
        s->eng.inst[slot].corr_mode = r->corr_mode;
        s->eng.inst[slot].inject_level = r->inject_level;
        rc = engine_alloc_observations(s->eng, slot, &dst);
        if (rc) engine_remove(s->eng, slot);
      }
      // the adjacency initialisation is on the engine stream: done before the
      // instance first runs
      std::lock_guard<std::mutex> lk(s->mu);
      r->t_admit = Clock::now();
      if (rc) {
        fail_request(s, *r, rc, hpc_last_error());
        s->cv_done.notify_all();
        continue;
      }
      r->slot = slot;
      r->dst = dst;
      r->state = 1;
      ++s->uploading;
      s->up_queue.push_back(r);
      s->cv_up.notify_one();
    }
    for (auto &r : ready) {
      std::lock_guard<std::mutex> lk(s->mu);
      --s->uploading;
      if (r->up_err != hipSuccess) {
        engine_remove(s->eng, r->slot);
        r->slot = -1;
        fail_request(s, *r, HPC_EHIP, std::string("observation upload failed: ") +
                                          hipGetErrorString(r->up_err));
        s->cv_done.notify_all();
        continue;
      }
      engine_observations_ready(s->eng, r->slot);
      s->eng.inst[r->slot].drop_obs = true;
      r->x.clear();
      r->x.shrink_to_fit();
      r->src = nullptr;
      s->running.push_back(r);
    }
    if (s->running.empty()) {
      // nothing to step: wait for an upload, a submission or the stop
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        return !s->up_done.empty() || s->submits != seen || (s->stop && s->uploading == 0);
      });
      continue;
    }
    std::vector<int> ids;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->tune_pending) {
        // checked by hpc_server_apply_tuning and hpc_server_set_tuning
        if (!s->tune.empty())
          (void)engine_apply_tuning(s->eng, s->tune.data(), (int)s->tune.size());
        for (const auto &kv : s->knobs) (void)set_engine_tuning(s->eng, kv.first, kv.second);
        s->tune.clear();
        s->knobs.clear();
        s->tune_pending = false;
      }
      for (auto &r : s->running) ids.push_back(r->slot);
      s->peak_active = std::max(s->peak_active, (int)ids.size());
      s->peak_bytes = std::max(s->peak_bytes, s->in_use);
    }
    if (ids.empty()) continue;
    // A lone request with nothing queued or uploading runs its remaining
    // levels in the one-instance loop (one wait per level, no table kernel).
    // A request that arrives meanwhile waits for it, as it would for a step.
    bool lone = false;
    if (ids.size() == 1) {
      std::lock_guard<std::mutex> lk(s->mu);
      const InstHost &h = s->eng.inst[ids[0]];
      lone = s->queue.empty() && s->uploading == 0 && s->up_done.empty() && h.have_cor &&
             h.next_level >= 1 && !h.finished;
    }
    auto t0 = Clock::now();
    const int rc = lone ? engine_run_single(s->eng, ids[0]) : engine_step(s->eng, ids);
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    const std::string msg = rc ? hpc_last_error() : "";
    // only the worker writes `running`, so it reads it without the lock
    std::vector<std::shared_ptr<Request>> done, keep;
    for (auto &r : s->running) {
      InstHost &h = s->eng.inst[r->slot];
      if (rc) {
        h.error = rc;
        r->message = msg;
      }
      (rc || h.finished ? done : keep).push_back(r);
    }
    std::vector<int> codes;
    if (!rc && done.size() > 1) {
      // the finished requests' results in one round of copies and one wait
      std::vector<int> slots;
      for (auto &r : done) slots.push_back(r->slot);
      (void)engine_prefetch_results(s->eng, slots);  // (a failure leaves the per-request fetches)
    }
    for (auto &r : done) codes.push_back(server_collect(s, *r));
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    s->busy_ms += ms;
    ++s->steps;
    for (size_t k = 0; k < done.size(); ++k) {
      Request &r = *done[k];
      s->in_use -= r.bytes;
      s->edges_in_use -= request_edges(r);
      r.error = codes[k];
      r.t_done = now;
      r.state = codes[k] ? 3 : 2;
    }
    s->running.swap(keep);
    s->cv_done.notify_all();
  }
}

extern "C" int hpc_server_create(int device, uint64_t hbm_budget, int max_active,
                                 hpc_server **out) {
  if (!out) return hpc_fail(HPC_EINVAL, "out is NULL");
  *out = nullptr;
  hpc_server *s = new hpc_server();
  s->eng.arena.slab_min = kSlabBytes;
  int rc = engine_init(s->eng, device);
  if (rc) { engine_free(s->eng); delete s; return rc; }
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)16 << 30;
    hbm_budget = (uint64_t)(free_b * 0.8);
  }
  s->budget = hbm_budget;
  s->max_active = max_active > 0 ? max_active : 1024;
  if (hipStreamCreateWithFlags(&s->up_stream, hipStreamNonBlocking) != hipSuccess) {
    engine_free(s->eng);
    delete s;
    return hpc_fail(HPC_EHIP, "stream creation failed");
  }
  s->uploader = std::thread(uploader_loop, s);
  s->worker = std::thread(server_loop, s);
  *out = s;
  return HPC_OK;
}

extern "C" int hpc_server_destroy(hpc_server *s) {
  if (!s) return HPC_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->queue.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_up = true;
  }
  s->cv_up.notify_all();
  if (s->uploader.joinable()) s->uploader.join();
  if (s->up_stream) (void)hipStreamDestroy(s->up_stream);
  engine_free(s->eng);
  delete s;
  return HPC_OK;
}

static int server_submit(hpc_server *s, uint64_t p, int n, double alpha, int max_level,
                         const double *x, bool borrow, int64_t *req) {
  if (!s || !x || !req) return hpc_fail(HPC_EINVAL, "NULL argument");
  if (p < 2 || p > (uint64_t)kMaxP || n < 1) return hpc_fail(HPC_EINVAL, "bad request shape");
  auto r = std::make_shared<Request>();
  r->p = (int)p;
  r->n = n;
  r->alpha = alpha;
  r->max_level = max_level;
  r->count = (size_t)p * n;
  r->corr_mode = s->corr_mode;
  if (borrow) {
    r->src = x;
  } else {
    r->x.assign(x, x + r->count);
    r->src = r->x.data();
  }
  r->bytes = engine_footprint((int)p, n, max_level);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->bytes > s->budget)
      return hpc_fail(HPC_ENOMEM, "request needs %zu bytes, budget is %zu", r->bytes, s->budget);
    r->id = s->next_id++;
    auto inj = s->inject.find(r->id);
    if (inj != s->inject.end()) {
      r->inject_level = inj->second;
      s->inject.erase(inj);
    }
    s->all[r->id] = r;
    s->queue.push_back(r);
    ++s->submits;
    *req = r->id;
  }
  s->cv_work.notify_all();
  return HPC_OK;
}

extern "C" int hpc_server_submit(hpc_server *s, uint64_t p, int n, double alpha, int max_level,
                                 const double *x, int64_t *req) {
  return server_submit(s, p, n, alpha, max_level, x, false, req);
}

extern "C" int hpc_server_submit_borrowed(hpc_server *s, uint64_t p, int n, double alpha,
                                          int max_level, const double *x, int64_t *req) {
  return server_submit(s, p, n, alpha, max_level, x, true, req);
}

extern "C" int hpc_server_wait(hpc_server *s, int64_t req, double timeout_s, int *done) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(req);
  if (it == s->all.end()) return hpc_fail(HPC_EINVAL, "no request %lld", (long long)req);
  auto r = it->second;
  auto pred = [&] { return r->state >= 2; };
  if (timeout_s < 0) s->cv_done.wait(lk, pred);
  else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (done) *done = r->state >= 2 ? 1 : 0;
  return HPC_OK;
}

extern "C" int hpc_server_result(hpc_server *s, int64_t req, int32_t *adj, int32_t *sep, int mc,
                                 hpc_request_info *info) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  if (sep && mc < 1) return hpc_fail(HPC_EINVAL, "sepset width must be >= 1");
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(req);
    if (it == s->all.end()) return hpc_fail(HPC_EINVAL, "no request %lld", (long long)req);
    r = it->second;
    if (r->state < 2) return hpc_fail(HPC_ESTATE, "request %lld not finished", (long long)req);
    if (r->state == 3)
      return hpc_fail(r->error ? r->error : HPC_EHIP, "request %lld failed: %s", (long long)req,
                      r->message.c_str());
  }
  // a finished request is no longer written by the worker: the dense
  // layouts are built without the lock
  if (adj) dense_adj(r->bits, r->p, adj);
  if (sep) dense_sepsets(r->bits, r->rec, r->p, sep, mc);
  if (info) {
    info->p = r->p;
    info->levels = r->levels;
    info->max_cond_size = r->mc;
    info->pad_ = 0;
    info->tests = r->tests;
    info->queue_ms = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
    info->latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  }
  return HPC_OK;
}

extern "C" int hpc_server_sepset_list(hpc_server *s, int64_t req, int32_t *out, int64_t cap,
                                      int64_t *words) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(req);
  if (it == s->all.end()) return hpc_fail(HPC_EINVAL, "no request %lld", (long long)req);
  auto &r = it->second;
  if (r->state != 2) return hpc_fail(HPC_ESTATE, "request %lld has no result", (long long)req);
  if (words) *words = (int64_t)r->rec.size();
  if (out) {
    if ((int64_t)r->rec.size() > cap) return hpc_fail(HPC_EINVAL, "cap too small");
    std::copy(r->rec.begin(), r->rec.end(), out);
  }
  return HPC_OK;
}

extern "C" int hpc_server_set_corr_mode(hpc_server *s, int mode) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  if (mode != HPC_CORR_CANON && mode != HPC_CORR_F32)
    return hpc_fail(HPC_EINVAL, "unknown correlation mode %d", mode);
  std::lock_guard<std::mutex> lk(s->mu);
  s->corr_mode = mode;
  return HPC_OK;
}

extern "C" int hpc_server_apply_tuning(hpc_server *s, const hpc_tune_entry *entries, int count) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  // checked here; the worker applies them before its next step
  const int rc = engine_check_tuning(entries, count);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  s->tune.insert(s->tune.end(), entries, entries + count);
  s->tune_pending = true;
  return HPC_OK;
}

extern "C" int hpc_server_set_tuning(hpc_server *s, int param, int value) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  const int rc = check_engine_tuning(param, value);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  s->knobs.push_back({param, value});
  s->tune_pending = true;
  return HPC_OK;
}

extern "C" int hpc_server_inject_failure(hpc_server *s, int64_t req, int level) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (req < s->next_id) {
    auto it = s->all.find(req);
    if (it == s->all.end()) return hpc_fail(HPC_EINVAL, "no request %lld", (long long)req);
    if (it->second->state != 0)
      return hpc_fail(HPC_ESTATE, "request %lld is already admitted", (long long)req);
    it->second->inject_level = level < 0 ? -1 : level;
  } else {
    s->inject[req] = level < 0 ? -1 : level;  // taken up when it is submitted
  }
  return HPC_OK;
}

extern "C" int hpc_server_release(hpc_server *s, int64_t req) {
  if (!s) return hpc_fail(HPC_EINVAL, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(req);
  if (it == s->all.end()) return hpc_fail(HPC_EINVAL, "no request %lld", (long long)req);
  if (it->second->state < 2) return hpc_fail(HPC_ESTATE, "request still running");
  s->all.erase(it);
  return HPC_OK;
}

extern "C" int hpc_server_stats(hpc_server *s, hpc_server_info *info) {
  if (!s || !info) return hpc_fail(HPC_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  info->steps = s->steps;
  info->peak_active = s->peak_active;
  info->peak_bytes = s->peak_bytes;
  info->busy_ms = s->busy_ms;
  info->budget = s->budget;
  info->queued = (int)s->queue.size();
  info->running = (int)s->running.size();
  return HPC_OK;
}
