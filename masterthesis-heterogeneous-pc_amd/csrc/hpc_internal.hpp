// Internal declarations shared by the translation units of libheterogpc_hip.so.
// Not part of the C ABI (include/heterogpc_hip.h).
//
// Every kernel works on a set of skeleton instances at once.  Instance k owns
// dense per-instance blocks (observations, z-scores, correlation, bit
// adjacency, removal levels, sepset planes) addressed with local variable
// ids 0..p_k-1; its variables also get global ids start_k..start_k+p_k-1 so
// that the per-level neighbour lists (CSR), first passing ranks and search
// tasks of all instances share one set of global arrays.  A single skeleton
// is the one-instance case.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/heterogpc_hip.h"

// Error reporting shared by the ABI translation units.
int hpc_fail(int code, const char *fmt, ...);

namespace hpc {

// Bound on |fast correlation - exact correlation| (HPC_CORR_F32): 64-sample
// FP32 accumulation (64 + 8 roundings of 2^-24), the float rounding of both
// operands, and 1% for the float-rounded standard deviation.
constexpr double kFastEps = 74.0 * 5.9604644775390625e-8 * 1.01;

// Largest p of an instance: p (p - 1) ordered edges still fit int32
// (the reference's level-0 index caps p the same way, gpuRowTests.cu:12).
constexpr int kMaxP = 46340;
// Ordered edges of all instances packed into one step: the neighbour lists
// are indexed with int32 row pointers.
constexpr int64_t kMaxPackEdges = 2147483647;

// Default longest host wait for a publication of device counts, in seconds
// (HPC_TUNE_WAIT_LIMIT).
constexpr double kWaitLimitS = 900.0;

// Largest level the device kernels accept (conditioning-set size).
constexpr int kMaxLevel = 16;
// Binomial table columns: C(n, k) for k <= kMaxLevel.
constexpr int kBinomK = kMaxLevel + 1;

// Device-wide counters of one phase; read back by the host.
struct Counters {
  unsigned long long n_edges;         // ordered edges of the snapshot (CSR size)
  unsigned long long n_tasks;         // search tasks of a level >= 2
  unsigned long long pinv_fallbacks;  // tests that took the pinv form
  unsigned long long n_rchk;          // queued exact rechecks of the current search
  unsigned long long rechecks;        // exact rechecks run in this phase
  int max_deg;                        // largest row length of the snapshot
  int overflow;                       // a recheck queue overflowed: the phase is void
  unsigned int done;                  // compaction workgroups finished (the last scans)
  unsigned int done2;                 // removal workgroups finished (the last publishes)
};

// A test whose fast-correlation statistic lies within its error bound of the
// threshold: ordered edge and rank, re-run on exact correlations.
struct Recheck {
  unsigned long long edge;
  unsigned long long rank;
};

// Level-0 counts are first added into kL0Slots slots per instance, each in its
// own 128-byte line (same-line atomics from thousands of workgroups
// serialise), then folded into InstCounters by one small kernel.
constexpr int kL0Slots = 32;
constexpr int kL0SlotWords = 16;

// Per-instance counters of one level.
struct InstCounters {
  unsigned long long tests;      // canonical CI tests T_l
  unsigned long long tests_max;  // all conditioning sets U_l
  unsigned long long edges;      // ordered edges searched
  unsigned long long removed;    // undirected edges removed
  unsigned long long rec_words;  // sepset record words appended
  int max_deg;                   // largest row length at level start
  int has_nan;                   // observations hold NaN cells (correlation phase);
                                 // the correlation matrix holds NaN (level 0)
};

enum InstFlags {
  kActive = 1,   // takes part in this phase
  kCorNaN = 2,   // the correlation matrix may hold NaN: tests check their inputs
};

// Class-major z-scores: slots per lane class, npad / 64 rounded up to a
// whole group of 16.
__host__ __device__ __forceinline__ int zc_slots(int npad) { return ((npad >> 6) + 15) & ~15; }
// The class-major z-scores are stored in blocks of 16 rows (layout in
// corr.hip): zc_block doubles per block, rows allocated in whole pairs of
// blocks (the matrix-core kernel reads 32 rows at a time; the rows past p
// are zeros).
__host__ __device__ __forceinline__ int zc_rows(int p) { return (p + 31) & ~31; }
__host__ __device__ __forceinline__ size_t zc_block(int npad) { return 1024 * (size_t)zc_slots(npad); }
__host__ __device__ __forceinline__ size_t zc_size(int p, int npad) {
  return (size_t)(zc_rows(p) / 16) * zc_block(npad);
}

// Device-side descriptor of one instance.
struct Inst {
  const double *x;        // [p][n] observations
  double *z;              // [p][npad] z-scores (the fast form's rechecks; or null)
  double *zc;             // class-major z-scores (the exact form; or null)
  uint8_t *nanm;          // [p][npad] NaN mask
  double *mean, *sd;      // [p]
  double *cor;            // [p][p]
  float *z32;             // [p][npad] z-scores rounded to float (fast correlation)
  unsigned long long *adj;  // [p][W] bit adjacency
  // sepset records of edges removed at levels >= 1, appended by the removal
  // kernel: i, j (i < j, local ids), level, then the level members
  int32_t *rec;
  unsigned long long rec_base;  // words written by earlier levels
  unsigned long long rec_cap;   // words the record buffer holds (writes beyond are dropped)
  double r_crit;          // Fisher-z decision boundary of (n, alpha)
  double r2_lo, r2_hi;    // square_bounds(r_crit): the squared decision filter
  double eps;             // bound on |cor - exact cor| when `fast`
  int start;              // first global variable id
  int p, n, npad, W;
  int level;              // level of this phase
  int flags;
  int fast;               // cor holds fast (MFMA) values: decisions need rechecks
  int grp0;               // first 8-row group of this instance in the pack (z-score kernel)
};

// What every kernel sees: the instance table and the global arrays.
struct Layout {
  const Inst *inst;             // [K]
  InstCounters *ictr;           // [K]
  const int32_t *row_inst;      // [p_total] instance of each global variable
  int K;
  int p_total;
  int max_deg;                  // largest row length over all instances
  int pmax;                     // largest instance p
  int32_t *deg;                 // [p_total]
  int32_t *rowptr;              // [p_total + 1]
  int32_t *nbr;                 // [E] global neighbour ids
  int32_t *erow;                // [E] row of each ordered edge
  int64_t *fp;                  // [E] first passing rank, -1 = none
  // [E] cor[i][nbr] of each ordered edge (i, nbr) of a row at level 1 (the
  // fill kernels write it for those rows; the level-1 search reads it)
  double *ccor;
  // [E] CSR index of the reverse edge (nbr, i), or -1: the fill kernels clear
  // it, the level-1 pivot search sets it for the edges it tests, the removal
  // reads it in place of a search of row nbr
  int32_t *erev;
  const unsigned long long *binom;  // [(pmax + 1) x kBinomK], saturating
  Counters *ctr;
  Recheck *rchk;                // recheck queue of a search
  unsigned long long rchk_cap;
  unsigned long long *l0slot;   // [K][kL0Slots][kL0SlotWords] level-0 partial counts, zero between uses
};

// ---- launchers (each defined in the .hip file that owns the kernel) ------
void launch_moments(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst);
// The same for active instances of at most max_n samples.
void launch_moments_n(hipStream_t s, const Layout &L, int max_n);
// pairwise sums: tiled kernel for NaN-free instances, masked kernel for the
// others; prefix arrays hold per-instance block counts (device and host)
void launch_corr(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                 long long *d_prefix, std::vector<long long> &h_prefix);
// The pairwise-sum kernels of launch_corr on block prefixes already in device
// memory: d_prefix holds the K + 1 exact-kernel prefixes (matrix-core blocks
// for instances with class-major z-scores, tiled slots for the others), then
// the K + 1 masked-block prefixes; tiled and masked are their totals, xmfma
// and tile whether any instance takes the matrix-core or the tiled kernel.
void launch_corr_sums(hipStream_t s, const Layout &L, const long long *d_prefix, long long tiled,
                      long long masked, bool xmfma = false, bool tile = true);
// Work units of one instance in the tiled and the masked pairwise kernels.
long long tiled_slots(int p);
long long masked_blocks(int p);
// Fast correlation of the packed instances with `fast` set: MFMA FP32 on the
// float z-scores, FP64 across 64-sample chunks (instances with NaN cells are
// left to launch_corr's masked kernel).
void launch_corr_fast(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                      long long *d_prefix, std::vector<long long> &h_prefix);
// Exact re-runs of the queued uncertain tests of `level`; then the queue is
// emptied, unless the caller's next kernel clears the counters anyway.
void launch_recheck(hipStream_t s, int level, const Layout &L, bool clear_queue = true);
void launch_adj_init(hipStream_t s, unsigned long long *adj, int p, int W);
// Instance tables of up to kTableArg entries travel as a kernel argument (one
// launch instead of a host-to-device copy); larger ones are copied.
constexpr int kTableArg = 8;
struct TableArg {
  Inst inst[kTableArg];
  unsigned long long *zero;  // words cleared by the same launch (or null)
  int zero_words;
  int K;
};
void launch_table(hipStream_t s, Inst *dst, const TableArg &arg);
// Copies `words` 8-byte words from device memory to fine-grained host memory
// (device-visible pointers), then stores `seq` to *flag: the host sees the
// words once it sees the flag.
// One copy of a gather of device blocks into one (host-mapped) buffer.
struct CopyOut {
  const void *src;
  unsigned long long bytes;  // a multiple of 4
  unsigned long long dst;    // byte offset, a multiple of 8
};
// Copies every entry's bytes to dst + entry.dst, one workgroup per entry.
void launch_copy_out(hipStream_t s, const CopyOut *entries, int count, char *dst);
void launch_publish(hipStream_t s, const unsigned long long *src, size_t words,
                    unsigned long long *dst, unsigned long long *flag, unsigned long long seq);
void launch_compact(hipStream_t s, const Layout &L);
// Counter publication by the last workgroup of a kernel: copies `words`
// 8-byte words from the counter block to fine-grained host memory `dst`,
// then stores `seq` to *flag (device-visible pointers).
struct Publish {
  unsigned long long *dst;
  unsigned long long *flag;
  unsigned long long seq;
  int words;
  // removal kernel of a one-instance run: after publishing, clear the counter
  // block and advance the instance's level and record base on the device
  int advance;
  // 1: dst is host memory, written behind a system-scope fence, then the
  // flag is set; 0: dst is device memory and no flag is set (the levels of a
  // captured loop, whose last publication copies every level's words out)
  int sys;
};
// Level 0 of the packed instances at level 0.  With `pub` (one instance),
// the kernel that folds the counts also publishes them, clears the counter
// block and sets the device table to level 1.
void launch_level0(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                   long long *d_prefix, std::vector<long long> &h_prefix,
                   const Publish *pub = nullptr);
// Compaction of every active instance (all at `level` >= 1) in two kernels
// that also plan the level's tasks (level >= 2, `rpt` ranks per task) and
// publish the counters (the first kernel's last workgroup, once the totals
// are known): row lengths and per-workgroup sums, then row pointers, task
// offsets and the neighbour lists with their first passing ranks reset.
// blk: 3 words per workgroup (compact_fused_blocks).
void launch_compact_fused(hipStream_t s, const Layout &L, int level, int rpt,
                          unsigned long long *blk, long long *row_tasks, long long *task_off,
                          size_t edge_cap, const Publish &pub);
// Rows per compaction workgroup.
constexpr int kCompRows = 16;
int compact_fused_blocks(int p_total);
// The compaction's second kernel alone, for counts that the removal kernel
// produced (NextCount).
void launch_compact_fill(hipStream_t s, const Layout &L, const unsigned long long *blk,
                         const long long *row_tasks, long long *task_off, size_t edge_cap);
// Removal kernel of a one-instance run that also counts the next level: each
// removal lowers both row lengths (L.deg), and the last workgroup turns the
// row lengths into what the compaction's first kernel computes (task plan of
// `level` with `rpt`, per-workgroup sums and their prefixes in blk, totals
// and longest row in the counters) before it publishes, so its publication
// carries the next level's sizes and launch_compact_fill completes the
// compaction.
struct NextCount {
  unsigned long long *blk;
  long long *row_tasks;
  int level;
  int rpt;
  int on;
};
// `fast`: the launch holds an instance with fast correlations (selects the
// variant with uncertainty bounds and the recheck queue)
// epb: ordered edges per workgroup of the LDS-staged form
// lds_limit: LDS bytes per workgroup for staging a row (search_lds_limit);
// a row whose staging needs more is searched by a gather kernel, launched
// after the staged one only when such a row exists
void launch_search_l1(hipStream_t s, const Layout &L, int64_t n_edges,
                      const int32_t *row_mask, bool fast, int epb, size_t lds_limit);
// Default (and largest) LDS bytes per workgroup for staging a row.
constexpr size_t kSearchLdsLimit = 64 * 1024;
// The staging limit in effect for a requested one (0 = the default).
size_t search_lds_limit(size_t requested);
void launch_task_plan(hipStream_t s, int level, const Layout &L, int rpt,
                      const int32_t *row_mask, long long *row_tasks, long long *task_off);
// The plan's per-row task counts and exclusive offsets define the tasks; the
// kernels read the task count from the device counters.
// tpb: tasks per workgroup of the LDS-staged form
// A task queue of the level >= 1 search kernels: `head` (device memory, 0
// between launches; the kernel leaves it so) is claimed a block of tasks at a
// time by the workgroups; with `done` (one int per block in host-visible
// memory) host threads take blocks from the other end and each side skips
// blocks the other has finished.  No head: blocks go to workgroups by id.
struct TaskQueue {
  unsigned long long *head;
  int *done;
  // the by-set search: each workgroup of a grid of resident size takes one
  // contiguous share of the level's tasks (no blocks, no claims)
  int balanced = 0;
};
// level 1 here only through a shared queue (exact instances)
void launch_search_ln(hipStream_t s, int level, const Layout &L, int rpt,
                      const long long *row_tasks, const long long *task_off, bool fast, int tpb,
                      size_t lds_limit, const TaskQueue &queue = TaskQueue{nullptr, nullptr});
// With pub.flag set, the last workgroup to finish publishes the counters.
void launch_commit(hipStream_t s, const Layout &L, int64_t n_edges, const Publish &pub,
                   const NextCount &next = NextCount{nullptr, nullptr, 0, 0, 0});
void launch_scatter_fp(hipStream_t s, const int64_t *idx, const int64_t *val,
                       long long count, int64_t *fp);
// fp[idx[t]] = min(fp[idx[t]], val[t]) (ranks >= 0; -1 counts as none).
void launch_scatter_min_fp(hipStream_t s, const int64_t *idx, const int64_t *val,
                           long long count, int64_t *fp);

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Host bookkeeping of one instance.
struct InstHost {
  bool used = false;
  int p = 0, n = 0, npad = 0, W = 0, mc = 1, max_level = 0;
  double alpha = 0.0, r_crit = 0.0;
  // device blocks (from the engine's arena)
  double *x = nullptr, *z = nullptr, *mean = nullptr, *sd = nullptr, *cor = nullptr;
  double *zc = nullptr;  // class-major z-scores of the exact form
  uint8_t *nanm = nullptr;
  unsigned long long *adj = nullptr;
  int32_t *rec = nullptr;
  size_t rec_cap = 0, rec_used = 0;  // sepset record words
  float *z32 = nullptr;
  int corr_mode = 0;       // HPC_CORR_* of the next correlation
  bool fast = false;       // cor holds fast values (mode F32 and no NaN cells)
  bool have_x = false, have_cor = false;
  // NaN in the correlation matrix: -1 unknown (tests check their inputs),
  // 0 none, 1 some; found by level 0, which reads every pair, or on upload
  int cor_nan = -1;
  bool drop_obs = false;   // free observations and z-scores once correlated
  // level state
  int start = 0;           // global id of local variable 0 in the last pack
  int next_level = 0;      // level the next step runs
  bool finished = false;   // no level left to run
  bool compacted = false;  // snapshot matches the adjacency
  bool ever_compacted = false;
  int max_deg = 0;         // of the last compaction
  int64_t undirected = 0;
  size_t bytes = 0;        // device footprint
  int error = 0;           // first error of a batched run (HPC_E*), 0 = none
  std::string message;     // its description
  int inject_level = -1;   // test hook: the host side of this level fails
  bool fresh = false;      // adjacency complete, no level run since (replayable)
  std::vector<hpc_level_stats> hist;  // one record per level run
  // host copies of the adjacency bits and the sorted sepset records, read by
  // the exports; dropped whenever device work may change the instance
  bool res_bits_valid = false, res_rec_valid = false;
  std::vector<unsigned long long> res_bits;
  std::vector<int32_t> res_rec;
};

// Device memory of instance blocks: slabs from hipMalloc, sub-allocated
// first fit with coalescing frees, returned to HIP only when the engine is
// freed.  Frees are stream-ordered safe because all device work that reads a
// block is on the engine's stream, ahead of any later reuse.
struct Arena {
  struct Slab {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length
  };
  std::vector<Slab> slabs;
  std::map<const void *, std::pair<int, size_t>> live;  // block -> (slab, length)
  size_t slab_min = 0;  // smallest slab (0: exactly the request)
  size_t reserved = 0, used = 0;
};

// Byte offset of the instance counters in the counter block.
constexpr size_t kIctrOffset = (sizeof(Counters) + 63) & ~(size_t)63;

// Sizes one level of a one-instance run read from the device: ordered
// edges and longest row of the layout and of the instance, and where they
// were published (0: the level's compaction, 1: the previous removal).
struct LevelShape {
  int64_t n_edges = 0;
  int max_deg = 0;
  int inst_max_deg = 0;
  int src = 0;
};
// The sizes of a whole fresh run, levels 0..max_level, and what else a
// capture of its launches bakes in.
struct RunShape {
  bool valid = false;
  int max_level = 0;
  // the last level searched, and with no level cap the level whose
  // compaction found nothing left to test (-1 with a cap)
  int last = 0, stop = -1;
  int cor_nan = -1;
  const int32_t *rec = nullptr;  // record buffer (its address is a kernel argument)
  LevelShape lv[kMaxLevel + 2];
};
// A captured one-instance level loop (engine_run_single replays it).
struct GraphRun {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int id = -1;
  RunShape shape;
  unsigned long long first_seq = 0, final_seq = 0;  // publication numbers it writes
  std::vector<long long> tuning;                     // launch shapes it bakes in
};
// Bytes per publication slot of a captured run (one counter block, K = 1).
constexpr size_t kSlotBytes = 256;
static_assert(kIctrOffset + sizeof(InstCounters) <= kSlotBytes, "slot too small");

// A set of instances on one device with their shared per-level arrays and
// one stream.
struct Engine {
  int device = 0;
  Arena arena;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;
  std::vector<InstHost> inst;  // slot per instance id
  // global arrays
  Inst *d_inst = nullptr;
  InstCounters *d_ictr = nullptr, *h_ictr = nullptr;
  int32_t *d_row_inst = nullptr;
  int32_t *d_deg = nullptr, *d_rowptr = nullptr, *d_nbr = nullptr, *d_erow = nullptr;
  int32_t *d_rowmask = nullptr;
  int64_t *d_fp = nullptr;
  double *d_ccor = nullptr;
  int32_t *d_erev = nullptr;
  long long *d_row_tasks = nullptr, *d_task_off = nullptr;
  // per-workgroup sums of the two-kernel compaction
  unsigned long long *d_blk = nullptr;
  unsigned long long *d_l0slot = nullptr;  // [cap_inst][kL0Slots][kL0SlotWords]
  unsigned long long *d_binom = nullptr;
  long long *d_prefix = nullptr;
  // device counters and the packed instances' counters in one block (and
  // its pinned host mirror), so one copy reads and one fill clears them all:
  // d_ctr at the start, d_ictr at kIctrOffset
  Counters *d_ctr = nullptr, *h_ctr = nullptr;
  Counters *h_ctr_dev = nullptr;  // h_ctr as the device addresses it
  // a second host mirror: a one-instance run's removal publishes there while
  // the next level's compaction publishes into h_ctr
  Counters *h_ctr2 = nullptr, *h_ctr2_dev = nullptr;
  // per-level events of a one-instance run: compaction, search, removal
  hipEvent_t lev_ev[kMaxLevel + 1][6] = {};
  bool lev_ev_ready = false;
  // publication flag (fine-grained host memory): the last sequence number a
  // counter publication wrote, its device address, and the next number
  unsigned long long *h_seq = nullptr, *d_seq = nullptr;
  unsigned long long seq = 0;
  // an asynchronous copy from a host vector (table, row map, prefixes) may
  // still read it: the next pack waits for the stream first
  bool host_upload_pending = false;
  size_t cap_inst = 0, cap_rows = 0, cap_edges = 0, cap_binom = 0;
  std::vector<unsigned long long> h_binom;
  std::vector<long long> h_prefix, h_prefix_fast;
  // the packed set of the current phase
  std::vector<Inst> h_inst;
  std::vector<int> packed;
  std::vector<int32_t> h_row_inst;
  std::vector<int32_t> h_row_inst_dev;  // what d_row_inst holds (when row_inst_valid)
  bool row_inst_valid = false;
  int p_total = 0;
  int pmax = 0;  // largest p of the packed instances
  int64_t n_edges = 0;
  int max_deg = 0;
  int ranks_per_task = 512;
  // launch shapes per level (set by the autotuner; 0 = defaults: ranks per
  // task `ranks_per_task`, 16 tasks and 64 ordered edges per workgroup)
  int rpt_level[kMaxLevel + 1] = {};
  int tpb_level[kMaxLevel + 1] = {};
  int queue_level[kMaxLevel + 1] = {};  // 0 = use_queue, 1 = device queue, 2 = workgroup order
  // batch tuning pass (engine_batch_autotune): levels tuned so far and their entries
  bool tune_pass = false;
  bool tuned[kMaxLevel + 1] = {};
  std::vector<hpc_tune_entry> tune_out;
  int epb_level1 = 0;
  hipEvent_t ev[8] = {};
  hipEvent_t sev[2 * (kMaxLevel + 1)] = {};  // search kernel brackets of one step
  hipEvent_t fp_ready = nullptr;  // recorded after the first-pass array is reset
  // host share scratch (single-instance use)
  std::vector<double> h_cor;
  bool h_cor_valid = false;
  std::vector<int32_t> h_rowptr, h_nbr;
  int64_t *d_scatter_idx = nullptr, *d_scatter_val = nullptr;
  size_t scatter_cap = 0;
  // task queue: the device head (0 between launches) and the per-block done
  // flags a queue shared with host threads uses (host-visible memory)
  unsigned long long *d_qhead = nullptr;
  int *h_qdone = nullptr, *d_qdone = nullptr;
  size_t cap_qdone = 0;
  std::vector<long long> h_task_off, h_row_tasks;
  bool use_queue = true;  // level >= 2 searches claim task blocks (HPC_TUNE_QUEUE)
  // one-instance level loop captured in a hipGraph (HPC_TUNE_GRAPH), and the
  // per-level publication slots its replays write (host-visible memory)
  bool use_graph = true;
  GraphRun graph;
  Counters *h_slots = nullptr, *d_slots = nullptr;
  unsigned long long *d_slots_local = nullptr;  // the captured levels' slots in device memory
  std::vector<char> h_stage;  // host staging of engine_prefetch_results
  char *d_stage = nullptr;    // and its device-side gather buffer
  size_t d_stage_cap = 0;
  CopyOut *d_copies = nullptr;  // the copy table of engine_prefetch_results
  size_t cap_copies = 0;
  int64_t graph_mismatches = 0;  // replays whose sizes left the recorded run's
  int64_t graph_replays = 0, graph_captures = 0;
  // recheck queues of fast instances
  Recheck *d_rchk = nullptr;
  size_t cap_rchk = 0;
  size_t rchk_limit = 0;  // queue length in use (0 = capacity); a tuning knob
  size_t lds_limit = 0;   // LDS bytes for staging a search row (0 = default); a tuning knob
  bool kernel_timing = true;  // level-loop device-time events (HPC_TUNE_KERNEL_TIMING)
  double wait_limit_s = kWaitLimitS;  // longest wait for a publication (HPC_TUNE_WAIT_LIMIT)
  unsigned long long rechecks = 0;  // exact rechecks of the last step or level
};

double fisher_r_crit(int n, double alpha);
// Saturating C(m, k) for m < rows, k <= kMaxLevel (row-major, kBinomK
// columns): 2^64 - 1 stands for every value that does not fit.
std::vector<unsigned long long> binom_table(size_t rows);

// Spins until *word == value (acquire loads).  Every 256 spins it calls
// poll(): a negative return ends the wait with that error, 1 says that the
// word will not change any more (one last look decides), 0 goes on.  After
// limit_s seconds (<= 0: no limit) the wait ends with HPC_EHIP.
template <class Poll>
int spin_wait(const unsigned long long *word, unsigned long long value, double limit_s, Poll poll) {
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1;; ++it) {
    if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == value) return HPC_OK;
    if ((it & 255u) == 0) {
      const int q = poll();
      if (q < 0) return q;
      if (q > 0) {
        if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == value) return HPC_OK;
        return hpc_fail(HPC_EHIP, "publication %llu did not arrive (the stream is idle)", value);
      }
      if (limit_s > 0.0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > limit_s)
        return hpc_fail(HPC_EHIP,
                        "publication %llu did not arrive within %.3f s (HPC_TUNE_WAIT_LIMIT)",
                        value, limit_s);
    }
    __builtin_ia32_pause();
  }
}
// spin_wait with nothing to poll (host words only).
int spin_wait_word(const unsigned long long *word, unsigned long long value, double limit_s);
int engine_init(Engine &e, int device);
void engine_free(Engine &e);
// New instance with device blocks and the initial adjacency; returns its id.
int engine_add(Engine &e, int p, int n, double alpha, int max_level, int *id);
void engine_remove(Engine &e, int id);
// Uploads [p][n] observations (NaN cells are detected on the device).
int engine_set_observations(Engine &e, int id, const double *h_x);
// A host-to-device copy on stream s: large copies through the device's ring
// of pinned staging buffers, filled by several host threads (copies on one
// device take turns), small ones directly.  With complete, the copy has ended
// when it returns; without, the source has been read (work later on s sees
// the data) and small copies are staged too.
hipError_t upload_h2d(int device, void *dst, const void *src, size_t bytes, hipStream_t s,
                      bool complete);
// The same for an array of doubles, sent at half width when FP32 holds every
// value exactly (the device receives the same bits either way).
hipError_t upload_f64_h2d(int device, double *dst, const double *src, size_t count,
                          hipStream_t s, bool complete);
// The two halves of an upload done elsewhere (another stream or thread):
// device blocks for the observations (*x receives the [p][n] block), then,
// once the copy has completed, the instance is marked ready.
int engine_alloc_observations(Engine &e, int id, double **x);
void engine_observations_ready(Engine &e, int id);
int engine_set_corr(Engine &e, int id, const double *h_cor);
int engine_reset(Engine &e, int id);
// Instance table of `ids` with the given per-instance levels; every phase
// below works on the packed set.
// zero_ctr: the same launch clears the counter block (small tables)
int engine_pack(Engine &e, const std::vector<int> &ids, const std::vector<int> &levels,
                bool zero_ctr = false);
int engine_corr(Engine &e, float *ms_moments, float *ms_corr);
// Before packing an instance for its correlation: the fast flag from its
// correlation mode, and its float z-score block.
// A device span of doubles to clear (pad rows gathered over instances).
struct ZeroSpan {
  double *ptr;
  size_t words;
};
// Clears the spans with one launch per 32 of them, on e.stream.
int engine_zero_spans(Engine &e, const std::vector<ZeroSpan> &spans);
// With clears, the instance's pad-row clear is appended there (the caller
// clears them all at once) instead of run here.
int engine_prepare_corr(Engine &e, int id, std::vector<ZeroSpan> *clears = nullptr);
// Exact correlations (from the device-resident z-scores) for the fast
// instances among ids, which then count as exact.
int engine_exact_from_z(Engine &e, const std::vector<int> &ids);
// Compaction of the packed instances; updates their max_deg and n_edges.
int engine_compact(Engine &e, float *ms);
// Search of the packed instances at `level` (rows: global ids, NULL = all).
int engine_search(Engine &e, int level, const int32_t *rows, int nrows, float *ms);
// The level's search on one task queue shared by the device and host threads
// (one exact instance packed at `level`).  _launch plans the tasks, copies the
// plan, snapshot and correlation to the host, clears the done flags and
// queues the persistent kernels; host threads then take blocks from the back
// (SharedPlan), and _finish lowers the device's first passing ranks with the
// host's results and waits for the stream.
struct SharedPlan {
  long long n_tasks = 0, n_blocks = 0;
  int rpt = 0, tpb = 0;
  int *done = nullptr;             // host view of the done flags
  hipEvent_t device_done = nullptr;  // recorded after the device's kernels
};
int engine_search_shared_launch(Engine &e, int level, SharedPlan &plan);
int engine_search_shared_finish(Engine &e, const std::vector<int64_t> &idx,
                                const std::vector<int64_t> &val, float *ms);
// Level 0 of the packed instances at level 0 (tests and removes).
int engine_level0(Engine &e, float *ms);
int engine_commit(Engine &e, float *ms);
int engine_read_inst_counters(Engine &e);
// Sepset width of an instance's exports covers `level`.
void engine_note_level(Engine &e, int id, int level);
// One level for every id (correlation first for those without one): level 0,
// compaction, the searches of every level present and the commit, with two
// host waits for published counters for the whole set.  Finished instances
// are marked; their per-level records go to InstHost::hist.
int engine_step(Engine &e, const std::vector<int> &ids);
// Whether a level can be planned for an instance whose longest row has
// max_deg neighbours: C(max_deg - 1, level) did not saturate and the row's
// task count stays within kMaxRowTasks (row_task_count; monotone in the row
// length).  Kernels give rows that cannot be planned no tasks; the host fails
// the level with HPC_ELEVEL.
bool engine_plannable(const Engine &e, int max_deg, int level);
// The device task queue of a level >= 2 search: the level's tuned choice,
// else HPC_TUNE_QUEUE (none when off).
TaskQueue device_queue(const Engine &e, int level);
// Launch shapes of a level's CI-test kernel.
int engine_rpt(const Engine &e, int level);
int engine_tpb(const Engine &e, int level);
int engine_epb(const Engine &e);
// Autotuner: runs instance id's levels 1..max_level, timing each level's
// CI-test kernel on the level's snapshot for every candidate launch shape
// (results do not depend on the shape), keeps the fastest per level in the
// engine and leaves the instance reset.
int engine_autotune(Engine &e, int id, int max_level, std::vector<hpc_tune_entry> &out);
// Autotuner of a batch: its unfinished (exact) instances run from their start
// once, and each level's CI-test kernel is timed with every candidate shape on
// the first step that meets the level (the step's own mix of instances); the
// fastest stays set per level and goes to `out`; every instance is then back
// at its start (correlations kept).
int engine_batch_autotune(Engine &e, std::vector<hpc_tune_entry> &out);
// Drops a captured level loop.
void graph_drop(Engine &e);
// hpc_run's level loop for one instance on exact correlations (levels from
// h.next_level to its cap), with one host wait per level.
int engine_run_single(Engine &e, int id);
// Per-level launch shapes (entries as hpc_autotune reports them): the check
// alone, and check-then-apply (nothing is applied unless every entry is valid).
int engine_check_tuning(const hpc_tune_entry *t, int count);
int engine_apply_tuning(Engine &e, const hpc_tune_entry *t, int count);
// Device bytes one instance needs (its blocks plus its share of the per-level
// arrays).
size_t engine_footprint(int p, int n, int max_level);
int export_cor(Engine &e, int id, double *cor);
int export_adj(Engine &e, int id, int32_t *adj);
int export_sepsets(Engine &e, int id, int32_t *sep, int mc);
int export_sepset_list(Engine &e, int id, int32_t *out, int64_t cap, int64_t *words);
// Host copies of an instance's bit adjacency ([p][W] words) and of its sepset
// records sorted by (i, j).
int fetch_adj_bits(Engine &e, int id, std::vector<unsigned long long> &bits);
int fetch_records(Engine &e, int id, std::vector<int32_t> &rec);
// Host copies of the results of every finished instance of `ids` that has
// none yet: all device-to-host copies in flight together, one wait.
int engine_prefetch_results(Engine &e, const std::vector<int> &ids);
void engine_drop_results(InstHost &h);
// Reference layouts from those host copies (sepSets: -1 empty, -2 removed
// at level 0, cpuUtil.cpp:3-48).
void dense_adj(const std::vector<unsigned long long> &bits, int p, int32_t *adj);
void dense_sepsets(const std::vector<unsigned long long> &bits, const std::vector<int32_t> &rec,
                   int p, int32_t *sep, int mc);

}  // namespace hpc

// One skeleton instance behind the single-instance ABI.
struct hpc_state {
  hpc::Engine eng;
  int id = 0;
  bool corr_timed = false;
  float ms_moments = 0.f, ms_corr = 0.f;
};

// Error reporting shared by the ABI translation units.
int hpc_fail(int code, const char *fmt, ...);

#define HPC_HIPCHK(call)                                                    \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return hpc_fail(HPC_EHIP, "%s failed: %s (%s:%d)", #call,             \
                      hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)
