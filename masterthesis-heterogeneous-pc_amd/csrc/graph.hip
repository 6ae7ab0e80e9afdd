// Adjacency kernels: initial state, the level-0 test and per-level compaction,
// for every active instance of the layout at once.
//
// The adjacency of an instance lives on the device as a bit matrix: row i is
// W = ceil(p/64) 64-bit words, bit j of word j/64 set while edge i-j is
// present.  p = 2000 takes 512 KB instead of the reference's 16 MB int32
// matrix (state.cu:21-38).
#pragma clang fp contract(off)

#include "exact_dot.hpp"
#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

__device__ __forceinline__ int find_inst(const long long *prefix, int K, long long idx) {
  int lo = 0, hi = K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// state.cu:31-38: every off-diagonal pair adjacent.
__global__ void adj_init_kernel(unsigned long long *adj, int p, int W) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)p * W) return;
  const int i = (int)(idx / W), w = (int)(idx % W);
  unsigned long long word = ~0ull;
  const int lo = w * 64;
  if (p - lo < 64) word = (p - lo <= 0) ? 0ull : ((1ull << (p - lo)) - 1ull);
  if (i >= lo && i < lo + 64) word &= ~(1ull << (i - lo));
  adj[idx] = word;
}

// Level 0 (testRowL0, gpuRowTests.cu:4-27; cpuRowTests.cpp:5-21): every
// adjacent pair j < i is tested on cor[i][j] (the lower triangle, as the
// reference reads it) and removed iff p >= alpha.
//
// One workgroup of kL0Waves wavefronts per 64 x 64 tile (I, J), I >= J, of an
// instance's lower block triangle; wavefront w takes rows 64I + r for r in
// [w * kL0Rows, (w + 1) * kL0Rows).  Iteration r reads row i = 64I + r,
// columns 64J..64J+63 (one coalesced 512 B load); the ballot of surviving
// pairs is row i's word J, and lane c collects bit r of row (64J + c)'s word
// I.  The column words of the wavefronts are OR-ed in LDS; each adjacency
// word has exactly one writer, so no atomics are needed.
//
// An instance with fast (MFMA) correlations first replaces every entry of the
// row whose fast value lies within eps of the threshold by the exact one
// (exact_cor, the whole wavefront per entry; both triangles are written), so
// level 0 decides on exact values.  Level 0 reads only the lower triangle and
// each of its entries belongs to one tile row, so the writes race with no
// read.
constexpr int kL0Waves = 8;
constexpr int kL0Rows = 64 / kL0Waves;

// prefix: per-instance tile offsets (K + 1), or null for one instance with
// `tiles` tiles
__global__ __launch_bounds__(64 * kL0Waves) void level0_kernel(Layout L,
                                                              const long long *__restrict__ prefix,
                                                              long long tiles) {
  __shared__ unsigned long long s_row[64];
  __shared__ unsigned long long s_col[kL0Waves][64];
  __shared__ unsigned long long s_cnt[kL0Waves][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long t = blockIdx.x;
  if (t >= (prefix ? prefix[L.K] : tiles)) return;
  const int k = prefix ? find_inst(prefix, L.K, t) : 0;
  const Inst In = L.inst[k];
  const int p = In.p, W = In.W;
  const long long tl = prefix ? t - prefix[k] : t;
  // lower block triangle, row-major: tile tl -> (I, J), J <= I
  int I = (int)((sqrt(8.0 * (double)tl + 1.0) - 1.0) * 0.5);
  while ((long long)I * (I + 1) / 2 > tl) --I;
  while ((long long)(I + 1) * (I + 2) / 2 <= tl) ++I;
  const int J = (int)(tl - (long long)I * (I + 1) / 2);
  const int j = J * 64 + lane;
  const bool diag = (I == J);
  unsigned long long colword = 0ull;   // bits r: pair (64I + r, j) kept
  unsigned long long rowword = 0ull;   // captured by lane r: row 64I + r, word J
  unsigned long long tested = 0ull, removed = 0ull;
  bool nan_seen = false;  // any pair of the tile, present or not
  const int r0 = wv * kL0Rows;
#pragma unroll
  for (int rr = 0; rr < kL0Rows; ++rr) {
    const int r = r0 + rr;
    const int i = I * 64 + r;
    bool present = false, keep = false;
    if (i < p) {
      const unsigned long long aw = In.adj[(size_t)i * W + J];
      const bool valid = (j < p) && (diag ? (lane < r) : true);
      present = valid && ((aw >> lane) & 1ull);
      double c = valid ? In.cor[(size_t)i * p + j] : 0.0;
      if (In.fast) {
        const double re = __builtin_fmin(kCutThr, __builtin_fabs(c));
        unsigned long long m =
            __ballot(valid && !__builtin_isnan(c) && __builtin_fabs(re - In.r_crit) <= In.eps);
        while (m) {
          const int bit = __ffsll((long long)m) - 1;
          m &= m - 1ull;
          const int jt = J * 64 + bit;
          const double ex = exact_cor(In, i, jt, lane);
          if (lane == bit) c = ex;
          if (lane == 0) {
            In.cor[(size_t)i * p + jt] = ex;
            In.cor[(size_t)jt * p + i] = ex;
          }
        }
      }
      if (valid) {
        nan_seen |= __builtin_isnan(c);
        if (present) keep = !removes(c, In.r_crit);
      }
    }
    const unsigned long long kb = __ballot(keep);
    tested += __popcll(__ballot(present));
    removed += __popcll(__ballot(present && !keep));
    if (lane == r) rowword = kb;
    colword |= (unsigned long long)(keep ? 1ull : 0ull) << r;
  }
  if (lane >= r0 && lane < r0 + kL0Rows) s_row[lane] = rowword;
  s_col[wv][lane] = colword;
  if (lane == 0) { s_cnt[wv][0] = tested; s_cnt[wv][1] = removed; }
  // the matrix is symmetric: the lower triangle decides whether tests at
  // later levels must check their inputs for NaN
  if (__any(nan_seen) && lane == 0) atomicOr(&L.ictr[k].has_nan, 1);
  __syncthreads();
  if (wv != 0) return;
  unsigned long long cw = 0ull;
#pragma unroll
  for (int w = 0; w < kL0Waves; ++w) cw |= s_col[w][lane];
  const unsigned long long rw = s_row[lane];
  const int i_lane = I * 64 + lane;
  if (diag) {
    if (i_lane < p) In.adj[(size_t)i_lane * W + J] = rw | cw;
  } else {
    if (i_lane < p) In.adj[(size_t)i_lane * W + J] = rw;
    if (j < p) In.adj[(size_t)j * W + I] = cw;
  }
  if (lane == 0) {
    unsigned long long te = 0, re = 0;
#pragma unroll
    for (int w = 0; w < kL0Waves; ++w) { te += s_cnt[w][0]; re += s_cnt[w][1]; }
    if (te) {
      unsigned long long *slot =
          L.l0slot + ((size_t)k * kL0Slots + (blockIdx.x & (kL0Slots - 1))) * kL0SlotWords;
      atomicAdd(&slot[0], te);
      atomicAdd(&slot[1], re);
    }
  }
}

// Folds the level-0 slots of instance blockIdx.x into its counters and
// clears them.  Tests, test bound and searched edges all equal the pairs
// tested at level 0.  With pub.flag set (one instance, one workgroup) lane 0
// then publishes the counter block, clears it and sets the device table to
// level 1, as the removal kernel does at the end of a level.
__global__ __launch_bounds__(64) void level0_fold_kernel(Layout L, Publish pub) {
  const int k = blockIdx.x, lane = threadIdx.x;
  unsigned long long te = 0, re = 0;
  if (lane < kL0Slots) {
    unsigned long long *slot = L.l0slot + ((size_t)k * kL0Slots + lane) * kL0SlotWords;
    te = slot[0];
    re = slot[1];
    slot[0] = 0ull;
    slot[1] = 0ull;
  }
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) {
    te += __shfl_xor(te, sh, 64);
    re += __shfl_xor(re, sh, 64);
  }
  if (lane == 0 && te) {
    InstCounters &c = L.ictr[k];
    c.tests += te;
    c.tests_max += te;
    c.edges += te;
    c.removed += re;
  }
  if (!pub.flag || lane != 0) return;
  unsigned long long *blk = reinterpret_cast<unsigned long long *>(L.ctr);
  for (int w = 0; w < pub.words; ++w) pub.dst[w] = blk[w];
  if (pub.sys) {
    __threadfence_system();
    __hip_atomic_store(pub.flag, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  } else {
    __threadfence();
  }
  if (pub.advance) {
    for (int w = 0; w < pub.words; ++w) blk[w] = 0ull;
    const_cast<Inst *>(L.inst)[0].level = 1;
  }
}

// Row lengths (compact.cu:80-88 stores them at adj_compact[i][p-1]); rows of
// inactive instances count as empty.  Per-instance maxima: the scan kernel
// takes the single-instance case; otherwise one atomic per instance and
// workgroup.
constexpr int kDegWaves = 16;
__global__ __launch_bounds__(64 * kDegWaves) void degree_kernel(Layout L) {
  __shared__ int s_c[kDegWaves], s_k[kDegWaves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = blockIdx.x * kDegWaves + wv;
  int c = 0, k = -1;
  if (g < L.p_total) {
    k = L.row_inst[g];
    const Inst In = L.inst[k];
    if ((In.flags & kActive) && In.level >= 1) {
      const unsigned long long *row = In.adj + (size_t)(g - In.start) * In.W;
      for (int w = lane; w < In.W; w += 64) c += __popcll(row[w]);
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s, 64);
    if (lane == 0) L.deg[g] = c;
  }
  if (L.K == 1) return;
  if (lane == 0) { s_c[wv] = c; s_k[wv] = c ? k : -1; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int cur = -1, m = 0;
    for (int w = 0; w <= kDegWaves; ++w) {
      const int kw = w < kDegWaves ? s_k[w] : -2;
      if (kw == -1) continue;
      if (kw != cur && cur >= 0) { atomicMax(&L.ictr[cur].max_deg, m); m = 0; }
      if (kw < 0) break;
      cur = kw;
      m = max(m, s_c[w]);
    }
  }
}

// Exclusive scan of row lengths -> CSR row pointers, total ordered edges and
// the maximum row length (compact.cu:81 atomicMax_system(max_adj)).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void scan_kernel(Layout L) {
  __shared__ long long part[kScanThreads];
  __shared__ int mx[kScanThreads];
  const int tid = threadIdx.x;
  const int p = L.p_total;
  const int per = (p + kScanThreads - 1) / kScanThreads;
  const int b = tid * per, e = min(p, b + per);
  long long s = 0;
  int m = 0;
  for (int i = b; i < e; ++i) { s += L.deg[i]; m = max(m, L.deg[i]); }
  part[tid] = s;
  mx[tid] = m;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    long long v = tid >= off ? part[tid - off] : 0;
    int mv = tid >= off ? mx[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    mx[tid] = max(mx[tid], mv);
    __syncthreads();
  }
  long long run = part[tid] - s;
  for (int i = b; i < e; ++i) { L.rowptr[i] = (int32_t)run; run += L.deg[i]; }
  if (tid == kScanThreads - 1) {
    L.rowptr[p] = (int32_t)part[tid];
    L.ctr->n_edges = (unsigned long long)part[tid];
    L.ctr->max_deg = mx[tid];
    if (L.K == 1) L.ictr[0].max_deg = mx[tid];
  }
}

// The entries of one lane's adjacency word w of row g, written from CSR
// offset `off` on: neighbour ids, the row, first passing rank -1 and, for a
// row at level 1, its correlations cor[g][nbr] (four at a time, their loads
// in flight together).
__device__ __forceinline__ void fill_word(const Layout &L, const Inst &In, int g, int w,
                                          unsigned long long word, int off) {
  // the fields in registers: the int stores below may alias the table
  const int start = In.start;
  const bool wcor = L.ccor && In.level == 1;
  const double *crow = In.cor + (size_t)(g - start) * In.p;
  while (word) {
    int col[4];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      col[u] = -1;
      if (word) {
        col[u] = w * 64 + __ffsll((long long)word) - 1;
        word &= word - 1ull;
        cnt = u + 1;
      }
    }
    double cv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) cv[u] = (wcor && col[u] >= 0) ? crow[col[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u < cnt) {
        L.nbr[off + u] = start + col[u];
        L.erow[off + u] = g;
        L.fp[off + u] = -1;  // no passing set found yet
        if (L.erev) L.erev[off + u] = -1;
        if (wcor) L.ccor[off + u] = cv[u];
      }
    }
    off += cnt;
  }
}

// compactCUTiled (compact.cu:4-89): the ascending neighbour ids of row g
// (global ids), written at rowptr[g] (CSR instead of a p-wide padded row),
// plus the row of every ordered edge, whose first passing rank starts at -1.
// A wave prefix sum of per-word popcounts gives each lane its output offset.
__global__ __launch_bounds__(256) void fill_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= L.p_total) return;
  const Inst In = L.inst[L.row_inst[g]];
  if (!(In.flags & kActive) || In.level < 1) return;
  const unsigned long long *row = In.adj + (size_t)(g - In.start) * In.W;
  int base = L.rowptr[g];
  for (int w0 = 0; w0 < In.W; w0 += 64) {
    const int w = w0 + lane;
    unsigned long long word = w < In.W ? row[w] : 0ull;
    int c = __popcll(word);
    int incl = c;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      int v = __shfl_up(incl, s, 64);
      if (lane >= s) incl += v;
    }
    fill_word(L, In, g, w, word, base + incl - c);
    base += __shfl(incl, 63, 64);
  }
}

__global__ void table_kernel(Inst *__restrict__ dst, TableArg arg) {
  const int k = threadIdx.x;
  if (k < arg.K) dst[k] = arg.inst[k];
  for (int w = k; w < arg.zero_words; w += blockDim.x) arg.zero[w] = 0ull;
}

// ---- two-kernel compaction ---------------------------------------------------
// Count: row lengths (compact.cu:80-88), task counts deg * ceil(C(deg-1, l) /
// rpt) and per-workgroup sums; the last workgroup to finish scans the sums,
// writes the totals and publishes the counters.  Fill: row pointers, the task
// plan and the neighbour lists of each workgroup's rows.  Nothing waits on
// another workgroup.
constexpr int kCompWaves = 4;
constexpr int kCompRowsPerWave = 4;
static_assert(kCompRows == kCompWaves * kCompRowsPerWave, "compaction rows per workgroup");

__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long *v) {
  return __hip_atomic_load(v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// blk[2b], blk[2b + 1]: edges and tasks of workgroup b; then, from the last
// workgroup, their exclusive prefixes.  blk[2 nblk + b]: the longest row of
// workgroup b, reduced by the last workgroup (per-instance maxima of several
// instances still take one atomic per workgroup and instance).
__global__ __launch_bounds__(64 * kCompWaves) void compact_count_kernel(
    Layout L, int level, int rpt, int nblk, unsigned long long *blk,
    long long *__restrict__ row_tasks, Publish pub) {
  __shared__ int s_deg[kCompRows], s_k[kCompRows];
  __shared__ long long s_tasks[kCompRows];
  __shared__ int s_last;
  __shared__ int s_mx[kCompWaves];
  __shared__ unsigned long long s_part[2 * 64 * kCompWaves];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int b = blockIdx.x;
  const int g0 = b * kCompRows;
  // the wave's rows side by side, so their word loads are in flight together
  // (a one-instance launch needs no row -> instance lookup)
  int kq[kCompRowsPerWave], Wq[kCompRowsPerWave], cq[kCompRowsPerWave];
  bool onq[kCompRowsPerWave];
  const unsigned long long *rowq[kCompRowsPerWave];
  int Wmax = 0;
#pragma unroll
  for (int rr = 0; rr < kCompRowsPerWave; ++rr) {
    const int g = g0 + wv * kCompRowsPerWave + rr;
    kq[rr] = -1;
    onq[rr] = false;
    Wq[rr] = 0;
    cq[rr] = 0;
    rowq[rr] = nullptr;
    if (g < L.p_total) {
      kq[rr] = L.K == 1 ? 0 : L.row_inst[g];
      const Inst &In = L.inst[kq[rr]];
      onq[rr] = (In.flags & kActive) && In.level >= 1;
      if (onq[rr]) {
        rowq[rr] = In.adj + (size_t)(g - In.start) * In.W;
        Wq[rr] = In.W;
        Wmax = max(Wmax, In.W);
      }
    }
  }
  for (int w0 = 0; w0 < Wmax; w0 += 64) {
    const int w = w0 + lane;
#pragma unroll
    for (int rr = 0; rr < kCompRowsPerWave; ++rr)
      if (w < Wq[rr]) cq[rr] += __popcll(rowq[rr][w]);
  }
#pragma unroll
  for (int rr = 0; rr < kCompRowsPerWave; ++rr) {
    const int r = wv * kCompRowsPerWave + rr;
    const int g = g0 + r;
    int c = cq[rr];
    const int k = kq[rr];
    long long tk = 0;
    if (g < L.p_total) {
#pragma unroll
      for (int sh = 32; sh > 0; sh >>= 1) c += __shfl_xor(c, sh, 64);
      if (onq[rr] && level >= 2 && L.inst[k].level == level && c > 0) {
        // a row that cannot be planned gets no tasks; the host fails its level
        tk = max(0ll, row_task_count(binom(L.binom, kBinomK, c - 1, level), c, rpt));
      }
    }
    if (lane == 0) {
      s_deg[r] = c;
      s_k[r] = c ? k : -1;
      s_tasks[r] = tk;
      if (g < L.p_total) {
        L.deg[g] = c;
        row_tasks[g] = tk;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long E = 0, T = 0;
    int mx = 0, cur = -1, m = 0;
    for (int r = 0; r < kCompRows; ++r) {
      E += (unsigned long long)s_deg[r];
      T += (unsigned long long)s_tasks[r];
      mx = max(mx, s_deg[r]);
      // per-instance maxima over runs of rows of one instance
      const int kr = s_k[r];
      if (kr < 0 || L.K == 1) continue;
      if (kr != cur && cur >= 0) { atomicMax(&L.ictr[cur].max_deg, m); m = 0; }
      cur = kr;
      m = max(m, s_deg[r]);
    }
    if (cur >= 0) atomicMax(&L.ictr[cur].max_deg, m);
    blk[2 * (size_t)b] = E;
    blk[2 * (size_t)b + 1] = T;
    blk[2 * (size_t)nblk + b] = (unsigned long long)mx;
    __threadfence();
    s_last = atomicAdd(&L.ctr->done, 1u) == (unsigned)(nblk - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // the last workgroup: exclusive scan of the workgroup sums
  const int nt = blockDim.x;
  const int per = (nblk + nt - 1) / nt;
  const int lo = tid * per, hi = min(nblk, lo + per);
  unsigned long long se = 0, st = 0;
  int mloc = 0;
  for (int q = lo; q < hi; ++q) {
    se += ld_agent(&blk[2 * (size_t)q]);
    st += ld_agent(&blk[2 * (size_t)q + 1]);
    mloc = max(mloc, (int)ld_agent(&blk[2 * (size_t)nblk + q]));
  }
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) mloc = max(mloc, __shfl_xor(mloc, sh, 64));
  if (lane == 0) s_mx[wv] = mloc;
  s_part[tid] = se;
  s_part[nt + tid] = st;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    const unsigned long long ve = tid >= off ? s_part[tid - off] : 0ull;
    const unsigned long long vt = tid >= off ? s_part[nt + tid - off] : 0ull;
    __syncthreads();
    s_part[tid] += ve;
    s_part[nt + tid] += vt;
    __syncthreads();
  }
  unsigned long long re = s_part[tid] - se, rt = s_part[nt + tid] - st;
  for (int q = lo; q < hi; ++q) {
    const unsigned long long e = ld_agent(&blk[2 * (size_t)q]);
    const unsigned long long t = ld_agent(&blk[2 * (size_t)q + 1]);
    blk[2 * (size_t)q] = re;
    blk[2 * (size_t)q + 1] = rt;
    re += e;
    rt += t;
  }
  if (tid == nt - 1) {
    L.rowptr[L.p_total] = (int32_t)s_part[tid];
    L.ctr->n_edges = s_part[tid];
    L.ctr->n_tasks = s_part[nt + tid];
    int mx = 0;
#pragma unroll
    for (int w = 0; w < kCompWaves; ++w) mx = max(mx, s_mx[w]);
    L.ctr->max_deg = max(L.ctr->max_deg, mx);
    if (L.K == 1) L.ictr[0].max_deg = max(L.ictr[0].max_deg, mx);
  }
  __threadfence();
  __syncthreads();
  if (!pub.flag) return;
  const unsigned long long *src = reinterpret_cast<const unsigned long long *>(L.ctr);
  for (int w = tid; w < pub.words; w += nt) pub.dst[w] = ld_agent(&src[w]);
  if (!pub.sys) return;  // device-memory slot: the kernel's end publishes it
  __threadfence_system();
  __syncthreads();
  if (tid == 0)
    __hip_atomic_store(pub.flag, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One wavefront per row of the workgroup's kCompRows, so the rows' dependent
// loads (row pointer, adjacency words, correlations) overlap.
__global__ __launch_bounds__(64 * kCompRows) void compact_fill_kernel(
    Layout L, const unsigned long long *__restrict__ blk, const long long *__restrict__ row_tasks,
    long long *__restrict__ task_off, unsigned long long edge_cap) {
  __shared__ int s_deg[kCompRows];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int b = blockIdx.x;
  const int g0 = b * kCompRows;
  const unsigned long long exe = blk[2 * (size_t)b], ext = blk[2 * (size_t)b + 1];
  if (tid < kCompRows) s_deg[tid] = g0 + tid < L.p_total ? L.deg[g0 + tid] : 0;
  __syncthreads();
  if (tid < kCompRows) {
    const int g = g0 + tid;
    if (g < L.p_total) {
      unsigned long long pe = 0, pt = 0;
      for (int r = 0; r < tid; ++r) {
        pe += (unsigned long long)s_deg[r];
        pt += (unsigned long long)row_tasks[g0 + r];
      }
      L.rowptr[g] = (int32_t)(exe + pe);
      task_off[g] = (long long)(ext + pt);
    }
  }
  {
    const int r = wv;
    const int g = g0 + r;
    if (g >= L.p_total || s_deg[r] == 0) return;
    unsigned long long base = exe;
    for (int q = 0; q < r; ++q) base += (unsigned long long)s_deg[q];
    if (base + (unsigned long long)s_deg[r] > edge_cap) return;  // never with a true prefix
    const Inst &In = L.inst[L.row_inst[g]];
    const unsigned long long *row = In.adj + (size_t)(g - In.start) * In.W;
    int wbase = (int)base;
    for (int w0 = 0; w0 < In.W; w0 += 64) {
      const int w = w0 + lane;
      unsigned long long word = w < In.W ? row[w] : 0ull;
      const int c = __popcll(word);
      int incl = c;
#pragma unroll
      for (int sh = 1; sh < 64; sh <<= 1) {
        const int v = __shfl_up(incl, sh, 64);
        if (lane >= sh) incl += v;
      }
      fill_word(L, In, g, w, word, wbase + incl - c);
      wbase += __shfl(incl, 63, 64);
    }
  }
}

// One workgroup: the words, a system-scope fence per thread, then the flag.
__global__ __launch_bounds__(256) void publish_kernel(const unsigned long long *__restrict__ src,
                                                      size_t words,
                                                      unsigned long long *__restrict__ dst,
                                                      unsigned long long *flag,
                                                      unsigned long long seq) {
  for (size_t w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void copy_out_kernel(const CopyOut *__restrict__ entries,
                                                      char *__restrict__ dst) {
  const CopyOut c = entries[blockIdx.x];
  const unsigned long long w8 = c.bytes / 8;
  const unsigned long long *s8 = static_cast<const unsigned long long *>(c.src);
  unsigned long long *d8 = reinterpret_cast<unsigned long long *>(dst + c.dst);
  if ((reinterpret_cast<uintptr_t>(c.src) & 7) == 0) {
    for (unsigned long long w = threadIdx.x; w < w8; w += blockDim.x) d8[w] = s8[w];
    if (threadIdx.x == 0 && (c.bytes & 4))
      reinterpret_cast<uint32_t *>(dst + c.dst)[2 * w8] = static_cast<const uint32_t *>(c.src)[2 * w8];
  } else {
    const uint32_t *s4 = static_cast<const uint32_t *>(c.src);
    uint32_t *d4 = reinterpret_cast<uint32_t *>(dst + c.dst);
    for (unsigned long long w = threadIdx.x; w < c.bytes / 4; w += blockDim.x) d4[w] = s4[w];
  }
}

}  // namespace

void launch_copy_out(hipStream_t s, const CopyOut *entries, int count, char *dst) {
  if (count > 0) hipLaunchKernelGGL(copy_out_kernel, dim3((unsigned)count), dim3(256), 0, s, entries, dst);
}

void launch_publish(hipStream_t s, const unsigned long long *src, size_t words,
                    unsigned long long *dst, unsigned long long *flag, unsigned long long seq) {
  hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(256), 0, s, src, words, dst, flag, seq);
}

int compact_fused_blocks(int p_total) { return (p_total + kCompRows - 1) / kCompRows; }

void launch_compact_fused(hipStream_t s, const Layout &L, int level, int rpt,
                          unsigned long long *blk, long long *row_tasks, long long *task_off,
                          size_t edge_cap, const Publish &pub) {
  const int nblk = compact_fused_blocks(L.p_total);
  hipLaunchKernelGGL(compact_count_kernel, dim3(nblk), dim3(64 * kCompWaves), 0, s, L, level, rpt,
                     nblk, blk, row_tasks, pub);
  hipLaunchKernelGGL(compact_fill_kernel, dim3(nblk), dim3(64 * kCompRows), 0, s, L,
                     (const unsigned long long *)blk, (const long long *)row_tasks, task_off,
                     (unsigned long long)edge_cap);
}

void launch_compact_fill(hipStream_t s, const Layout &L, const unsigned long long *blk,
                         const long long *row_tasks, long long *task_off, size_t edge_cap) {
  const int nblk = compact_fused_blocks(L.p_total);
  hipLaunchKernelGGL(compact_fill_kernel, dim3(nblk), dim3(64 * kCompRows), 0, s, L, blk,
                     row_tasks, task_off, (unsigned long long)edge_cap);
}

void launch_table(hipStream_t s, Inst *dst, const TableArg &arg) {
  hipLaunchKernelGGL(table_kernel, dim3(1), dim3(64), 0, s, dst, arg);
}

void launch_adj_init(hipStream_t s, unsigned long long *adj, int p, int W) {
  long long n = (long long)p * W;
  hipLaunchKernelGGL(adj_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, adj, p, W);
}

void launch_level0(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                   long long *d_prefix, std::vector<long long> &h_prefix, const Publish *pub) {
  const int K = L.K;
  h_prefix.assign(K + 1, 0);
  for (int k = 0; k < K; ++k) {
    const Inst &I = host_inst[k];
    const bool on = (I.flags & kActive) && I.level == 0;
    h_prefix[k + 1] = h_prefix[k] + (on ? (long long)I.W * (I.W + 1) / 2 : 0);
  }
  Publish none{};
  const Publish fold_pub = (pub && K == 1) ? *pub : none;
  if (h_prefix[K] == 0) {
    // nothing to test; a requested publication still happens
    if (fold_pub.flag) hipLaunchKernelGGL(level0_fold_kernel, dim3(1), dim3(64), 0, s, L, fold_pub);
    return;
  }
  if (K == 1) {
    // one instance: no offsets to upload
    hipLaunchKernelGGL(level0_kernel, dim3((unsigned)h_prefix[K]), dim3(64 * kL0Waves), 0, s, L,
                       (const long long *)nullptr, h_prefix[K]);
  } else {
    (void)hipMemcpyAsync(d_prefix, h_prefix.data(), h_prefix.size() * sizeof(long long),
                         hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(level0_kernel, dim3((unsigned)h_prefix[K]), dim3(64 * kL0Waves), 0, s, L,
                       (const long long *)d_prefix, h_prefix[K]);
  }
  hipLaunchKernelGGL(level0_fold_kernel, dim3((unsigned)K), dim3(64), 0, s, L, fold_pub);
}

void launch_compact(hipStream_t s, const Layout &L) {
  hipLaunchKernelGGL(degree_kernel, dim3((L.p_total + kDegWaves - 1) / kDegWaves),
                     dim3(64 * kDegWaves), 0, s, L);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, L);
  const int wpb = 4;
  hipLaunchKernelGGL(fill_kernel, dim3((L.p_total + wpb - 1) / wpb), dim3(64 * wpb), 0, s, L);
}

}  // namespace hpc
