// Level >= 1 CI tests (replaces testRowL1 / testRowLN<> of gpuRowTests.cu:29-205
// and the workstealing variants of gpuWorkstealingTests.cu) and the level-end
// edge removal.
//
// Search: for every ordered edge (i, j) of the level-start snapshot, the
// lowest rank whose conditioning set S (an l-subset of N(i)\{j} in
// lexicographic order of positions, cuda_util.cu:34-50 and
// gpuRowTests.cu:128-142) gives p >= alpha.  A 64-wide wavefront tests 64
// consecutive ranks at a time, lane t rank base + t, and a ballot picks the
// lowest passing lane.  Level 1 runs one wavefront per edge.  Levels >= 2
// cut each edge's ranks into tasks of `ranks_per_task` (a multiple of 64) so
// that edges with many sets spread over many wavefronts; a task skips ranks
// above a pass another task of its edge already found, and the lowest rank
// wins through atomicMin.  Edges do not cut each other's searches short, so
// the result is independent of scheduling.
//
// Commit: edge i-j (i < j) is removed when either side found a set; the
// sepset is the lower-id row's lowest passing set, else the higher-id row's.
// It also counts the canonical tests (per ordered edge: first passing rank +
// 1, else all its sets), which depend only on the first passing ranks.
#pragma clang fp contract(off)

#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ void block_add(Counters *ctr, unsigned long long t,
                                          unsigned long long u,
                                          unsigned long long e,
                                          unsigned long long f) {
  __shared__ unsigned long long red[4][kWavesPerBlock];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = t; red[1][wv] = u; red[2][wv] = e; red[3][wv] = f; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) {
      s0 += red[0][k]; s1 += red[1][k]; s2 += red[2][k]; s3 += red[3][k];
    }
    if (s0) atomicAdd(&ctr->tests, s0);
    if (s1) atomicAdd(&ctr->tests_max, s1);
    if (s2) atomicAdd(&ctr->edges, s2);
    if (s3) atomicAdd(&ctr->pinv_fallbacks, s3);
  }
}

// Level 1 (testRowL1, gpuRowTests.cu:29-100; cpuRowTests.cpp:23-56).
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_l1_kernel(
    Views v, long long n_edges, const int32_t *__restrict__ row_mask) {
  const int lane = threadIdx.x & 63;
  const long long wave0 = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const long long nw = (long long)gridDim.x * kWavesPerBlock;
  const int p = v.p;
  for (long long e = wave0; e < n_edges; e += nw) {
    const int i = v.erow[e];
    if (row_mask && !row_mask[i]) continue;
    const int base = v.rowptr[i];
    const int nn = v.rowptr[i + 1] - base - 1;
    const int y = (int)(e - base);
    if (nn < 1) continue;
    const int j = v.nbr[e];
    const double x1 = v.cor[(size_t)i * p + j];
    long long found = -1;
    for (int q0 = 0; q0 < nn; q0 += 64) {
      const int q = q0 + lane;
      bool pass = false;
      if (q < nn) {
        const int k = v.nbr[base + (q < y ? q : q + 1)];
        const double x2 = v.cor[(size_t)i * p + k];
        const double x3 = v.cor[(size_t)j * p + k];
        if (!(isnan(x1) || isnan(x2) || isnan(x3)))
          pass = removes(pcor_l1(x1, x2, x3), v.r_crit);
      }
      const unsigned long long m = __ballot(pass);
      if (m) { found = q0 + __ffsll((long long)m) - 1; break; }
    }
    if (lane == 0) v.fp[e] = found;
  }
}

// Level L >= 2 (testRowLN<>, gpuRowTests.cu:102-205; the general-level form of
// cpuWorkstealingTests.cpp:39-112).  tasks[t] = (ordered edge << 32) | chunk;
// chunk c covers ranks [c * rpt, (c + 1) * rpt).
template <int L>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_ln_kernel(
    Views v, const unsigned long long *__restrict__ tasks, long long n_tasks, int rpt,
    int lds_binom) {
  // C(m, k) for m <= max_deg, k <= L staged in LDS (columns 0..L)
  extern __shared__ unsigned long long s_binom[];
  const unsigned long long *tab = v.binom;
  int kc = kBinomK;
  if (lds_binom) {
    const int rows = v.max_deg + 1;
    for (int t = threadIdx.x; t < rows * (L + 1); t += blockDim.x) {
      const int m = t / (L + 1), k = t % (L + 1);
      s_binom[t] = v.binom[(size_t)m * kBinomK + k];
    }
    __syncthreads();
    tab = s_binom;
    kc = L + 1;
  }
  const int lane = threadIdx.x & 63;
  const long long wave0 = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const long long nw = (long long)gridDim.x * kWavesPerBlock;
  unsigned long long F = 0;
  const int p = v.p;
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(v.fp);
  for (long long t = wave0; t < n_tasks; t += nw) {
    const unsigned long long task = tasks[t];
    const long long e = (long long)(task >> 32);
    const unsigned long long r_begin = (task & 0xffffffffull) * (unsigned long long)rpt;
    // a lower rank of this edge already passed: nothing here can win
    if (__atomic_load_n(&fpu[e], __ATOMIC_RELAXED) < r_begin) continue;
    const int i = v.erow[e];
    const int base = v.rowptr[i];
    const int nn = v.rowptr[i + 1] - base - 1;
    const int y = (int)(e - base);
    const unsigned long long total = binom(tab, kc, nn, L);
    const unsigned long long r_end = min(total, r_begin + (unsigned long long)rpt);
    const int j = v.nbr[e];
    const double a = v.cor[(size_t)i * p + j];
    for (unsigned long long r0 = r_begin; r0 < r_end; r0 += 64) {
      const unsigned long long rank = r0 + lane;
      bool pass = false;
      if (rank < r_end) {
        int q[L], S[L];
        unrank<L>(tab, kc, rank, nn, q);
#pragma unroll
        for (int c = 0; c < L; ++c) S[c] = v.nbr[base + (q[c] < y ? q[c] : q[c] + 1)];
        double x[L], yv[L], C[L][L];
        bool anynan = __builtin_isnan(a);
#pragma unroll
        for (int c = 0; c < L; ++c) {
          x[c] = v.cor[(size_t)i * p + S[c]];
          yv[c] = v.cor[(size_t)j * p + S[c]];
          anynan |= __builtin_isnan(x[c]) || __builtin_isnan(yv[c]);
        }
#pragma unroll
        for (int c = 0; c < L; ++c)
#pragma unroll
          for (int r = c + 1; r < L; ++r) {
            C[r][c] = v.cor[(size_t)S[c] * p + S[r]];
            anynan |= __builtin_isnan(C[r][c]);
          }
        if (!anynan) {
          bool fb;
          const double r = pcor_canon<L>(a, x, yv, C, fb);
          pass = removes(r, v.r_crit);
          F += fb ? 1 : 0;
        }
      }
      const unsigned long long m = __ballot(pass);
      if (m) {
        if (lane == 0) atomicMin(&fpu[e], r0 + (unsigned long long)(__ffsll((long long)m) - 1));
        break;
      }
    }
  }
  // per-lane fallback counts -> wave sum
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) F += __shfl_xor(F, s, 64);
  block_add(v.ctr, 0, 0, 0, F);
}

// Tasks per row: deg * ceil(C(deg - 1, L) / rpt) (0 for masked-out rows).
__global__ void task_count_kernel(Views v, int level, int rpt, const int32_t *__restrict__ row_mask,
                                  long long *__restrict__ row_tasks) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.p) return;
  const int deg = v.rowptr[i + 1] - v.rowptr[i];
  const unsigned long long total = binom(v.binom, kBinomK, deg - 1, level);
  long long n = 0;
  if ((!row_mask || row_mask[i]) && total > 0)
    n = (long long)deg * (long long)((total + rpt - 1) / rpt);
  row_tasks[i] = n;
}

// Exclusive scan of the per-row task counts (one workgroup), total to
// ctr->n_tasks.
constexpr int kScan = 1024;
__global__ __launch_bounds__(kScan) void task_scan_kernel(const long long *__restrict__ cnt, int p,
                                                          long long *__restrict__ off,
                                                          Counters *__restrict__ ctr) {
  __shared__ long long part[kScan];
  const int tid = threadIdx.x;
  const int per = (p + kScan - 1) / kScan;
  const int b = tid * per, e = min(p, b + per);
  long long s = 0;
  for (int i = b; i < e; ++i) s += cnt[i];
  part[tid] = s;
  __syncthreads();
  for (int o = 1; o < kScan; o <<= 1) {
    long long x = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += x;
    __syncthreads();
  }
  long long run = part[tid] - s;
  for (int i = b; i < e; ++i) { off[i] = run; run += cnt[i]; }
  if (tid == kScan - 1) ctr->n_tasks = (unsigned long long)part[tid];
}

// One wavefront per row writes its tasks, edge-major.
__global__ __launch_bounds__(256) void task_fill_kernel(Views v, int level, int rpt,
                                                        const long long *__restrict__ row_tasks,
                                                        const long long *__restrict__ off,
                                                        unsigned long long *__restrict__ tasks) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= v.p) return;
  const long long n = row_tasks[i];
  if (n == 0) return;
  const int base = v.rowptr[i];
  const int deg = v.rowptr[i + 1] - base;
  const long long nch = n / deg;
  const long long o = off[i];
  for (long long t = lane; t < n; t += 64) {
    const long long y = t / nch, c = t % nch;
    tasks[o + t] = ((unsigned long long)(base + y) << 32) | (unsigned long long)c;
  }
}

__device__ __forceinline__ int find_pos(const int32_t *row, int len, int node) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (row[mid] < node) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Removal and sepset recording (cpuUtil.cpp:23-48).  The device keeps sepset
// member k of pair (min, max) at sep[k*p*p + min*p + max] (one plane per
// member, so the width can grow by appending planes) and the removal level
// at rmlev[min*p + max]; the host export rebuilds the reference layout.
__global__ __launch_bounds__(256) void commit_kernel(
    Views v, long long n_edges, int level, unsigned long long *__restrict__ adj,
    int W, int32_t *__restrict__ sep, int8_t *__restrict__ rmlev) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long T = 0, U = 0, E = 0;
  if (e < n_edges) {
    const int i = v.erow[e];
    const int j = v.nbr[e];
    const int bi = v.rowptr[i];
    const int nni = v.rowptr[i + 1] - bi - 1;
    const unsigned long long total = binom(v.binom, kBinomK, nni, level);
    const long long fi = v.fp[e];
    if (total > 0) {
      T = fi >= 0 ? (unsigned long long)fi + 1 : total;
      U = total;
      E = 1;
    }
    if (j > i) {
      const int bj = v.rowptr[j];
      const int yi = (int)(e - bi);
      const int lenj = v.rowptr[j + 1] - bj;
      const int yj = find_pos(v.nbr + bj, lenj, i);
      // the adjacency is symmetric, so i is in N(j); guard anyway
      const long long fj = (yj < lenj && v.nbr[bj + yj] == i) ? v.fp[bj + yj] : -1;
      int row = -1, y = 0, base = 0;
      unsigned long long rank = 0;
      if (fi >= 0) { row = i; y = yi; base = bi; rank = (unsigned long long)fi; }
      else if (fj >= 0) { row = j; y = yj; base = bj; rank = (unsigned long long)fj; }
      if (row >= 0) {
        const int nn = v.rowptr[row + 1] - base - 1;
        atomicAnd(&adj[(size_t)i * W + (j >> 6)], ~(1ull << (j & 63)));
        atomicAnd(&adj[(size_t)j * W + (i >> 6)], ~(1ull << (i & 63)));
        rmlev[(size_t)i * v.p + j] = (int8_t)level;
        const size_t plane = (size_t)v.p * v.p;
        int32_t *out = sep + (size_t)i * v.p + j;
        int val = 0;
        for (int k = 0; k < level; ++k) {
          for (;;) {
            unsigned long long c = binom(v.binom, kBinomK, nn - 1 - val, level - 1 - k);
            if (rank < c) break;
            rank -= c;
            ++val;
          }
          const int q = val++;
          out[(size_t)k * plane] = v.nbr[base + (q < y ? q : q + 1)];
        }
        atomicAdd(&v.ctr->removed, 1ull);
      }
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    T += __shfl_xor(T, s, 64);
    U += __shfl_xor(U, s, 64);
    E += __shfl_xor(E, s, 64);
  }
  if ((threadIdx.x & 63) == 0 && E) {
    atomicAdd(&v.ctr->tests, T);
    atomicAdd(&v.ctr->tests_max, U);
    atomicAdd(&v.ctr->edges, E);
  }
}

template <int L>
void launch_ln(hipStream_t s, const Views &v, const unsigned long long *tasks,
               long long n_tasks, int rpt) {
  long long blocks = (n_tasks + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks > 8192) blocks = 8192;
  const size_t lds = (size_t)(v.max_deg + 1) * (L + 1) * sizeof(unsigned long long);
  const int use = lds <= 48 * 1024;
  hipLaunchKernelGGL(search_ln_kernel<L>, dim3((unsigned)blocks), dim3(64 * kWavesPerBlock),
                     use ? lds : 0, s, v, tasks, n_tasks, rpt, use);
}

}  // namespace

void launch_search_l1(hipStream_t s, const Views &v, int64_t n_edges,
                      const int32_t *row_mask) {
  if (n_edges <= 0) return;
  long long blocks = (n_edges + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(search_l1_kernel, dim3((unsigned)blocks), dim3(64 * kWavesPerBlock), 0, s,
                     v, (long long)n_edges, row_mask);
}

void launch_task_plan(hipStream_t s, int level, const Views &v, int rpt,
                      const int32_t *row_mask, long long *row_tasks, long long *task_off) {
  hipLaunchKernelGGL(task_count_kernel, dim3((v.p + 255) / 256), dim3(256), 0, s, v, level, rpt,
                     row_mask, row_tasks);
  hipLaunchKernelGGL(task_scan_kernel, dim3(1), dim3(kScan), 0, s, row_tasks, v.p, task_off,
                     v.ctr);
}

void launch_task_fill(hipStream_t s, int level, const Views &v, int rpt,
                      const long long *row_tasks, const long long *task_off,
                      unsigned long long *tasks) {
  hipLaunchKernelGGL(task_fill_kernel, dim3((v.p + 3) / 4), dim3(256), 0, s, v, level, rpt,
                     row_tasks, task_off, tasks);
}

void launch_search_ln(hipStream_t s, int level, const Views &v,
                      const unsigned long long *tasks, long long n_tasks, int rpt) {
  if (n_tasks <= 0) return;
  switch (level) {
    case 2: launch_ln<2>(s, v, tasks, n_tasks, rpt); break;
    case 3: launch_ln<3>(s, v, tasks, n_tasks, rpt); break;
    case 4: launch_ln<4>(s, v, tasks, n_tasks, rpt); break;
    case 5: launch_ln<5>(s, v, tasks, n_tasks, rpt); break;
    case 6: launch_ln<6>(s, v, tasks, n_tasks, rpt); break;
    case 7: launch_ln<7>(s, v, tasks, n_tasks, rpt); break;
    case 8: launch_ln<8>(s, v, tasks, n_tasks, rpt); break;
    case 9: launch_ln<9>(s, v, tasks, n_tasks, rpt); break;
    case 10: launch_ln<10>(s, v, tasks, n_tasks, rpt); break;
    case 11: launch_ln<11>(s, v, tasks, n_tasks, rpt); break;
    case 12: launch_ln<12>(s, v, tasks, n_tasks, rpt); break;
    case 13: launch_ln<13>(s, v, tasks, n_tasks, rpt); break;
    case 14: launch_ln<14>(s, v, tasks, n_tasks, rpt); break;
    case 15: launch_ln<15>(s, v, tasks, n_tasks, rpt); break;
    case 16: launch_ln<16>(s, v, tasks, n_tasks, rpt); break;
    default: break;  // rejected by the caller (kMaxLevel)
  }
}

void launch_commit(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   unsigned long long *adj, int W, int32_t *sep, int8_t *rmlev) {
  if (n_edges <= 0) return;
  dim3 g((unsigned)((n_edges + 255) / 256)), b(256);
  hipLaunchKernelGGL(commit_kernel, g, b, 0, s, v, (long long)n_edges, level, adj, W, sep,
                     rmlev);
}

}  // namespace hpc
