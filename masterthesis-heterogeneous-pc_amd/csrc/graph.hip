// Adjacency kernels: initial state, the level-0 test and per-level compaction.
//
// The adjacency lives on the device as a bit matrix: row i is W = ceil(p/64)
// 64-bit words, bit j of word j/64 set while edge i-j is present.  p = 2000
// takes 512 KB instead of the reference's 16 MB int32 matrix (state.cu:21-38).
#pragma clang fp contract(off)

#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

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
// One wavefront per 64 x 64 tile (I, J) with I >= J of the lower block
// triangle.  Iteration r reads row i = 64I + r, columns 64J..64J+63 (one
// coalesced 512 B load); the ballot of surviving pairs is row i's word J, and
// lane c collects bit r of row (64J + c)'s word I.  Each adjacency word has
// exactly one owner, so no atomics are needed.
__global__ __launch_bounds__(256) void level0_kernel(
    const double *__restrict__ cor, unsigned long long *__restrict__ adj, int p,
    int W, double r_crit, long long ntiles, Counters *__restrict__ ctr) {
  const int lane = threadIdx.x & 63;
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= ntiles) return;
  // lower block triangle, row-major: tile t -> (I, J), J <= I
  int I = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((long long)I * (I + 1) / 2 > t) --I;
  while ((long long)(I + 1) * (I + 2) / 2 <= t) ++I;
  const int J = (int)(t - (long long)I * (I + 1) / 2);
  const int j = J * 64 + lane;
  const bool diag = (I == J);
  unsigned long long colword = 0ull;   // bits r: pair (64I + r, j) kept
  unsigned long long rowword = 0ull;   // captured by lane r: row 64I + r, word J
  unsigned long long tested = 0ull, removed = 0ull;
  for (int r = 0; r < 64; ++r) {
    const int i = I * 64 + r;
    if (i >= p) break;
    const unsigned long long aw = adj[(size_t)i * W + J];
    bool valid = (j < p) && (diag ? (lane < r) : true);
    bool present = valid && ((aw >> lane) & 1ull);
    bool keep = false;
    if (present) {
      double c = cor[(size_t)i * p + j];
      keep = !removes(c, r_crit);
    }
    const unsigned long long kb = __ballot(keep);
    tested += __popcll(__ballot(present));
    removed += __popcll(__ballot(present && !keep));
    if (lane == r) rowword = kb;
    colword |= (unsigned long long)(keep ? 1ull : 0ull) << r;
  }
  const int i_lane = I * 64 + lane;
  if (diag) {
    if (i_lane < p) adj[(size_t)i_lane * W + J] = rowword | colword;
  } else {
    if (i_lane < p) adj[(size_t)i_lane * W + J] = rowword;
    if (j < p) adj[(size_t)j * W + I] = colword;
  }
  if (lane == 0) {
    atomicAdd(&ctr->tests, tested);
    atomicAdd(&ctr->tests_max, tested);
    atomicAdd(&ctr->edges, tested);
    atomicAdd(&ctr->removed, removed);
  }
}

// Row lengths (compact.cu:80-88 stores them at adj_compact[i][p-1]).
__global__ __launch_bounds__(256) void degree_kernel(
    const unsigned long long *__restrict__ adj, int p, int W,
    int32_t *__restrict__ deg) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= p) return;
  int c = 0;
  for (int w = lane; w < W; w += 64) c += __popcll(adj[(size_t)i * W + w]);
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s, 64);
  if (lane == 0) deg[i] = c;
}

// Exclusive scan of row lengths -> CSR row pointers, total ordered edges and
// the maximum row length (compact.cu:81 atomicMax_system(max_adj)).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void scan_kernel(
    const int32_t *__restrict__ deg, int p, int32_t *__restrict__ rowptr,
    Counters *__restrict__ ctr) {
  __shared__ long long part[kScanThreads];
  __shared__ int mx[kScanThreads];
  const int tid = threadIdx.x;
  const int per = (p + kScanThreads - 1) / kScanThreads;
  const int b = tid * per, e = min(p, b + per);
  long long s = 0;
  int m = 0;
  for (int i = b; i < e; ++i) { s += deg[i]; m = max(m, deg[i]); }
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
  for (int i = b; i < e; ++i) { rowptr[i] = (int32_t)run; run += deg[i]; }
  if (tid == kScanThreads - 1) {
    rowptr[p] = (int32_t)part[tid];
    ctr->n_edges = (unsigned long long)part[tid];
    ctr->max_deg = mx[tid];
  }
}

// compactCUTiled (compact.cu:4-89): the ascending neighbour ids of row i,
// written at rowptr[i] (CSR instead of a p-wide padded row), plus the row id
// of every ordered edge.  Wave prefix sum of per-word popcounts gives each
// lane its output offset.
__global__ __launch_bounds__(256) void fill_kernel(
    const unsigned long long *__restrict__ adj, int p, int W,
    const int32_t *__restrict__ rowptr, int32_t *__restrict__ nbr,
    int32_t *__restrict__ erow) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= p) return;
  int base = rowptr[i];
  for (int w0 = 0; w0 < W; w0 += 64) {
    const int w = w0 + lane;
    unsigned long long word = w < W ? adj[(size_t)i * W + w] : 0ull;
    int c = __popcll(word);
    int incl = c;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      int v = __shfl_up(incl, s, 64);
      if (lane >= s) incl += v;
    }
    int off = base + incl - c;
    while (word) {
      int b = __ffsll((long long)word) - 1;
      word &= word - 1ull;
      nbr[off] = w * 64 + b;
      erow[off] = i;
      ++off;
    }
    base += __shfl(incl, 63, 64);
  }
}

}  // namespace

void launch_adj_init(hipStream_t s, unsigned long long *adj, int p, int W) {
  long long n = (long long)p * W;
  dim3 grid((unsigned)((n + 255) / 256)), block(256);
  hipLaunchKernelGGL(adj_init_kernel, grid, block, 0, s, adj, p, W);
}

void launch_level0(hipStream_t s, const double *cor, unsigned long long *adj,
                   int p, int W, double r_crit, Counters *ctr) {
  const long long nt = (long long)W * (W + 1) / 2;
  const int wpb = 4;
  dim3 grid((unsigned)((nt + wpb - 1) / wpb)), block(64 * wpb);
  hipLaunchKernelGGL(level0_kernel, grid, block, 0, s, cor, adj, p, W, r_crit,
                     nt, ctr);
}

void launch_compact(hipStream_t s, const unsigned long long *adj, int p, int W,
                    int32_t *deg, int32_t *rowptr, int32_t *nbr, int32_t *erow,
                    Counters *ctr) {
  const int wpb = 4;
  dim3 grid((p + wpb - 1) / wpb), block(64 * wpb);
  hipLaunchKernelGGL(degree_kernel, grid, block, 0, s, adj, p, W, deg);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, deg, p,
                     rowptr, ctr);
  hipLaunchKernelGGL(fill_kernel, grid, block, 0, s, adj, p, W, rowptr, nbr,
                     erow);
}

}  // namespace hpc
