// Level >= 1 CI tests (replaces testRowL1 / testRowLN<> of gpuRowTests.cu:29-205
// and the workstealing variants of gpuWorkstealingTests.cu) and the level-end
// edge removal.
//
// Search: for every ordered edge (i, j) of the level-start snapshot, the
// lowest rank whose conditioning set S (an l-subset of N(i)\{j} in
// lexicographic order of positions, cuda_util.cu:34-50 and
// gpuRowTests.cu:128-142) gives p >= alpha.  One 64-wide wavefront per edge;
// lane t tests rank base + t, a ballot picks the lowest passing lane.  Edges
// do not cut each other's searches short, so the result and the test count
// are independent of scheduling.
//
// Commit: edge i-j (i < j) is removed when either side found a set; the
// sepset is the lower-id row's lowest passing set, else the higher-id row's.
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
  unsigned long long T = 0, U = 0, E = 0;
  const int p = v.p;
  for (long long e = wave0; e < n_edges; e += nw) {
    const int i = v.erow[e];
    if (row_mask && !row_mask[i]) continue;
    const int base = v.rowptr[i];
    const int nn = v.rowptr[i + 1] - base - 1;
    const int y = (int)(e - base);
    if (nn < 1) { if (lane == 0) v.fp[e] = -1; continue; }
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
    if (lane == 0) {
      v.fp[e] = found;
      T += found >= 0 ? (unsigned long long)found + 1 : (unsigned long long)nn;
      U += nn;
      E += 1;
    }
  }
  block_add(v.ctr, T, U, E, 0);
}

// Level L >= 2 (testRowLN<>, gpuRowTests.cu:102-205; the general-level form of
// cpuWorkstealingTests.cpp:39-112).
template <int L>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_ln_kernel(
    Views v, long long n_edges, const int32_t *__restrict__ row_mask, int lds_binom) {
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
  unsigned long long T = 0, U = 0, E = 0, F = 0;
  const int p = v.p;
  for (long long e = wave0; e < n_edges; e += nw) {
    const int i = v.erow[e];
    if (row_mask && !row_mask[i]) continue;
    const int base = v.rowptr[i];
    const int nn = v.rowptr[i + 1] - base - 1;
    const int y = (int)(e - base);
    const unsigned long long total = binom(tab, kc, nn, L);
    if (total == 0) { if (lane == 0) v.fp[e] = -1; continue; }
    const int j = v.nbr[e];
    const double a = v.cor[(size_t)i * p + j];
    long long found = -1;
    for (unsigned long long r0 = 0; r0 < total; r0 += 64) {
      const unsigned long long rank = r0 + lane;
      bool pass = false;
      if (rank < total) {
        int q[L], S[L];
        unrank<L>(tab, kc, rank, nn, q);
#pragma unroll
        for (int c = 0; c < L; ++c) S[c] = v.nbr[base + (q[c] < y ? q[c] : q[c] + 1)];
        double x[L], yv[L], C[L][L];
        bool anynan = isnan(a);
#pragma unroll
        for (int c = 0; c < L; ++c) {
          x[c] = v.cor[(size_t)i * p + S[c]];
          yv[c] = v.cor[(size_t)j * p + S[c]];
          anynan |= isnan(x[c]) || isnan(yv[c]);
        }
#pragma unroll
        for (int c = 0; c < L; ++c)
#pragma unroll
          for (int r = c + 1; r < L; ++r) {
            C[r][c] = v.cor[(size_t)S[c] * p + S[r]];
            anynan |= isnan(C[r][c]);
          }
        if (!anynan) {
          bool fb;
          double r = pcor_canon<L>(a, x, yv, C, fb);
          pass = removes(r, v.r_crit);
          F += fb ? 1 : 0;
        }
      }
      const unsigned long long m = __ballot(pass);
      if (m) { found = (long long)(r0 + __ffsll((long long)m) - 1); break; }
    }
    if (lane == 0) {
      v.fp[e] = found;
      T += found >= 0 ? (unsigned long long)found + 1 : total;
      U += total;
      E += 1;
    }
  }
  // per-lane fallback counts -> wave sum
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) F += __shfl_xor(F, s, 64);
  block_add(v.ctr, T, U, E, F);
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
    int W, int32_t *__restrict__ sep, int mc, int8_t *__restrict__ rmlev) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edges) return;
  const int i = v.erow[e];
  const int j = v.nbr[e];
  if (j < i) return;
  const int bi = v.rowptr[i], bj = v.rowptr[j];
  const int yi = (int)(e - bi);
  const int lenj = v.rowptr[j + 1] - bj;
  const int yj = find_pos(v.nbr + bj, lenj, i);
  const long long fi = v.fp[e];
  // the adjacency is symmetric, so i is in N(j); guard anyway
  const long long fj = (yj < lenj && v.nbr[bj + yj] == i) ? v.fp[bj + yj] : -1;
  int row, y, base, nn;
  unsigned long long rank;
  if (fi >= 0) { row = i; y = yi; base = bi; rank = (unsigned long long)fi; }
  else if (fj >= 0) { row = j; y = yj; base = bj; rank = (unsigned long long)fj; }
  else return;
  nn = v.rowptr[row + 1] - base - 1;
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

template <int L>
void launch_ln(hipStream_t s, dim3 g, dim3 b, const Views &v, long long ne,
               const int32_t *mask) {
  const size_t lds = (size_t)(v.max_deg + 1) * (L + 1) * sizeof(unsigned long long);
  const int use = lds <= 48 * 1024;
  hipLaunchKernelGGL(search_ln_kernel<L>, g, b, use ? lds : 0, s, v, ne, mask, use);
}

}  // namespace

void launch_search(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   const int32_t *row_mask) {
  if (n_edges <= 0) return;
  long long blocks = (n_edges + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks > 16384) blocks = 16384;
  dim3 g((unsigned)blocks), b(64 * kWavesPerBlock);
  switch (level) {
    case 1: hipLaunchKernelGGL(search_l1_kernel, g, b, 0, s, v, (long long)n_edges, row_mask); break;
    case 2: launch_ln<2>(s, g, b, v, n_edges, row_mask); break;
    case 3: launch_ln<3>(s, g, b, v, n_edges, row_mask); break;
    case 4: launch_ln<4>(s, g, b, v, n_edges, row_mask); break;
    case 5: launch_ln<5>(s, g, b, v, n_edges, row_mask); break;
    case 6: launch_ln<6>(s, g, b, v, n_edges, row_mask); break;
    case 7: launch_ln<7>(s, g, b, v, n_edges, row_mask); break;
    case 8: launch_ln<8>(s, g, b, v, n_edges, row_mask); break;
    case 9: launch_ln<9>(s, g, b, v, n_edges, row_mask); break;
    case 10: launch_ln<10>(s, g, b, v, n_edges, row_mask); break;
    case 11: launch_ln<11>(s, g, b, v, n_edges, row_mask); break;
    case 12: launch_ln<12>(s, g, b, v, n_edges, row_mask); break;
    case 13: launch_ln<13>(s, g, b, v, n_edges, row_mask); break;
    case 14: launch_ln<14>(s, g, b, v, n_edges, row_mask); break;
    case 15: launch_ln<15>(s, g, b, v, n_edges, row_mask); break;
    case 16: launch_ln<16>(s, g, b, v, n_edges, row_mask); break;
    default: break;  // rejected by the caller (kMaxLevel)
  }
}

void launch_commit(hipStream_t s, int level, const Views &v, int64_t n_edges,
                   unsigned long long *adj, int W, int32_t *sep, int mc,
                   int8_t *rmlev) {
  if (n_edges <= 0) return;
  dim3 g((unsigned)((n_edges + 255) / 256)), b(256);
  hipLaunchKernelGGL(commit_kernel, g, b, 0, s, v, (long long)n_edges, level,
                     adj, W, sep, mc, rmlev);
}

}  // namespace hpc
