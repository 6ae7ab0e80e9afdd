// Exact rechecks for instances whose correlation matrix holds fast (MFMA)
// values.  A fast entry differs from the exact one by at most Inst::eps, so
// a test decides the same on both unless its statistic lies within its error
// bound of the threshold.  Those tests are re-run here on exact correlations,
// each computed in the canonical order of the FP64 correlation kernels
// (lane-strided FMA sums over the z-scores, halving tree), so decisions,
// first passing ranks, sepsets and test counts equal the exact path's.
#pragma clang fp contract(off)

#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

__device__ __forceinline__ double tree64(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

// cor[a][b] of instance In (local ids) as corr_tile_kernel computes it; the
// value is returned on every lane.  Instances with fast values have no NaN
// cells, so no sample is masked.
__device__ __forceinline__ double exact_cor(const Inst &In, int a, int b, int lane) {
  const double *za = In.z + (size_t)a * In.npad;
  const double *zb = In.z + (size_t)b * In.npad;
  double acc = 0.0;
  for (int k = lane; k < In.npad; k += 64) acc = __builtin_fma(za[k], zb[k], acc);
  acc = tree64(acc);
  const double div = (double)((float)In.n - 1.f);
  return __shfl(acc, 0, 64) / div;
}

// Level-0 pairs (i > j, the triangle level 0 reads) of fast instances whose
// fast correlation lies within eps of the threshold.  One wavefront per row.
__global__ __launch_bounds__(256) void refine_scan_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= L.p_total) return;
  const Inst &In = L.inst[L.row_inst[g]];
  if (!In.fast || In.level != 0) return;
  const int i = g - In.start, p = In.p;
  const double r_crit = In.r_crit, eps = In.eps;
  for (int j0 = 0; j0 < i; j0 += 64) {
    const int j = j0 + lane;
    bool want = false;
    if (j < i) {
      const double c = In.cor[(size_t)i * p + j];
      if (!__builtin_isnan(c)) {
        const double re = __builtin_fmin(kCutThr, __builtin_fabs(c));
        want = __builtin_fabs(re - r_crit) <= eps;
      }
    }
    const unsigned long long m = __ballot(want);
    if (!m) continue;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&L.ctr->n_pairs, (unsigned long long)__popcll(m));
    base = __shfl(base, 0, 64);
    if (want) {
      const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1ull));
      if (slot < L.pairs_cap) L.pairs[slot] = make_int2(g, In.start + j);
      else L.ctr->overflow = 1;
    }
  }
}

// Exact correlation of every queued pair, written to both triangles.
__global__ __launch_bounds__(256) void exact_pairs_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const unsigned long long n = min(L.ctr->n_pairs, L.pairs_cap);
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       w < (long long)n; w += nw) {
    const int2 pr = L.pairs[w];
    const Inst &In = L.inst[L.row_inst[pr.x]];
    const int a = pr.x - In.start, b = pr.y - In.start;
    const double c = exact_cor(In, a, b, lane);
    if (lane == 0) {
      In.cor[(size_t)a * In.p + b] = c;
      In.cor[(size_t)b * In.p + a] = c;
    }
  }
}

// Exact re-run of each queued (edge, rank) of a level-LV search whose rank is
// still below the edge's first passing rank; a pass lowers it.
template <int LV>
__global__ __launch_bounds__(256) void recheck_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const unsigned long long n = min(L.ctr->n_rchk, L.rchk_cap);
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(L.fp);
  unsigned long long done = 0;
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       w < (long long)n; w += nw) {
    const Recheck item = L.rchk[w];
    const long long e = (long long)item.edge;
    const unsigned long long rank = item.rank;
    if (__atomic_load_n(&fpu[e], __ATOMIC_RELAXED) <= rank) continue;
    const int i = L.erow[e];
    const Inst &In = L.inst[L.row_inst[i]];
    const int s0 = In.start;
    const int base = L.rowptr[i];
    const int nn = L.rowptr[i + 1] - base - 1;
    const int y = (int)(e - base);
    int q[LV];
    if (LV == 1) q[0] = (int)rank;
    else unrank<LV>(L.binom, kBinomK, rank, nn, q);
    int S[LV];
#pragma unroll
    for (int c = 0; c < LV; ++c) S[c] = L.nbr[base + (q[c] < y ? q[c] : q[c] + 1)] - s0;
    const int li = i - s0, lj = L.nbr[e] - s0;
    const double a = exact_cor(In, li, lj, lane);
    double x[LV], yv[LV], C[LV][LV];
#pragma unroll
    for (int c = 0; c < LV; ++c) {
      x[c] = exact_cor(In, li, S[c], lane);
      yv[c] = exact_cor(In, lj, S[c], lane);
    }
#pragma unroll
    for (int c = 0; c < LV; ++c)
#pragma unroll
      for (int r = c + 1; r < LV; ++r) C[r][c] = exact_cor(In, S[c], S[r], lane);
    ++done;
    if (lane != 0) continue;
    bool anynan = __builtin_isnan(a);
#pragma unroll
    for (int c = 0; c < LV; ++c) anynan |= __builtin_isnan(x[c]) || __builtin_isnan(yv[c]);
#pragma unroll
    for (int c = 0; c < LV; ++c)
#pragma unroll
      for (int r = c + 1; r < LV; ++r) anynan |= __builtin_isnan(C[r][c]);
    bool pass = false;
    if (!anynan) {
      if (LV == 1) {
        pass = removes(pcor_l1(a, x[0], yv[0]), In.r_crit);
      } else {
        bool fb;
        pass = removes(pcor_canon<LV>(a, x, yv, C, fb), In.r_crit);
      }
    }
    if (pass) atomicMin(&fpu[e], rank);
  }
  if (lane == 0 && done) atomicAdd(&L.ctr->rechecks, done);
}

template <int LV>
void launch_rk(hipStream_t s, const Layout &L) {
  hipLaunchKernelGGL(recheck_kernel<LV>, dim3(1024), dim3(256), 0, s, L);
}

}  // namespace

void launch_refine_l0(hipStream_t s, const Layout &L) {
  const int wpb = 4;
  hipLaunchKernelGGL(refine_scan_kernel, dim3((L.p_total + wpb - 1) / wpb), dim3(64 * wpb), 0, s,
                     L);
  hipLaunchKernelGGL(exact_pairs_kernel, dim3(1024), dim3(256), 0, s, L);
}

void launch_recheck(hipStream_t s, int level, const Layout &L) {
  switch (level) {
    case 1: launch_rk<1>(s, L); break;
    case 2: launch_rk<2>(s, L); break;
    case 3: launch_rk<3>(s, L); break;
    case 4: launch_rk<4>(s, L); break;
    case 5: launch_rk<5>(s, L); break;
    case 6: launch_rk<6>(s, L); break;
    case 7: launch_rk<7>(s, L); break;
    case 8: launch_rk<8>(s, L); break;
    case 9: launch_rk<9>(s, L); break;
    case 10: launch_rk<10>(s, L); break;
    case 11: launch_rk<11>(s, L); break;
    case 12: launch_rk<12>(s, L); break;
    case 13: launch_rk<13>(s, L); break;
    case 14: launch_rk<14>(s, L); break;
    case 15: launch_rk<15>(s, L); break;
    case 16: launch_rk<16>(s, L); break;
    default: return;
  }
  // the queue starts empty for the next search
  (void)hipMemsetAsync(&L.ctr->n_rchk, 0, sizeof(unsigned long long), s);
}

}  // namespace hpc
