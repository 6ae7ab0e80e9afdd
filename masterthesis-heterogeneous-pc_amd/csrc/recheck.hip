// Exact rechecks for instances whose correlation matrix holds fast (MFMA)
// values.  A fast entry differs from the exact one by at most Inst::eps, so
// a test decides the same on both unless its statistic lies within its error
// bound of the threshold.  Those tests are re-run here on exact correlations,
// each computed in the canonical order of the FP64 correlation kernels
// (lane-strided FMA sums over the z-scores, halving tree), so decisions,
// first passing ranks, sepsets and test counts equal the exact path's.
#pragma clang fp contract(off)

#include "exact_dot.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

constexpr int kRkWaves = 4;    // wavefronts per workgroup, one item each
constexpr int kRkJoint = 6;    // levels whose items load their rows jointly

// Exact re-run of each queued (edge, rank) of a level-LV search whose rank is
// still below the edge's first passing rank; a pass lowers it.  One
// wavefront per item.  The test needs every entry among i, j and the LV
// members of S: up to level kRkJoint they come from one pass over the
// LV + 2 z-score rows (exact_cor_all), above it one pair at a time.
template <int LV>
__global__ __launch_bounds__(64 * kRkWaves) void recheck_kernel(Layout L) {
  constexpr int V = LV + 2, D = V * (V - 1) / 2;
  __shared__ double s_dot[kRkWaves][D];
  __shared__ int s_vid[kRkWaves][V];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long n = min(L.ctr->n_rchk, L.rchk_cap);
  const long long nw = (long long)gridDim.x * kRkWaves;
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(L.fp);
  unsigned long long done = 0;
  for (long long w = (long long)blockIdx.x * kRkWaves + wave; w < (long long)n; w += nw) {
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
    int vid[V];
    vid[0] = i - s0;
    vid[1] = L.nbr[e] - s0;
#pragma unroll
    for (int c = 0; c < LV; ++c) vid[2 + c] = L.nbr[base + (q[c] < y ? q[c] : q[c] + 1)] - s0;
    if constexpr (LV <= kRkJoint) {
      double dots[D];
      exact_cor_all<V>(In, vid, lane, dots);
      if (lane == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d) s_dot[wave][d] = dots[d];
      }
    } else {
      if (lane < V) {
#pragma unroll
        for (int a = 0; a < V; ++a)
          if (a == lane) s_vid[wave][a] = vid[a];
      }
      __builtin_amdgcn_wave_barrier();
      int d = 0;
#pragma unroll 1
      for (int a = 0; a < V; ++a)
#pragma unroll 1
        for (int b = a + 1; b < V; ++b, ++d) {
          const double c = exact_cor(In, s_vid[wave][a], s_vid[wave][b], lane);
          if (lane == 0) s_dot[wave][d] = c;
        }
    }
    ++done;
    if (lane != 0) continue;
    // pair (a, b), a < b, of the vertex list [i, j, S]
    auto dot = [&](int a, int b) { return s_dot[wave][a * V - a * (a + 1) / 2 + b - a - 1]; };
    const double a = dot(0, 1);
    double x[LV], yv[LV], C[LV][LV];
#pragma unroll
    for (int c = 0; c < LV; ++c) {
      x[c] = dot(0, 2 + c);
      yv[c] = dot(1, 2 + c);
    }
#pragma unroll
    for (int c = 0; c < LV; ++c)
#pragma unroll
      for (int r = c + 1; r < LV; ++r) C[r][c] = dot(2 + c, 2 + r);
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
  hipLaunchKernelGGL(recheck_kernel<LV>, dim3(1024), dim3(64 * kRkWaves), 0, s, L);
}

}  // namespace

void launch_recheck(hipStream_t s, int level, const Layout &L, bool clear_queue) {
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
  if (clear_queue) (void)hipMemsetAsync(&L.ctr->n_rchk, 0, sizeof(unsigned long long), s);
}

}  // namespace hpc
