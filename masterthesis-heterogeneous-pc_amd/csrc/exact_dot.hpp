// Exact correlation entries recomputed by one wavefront from the z-scores, in
// the canonical order of the FP64 correlation kernels: lane l sums
// z_a[k] z_b[k] for k = l, l + 64, ... with one fma per sample, then the 64
// partial sums meet in a halving tree, and the total is divided by n - 1.
// fma(x, y, c) == fma(y, x, c), so cor[a][b] and cor[b][a] are the same bits.
// Used where a correlation matrix holds fast (MFMA) values: such instances
// have no NaN cells, so no sample is masked.
#pragma once
#pragma clang fp contract(off)

#include "hpc_internal.hpp"

namespace hpc {

__device__ __forceinline__ double tree64(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

__device__ __forceinline__ double cor_divisor(const Inst &In) {
  return (double)((float)In.n - 1.f);
}

// cor[a][b] of instance In (local ids), on every lane.  The sample loop is
// unrolled so that eight loads per operand are in flight.
__device__ __forceinline__ double exact_cor(const Inst &In, int a, int b, int lane) {
  const double *za = In.z + (size_t)a * In.npad;
  const double *zb = In.z + (size_t)b * In.npad;
  double acc = 0.0;
#pragma unroll 8
  for (int k = lane; k < In.npad; k += 64) acc = __builtin_fma(za[k], zb[k], acc);
  acc = tree64(acc);
  return __shfl(acc, 0, 64) / cor_divisor(In);
}

// Every pairwise entry of the V vertices v[0..V-1] at once: each sample of
// the V rows is loaded once and feeds all V (V - 1) / 2 running sums, each
// still in its own canonical order.  out[d] for pair (a, b), a < b, at
// d = a V - a (a + 1) / 2 + b - a - 1 (row-major over the upper triangle);
// the values are valid on lane 0.
template <int V>
__device__ __forceinline__ void exact_cor_all(const Inst &In, const int (&v)[V], int lane,
                                              double (&out)[V * (V - 1) / 2]) {
  constexpr int D = V * (V - 1) / 2;
  const double *zr[V];
#pragma unroll
  for (int a = 0; a < V; ++a) zr[a] = In.z + (size_t)v[a] * In.npad;
  double acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.0;
#pragma unroll 4
  for (int k = lane; k < In.npad; k += 64) {
    double zv[V];
#pragma unroll
    for (int a = 0; a < V; ++a) zv[a] = zr[a][k];
    int d = 0;
#pragma unroll
    for (int a = 0; a < V; ++a)
#pragma unroll
      for (int b = a + 1; b < V; ++b, ++d) acc[d] = __builtin_fma(zv[a], zv[b], acc[d]);
  }
  const double div = cor_divisor(In);
#pragma unroll
  for (int d = 0; d < D; ++d) out[d] = tree64(acc[d]) / div;
}

}  // namespace hpc
