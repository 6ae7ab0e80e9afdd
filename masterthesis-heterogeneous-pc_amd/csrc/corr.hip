// Correlation matrix of the observations (replaces gpuMeans / gpuSD /
// gpuPMCC, corHelper.cu:8-116, driven by gpuPMCC in corOwn.cu:10-48).
//
// Summation order is the reference's: 64 lanes, lane t owning samples
// k = t, t+64, ..., accumulated in order with fma (means: plain adds), then a
// halving tree (t += t + 32, ..., t += t + 1).  One 64-wide wavefront holds
// exactly those 64 partial sums, so the device result is bit-identical to a
// CPU evaluation of the same order.
#pragma clang fp contract(off)

#include "hpc_internal.hpp"

namespace hpc {
namespace {

__device__ __forceinline__ double tree64(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

__device__ __forceinline__ int tree64i(int v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

// One wavefront per variable: mean over non-NaN samples, sd rounded through
// float (corHelper.cu:75), z-scores (x - mean) / sd padded with zeros to npad,
// NaN mask.
__global__ __launch_bounds__(256) void moments_kernel(
    const double *__restrict__ x, int p, int n, int npad,
    double *__restrict__ mean_out, double *__restrict__ sd_out,
    double *__restrict__ z, uint8_t *__restrict__ nanm,
    int32_t *__restrict__ hasnan) {
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (a >= p) return;
  const double *xa = x + (size_t)a * n;
  double s = 0.0;
  int c = 0;
  for (int k = lane; k < n; k += 64) {
    double v = xa[k];
    if (!isnan(v)) { s = s + v; c += 1; }
  }
  s = tree64(s);
  c = tree64i(c);
  const double mu = __shfl(s, 0, 64) / (double)__shfl(c, 0, 64);
  double q = 0.0;
  int anynan = 0;
  for (int k = lane; k < n; k += 64) {
    double v = xa[k];
    if (!isnan(v)) {
      double d = v - mu;
      q = __builtin_fma(d, d, q);
    } else {
      anynan = 1;
    }
  }
  q = tree64(q);
  const float nm1 = (float)n - 1.f;
  const double sd = (double)sqrtf((float)(__shfl(q, 0, 64) / (double)nm1));
  const bool nan_any = __any(anynan);
  if (lane == 0) {
    mean_out[a] = mu;
    sd_out[a] = sd;
    hasnan[a] = nan_any ? 1 : 0;
  }
  double *za = z + (size_t)a * npad;
  for (int k = lane; k < npad; k += 64) {
    double v = k < n ? xa[k] : 0.0;
    za[k] = k < n ? (v - mu) / sd : 0.0;
    if (nanm) nanm[(size_t)a * npad + k] = (k < n && isnan(v)) ? 1 : 0;
  }
}

constexpr int TB = 4;  // pairs per wavefront: TB x TB

// One wavefront per TB x TB block of variable pairs (upper block triangle,
// mirrored).  Lane t accumulates sum over k = t (mod 64) of z_a[k] * z_b[k]
// with fma, skipping pairs where either raw sample is NaN.
__global__ __launch_bounds__(256) void corr_canon_kernel(
    const double *__restrict__ z, const uint8_t *__restrict__ nanm,
    const int32_t *__restrict__ hasnan, int p, int n, int npad, int nb,
    long long nblocks, double *__restrict__ cor) {
  const int lane = threadIdx.x & 63;
  const long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nblocks) return;
  // unrank w into an upper-triangle block pair (A <= B), row-major over A
  long long rem = w;
  int A = 0;
  while (rem >= nb - A) { rem -= nb - A; ++A; }
  const int B = A + (int)rem;
  const int a0 = A * TB, b0 = B * TB;
  int ia[TB], ib[TB];
  bool anynan = false;
#pragma unroll
  for (int u = 0; u < TB; ++u) {
    ia[u] = min(a0 + u, p - 1);
    ib[u] = min(b0 + u, p - 1);
    anynan |= hasnan[ia[u]] != 0 || hasnan[ib[u]] != 0;
  }
  double acc[TB][TB];
#pragma unroll
  for (int u = 0; u < TB; ++u)
#pragma unroll
    for (int v = 0; v < TB; ++v) acc[u][v] = 0.0;
  if (!anynan) {
    for (int k = lane; k < npad; k += 64) {
      double za[TB], zb[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) za[u] = z[(size_t)ia[u] * npad + k];
#pragma unroll
      for (int v = 0; v < TB; ++v) zb[v] = z[(size_t)ib[v] * npad + k];
#pragma unroll
      for (int u = 0; u < TB; ++u)
#pragma unroll
        for (int v = 0; v < TB; ++v) acc[u][v] = __builtin_fma(za[u], zb[v], acc[u][v]);
    }
  } else {
    for (int k = lane; k < npad; k += 64) {
      double za[TB], zb[TB];
      bool na[TB], nbm[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        za[u] = z[(size_t)ia[u] * npad + k];
        na[u] = nanm[(size_t)ia[u] * npad + k] != 0;
      }
#pragma unroll
      for (int v = 0; v < TB; ++v) {
        zb[v] = z[(size_t)ib[v] * npad + k];
        nbm[v] = nanm[(size_t)ib[v] * npad + k] != 0;
      }
#pragma unroll
      for (int u = 0; u < TB; ++u)
#pragma unroll
        for (int v = 0; v < TB; ++v)
          if (!(na[u] || nbm[v])) acc[u][v] = __builtin_fma(za[u], zb[v], acc[u][v]);
    }
  }
  const double div = (double)((float)n - 1.f);
#pragma unroll
  for (int u = 0; u < TB; ++u)
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      double r = tree64(acc[u][v]);
      if (lane == 0 && a0 + u < p && b0 + v < p) {
        r = r / div;
        cor[(size_t)(a0 + u) * p + (b0 + v)] = r;
        cor[(size_t)(b0 + v) * p + (a0 + u)] = r;
      }
    }
}

}  // namespace

void launch_moments(hipStream_t s, const double *x, int p, int n, int npad,
                    double *mean, double *sd, double *z, uint8_t *nanm,
                    int32_t *hasnan) {
  const int wpb = 4;
  dim3 grid((p + wpb - 1) / wpb), block(64 * wpb);
  hipLaunchKernelGGL(moments_kernel, grid, block, 0, s, x, p, n, npad, mean, sd,
                     z, nanm, hasnan);
}

void launch_corr_canon(hipStream_t s, const double *z, const uint8_t *nanm,
                       const int32_t *hasnan, int p, int n, int npad,
                       double *cor) {
  const int nb = (p + TB - 1) / TB;
  const long long nblocks = (long long)nb * (nb + 1) / 2;
  const int wpb = 4;
  dim3 grid((unsigned)((nblocks + wpb - 1) / wpb)), block(64 * wpb);
  hipLaunchKernelGGL(corr_canon_kernel, grid, block, 0, s, z, nanm, hasnan, p,
                     n, npad, nb, nblocks, cor);
}

}  // namespace hpc
