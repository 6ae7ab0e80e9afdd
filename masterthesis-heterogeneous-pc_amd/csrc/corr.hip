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

// ---------------------------------------------------------------------------
// Tiled form for NaN-free data: a 256-thread workgroup owns a 16 x 16 tile of
// variable pairs; each wavefront an 8 x 8 sub-tile (64 accumulators per lane,
// lane t = residue class t).  Samples stream through LDS 64 at a time, double
// buffered, so each z element is read from L2 once per workgroup.  Tiles are
// visited in 8 x 8 super-tiles of the upper triangle and every XCD (workgroup
// id mod 8) walks a contiguous range of super-tiles, keeping the rows it
// re-reads in its own L2.
// ---------------------------------------------------------------------------
constexpr int CT = 16;  // variables per workgroup tile side
constexpr int ST = 8;   // tiles per super-tile side
constexpr int kXcds = 8;

__global__ __launch_bounds__(256) void corr_tile_kernel(
    const double *__restrict__ z, int p, int n, int npad, int nt, int ns,
    long long per_xcd, double *__restrict__ cor) {
  __shared__ double sa[2][CT][64];
  __shared__ double sb[2][CT][64];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wa = wave >> 1, wb = wave & 1;
  // XCD-aware logical tile index
  const long long g = blockIdx.x;
  const long long L = (g % kXcds) * per_xcd + g / kXcds;
  const long long sb_idx = L / (ST * ST);
  const int within = (int)(L % (ST * ST));
  const long long nsb = (long long)ns * (ns + 1) / 2;
  if (sb_idx >= nsb) return;
  // upper-triangle super-tile (SA <= SB), row-major over SA
  int SA = (int)((2.0 * ns + 1.0 - sqrt((2.0 * ns + 1.0) * (2.0 * ns + 1.0) - 8.0 * (double)sb_idx)) * 0.5);
  if (SA < 0) SA = 0;
  while (SA > 0 && (long long)SA * ns - (long long)SA * (SA - 1) / 2 > sb_idx) --SA;
  while ((long long)(SA + 1) * ns - (long long)(SA + 1) * SA / 2 <= sb_idx) ++SA;
  const int SB = SA + (int)(sb_idx - ((long long)SA * ns - (long long)SA * (SA - 1) / 2));
  const int TA = SA * ST + within / ST, TB = SB * ST + within % ST;
  if (TA >= nt || TB >= nt || TA > TB) return;
  const int a0 = TA * CT, b0 = TB * CT;

  // loader mapping: 16 rows x 64 samples per operand = 256 threads x 4 doubles
  const int lr = tid >> 4, lc = (tid & 15) * 4;
  const double *ga = z + (size_t)min(a0 + lr, p - 1) * npad + lc;
  const double *gb = z + (size_t)min(b0 + lr, p - 1) * npad + lc;
  const int nsteps = npad >> 6;

  double acc[8][8];
#pragma unroll
  for (int u = 0; u < 8; ++u)
#pragma unroll
    for (int v = 0; v < 8; ++v) acc[u][v] = 0.0;

  double4 pa = *reinterpret_cast<const double4 *>(ga);
  double4 pb = *reinterpret_cast<const double4 *>(gb);
  *reinterpret_cast<double4 *>(&sa[0][lr][lc]) = pa;
  *reinterpret_cast<double4 *>(&sb[0][lr][lc]) = pb;
  __syncthreads();
  for (int m = 0; m < nsteps; ++m) {
    const int cur = m & 1;
    if (m + 1 < nsteps) {
      pa = *reinterpret_cast<const double4 *>(ga + (size_t)(m + 1) * 64);
      pb = *reinterpret_cast<const double4 *>(gb + (size_t)(m + 1) * 64);
    }
    double za[8], zb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) za[u] = sa[cur][wa * 8 + u][lane];
#pragma unroll
    for (int v = 0; v < 8; ++v) zb[v] = sb[cur][wb * 8 + v][lane];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < 8; ++v) acc[u][v] = __builtin_fma(za[u], zb[v], acc[u][v]);
    if (m + 1 < nsteps) {
      *reinterpret_cast<double4 *>(&sa[cur ^ 1][lr][lc]) = pa;
      *reinterpret_cast<double4 *>(&sb[cur ^ 1][lr][lc]) = pb;
    }
    __syncthreads();
  }
  const double div = (double)((float)n - 1.f);
#pragma unroll
  for (int u = 0; u < 8; ++u)
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      double r = tree64(acc[u][v]);
      const int a = a0 + wa * 8 + u, b = b0 + wb * 8 + v;
      if (lane == 0 && a < p && b < p) {
        r = r / div;
        cor[(size_t)a * p + b] = r;
        cor[(size_t)b * p + a] = r;
      }
    }
}

}  // namespace

void launch_corr_tiled(hipStream_t s, const double *z, int p, int n, int npad,
                       double *cor) {
  const int nt = (p + CT - 1) / CT;
  const int ns = (nt + ST - 1) / ST;
  const long long nsb = (long long)ns * (ns + 1) / 2;
  const long long total = nsb * ST * ST;
  const long long per_xcd = (total + kXcds - 1) / kXcds;
  dim3 grid((unsigned)(per_xcd * kXcds)), block(256);
  hipLaunchKernelGGL(corr_tile_kernel, grid, block, 0, s, z, p, n, npad, nt, ns,
                     per_xcd, cor);
}

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
