// Correlation matrix of the observations (replaces gpuMeans / gpuSD /
// gpuPMCC, corHelper.cu:8-116, driven by gpuPMCC in corOwn.cu:10-48), for
// every active instance of the layout at once.
//
// Summation order is the reference's: 64 lanes, lane t owning samples
// k = t, t+64, ..., accumulated in order with fma (means: plain adds), then a
// halving tree (t += t + 32, ..., t += t + 1).  One 64-wide wavefront holds
// exactly those 64 partial sums, so the device result is bit-identical to a
// CPU evaluation of the same order.
#pragma clang fp contract(off)

#include <algorithm>

#include "hpc_internal.hpp"

namespace hpc {
namespace {

__device__ __forceinline__ double tree64(double v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

// All 64 accumulators of a wavefront at once, with tree64's additions: v[a]
// holds lane t's partial sum of accumulator a; on return v[0] of lane t holds
// accumulator t's total.  Each stage s = 32, 16, ..., 1 adds the partial sums
// of lanes t and t + s of an accumulator, as tree64 does, but in whichever of
// the two lanes keeps that accumulator (bit s of the lane picks the upper or
// lower half of the values left); addition is commutative, so the sum has the
// same bits.  63 exchanges per lane instead of tree64's 6 per accumulator.
// One stage: 2 H values per lane in, H out; the partner is lane ^ S.
template <int H, int S>
__device__ __forceinline__ void halve(double *v, int lane) {
  const bool up = (lane & S) != 0;
#pragma unroll
  for (int a = 0; a < H; ++a) {
    const double lo = v[a], hi = v[a + H];
    const double keep = up ? hi : lo;
    const double send = up ? lo : hi;
    v[a] = keep + __shfl_xor(send, S, 64);
  }
}

__device__ __forceinline__ void tree64_all(double (&v)[64], int lane) {
  halve<32, 32>(v, lane);
  halve<16, 16>(v, lane);
  halve<8, 8>(v, lane);
  halve<4, 4>(v, lane);
  halve<2, 2>(v, lane);
  halve<1, 1>(v, lane);
}

// The same for 16 accumulators: stages 32, 16, 8, 4 halve the values kept,
// stages 2 and 1 add the one left in both lanes of a pair.  Lane t ends with
// the total of accumulator (t >> 2) & 15 (its four lanes t & ~3 .. t | 3 all
// hold it).
__device__ __forceinline__ void tree64_16(double (&v)[16], int lane) {
  halve<8, 32>(v, lane);
  halve<4, 16>(v, lane);
  halve<2, 8>(v, lane);
  halve<1, 4>(v, lane);
  v[0] = v[0] + __shfl_xor(v[0], 2, 64);
  v[0] = v[0] + __shfl_xor(v[0], 1, 64);
}

__device__ __forceinline__ int tree64i(int v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = v + __shfl_down(v, s, 64);
  return v;
}

// k with prefix[k] <= idx < prefix[k+1] (empty instances skipped)
__device__ __forceinline__ int find_inst(const long long *prefix, int K, long long idx) {
  int lo = 0, hi = K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One wavefront per variable: mean over non-NaN samples, sd rounded through
// float (corHelper.cu:75), z-scores (x - mean) / sd padded with zeros to npad,
// NaN mask.  A variable with NaN cells flags its instance (ictr.has_nan),
// which selects the masked pairwise kernel for it.  With J > 0 (every
// instance's n <= 64 J) lane t loads its samples t, t + 64, ... once into
// registers and the three passes run from there; J = 0 re-reads the row for
// each pass.  Both do the same operations in the same order.
template <int J>
__global__ __launch_bounds__(256) void moments_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= L.p_total) return;
  const int kin = L.row_inst[g];
  const Inst I = L.inst[kin];
  if (!(I.flags & kActive)) return;
  const int a = g - I.start, n = I.n, npad = I.npad;
  const double *xa = I.x + (size_t)a * n;
  double xv[J > 0 ? J : 1];
  if constexpr (J > 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k = lane + 64 * j;
      xv[j] = k < n ? xa[k] : 0.0;
    }
  }
  double s = 0.0;
  int c = 0;
  if constexpr (J > 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double v = xv[j];
      if (lane + 64 * j < n && !__builtin_isnan(v)) { s = s + v; c += 1; }
    }
  } else {
    for (int k = lane; k < n; k += 64) {
      double v = xa[k];
      if (!__builtin_isnan(v)) { s = s + v; c += 1; }
    }
  }
  s = tree64(s);
  c = tree64i(c);
  if (lane == 0 && c < n) atomicOr(&L.ictr[kin].has_nan, 1);
  const double mu = __shfl(s, 0, 64) / (double)__shfl(c, 0, 64);
  double q = 0.0;
  if constexpr (J > 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double v = xv[j];
      if (lane + 64 * j < n && !__builtin_isnan(v)) {
        double d = v - mu;
        q = __builtin_fma(d, d, q);
      }
    }
  } else {
    for (int k = lane; k < n; k += 64) {
      double v = xa[k];
      if (!__builtin_isnan(v)) {
        double d = v - mu;
        q = __builtin_fma(d, d, q);
      }
    }
  }
  q = tree64(q);
  const float nm1 = (float)n - 1.f;
  const double sd = (double)sqrtf((float)(__shfl(q, 0, 64) / (double)nm1));
  if (lane == 0) {
    I.mean[a] = mu;
    I.sd[a] = sd;
  }
  double *za = I.z + (size_t)a * npad;
  uint8_t *na = I.nanm + (size_t)a * npad;
  float *z32 = I.z32 ? I.z32 + (size_t)a * npad : nullptr;
  auto put = [&](int k, double v) {
    const double zk = k < n ? (v - mu) / sd : 0.0;
    za[k] = zk;
    na[k] = (k < n && __builtin_isnan(v)) ? 1 : 0;
    if (z32) z32[k] = (float)zk;
  };
  if constexpr (J > 0) {
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k = lane + 64 * j;
      if (k < npad) put(k, xv[j]);
    }
  } else {
    for (int k = lane; k < npad; k += 64) put(k, k < n ? xa[k] : 0.0);
  }
}

// ---------------------------------------------------------------------------
// Masked form (instances with NaN cells): one wavefront per 4 x 4 block of
// variable pairs of the upper block triangle, mirrored.  Lane t accumulates
// the sum over k = t (mod 64) of z_a[k] * z_b[k] with fma, skipping pairs
// where either raw sample is NaN (corHelper.cu:101).
// ---------------------------------------------------------------------------
constexpr int TB = 4;

// One 4 x 4 block (index `rem` of instance k's upper block triangle) of the
// masked form.
__device__ __forceinline__ void corr_masked_block(Layout L, int k, long long rem, int lane) {
  const Inst I = L.inst[k];
  const int p = I.p, npad = I.npad;
  const int nb = (p + TB - 1) / TB;
  int A = 0;
  while (rem >= nb - A) { rem -= nb - A; ++A; }
  const int B = A + (int)rem;
  const int a0 = A * TB, b0 = B * TB;
  int ia[TB], ib[TB];
#pragma unroll
  for (int u = 0; u < TB; ++u) {
    ia[u] = min(a0 + u, p - 1);
    ib[u] = min(b0 + u, p - 1);
  }
  static_assert(TB * TB == 16, "tree64_16 reduces 16 accumulators");
  double acc[TB * TB];  // acc[TB u + v]: pair (a0 + u, b0 + v)
#pragma unroll
  for (int a = 0; a < TB * TB; ++a) acc[a] = 0.0;
  for (int kk = lane; kk < npad; kk += 64) {
    double za[TB], zb[TB];
    bool na[TB], nbm[TB];
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      za[u] = I.z[(size_t)ia[u] * npad + kk];
      na[u] = I.nanm[(size_t)ia[u] * npad + kk] != 0;
    }
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      zb[v] = I.z[(size_t)ib[v] * npad + kk];
      nbm[v] = I.nanm[(size_t)ib[v] * npad + kk] != 0;
    }
#pragma unroll
    for (int u = 0; u < TB; ++u)
#pragma unroll
      for (int v = 0; v < TB; ++v)
        if (!(na[u] || nbm[v])) acc[TB * u + v] = __builtin_fma(za[u], zb[v], acc[TB * u + v]);
  }
  const double div = (double)((float)I.n - 1.f);
  tree64_16(acc, lane);
  const int idx = (lane >> 2) & 15;
  const int a = a0 + idx / TB, b = b0 + idx % TB;
  if ((lane & 3) == 0 && a < p && b < p) {
    const double r = acc[0] / div;
    I.cor[(size_t)a * p + b] = r;
    I.cor[(size_t)b * p + a] = r;
  }
}

__global__ __launch_bounds__(256) void corr_canon_kernel(Layout L, const long long *__restrict__ prefix) {
  const int lane = threadIdx.x & 63;
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       w < prefix[L.K]; w += nw) {
    const int k = find_inst(prefix, L.K, w);
    if (!L.ictr[k].has_nan) {
      // NaN-free instances go to the tiled kernel: jump past this one's blocks
      const long long end = prefix[k + 1];
      w += ((end - 1 - w) / nw) * nw;
      continue;
    }
    corr_masked_block(L, k, w - prefix[k], lane);
  }
}

// ---------------------------------------------------------------------------
// Tiled form for NaN-free instances: a 256-thread workgroup owns a 16 x 16
// tile of variable pairs; each wavefront an 8 x 8 sub-tile (64 accumulators
// per lane, lane t = residue class t).  Samples stream through LDS 64 at a
// time in a four-slot ring, so each z element is read from L2 once per
// workgroup.  Tiles are visited in 8 x 8 super-tiles of each instance's
// upper triangle; consecutive logical indices go to one XCD (workgroup id
// mod 8), so every XCD walks a contiguous range of super-tiles and keeps the
// rows it re-reads in its own L2.
// ---------------------------------------------------------------------------
constexpr int CT = 16;  // variables per workgroup tile side
constexpr int ST = 8;   // tiles per super-tile side
constexpr int kXcds = 8;

__global__ __launch_bounds__(256) void corr_tile_kernel(Layout L, const long long *__restrict__ prefix,
                                                        long long per_xcd) {
  __shared__ double sa[4][CT][64];
  __shared__ double sb[4][CT][64];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wa = wave >> 1, wb = wave & 1;
  const long long g = blockIdx.x;
  const long long Lg = (g % kXcds) * per_xcd + g / kXcds;
  if (Lg >= prefix[L.K]) return;
  const int k = find_inst(prefix, L.K, Lg);
  // instances with NaN cells go to the masked kernel, fast ones to the MFMA
  // kernel
  if (L.ictr[k].has_nan || L.inst[k].fast) return;
  const Inst I = L.inst[k];
  const int p = I.p, npad = I.npad;
  const int nt = (p + CT - 1) / CT;
  const int ns = (nt + ST - 1) / ST;
  const long long local = Lg - prefix[k];
  const long long sb_idx = local / (ST * ST);
  const int within = (int)(local % (ST * ST));
  // upper-triangle super-tile (SA <= SB), row-major over SA
  int SA = (int)((2.0 * ns + 1.0 - sqrt((2.0 * ns + 1.0) * (2.0 * ns + 1.0) - 8.0 * (double)sb_idx)) * 0.5);
  if (SA < 0) SA = 0;
  while (SA > 0 && (long long)SA * ns - (long long)SA * (SA - 1) / 2 > sb_idx) --SA;
  while ((long long)(SA + 1) * ns - (long long)(SA + 1) * SA / 2 <= sb_idx) ++SA;
  const int SB = SA + (int)(sb_idx - ((long long)SA * ns - (long long)SA * (SA - 1) / 2));
  const int TA = SA * ST + within / ST, TBt = SB * ST + within % ST;
  if (TA >= nt || TBt >= nt || TA > TBt) return;
  const int a0 = TA * CT, b0 = TBt * CT;

  // loader mapping: 16 rows x 64 samples per operand = 512 chunks of 16 B;
  // thread tid moves chunks tid and tid + 256 (row c / 32, column 2 (c % 32)),
  // so the lanes of one LDS store write consecutive 16 B (no bank conflicts)
  const int lr0 = tid >> 5, lr1 = lr0 + 8, lc = (tid & 31) * 2;
  const double *ga0 = I.z + (size_t)min(a0 + lr0, p - 1) * npad + lc;
  const double *ga1 = I.z + (size_t)min(a0 + lr1, p - 1) * npad + lc;
  const double *gb0 = I.z + (size_t)min(b0 + lr0, p - 1) * npad + lc;
  const double *gb1 = I.z + (size_t)min(b0 + lr1, p - 1) * npad + lc;
  const int nsteps = npad >> 6;
  double acc[64];  // acc[8 u + v]: pair (wa * 8 + u, wb * 8 + v) of the tile
#pragma unroll
  for (int a = 0; a < 64; ++a) acc[a] = 0.0;
  // Step s of the samples lives in LDS slot s % 4.  While a step's FMAs run,
  // the wave's next operands are read from LDS, step s + 3 is stored from
  // registers and the global loads of step s + 4 are in flight; the loop body
  // covers two steps with two sets of load registers.  One barrier per two
  // steps suffices: a slot is rewritten (at step s + 1 for the data of step
  // s + 4) only after the barrier that follows every read of its old data (at
  // step s - 1), and a slot is read (at step s - 1) only after the barrier
  // that follows its store (at step s - 3).
  double za[8], zb[8], zn[8], zm[8];
  auto lds_read = [&](int buf, double (&xa)[8], double (&xb)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) xa[u] = sa[buf][wa * 8 + u][lane];
#pragma unroll
    for (int v = 0; v < 8; ++v) xb[v] = sb[buf][wb * 8 + v][lane];
  };
  auto fmas = [&]() {
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < 8; ++v) acc[8 * u + v] = __builtin_fma(za[u], zb[v], acc[8 * u + v]);
  };
  const int last = nsteps - 1;
#define HPC_LOAD(st, A0, A1, B0, B1)                                       \
  do {                                                                     \
    const size_t o_ = (size_t)(st) * 64;                                   \
    A0 = *reinterpret_cast<const double2 *>(ga0 + o_);                     \
    A1 = *reinterpret_cast<const double2 *>(ga1 + o_);                     \
    B0 = *reinterpret_cast<const double2 *>(gb0 + o_);                     \
    B1 = *reinterpret_cast<const double2 *>(gb1 + o_);                     \
  } while (0)
#define HPC_STORE(buf, A0, A1, B0, B1)                                     \
  do {                                                                     \
    *reinterpret_cast<double2 *>(&sa[buf][lr0][lc]) = A0;                  \
    *reinterpret_cast<double2 *>(&sa[buf][lr1][lc]) = A1;                  \
    *reinterpret_cast<double2 *>(&sb[buf][lr0][lc]) = B0;                  \
    *reinterpret_cast<double2 *>(&sb[buf][lr1][lc]) = B1;                  \
  } while (0)
  // one step: FMAs on (za, zb) = step m while reading step m + 1 into
  // (zn, zm) and storing step m + 3 from the given registers
#define HPC_STEP(m, A0, A1, B0, B1, SYNC)                                  \
  do {                                                                     \
    lds_read(((m) + 1) & 3, zn, zm);                                       \
    fmas();                                                                \
    __builtin_amdgcn_sched_barrier(0);                                     \
    HPC_STORE(((m) + 3) & 3, A0, A1, B0, B1);                              \
    if (SYNC) __syncthreads();                                             \
    _Pragma("unroll") for (int u_ = 0; u_ < 8; ++u_) {                     \
      za[u_] = zn[u_];                                                     \
      zb[u_] = zm[u_];                                                     \
    }                                                                      \
  } while (0)
  double2 pa0, pa1, pb0, pb1, qa0, qa1, qb0, qb1;
  // steps 0 and 1 in flight together, then step 2; step 3 in registers
  HPC_LOAD(0, pa0, pa1, pb0, pb1);
  HPC_LOAD(min(1, last), qa0, qa1, qb0, qb1);
  HPC_STORE(0, pa0, pa1, pb0, pb1);
  HPC_STORE(1, qa0, qa1, qb0, qb1);
  HPC_LOAD(min(2, last), pa0, pa1, pb0, pb1);
  HPC_STORE(2, pa0, pa1, pb0, pb1);
  HPC_LOAD(min(3, last), pa0, pa1, pb0, pb1);
  __syncthreads();
  lds_read(0, za, zb);
  __syncthreads();  // slot 0 is rewritten at step 1
  int m = 0;
  for (; m + 1 < nsteps; m += 2) {
    // the p registers hold step m + 3
    HPC_LOAD(min(m + 4, last), qa0, qa1, qb0, qb1);
    HPC_STEP(m, pa0, pa1, pb0, pb1, false);
    // the q registers hold step m + 4
    HPC_LOAD(min(m + 5, last), pa0, pa1, pb0, pb1);
    HPC_STEP(m + 1, qa0, qa1, qb0, qb1, true);
  }
  if (m < nsteps) fmas();  // odd step count: the last step
#undef HPC_STEP
#undef HPC_LOAD
#undef HPC_STORE
  const double div = (double)((float)I.n - 1.f);
  // lane t: the total of pair t of the wave's 8 x 8 sub-tile
  tree64_all(acc, lane);
  const int a = a0 + wa * 8 + (lane >> 3), b = b0 + wb * 8 + (lane & 7);
  if (a < p && b < p) {
    const double r = acc[0] / div;
    I.cor[(size_t)a * p + b] = r;
    I.cor[(size_t)b * p + a] = r;
  }
}

// ---------------------------------------------------------------------------
// Fast form (HPC_CORR_F32): z-scores rounded to float, 64 x 64 output tiles
// of the upper triangle per workgroup, each wavefront a 32 x 32 quarter as
// 2 x 2 MFMA 16x16x4 FP32 blocks.  Samples stream through LDS 64 at a time;
// the FP32 accumulators are added into FP64 ones after every 64 samples, so
// the error of an entry is at most (64 + 8 + 2) * 2^-24 of sum |z_a z_b| /
// (n - 1) <= 1.01 (the +2: rounding of the two operands).
//
// v_mfma_f32_16x16x4f32 lane layout: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; it receives D[4 * (l / 16) + r][l % 16] in register r.
// ---------------------------------------------------------------------------
constexpr int FT = 64;   // output tile side
constexpr int FK = 64;   // samples per LDS stage and per FP32 accumulation

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void corr_mfma_kernel(Layout L, const long long *__restrict__ prefix,
                                                        long long per_xcd) {
  __shared__ float sa[2][FT][FK + 4];
  __shared__ float sb[2][FT][FK + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const long long g = blockIdx.x;
  const long long Lg = (g % kXcds) * per_xcd + g / kXcds;
  if (Lg >= prefix[L.K]) return;
  const int k = find_inst(prefix, L.K, Lg);
  if (L.ictr[k].has_nan) return;  // the masked kernel takes it
  const Inst I = L.inst[k];
  const int p = I.p, npad = I.npad;
  const int nt = (p + FT - 1) / FT;
  // upper-triangle tile (TA <= TB), row-major over TA
  const long long t = Lg - prefix[k];
  int TA = (int)((2.0 * nt + 1.0 - sqrt((2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * (double)t)) * 0.5);
  if (TA < 0) TA = 0;
  while (TA > 0 && (long long)TA * nt - (long long)TA * (TA - 1) / 2 > t) --TA;
  while ((long long)(TA + 1) * nt - (long long)(TA + 1) * TA / 2 <= t) ++TA;
  const int TBi = TA + (int)(t - ((long long)TA * nt - (long long)TA * (TA - 1) / 2));
  const int a0 = TA * FT, b0 = TBi * FT;
  f4v acc32[2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[u][v][r] = 0.0;
  const int li = lane & 15, lk = lane >> 4;
  // staging: thread tid moves float4 number tid + 256 j (j < 4) of each
  // operand's 64 x 64 chunk: row (tid + 256 j) / 16, column 4 ((tid) % 16)
  const int sr = tid >> 4, sc = (tid & 15) * 4;
  const float *ga[4], *gb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    ga[j] = I.z32 + (size_t)min(a0 + sr + 16 * j, p - 1) * npad + sc;
    gb[j] = I.z32 + (size_t)min(b0 + sr + 16 * j, p - 1) * npad + sc;
  }
  const int nsteps = npad / FK, last = nsteps - 1;
  auto load = [&](int st, f4v (&ra)[4], f4v (&rb)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ra[j] = *reinterpret_cast<const f4v *>(ga[j] + (size_t)st * FK);
      rb[j] = *reinterpret_cast<const f4v *>(gb[j] + (size_t)st * FK);
    }
  };
  auto store = [&](int buf, const f4v (&ra)[4], const f4v (&rb)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      *reinterpret_cast<f4v *>(&sa[buf][sr + 16 * j][sc]) = ra[j];
      *reinterpret_cast<f4v *>(&sb[buf][sr + 16 * j][sc]) = rb[j];
    }
  };
  auto chunk = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v) acc32[u][v] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int kk = 0; kk < FK; kk += 4) {
      const float x0 = sa[buf][wr * 32 + li][kk + lk];
      const float x1 = sa[buf][wr * 32 + 16 + li][kk + lk];
      const float y0 = sb[buf][wc * 32 + li][kk + lk];
      const float y1 = sb[buf][wc * 32 + 16 + li][kk + lk];
      acc32[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, y0, acc32[0][0], 0, 0, 0);
      acc32[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(x0, y1, acc32[0][1], 0, 0, 0);
      acc32[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, y0, acc32[1][0], 0, 0, 0);
      acc32[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(x1, y1, acc32[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[u][v][r] += (double)acc32[u][v][r];
  };
  // chunk s in LDS buffer s & 1, the next one or two in registers (loaded two
  // chunks ahead; the body covers two chunks with two register sets)
  f4v pa[4], pb[4], qa[4], qb[4];
  load(0, pa, pb);
  store(0, pa, pb);
  load(min(1, last), pa, pb);
  __syncthreads();
  int m = 0;
  for (; m + 1 < nsteps; m += 2) {
    load(min(m + 2, last), qa, qb);
    chunk(0);
    __builtin_amdgcn_sched_barrier(0);
    store(1, pa, pb);
    __syncthreads();
    load(min(m + 3, last), pa, pb);
    chunk(1);
    __builtin_amdgcn_sched_barrier(0);
    store(0, qa, qb);
    __syncthreads();
  }
  if (m < nsteps) chunk(0);
  const double div = (double)((float)I.n - 1.f);
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = a0 + wr * 32 + u * 16 + 4 * lk + r;
        const int b = b0 + wc * 32 + v * 16 + li;
        if (a < p && b < p && a <= b) {
          const double c = acc64[u][v][r] / div;
          I.cor[(size_t)a * p + b] = c;
          I.cor[(size_t)b * p + a] = c;
        }
      }
}

long long mfma_tiles(int p) {
  const long long nt = (p + FT - 1) / FT;
  return nt * (nt + 1) / 2;
}

}  // namespace

long long tiled_slots(int p) {
  const int nt = (p + CT - 1) / CT;
  const int ns = (nt + ST - 1) / ST;
  return (long long)ns * (ns + 1) / 2 * ST * ST;
}

long long masked_blocks(int p) {
  const long long nb = (p + TB - 1) / TB;
  return nb * (nb + 1) / 2;
}

void launch_moments(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst) {
  int max_n = 0;
  for (const Inst &I : host_inst)
    if (I.flags & kActive) max_n = std::max(max_n, I.n);
  launch_moments_n(s, L, max_n);
}

void launch_moments_n(hipStream_t s, const Layout &L, int max_n) {
  const int wpb = 4;
  const dim3 grid((L.p_total + wpb - 1) / wpb), block(64 * wpb);
  // the samples in registers when every active instance's row fits
  if (max_n <= 64 * 32)
    hipLaunchKernelGGL(moments_kernel<32>, grid, block, 0, s, L);
  else if (max_n <= 64 * 80)
    hipLaunchKernelGGL(moments_kernel<80>, grid, block, 0, s, L);
  else
    hipLaunchKernelGGL(moments_kernel<0>, grid, block, 0, s, L);
}

void launch_corr(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                 long long *d_prefix, std::vector<long long> &h_prefix) {
  const int K = L.K;
  // two prefix arrays in one upload: tiled slots over the active exact
  // instances, then masked blocks over every active instance; the moments
  // kernel's NaN flag decides which kernel computes an instance (fast ones
  // without NaN cells go to the MFMA kernel)
  h_prefix.assign(2 * (K + 1), 0);
  for (int k = 0; k < K; ++k) {
    const Inst &I = host_inst[k];
    const bool on = I.flags & kActive;
    h_prefix[k + 1] = h_prefix[k] + (on && !I.fast ? tiled_slots(I.p) : 0);
    h_prefix[K + 1 + k + 1] = h_prefix[K + 1 + k] + (on ? masked_blocks(I.p) : 0);
  }
  (void)hipMemcpyAsync(d_prefix, h_prefix.data(), h_prefix.size() * sizeof(long long),
                       hipMemcpyHostToDevice, s);
  launch_corr_sums(s, L, d_prefix, h_prefix[K], h_prefix[2 * K + 1]);
}

void launch_corr_sums(hipStream_t s, const Layout &L, const long long *d_prefix, long long tiled,
                      long long masked) {
  if (tiled > 0) {
    const long long per_xcd = (tiled + kXcds - 1) / kXcds;
    hipLaunchKernelGGL(corr_tile_kernel, dim3((unsigned)(per_xcd * kXcds)), dim3(256), 0, s, L,
                       d_prefix, per_xcd);
  }
  if (masked > 0) {
    const int wpb = 4;
    const long long blocks = std::min<long long>((masked + wpb - 1) / wpb, 4096);
    hipLaunchKernelGGL(corr_canon_kernel, dim3((unsigned)blocks), dim3(64 * wpb), 0, s, L,
                       d_prefix + L.K + 1);
  }
}

void launch_corr_fast(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                      long long *d_prefix, std::vector<long long> &h_prefix) {
  const int K = L.K;
  h_prefix.assign(K + 1, 0);
  for (int k = 0; k < K; ++k) {
    const Inst &I = host_inst[k];
    const bool on = (I.flags & kActive) && I.fast;
    h_prefix[k + 1] = h_prefix[k] + (on ? mfma_tiles(I.p) : 0);
  }
  const long long tiles = h_prefix[K];
  if (tiles == 0) return;
  (void)hipMemcpyAsync(d_prefix, h_prefix.data(), h_prefix.size() * sizeof(long long),
                       hipMemcpyHostToDevice, s);
  const long long per_xcd = (tiles + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(corr_mfma_kernel, dim3((unsigned)(per_xcd * kXcds)), dim3(256), 0, s, L,
                     d_prefix, per_xcd);
}

}  // namespace hpc
