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

// Class-major z-scores (zc, read by the matrix-core kernel).  Lane class t
// holds the samples t, t + 64, ... (m = 0 .. M - 1, M = npad / 64) after
// P0 = MS - M leading zeros, MS = M rounded up to 16 (zc_slots); slot
// m' = m + P0 = 16 g + j, j = 8 h + 4 b + kk (h, b in 0..1, kk in 0..3).
// Rows are stored in blocks of 16: block R holds rows 16 R .. 16 R + 15 as
// [g][t][h][kk][row % 16][b], 256 doubles per (g, t).  For one (g, t) and
// half h, the 64 lanes of a wavefront (lane = 16 kk + row % 16) load 16 bytes
// each from 1 KB of consecutive memory, and get m' = 16 g + 8 h + kk and
// 16 g + 8 h + 4 + kk of their row: the k-th elements of two consecutive
// 4-step MFMA operands of 16 rows.  (One load instruction then touches 8
// whole cache lines; a row-major layout, 16 rows per instruction, measured at
// half the load throughput.)
__host__ __device__ __forceinline__ size_t zc_off(int g, int t, int j, int r) {
  return (size_t)16384 * g + 256 * t + 32 * (4 * (j >> 3) + (j & 3)) + 2 * r + ((j >> 2) & 1);
}

__device__ __forceinline__ size_t zc_index(size_t blk, int p0, int a, int k) {
  const int t = k & 63, mp = (k >> 6) + p0;
  return (size_t)(a >> 4) * blk + zc_off(mp >> 4, t, mp & 15, a & 15);
}

// z-score of row a, sample k, from whichever layout the instance holds
__device__ __forceinline__ double zval(const Inst &I, int a, int k) {
  if (I.z) return I.z[(size_t)a * I.npad + k];
  const int ms = zc_slots(I.npad);
  return I.zc[zc_index(zc_block(I.npad), ms - (I.npad >> 6), a, k)];
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

// Means and standard deviations, one wavefront per variable: mean over
// non-NaN samples, sd rounded through float (corHelper.cu:75); a variable
// with NaN cells flags its instance (ictr.has_nan), which selects the masked
// pairwise kernel for it.  With J > 0 (every instance's n <= 64 J) lane t
// loads its samples t, t + 64, ... once into registers and both passes run
// from there; J = 0 re-reads the row for each pass.  Both do the same
// operations in the same order.
template <int J>
constexpr int stats_waves() { return J > 32 ? 2 : 4; }

template <int J>
__global__ __launch_bounds__(64 * stats_waves<J>()) void stats_kernel(Layout L) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= L.p_total) return;
  const int kin = L.row_inst[g];
  // a copy: the stores below could alias the table for the compiler
  const Inst I = L.inst[kin];
  if (!(I.flags & kActive)) return;
  const int a = g - I.start, n = I.n;
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
}

// 16-byte unit of piece (t, kk, r8) in the z-score kernel's LDS tile: the
// low four bits of (kk & 1, r8) are XORed with t, so the 16 lanes of a
// quarter wave hit distinct bank groups both when the pieces are written (t
// varies) and when a line's pieces are read for the store (kk, r8 vary).
__device__ __forceinline__ int zpiece(int t, int kk, int r8) {
  return t * 32 + (kk >> 1) * 16 + ((((kk & 1) << 3) | r8) ^ (t & 15));
}

// z-scores (x - mean) / sd padded with zeros to npad, in every layout the
// instance holds (canonical FP64, class-major FP64, FP32 copy), and the NaN
// mask (instances with NaN cells only: nothing reads it otherwise).
// Workgroup (b, gh) takes the eight rows 8 (b - grp0) .. + 7 of the instance
// whose 8-row groups start at grp0 <= b, and the class-major slots
// 8 gh .. 8 gh + 7: half h = gh % 2 of slot group gg = gh / 2 (slot m of
// lane class t is sample t + 64 (m - P0), the first P0 slots zero).  Each
// (row, slot) is a 512-byte run of samples, one coalesced load; the
// class-major values go through a 32 KB LDS tile so that the stores write
// whole 128-byte lines (eight rows' 16-byte pieces of one (gg, t, h, kk)),
// where a row on its own writes 16 bytes per line.
__global__ __launch_bounds__(256) void zscore_kernel(Layout L) {
  __shared__ double2 tile[2048];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int b = blockIdx.x, gh = blockIdx.y, gg = gh >> 1, h = gh & 1;
  int k = 0;
  if (L.K > 1) {
    int lo = 0, hi = L.K - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (L.inst[mid].grp0 <= b) lo = mid; else hi = mid - 1;
    }
    k = lo;
  }
  // a copy: the stores below could alias the table for the compiler
  const Inst I = L.inst[k];
  const int p = I.p, n = I.n, npad = I.npad;
  const int a0 = 8 * (b - I.grp0);
  const int ms = zc_slots(npad), p0 = ms - (npad >> 6);
  // (uniform) an inactive instance, rows past the instance, slots past the row
  if (!(I.flags & kActive) || a0 >= p || 8 * gh >= ms) return;
  const bool mask = L.ictr[k].has_nan != 0;  // raised by the statistics kernel
  // wave wv: rows a0 + 2 wv and a0 + 2 wv + 1, the half group's 8 slots each
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const int r8 = 2 * wv + rr, a = a0 + r8;
    if (a >= p) continue;
    const double *xa = I.x + (size_t)a * n;
    const double mu = I.mean[a], sd = I.sd[a];
    double xs[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int m = 8 * gh + i, kx = lane + 64 * (m - p0);
      xs[i] = (m >= p0 && kx < n) ? xa[kx] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int m = 8 * gh + i, kx = lane + 64 * (m - p0);
      const bool in = m >= p0 && kx < n;
      const double zk = in ? (xs[i] - mu) / sd : 0.0;
      if (m >= p0) {  // a sample slot: kx < npad
        if (mask) I.nanm[(size_t)a * npad + kx] = (in && __builtin_isnan(xs[i])) ? 1 : 0;
        if (I.z) I.z[(size_t)a * npad + kx] = zk;
        if (I.z32) I.z32[(size_t)a * npad + kx] = (float)zk;
      }
      // slot 8 h + 4 bb + kk of the group: piece (t, kk), element bb
      double *pc = reinterpret_cast<double *>(&tile[zpiece(lane, i & 3, r8)]);
      pc[i >> 2] = zk;
    }
  }
  if (!I.zc) return;  // (uniform)
  __syncthreads();
  // piece (t, kk, rr) of row a0 + rr goes to zc_off(gg, t, 8 h + kk, a % 16);
  // a wave stores 8 whole lines
  double *blk = I.zc + (size_t)(a0 >> 4) * zc_block(npad) + (size_t)16384 * gg + 128 * h +
                16 * ((a0 >> 3) & 1);
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int P = it * 256 + tid;
    const int t = P >> 5, kk = (P >> 3) & 3, rr = P & 7;
    if (a0 + rr < p)
      *reinterpret_cast<double2 *>(blk + 256 * t + 32 * kk + 2 * rr) = tile[zpiece(t, kk, rr)];
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
      za[u] = zval(I, ia[u], kk);
      na[u] = I.nanm[(size_t)ia[u] * npad + kk] != 0;
    }
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      zb[v] = zval(I, ib[v], kk);
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
  // instances with NaN cells go to the masked kernel, fast ones to the FP32
  // MFMA kernel, those with class-major z-scores to the FP64 MFMA kernel
  if (L.ictr[k].has_nan || L.inst[k].fast || L.inst[k].zc) return;
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
// Exact form on the FP64 matrix cores, for NaN-free instances that hold the
// class-major z-scores.  v_mfma_f64_16x16x4f64 computes every output element
// as four fused multiply-adds in k order, each rounded (checked bit for bit
// against fma chains on the device, ties, subnormals and signed zeros
// included), so one instruction advances the canonical sums of 16 x 16 pairs
// by four consecutive samples of one lane class t: the same fma chain, in the
// same order, as lane t of the tiled kernel.  A class's P0 leading zero steps
// are exact too: fma(0, 0, +0) = +0, and every chain starts at +0.
//
// A 256-thread workgroup owns a 32 x 32 block of pairs as 2 x 2 MFMA tiles
// (loads issued XD steps ahead).  Wave c (0..3) runs the classes t = c + 4 u, which form the
// subtree N(c, 4) of the halving tree (N(t, s) = N(t, 2 s) + N(t + s, 2 s),
// N(t, 64) = class t's sum, the total is N(0, 1)): it visits the leaves depth
// first (u = the 4-bit reversal of 0, 1, ..., 15) and adds finished subtrees
// on a four-entry stack (levels 0 and 1 in registers, 2 and 3 in LDS), so
// each addition is the tree's.  The four subtrees then meet in LDS as
// (N0 + N2) + (N1 + N3).  Operands come straight from L2/HBM into registers;
// no LDS staging and no barrier in the loop.  Two workgroups share a CU, so
// one's start and end overlap the other's matrix work.
//
// v_mfma_f64_16x16x4f64 lane layout: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; it receives D[4 r + l / 16][l % 16] in register r.
// ---------------------------------------------------------------------------
constexpr int XS = 32;  // pairs per block side
#ifndef HPC_XMFMA_XD
#define HPC_XMFMA_XD 2
#endif
constexpr int XD = HPC_XMFMA_XD;  // steps (16 samples of a class each) loaded ahead

typedef double d4v __attribute__((ext_vector_type(4)));

// Valid block v (0 <= v < nb (nb + 1) / 2) of an instance's upper block
// triangle, in super-tile order: XT x XT super-tiles row-major over SA <= SB,
// blocks row-major within each.  Only valid blocks are numbered, so equal
// ranges of v are equal work (the XCDs each take one).  P(SA) counts the
// blocks of super-tile rows before SA, all of them full (XT block rows).
constexpr int XT = 8;

__device__ __forceinline__ void xblock(int nb, long long v, int &BA, int &BB) {
  auto P = [&](long long a) {
    return (long long)XT * (XT + 1) / 2 * a + (long long)XT * (a * (nb - XT) - (long long)XT * a * (a - 1) / 2);
  };
  const int ns = (nb + XT - 1) / XT;
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P(mid) <= v) lo = mid; else hi = mid - 1;
  }
  const int SA = lo, r = min(XT, nb - XT * SA);
  long long o = v - P(SA);
  const int D = r * (r + 1) / 2;
  if (o < D) {  // the diagonal super-tile's upper triangle
    int i = 0;
    while (o >= r - i) { o -= r - i; ++i; }
    BA = XT * SA + i;
    BB = XT * SA + i + (int)o;
    return;
  }
  o -= D;
  const long long k = o / ((long long)XT * r);
  const int SB = SA + 1 + (int)k;
  o -= k * XT * r;
  const int cols = min(XT, nb - XT * SB);
  BA = XT * SA + (int)(o / cols);
  BB = XT * SB + (int)(o % cols);
}

__global__ __launch_bounds__(256, 2) void corr_xmfma_kernel(Layout L, const long long *__restrict__ prefix,
                                                         long long per_xcd) {
  // merge stack levels 2 and 3: wave, level - 2, tile, register, lane (the
  // first also takes the wave's subtree at the end)
  __shared__ double stk[4][2][4][4][64];
  __shared__ double fin[XS][XS + 1];
  const int tid = threadIdx.x, lane = tid & 63, c = tid >> 6;
  const long long g = blockIdx.x;
  const long long Lg = (g % kXcds) * per_xcd + g / kXcds;
  if (Lg >= prefix[L.K]) return;
  const int k = find_inst(prefix, L.K, Lg);
  // the masked kernel takes instances with NaN cells, the tiled one those
  // without class-major z-scores
  if (L.ictr[k].has_nan || L.inst[k].fast || !L.inst[k].zc) return;
  const Inst I = L.inst[k];
  const int p = I.p;
  const int nb = (p + XS - 1) / XS;
  int BA, BB;
  xblock(nb, Lg - prefix[k], BA, BB);
  const int a0 = BA * XS, b0 = BB * XS;

  const int G = zc_slots(I.npad) >> 4;
  // the four 16-row blocks (rows past p are the zero pad rows)
  const size_t zblk = zc_block(I.npad);
  const double *r0 = I.zc + (size_t)(a0 >> 4) * zblk + 2 * lane;
  const double *r1 = r0 + zblk;
  const double *r2 = I.zc + (size_t)(b0 >> 4) * zblk + 2 * lane;
  const double *r3 = r2 + zblk;
  // step s: leaf ci = s / G of the subtree (depth first), group s % G of its
  // class: 16 samples, four MFMA k-steps per tile
  const int nsteps = 16 * G;
  auto offset = [&](int st) -> size_t {
    const int ci = st / G, gg = st - ci * G;
    const int u = ((ci & 1) << 3) | ((ci & 2) << 1) | ((ci >> 1) & 2) | ((ci >> 3) & 1);
    return (size_t)16384 * gg + 256 * (c + 4 * u);
  };
  // slot: half h = 0 of the four row blocks (k-steps 0, 1), then h = 1
  // (k-steps 2, 3); each load is 1 KB of consecutive memory
  double2 buf[XD][8];
  auto load = [&](double2 (&b)[8], int st) {
    const size_t o = offset(min(st, nsteps - 1));
    b[0] = *reinterpret_cast<const double2 *>(r0 + o);
    b[1] = *reinterpret_cast<const double2 *>(r1 + o);
    b[2] = *reinterpret_cast<const double2 *>(r2 + o);
    b[3] = *reinterpret_cast<const double2 *>(r3 + o);
    b[4] = *reinterpret_cast<const double2 *>(r0 + o + 128);
    b[5] = *reinterpret_cast<const double2 *>(r1 + o + 128);
    b[6] = *reinterpret_cast<const double2 *>(r2 + o + 128);
    b[7] = *reinterpret_cast<const double2 *>(r3 + o + 128);
  };
  // slots loaded in order (the barriers keep the scheduler from reordering
  // them), so the loop's wait for slot d is for its own loads only
#pragma unroll
  for (int d = 0; d < XD; ++d) {
    load(buf[d], d);
    __builtin_amdgcn_sched_barrier(0);
  }
  const d4v zero4 = {0.0, 0.0, 0.0, 0.0};
  // acc[2 u + v]: tile (rows a0 + 16 u.., columns b0 + 16 v..)
  d4v acc[4] = {zero4, zero4, zero4, zero4};
  d4v st0[4] = {zero4, zero4, zero4, zero4}, st1[4] = {zero4, zero4, zero4, zero4};
  // step s + d from slot d: the slot's MFMAs in k order, then its reload
  // (ahead by XD steps); the loaded registers are never copied, so each wait
  // is for one slot's loads
  auto step = [&](const int d, const int s) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double2 *b = buf[d] + 4 * h;
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[0].x, b[2].x, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[0].x, b[3].x, acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[1].x, b[2].x, acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[1].x, b[3].x, acc[3], 0, 0, 0);
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[0].y, b[2].y, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[0].y, b[3].y, acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[1].y, b[2].y, acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[1].y, b[3].y, acc[3], 0, 0, 0);
    }
    load(buf[d], s + d + XD);
    const int st = s + d;
    if ((st + 1) % G == 0) {
      // leaf ci done: binary-counter merge, the earlier subtree (from the
      // stack) the left operand, as in the tree
      const int ci = st / G;
      if (!(ci & 1)) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          st0[q] = acc[q];
          acc[q] = zero4;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = st0[q] + acc[q];
        if (!(ci & 2)) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            st1[q] = acc[q];
            acc[q] = zero4;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = st1[q] + acc[q];
          int lvl = 2;
          for (; lvl < 4 && ((ci >> lvl) & 1); ++lvl) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
              for (int r = 0; r < 4; ++r) acc[q][r] = stk[c][lvl - 2][q][r][lane] + acc[q][r];
          }
          if (lvl < 4) {  // else ci == 15: acc holds N(c, 4)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
              for (int r = 0; r < 4; ++r) stk[c][lvl - 2][q][r][lane] = acc[q][r];
              acc[q] = zero4;
            }
          }
        }
      }
    }
  };
  // whole rounds of XD steps, then the last nsteps % XD (nsteps is a
  // multiple of 16) outside the loop, so the loop's waits stay per slot
  const int nmain = nsteps - nsteps % XD;
  for (int s = 0; s < nmain; s += XD) {
#pragma unroll
    for (int d = 0; d < XD; ++d) step(d, s);
  }
  if (nmain < nsteps) {
#pragma unroll
    for (int d = 0; d < XD; ++d)
      if (nmain + d < nsteps) step(d, nmain);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) stk[c][0][q][r][lane] = acc[q][r];
  __syncthreads();
  const double div = (double)((float)I.n - 1.f);
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int e = tid + 256 * h;
    const int q = e >> 8, r = (e >> 6) & 3, l = e & 63;
    const double tot = (stk[0][0][q][r][l] + stk[2][0][q][r][l]) + (stk[1][0][q][r][l] + stk[3][0][q][r][l]);
    fin[16 * (q >> 1) + 4 * r + (l >> 4)][16 * (q & 1) + (l & 15)] = tot / div;
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int e = tid + 256 * h;
    const int x = e >> 5, y = e & 31;
    // row a0 + x, columns b0 + y; and row b0 + x, columns a0 + y
    const int a = a0 + x, b = b0 + y;
    if (a < p && b < p && (BA != BB || a <= b)) I.cor[(size_t)a * p + b] = fin[x][y];
    const int bt = b0 + x, at = a0 + y;
    if (at < p && bt < p && (BA != BB || at <= bt)) I.cor[(size_t)bt * p + at] = fin[y][x];
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

long long xmfma_slots(int p) {
  const long long nb = (p + XS - 1) / XS;
  return nb * (nb + 1) / 2;
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

template <int J>
static void launch_stats_j(hipStream_t s, const Layout &L) {
  const int wpb = stats_waves<J>();
  hipLaunchKernelGGL(stats_kernel<J>, dim3((L.p_total + wpb - 1) / wpb), dim3(64 * wpb), 0, s, L);
}

void launch_moments_n(hipStream_t s, const Layout &L, int max_n) {
  if (max_n <= 0 || L.p_total <= 0) return;
  // the samples in registers when every active instance's row fits
  if (max_n <= 64 * 32)
    launch_stats_j<32>(s, L);
  else if (max_n <= 64 * 80)
    launch_stats_j<80>(s, L);
  else
    launch_stats_j<0>(s, L);
  // every instance has at most ceil(p / 8) <= p / 8 + 1 row groups, and at
  // most zc_slots(npad) / 8 half slot groups; workgroups past either exit
  const long long groups = (long long)L.p_total / 8 + L.K;
  const int npad = (max_n + 63) & ~63;
  const int halves = zc_slots(npad) / 8;
  hipLaunchKernelGGL(zscore_kernel, dim3((unsigned)groups, (unsigned)halves), dim3(256), 0, s, L);
}

void launch_corr(hipStream_t s, const Layout &L, const std::vector<Inst> &host_inst,
                 long long *d_prefix, std::vector<long long> &h_prefix) {
  const int K = L.K;
  // two prefix arrays in one upload: tiled slots over the active exact
  // instances, then masked blocks over every active instance; the moments
  // kernel's NaN flag decides which kernel computes an instance (fast ones
  // without NaN cells go to the MFMA kernel)
  h_prefix.assign(2 * (K + 1), 0);
  bool xm = false, tile = false;
  for (int k = 0; k < K; ++k) {
    const Inst &I = host_inst[k];
    const bool on = I.flags & kActive;
    const bool exact = on && !I.fast;
    // one prefix for both exact kernels, in each instance's own unit; each
    // kernel skips the other's instances
    h_prefix[k + 1] = h_prefix[k] + (exact ? (I.zc ? xmfma_slots(I.p) : tiled_slots(I.p)) : 0);
    h_prefix[K + 1 + k + 1] = h_prefix[K + 1 + k] + (on ? masked_blocks(I.p) : 0);
    xm |= exact && I.zc;
    tile |= exact && !I.zc;
  }
  (void)hipMemcpyAsync(d_prefix, h_prefix.data(), h_prefix.size() * sizeof(long long),
                       hipMemcpyHostToDevice, s);
  launch_corr_sums(s, L, d_prefix, h_prefix[K], h_prefix[2 * K + 1], xm, tile);
}

void launch_corr_sums(hipStream_t s, const Layout &L, const long long *d_prefix, long long tiled,
                      long long masked, bool xmfma, bool tile) {
  if (tiled > 0) {
    const long long per_xcd = (tiled + kXcds - 1) / kXcds;
    if (xmfma)
      hipLaunchKernelGGL(corr_xmfma_kernel, dim3((unsigned)(per_xcd * kXcds)), dim3(256), 0, s, L,
                         d_prefix, per_xcd);
    if (tile)
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
