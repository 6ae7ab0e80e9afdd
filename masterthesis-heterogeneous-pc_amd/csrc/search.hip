// Level >= 1 CI tests (replaces testRowL1 / testRowLN<> of gpuRowTests.cu:29-205
// and the workstealing variants of gpuWorkstealingTests.cu) and the level-end
// edge removal, for every active instance of the layout at once.
//
// Search: for every ordered edge (i, j) of the level-start snapshot, the
// lowest rank whose conditioning set S (an l-subset of N(i)\{j} in
// lexicographic order of positions, cuda_util.cu:34-50 and
// gpuRowTests.cu:128-142) gives p >= alpha.  A 64-wide wavefront tests a
// chunk of consecutive ranks at a time and a ballot picks the lowest passing
// one.  Level 1: a wavefront takes four edges, ranks 0..15 of each first,
// then 64 at a time for the edges still open.  Levels >= 2 cut each edge's
// ranks into tasks of `ranks_per_task` (a multiple of 64) so that edges with
// many sets spread over many wavefronts; within a task each lane tests a run
// of consecutive ranks.  A task skips ranks above a pass another task of its
// edge already found, and the lowest rank wins through atomicMin.  Edges do
// not cut each other's searches short, so the result is independent of
// scheduling.
//
// Commit: edge i-j (i < j) is removed when either side found a set; the
// sepset is the lower-id row's lowest passing set, else the higher-id row's.
// It also counts the canonical tests (per ordered edge: first passing rank +
// 1, else all its sets), which depend only on the first passing ranks.
//
// Neighbour lists hold global variable ids; an instance's dense blocks are
// indexed with local ids (global - start).
#pragma clang fp contract(off)

#include <algorithm>

#include "hpc_internal.hpp"
#include "pc_math.hpp"

namespace hpc {
namespace {

constexpr int kWavesPerBlock = 4;
// consecutive ranks per lane in the LDS-staged level >= 2 search: longer
// runs share one unrank among more tests (a task shorter than a chunk
// shortens them, so its ranks still spread over every lane)
template <int LV>
constexpr int run_len() { return LV >= 4 ? 8 : 4; }
// wavefronts per workgroup of the removal kernel (one record reservation and
// one set of counter atomics per workgroup).  Eight spread a short level's
// few thousand edges over twice the CUs of sixteen: the p = 1000 run without
// a level cap (levels 5..15 remove from a few thousand edges each) took
// 1.09 ms with sixteen, 0.99 ms with eight; four cost the p = 2000 levels
// 0..4 5% (more workgroups for the last one to wait on).
constexpr int kCommitWaves = 8;

// Binomial table staged in LDS, addressed as LDS (32-bit addresses, ds_read)
// rather than through generic pointers.
typedef __attribute__((address_space(3))) const unsigned long long lds_u64;

// unrank() (the rank-th L-subset of {0..nn-1} in lexicographic order, L >= 2)
// on a table in LDS with kcols columns: a binary search per member but the
// last, which is the remaining rank past the previous member.
template <int L>
__device__ __forceinline__ void unrank_lds(const lds_u64 *tab, int kcols, unsigned long long rank,
                                           int nn, int (&q)[L]) {
  int v0 = 0;
#pragma unroll
  for (int i = 0; i < L - 1; ++i) {
    const int k = L - i;
    const unsigned long long top = tab[(nn - v0) * kcols + k];
    const unsigned long long X = top - rank;
    int lo = k - 1, hi = nn - 1 - v0;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[mid * kcols + k] < X) lo = mid; else hi = mid - 1;
    }
    const int v = nn - 1 - lo;
    q[i] = v;
    rank = rank + tab[(lo + 1) * kcols + k] - top;
    v0 = v + 1;
  }
  q[L - 1] = v0 + (int)rank;
}

// Uncertain (edge, rank) pairs of a fast instance, one per lane where `unc`
// holds, appended to the recheck queue.  Wave-uniform call.
__device__ __forceinline__ void queue_rank(const Layout &L, bool unc, long long e,
                                           unsigned long long rank, int lane) {
  const unsigned long long m = __ballot(unc);
  if (!m) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(&L.ctr->n_rchk, (unsigned long long)__popcll(m));
  base = __shfl(base, 0, 64);
  if (unc) {
    const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1ull));
    if (slot < L.rchk_cap) L.rchk[slot] = Recheck{(unsigned long long)e, rank};
    else L.ctr->overflow = 1;
  }
}

// Lanes of a fast instance's test chunk whose decision is uncertain and whose
// rank lies below the chunk's lowest certain pass (`certain_passes`) are
// queued for an exact re-run.  Wave-uniform call.
__device__ __forceinline__ void queue_uncertain(const Layout &L, bool unc,
                                                unsigned long long certain_passes, long long e,
                                                unsigned long long r0, int lane) {
  unsigned long long m = __ballot(unc);
  if (certain_passes) m &= (1ull << (__ffsll((long long)certain_passes) - 1)) - 1ull;
  if (!m) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(&L.ctr->n_rchk, (unsigned long long)__popcll(m));
  base = __shfl(base, 0, 64);
  if ((m >> lane) & 1ull) {
    const unsigned long long slot = base + __popcll(m & ((1ull << lane) - 1ull));
    if (slot < L.rchk_cap) L.rchk[slot] = Recheck{(unsigned long long)e, r0 + lane};
    else L.ctr->overflow = 1;
  }
}

// Level 1 (testRowL1, gpuRowTests.cu:29-100; cpuRowTests.cpp:23-56) for the
// rows longer than `cap` neighbours, whose ids and correlations do not fit
// the LDS-staged kernel below: the operands are gathered from the
// correlation matrix.  Workgroups take blocks of `epb` consecutive ordered
// edges (grid-stride); a block's segment on a row that the staged kernel
// takes is skipped whole, and the four wavefronts share the edges of a long
// row's segment, one edge at a time.  Same tests, order and arithmetic as
// the staged kernel.
template <bool FAST>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_l1_kernel(
    Layout L, long long n_edges, const int32_t *__restrict__ row_mask, int cap, int epb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the snapshot's own size bounds the count a launch was sized with
  n_edges = min(n_edges, (long long)L.rowptr[L.p_total]);
  for (long long b0 = (long long)blockIdx.x * epb; b0 < n_edges;
       b0 += (long long)gridDim.x * epb) {
    const long long b1 = min(n_edges, b0 + epb);
    long long e0 = b0;
    while (e0 < b1) {
      const int i = L.erow[e0];
      const int base = L.rowptr[i];
      const int deg = L.rowptr[i + 1] - base;
      const long long seg_end = min(b1, (long long)base + deg);
      if (seg_end <= e0) break;  // (a consistent snapshot always advances)
      const Inst &In = L.inst[L.row_inst[i]];
      // uniform across the workgroup
      if (deg > cap && In.level == 1 && (!row_mask || row_mask[i])) {
        const int nn = deg - 1;
        const int p = In.p, s0 = In.start;
        const double *cor = In.cor;
        const double r_crit = In.r_crit;
        const bool fast = FAST && In.fast;
        const bool chk = In.flags & kCorNaN;  // inputs may be NaN
        const int li = i - s0;
        for (long long e = e0 + wave; e < seg_end; e += kWavesPerBlock) {
          const int y = (int)(e - base);
          const int lj = L.nbr[e] - s0;
          const double x1 = cor[(size_t)li * p + lj];
          long long found = -1;
          for (int q0 = 0; q0 < nn; q0 += 64) {
            const int q = q0 + lane;
            bool pass = false, unc = false;
            if (q < nn) {
              const int lk = L.nbr[base + (q < y ? q : q + 1)] - s0;
              const double x2 = cor[(size_t)li * p + lk];
              const double x3 = cor[(size_t)lj * p + lk];
              if (!(chk && (__builtin_isnan(x1) || __builtin_isnan(x2) || __builtin_isnan(x3)))) {
                const double r = pcor_l1(x1, x2, x3);
                pass = removes(r, r_crit);
                if (fast) {
                  const double xa[1] = {x2}, ya[1] = {x3}, Cz[1][1] = {{1.0}};
                  unc = !fast_certain(r, r_crit, pcor_sensitivity<1>(x1, xa, ya, Cz), In.eps);
                  pass &= !unc;
                }
              }
            }
            const unsigned long long m = __ballot(pass);
            if (fast) queue_uncertain(L, unc, m, e, (unsigned long long)q0, lane);
            if (m) { found = q0 + __ffsll((long long)m) - 1; break; }
          }
          if (lane == 0) L.fp[e] = found;
        }
      }
      e0 = seg_end;
    }
  }
}

// Level 1 with row i's correlations staged in LDS: copies ids and
// cor[i][N(i)] of each row a workgroup meets into LDS and gathers only
// cor[j][k] per test.  Same tests, order and arithmetic as search_l1_kernel.
//
// Every test of edge (i, j) gathers from row j, so the edges are split by the
// class of j (local id mod 8) over the eight XCDs: workgroup b (on XCD b mod
// 8 under round-robin dispatch) takes the edges of class b mod 8 among the
// ordered edges [c * 8 epb, (c + 1) * 8 epb), c = b / 8 (CSR order, grouped
// by row).  An XCD then gathers from an eighth of the rows, which its L2 can
// hold, instead of from all of them.
constexpr int kL1Classes = 8;

//
// Rows longer than `cap` neighbours (the staged capacity) are left to
// search_l1_kernel.
template <bool FAST>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_l1_lds_kernel(
    Layout L, long long n_edges, const int32_t *__restrict__ row_mask, int epb, int cap) {
  extern __shared__ unsigned long long smem[];
  double *Mi = reinterpret_cast<double *>(smem);                 // cap
  int *ids = reinterpret_cast<int *>(Mi + cap);                   // cap
  int *sel = ids + cap;                                           // cap: the class's edges
  __shared__ int s_nsel;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cls = blockIdx.x % kL1Classes;
  const long long span = (long long)kL1Classes * epb;
  // the snapshot's own size bounds the count a launch was sized with
  n_edges = min(n_edges, (long long)L.rowptr[L.p_total]);
  long long e0 = (long long)(blockIdx.x / kL1Classes) * span;
  const long long e1 = min(n_edges, e0 + span);
  while (e0 < e1) {
    const int i = L.erow[e0];
    const int base = L.rowptr[i];
    const int deg = L.rowptr[i + 1] - base;
    const long long seg_end = min(e1, (long long)base + deg);
    if (seg_end <= e0) break;  // (a consistent snapshot always advances)
    const Inst &In = L.inst[L.row_inst[i]];
    // uniform across the workgroup: the row is skipped or staged by all
    if (In.level == 1 && (!row_mask || row_mask[i]) && deg >= 2 && deg <= cap) {
      const int p = In.p, s0 = In.start, li = i - s0;
      __syncthreads();  // the previous row is done with the staged values
      for (int a = tid; a < deg; a += blockDim.x) {
        const int lk = L.nbr[base + a] - s0;
        ids[a] = lk;
        Mi[a] = In.cor[(size_t)li * p + lk];
      }
      __syncthreads();
      // positions y (edges base + y) of the segment whose neighbour is in
      // this workgroup's class, in order
      if (wave == 0) {
        const int y_lo = (int)(e0 - base), y_hi = (int)(seg_end - base);
        int cnt = 0;
        for (int y0 = y_lo; y0 < y_hi; y0 += 64) {
          const int y = y0 + lane;
          const bool mine = y < y_hi && (ids[y] % kL1Classes) == cls;
          const unsigned long long m = __ballot(mine);
          if (mine) sel[cnt + __popcll(m & ((1ull << lane) - 1ull))] = y;
          cnt += __popcll(m);
        }
        if (lane == 0) s_nsel = cnt;
      }
      __syncthreads();
      const int nsel = s_nsel;
      const double r_crit = In.r_crit;
      const int nn = deg - 1;
      const bool fast = FAST && In.fast;
      const bool chk = In.flags & kCorNaN;  // inputs may be NaN
      // A wave takes 4 edges: first ranks 0..15 of all 4 (16 lanes each),
      // then 64-rank chunks of each edge that has not passed yet, so an
      // early pass costs 16 gathers instead of 64.
      auto test = [&](int y, double x1, const double *cj, int q, bool &unc) -> bool {
        unc = false;
        const int pos = q < y ? q : q + 1;
        const double x2 = Mi[pos];
        const double x3 = cj[ids[pos]];
        if (chk && (__builtin_isnan(x1) || __builtin_isnan(x2) || __builtin_isnan(x3))) return false;
        const double r = pcor_l1(x1, x2, x3);
        bool pass = removes(r, r_crit);
        if (fast) {
          const double xa[1] = {x2}, ya[1] = {x3}, Cz[1][1] = {{1.0}};
          unc = !fast_certain(r, r_crit, pcor_sensitivity<1>(x1, xa, ya, Cz), In.eps);
          pass &= !unc;
        }
        return pass;
      };
      for (int kb = 4 * wave; kb < nsel; kb += 4 * kWavesPerBlock) {
        const int grp = lane >> 4, t = lane & 15;
        const bool live = kb + grp < nsel;
        const int yg = live ? sel[kb + grp] : 0;
        const long long eg = (long long)base + yg;
        const double *cjg = In.cor + (size_t)ids[yg] * p;
        bool pass = false, unc = false;
        if (live && t < nn) pass = test(yg, Mi[yg], cjg, t, unc);
        const unsigned long long m = __ballot(pass);
        const unsigned gm = (unsigned)(m >> (16 * grp)) & 0xFFFFu;
        // this group's uncertain tests below its first certain pass
        if (fast) queue_rank(L, unc && (gm == 0u || t < __ffs(gm) - 1), eg, (unsigned long long)t, lane);
        long long found = gm ? (long long)(__ffs(gm) - 1) : -1;
        if (live && t == 0 && (found >= 0 || nn <= 16)) L.fp[eg] = found;
        // the edges still open, one after the other with 64-rank chunks
        const unsigned long long open = __ballot(live && t == 0 && found < 0 && nn > 16);
#pragma unroll 1
        for (int g = 0; g < 4; ++g) {
          if (!((open >> (16 * g)) & 1ull)) continue;
          const int y = sel[kb + g];
          const long long e = (long long)base + y;
          const double x1 = Mi[y];
          const double *cj = In.cor + (size_t)ids[y] * p;
          long long f = -1;
          for (int q0 = 16; q0 < nn; q0 += 64) {
            const int q = q0 + lane;
            bool ps = false, un = false;
            if (q < nn) ps = test(y, x1, cj, q, un);
            const unsigned long long mm = __ballot(ps);
            if (fast) queue_uncertain(L, un, mm, e, (unsigned long long)q0, lane);
            if (mm) { f = q0 + __ffsll((long long)mm) - 1; break; }
          }
          if (lane == 0) L.fp[e] = f;
        }
      }
    }
    e0 = seg_end;
  }
}

// The level-1 decision of (x1, x2, x3) = (r_ij, r_ik, r_jk), as removes(
// pcor_l1(...)) decides it: exact instances through the squared filter
// (removes_sq, the same quotient when it cannot decide), fast ones through r
// and the certainty test of their recheck.
template <bool FAST>
__device__ __forceinline__ bool l1_decide(const Inst &In, bool fast, bool chk, double r_crit,
                                          double lo, double hi, double x1, double x2, double x3,
                                          bool &unc) {
  unc = false;
  if (chk && (__builtin_isnan(x1) || __builtin_isnan(x2) || __builtin_isnan(x3))) return false;
  if (!FAST || !fast) {
    // pcor_l1's numerator and radicand, in its order of operations
    const double s01 = x1 - x2 * x3;
    const double q = (1.0 - x3 * x3) * (1.0 - x2 * x2);
    if (q > 0.0 && __builtin_isfinite(s01)) return removes_sq(s01, q, r_crit, lo, hi);
    return removes(s01 / __builtin_sqrt(q), r_crit);
  }
  const double r = pcor_l1(x1, x2, x3);
  const double xa[1] = {x2}, ya[1] = {x3}, Cz[1][1] = {{1.0}};
  unc = !fast_certain(r, r_crit, pcor_sensitivity<1>(x1, xa, ya, Cz), In.eps);
  return removes(r, r_crit) && !unc;
}

// Level 1 by pivot row.  The tests of ordered edge (i, j) read cor[j][k] for
// every k of N(i)\{j}, scattered over row j, and cor[i][k], contiguous in the
// compacted correlations (ccor, written with the neighbour lists).  A
// workgroup stages whole correlation rows j in LDS and runs the tests of the
// edges (i, j) of the neighbours i of j: workgroup block b takes the CSR
// edges (j, i) [b epb, (b + 1) epb), grouped by pivot row j, and the XCDs each
// take a contiguous eighth of the blocks, so a pivot row's workgroups share an
// L2.  N(i) and ccor (a few MB in all) stay in L2; HBM sees about one read of
// each correlation row.  Same tests, ranks and decisions as search_l1_kernel.
//
// Ranks skip j: position pos of N(i) has rank pos - (pos > y), y the position
// of j, which is the number of neighbours of i below j in the bit adjacency
// (the level-start snapshot while the search runs).  A wave takes four edges,
// positions 0..15 of each first (16 lanes per edge), then 64-position chunks
// of each edge that has not passed yet.
template <bool FAST>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_l1_pivot_kernel(
    Layout L, long long n_edges, const int32_t *__restrict__ row_mask, int epb, long long per_xcd) {
  extern __shared__ unsigned long long smem[];
  double *Rj = reinterpret_cast<double *>(smem);  // [p]: row j of its instance
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int grp = lane >> 4, q = lane & 15;
  const long long g = blockIdx.x;
  const long long blk = (g % kL1Classes) * per_xcd + g / kL1Classes;
  // the snapshot's own size bounds the count a launch was sized with
  n_edges = min(n_edges, (long long)L.rowptr[L.p_total]);
  long long e0 = blk * epb;
  const long long e1 = min(n_edges, e0 + epb);
  while (e0 < e1) {
    const int j = L.erow[e0];
    const int bj = L.rowptr[j];
    const long long seg_end = min(e1, (long long)L.rowptr[j + 1]);
    if (seg_end <= e0) break;  // (a consistent snapshot always advances)
    const Inst &In = L.inst[L.row_inst[j]];
    // uniform across the workgroup: the pivot row is staged or skipped by all
    if (In.level == 1) {
      const int p = In.p, s0 = In.start, lj = j - s0, W = In.W;
      const double *crow = In.cor + (size_t)lj * p;
      __syncthreads();  // the previous pivot's tests are done with Rj
      for (int a0 = 0; a0 < p; a0 += 8 * 64 * kWavesPerBlock) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int a = a0 + u * 64 * kWavesPerBlock + tid;
          v[u] = a < p ? crow[a] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int a = a0 + u * 64 * kWavesPerBlock + tid;
          if (a < p) Rj[a] = v[u];
        }
      }
      __syncthreads();
      const double r_crit = In.r_crit, lo = In.r2_lo, hi = In.r2_hi;
      const bool fast = FAST && In.fast;
      const bool chk = In.flags & kCorNaN;  // inputs may be NaN
      const int wj = lj >> 6;
      const unsigned long long below = (1ull << (lj & 63)) - 1ull;
      const int t_lo = (int)(e0 - bj), t_hi = (int)(seg_end - bj);
      for (int tb = t_lo + 4 * wave; tb < t_hi; tb += 4 * kWavesPerBlock) {
        const int t = tb + grp;
        bool live = t < t_hi;
        const int i = live ? L.nbr[bj + t] : j;
        if (live && row_mask && !row_mask[i]) live = false;
        int bi = 0, degi = 0, c = 0;
        double x1 = 0.0;
        if (live) {
          const int li = i - s0;
          bi = L.rowptr[i];
          degi = L.rowptr[i + 1] - bi;
          x1 = In.cor[(size_t)li * p + lj];
          const unsigned long long *arow = In.adj + (size_t)li * W;
          for (int w = q; w <= wj; w += 16) c += __popcll(w < wj ? arow[w] : arow[w] & below);
        }
#pragma unroll
        for (int sh = 8; sh > 0; sh >>= 1) c += __shfl_xor(c, sh, 64);
        const int y = c;  // j's position in N(i)
        live = live && degi >= 2 && y < degi;
        const long long eg = (long long)bi + y;
        if (live && q == 0) L.erev[eg] = bj + t;  // (i, j) and its reverse (j, i)
        bool pass = false, unc = false;
        if (live && q < degi && q != y) {
          const int k = L.nbr[bi + q];
          pass = l1_decide<FAST>(In, fast, chk, r_crit, lo, hi, x1, L.ccor[bi + q], Rj[k - s0], unc);
        }
        const unsigned long long m = __ballot(pass);
        const unsigned gm = (unsigned)(m >> (16 * grp)) & 0xFFFFu;
        // this group's uncertain tests below its first certain pass
        if (FAST && fast)
          queue_rank(L, unc && (gm == 0u || q < __ffs(gm) - 1), eg,
                     (unsigned long long)(q - (q > y ? 1 : 0)), lane);
        long long found = -1;
        if (gm) {
          const int pf = __ffs(gm) - 1;
          found = pf - (pf > y ? 1 : 0);
        }
        if (live && q == 0 && (found >= 0 || degi <= 16)) L.fp[eg] = found;
        // the edges still open, one after the other with 64-position chunks
        const unsigned long long open = __ballot(live && q == 0 && found < 0 && degi > 16);
#pragma unroll 1
        for (int gg = 0; gg < 4; ++gg) {
          if (!((open >> (16 * gg)) & 1ull)) continue;
          const int b = __shfl(bi, 16 * gg, 64), d = __shfl(degi, 16 * gg, 64);
          const int yy = __shfl(y, 16 * gg, 64);
          const double a1 = __shfl(x1, 16 * gg, 64);
          const long long e = (long long)b + yy;
          long long f = -1;
          for (int q0 = 16; q0 < d; q0 += 64) {
            const int ps = q0 + lane;
            bool ok = false, un = false;
            if (ps < d && ps != yy) {
              const int k = L.nbr[b + ps];
              ok = l1_decide<FAST>(In, fast, chk, r_crit, lo, hi, a1, L.ccor[b + ps], Rj[k - s0], un);
            }
            const unsigned long long mm = __ballot(ok);
            if (FAST && fast)
              queue_rank(L, un && (mm == 0ull || lane < __ffsll((long long)mm) - 1), e,
                         (unsigned long long)(ps - (ps > yy ? 1 : 0)), lane);
            if (mm) {
              const int pf = q0 + __ffsll((long long)mm) - 1;
              f = pf - (pf > yy ? 1 : 0);
              break;
            }
          }
          if (lane == 0) L.fp[e] = f;
        }
      }
    }
    e0 = seg_end;
  }
}

// Task t of a level's plan: row g with task_off[g] <= t < task_off[g] +
// row_tasks[g] (the last row whose exclusive offset is <= t), then ordered
// edge rowptr[g] + local / nch and rank chunk local % nch, where
// nch = row_tasks[g] / deg.  Found by a whole wavefront (every lane takes
// part and gets the same row), given task_off[lo] <= t: with lo > 0 (the row
// of the block's previous segment) a window of the 64 rows from lo first,
// which holds the next row unless 63 or more rows in between have no tasks;
// then 64-ary rounds, one dependent load each (two for a few thousand rows,
// where a binary search takes a dozen).
__device__ __forceinline__ int task_row_wave(const long long *__restrict__ task_off, int rows,
                                             int lo, long long t) {
  const int lane = threadIdx.x & 63;
  if (lo > 0) {
    const int c = lo + lane;
    const bool ok = c < rows && task_off[c] <= t;
    const unsigned long long m = __ballot(ok) | 1ull;
    const int last = 63 - __clzll(m);
    if (last < 63 || lo + 64 >= rows) return lo + last;
    lo += 63;
  }
  int n = rows - lo;  // the row is in [lo, lo + n)
  while (n > 1) {
    const int step = (n + 63) >> 6;
    const bool ok = lane * step < n && task_off[lo + lane * step] <= t;
    const unsigned long long m = __ballot(ok) | 1ull;
    const int nlo = lo + (63 - __clzll(m)) * step;
    n = min(step, lo + n - nlo);
    lo = nlo;
  }
  return lo;
}

// Next block of a task queue for the calling workgroup (every thread gets
// it).  Each workgroup makes exactly one claim past the last block (its
// last), so the claim that returns n_blocks + gridDim.x - 1 is the grid's
// final one and resets the head for the next launch on the stream.
__device__ __forceinline__ long long claim_block(const TaskQueue &Q, long long n_blocks,
                                                 long long &s_blk) {
  __syncthreads();  // every thread has read the previous claim
  if (threadIdx.x == 0) {
    const unsigned long long b = atomicAdd(Q.head, 1ull);
    if (b == (unsigned long long)n_blocks + gridDim.x - 1)
      __hip_atomic_store(Q.head, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_blk = (long long)b;
  }
  __syncthreads();
  return s_blk;
}

// Level L >= 2 (testRowLN<>, gpuRowTests.cu:102-205; the general-level form of
// cpuWorkstealingTests.cpp:39-112) for the rows whose neighbourhood block
// does not fit the LDS-staged kernel below (row length + 1 > dmin), gathering
// the operands from the correlation matrix.  Workgroups take blocks of tpb
// consecutive tasks (grid-stride; the task count is read on the device); a
// block's segment on a row that the staged kernel takes is skipped whole,
// and the four wavefronts share the tasks of a long row's segment.  Task t
// covers ranks [chunk * rpt, (chunk + 1) * rpt) of its edge, 64 at a time,
// one rank per lane.  Same ranks, arithmetic and orientation of every
// correlation read as the staged kernel.
template <int LV, bool FAST>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_ln_kernel(
    Layout L, int rpt, const long long *__restrict__ task_off,
    const long long *__restrict__ row_tasks, int tpb, int dmin, int lds_binom) {
  // C(m, k) for m <= max_deg, k <= LV staged in LDS (columns 0..LV)
  extern __shared__ unsigned long long s_binom[];
  const unsigned long long *tab = L.binom;
  int kc = kBinomK;
  if (lds_binom) {
    const int rows = L.max_deg + 1;
    for (int t = threadIdx.x; t < rows * (LV + 1); t += blockDim.x) {
      const int m = t / (LV + 1), k = t % (LV + 1);
      s_binom[t] = L.binom[(size_t)m * kBinomK + k];
    }
    __syncthreads();
    tab = s_binom;
    kc = LV + 1;
  }
  const long long n_tasks = (long long)L.ctr->n_tasks;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long F = 0;
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(L.fp);
  for (long long blk = blockIdx.x; blk * tpb < n_tasks; blk += gridDim.x) {
    long long t0 = blk * tpb;
    const long long t1 = min(n_tasks, t0 + tpb);
    int prev = 0;  // the row of the block's previous segment
    while (t0 < t1) {
      // the segment of this block's tasks on row i (uniform across the block)
      const int i = task_row_wave(task_off, L.p_total, prev, t0);
      prev = i;
      const long long off = task_off[i];
      const long long seg_end = min(t1, off + row_tasks[i]);
      if (seg_end <= t0) break;  // (a consistent plan always advances)
      const int base = L.rowptr[i];
      const int deg = L.rowptr[i + 1] - base;
      // the staged kernel's row; or one longer than the layout's longest
      // (only in a launch sized for other data, whose results are discarded)
      if (deg + 1 <= dmin || deg > L.max_deg) {
        t0 = seg_end;
        continue;
      }
      const Inst &In = L.inst[L.row_inst[i]];
      const int p = In.p, s0 = In.start;
      const double *cor = In.cor;
      const double r_crit = In.r_crit;
      const bool fast = FAST && In.fast;
      const bool chk = In.flags & kCorNaN;  // inputs may be NaN
      const int nn = deg - 1;
      const long long nch = row_tasks[i] / deg;
      const unsigned long long total = binom(tab, kc, nn, LV);
      const int li = i - s0;
      for (long long t = t0 + wave; t < seg_end; t += kWavesPerBlock) {
        const long long local = t - off;
        const int y = (int)(local / nch);
        const long long e = base + y;
        const unsigned long long r_begin = (unsigned long long)(local % nch) * (unsigned long long)rpt;
        // a lower rank of this edge already passed: nothing here can win
        if (__atomic_load_n(&fpu[e], __ATOMIC_RELAXED) < r_begin) continue;
        const unsigned long long r_end = min(total, r_begin + (unsigned long long)rpt);
        const int lj = L.nbr[e] - s0;
        const double a = cor[(size_t)li * p + lj];
        for (unsigned long long r0 = r_begin; r0 < r_end; r0 += 64) {
          const unsigned long long rank = r0 + lane;
          bool pass = false, unc = false;
          if (rank < r_end) {
            int q[LV], S[LV];
            unrank<LV>(tab, kc, rank, nn, q);
#pragma unroll
            for (int c = 0; c < LV; ++c) S[c] = L.nbr[base + (q[c] < y ? q[c] : q[c] + 1)] - s0;
            double x[LV], yv[LV], C[LV][LV];
#pragma unroll
            for (int c = 0; c < LV; ++c) {
              x[c] = cor[(size_t)li * p + S[c]];
              yv[c] = cor[(size_t)lj * p + S[c]];
            }
#pragma unroll
            for (int c = 0; c < LV; ++c)
#pragma unroll
              for (int r = c + 1; r < LV; ++r) C[r][c] = cor[(size_t)S[c] * p + S[r]];
            bool anynan = false;
            if (chk) {
              anynan = __builtin_isnan(a);
#pragma unroll
              for (int c = 0; c < LV; ++c)
                anynan |= __builtin_isnan(x[c]) || __builtin_isnan(yv[c]);
#pragma unroll
              for (int c = 0; c < LV; ++c)
#pragma unroll
                for (int r = c + 1; r < LV; ++r) anynan |= __builtin_isnan(C[r][c]);
            }
            if (!anynan) {
              bool fb = false;
              double r;
              if constexpr (LV == 1) {
                r = pcor_l1(a, x[0], yv[0]);
              } else if constexpr (FAST) {
                double sens;
                r = pcor_canon_sens<LV>(a, x, yv, C, fb, sens);
                unc = fast && (fb || !fast_certain(r, r_crit, sens, In.eps));
              } else {
                r = pcor_canon<LV>(a, x, yv, C, fb);
              }
              pass = removes(r, r_crit) && !unc;
              F += fb ? 1 : 0;
            }
          }
          const unsigned long long m = __ballot(pass);
          if (fast) queue_uncertain(L, unc, m, e, r0, lane);
          if (m) {
            if (lane == 0) atomicMin(&fpu[e], r0 + (unsigned long long)(__ffsll((long long)m) - 1));
            break;
          }
        }
      }
      t0 = seg_end;
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) F += __shfl_xor(F, s, 64);
  if (lane == 0 && F) atomicAdd(&L.ctr->pinv_fallbacks, F);
}

// Level L >= 2 with the row's neighbourhood staged in LDS.  Every operand of
// a test on ordered edge (i, j) lies in cor[{i} u N(i), {i} u N(i)] (j and S
// are neighbours of i), so a workgroup copies that (deg + 1)^2 block of the
// correlation matrix into LDS once -- position a < deg is neighbour N(i)[a],
// position deg is i -- and runs the row's tasks against it.  Workgroups take
// blocks of tpb consecutive tasks (grid-stride; the task count is read on
// the device), one staging per row a block meets.  Rows whose block is larger
// than dmax x dmax are left to search_ln_kernel.  Same ranks, same arithmetic
// and the same orientation of every correlation read as search_ln_kernel.
//
// With a task queue (Q.head set) the kernel is persistent: each workgroup
// claims the next block of tpb tasks with one device atomic until the blocks
// run out, instead of taking blocks blockIdx.x, + gridDim.x, ...  With a
// shared queue (Q.done set) host threads take blocks from the other end: a
// workgroup skips a block the host has finished and marks each block it
// finishes.  Level 1 runs here only in that shared form (one rank = one
// neighbour k, the closed form of pcor_l1).
template <int LV, bool FAST>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_ln_lds_kernel(
    Layout L, int rpt, const long long *__restrict__ task_off,
    const long long *__restrict__ row_tasks, int tpb, int dmax, TaskQueue Q) {
  static_assert(LV >= 2 || !FAST, "level 1 takes this kernel in exact mode only");
  extern __shared__ unsigned long long smem[];
  // binomial rows 0..dmax-1 (a staged row's tests reach row deg - 1 < dmax)
  unsigned long long *bt = smem;                                   // dmax x (LV+1)
  const int brows = dmax;
  double *M = reinterpret_cast<double *>(smem + (size_t)brows * (LV + 1));  // dmax x dmax
  int *ids = reinterpret_cast<int *>(M + (size_t)dmax * dmax);     // dmax local ids
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int kRun = run_len<LV>();
  for (int t = tid; t < brows * (LV + 1); t += blockDim.x) {
    const int m = t / (LV + 1), k = t % (LV + 1);
    bt[t] = L.binom[(size_t)m * kBinomK + k];
  }
  const long long n_tasks = (long long)L.ctr->n_tasks;
  const long long n_blocks = (n_tasks + tpb - 1) / tpb;
  unsigned long long F = 0;
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(L.fp);
  __shared__ long long s_blk;
  __shared__ int s_done;
  for (long long blk = Q.head ? claim_block(Q, n_blocks, s_blk) : blockIdx.x; blk < n_blocks;
       blk = Q.head ? claim_block(Q, n_blocks, s_blk) : blk + gridDim.x) {
    if (Q.done) {
      // the host's end of a shared queue: skip a block it has finished (the
      // next claim's barrier orders this read before s_done is rewritten)
      if (tid == 0)
        s_done = __hip_atomic_load(&Q.done[blk], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
      __syncthreads();
      if (s_done) continue;
    }
    long long t0 = blk * tpb;
    const long long t1 = min(n_tasks, t0 + tpb);
    int prev = 0;  // the row of the block's previous segment
    while (t0 < t1) {
      // the segment of this block's tasks on row i (uniform across the block)
      const int i = task_row_wave(task_off, L.p_total, prev, t0);
      prev = i;
      const long long off = task_off[i];
      const long long seg_end = min(t1, off + row_tasks[i]);
      if (seg_end <= t0) break;  // (a consistent plan always advances)
      const int base = L.rowptr[i];
      const int deg = L.rowptr[i + 1] - base;
      const int D = deg + 1;
      if (D > dmax) {  // search_ln_kernel's row
        t0 = seg_end;
        continue;
      }
      const Inst &In = L.inst[L.row_inst[i]];
      const int p = In.p, s0 = In.start;
      const long long nch = row_tasks[i] / deg;
      __syncthreads();  // the previous segment is done with M (and bt is staged)
      for (int a = tid; a < D; a += blockDim.x) ids[a] = (a < deg ? L.nbr[base + a] : i) - s0;
      __syncthreads();
      for (int idx = tid; idx < D * D; idx += blockDim.x) {
        const int a = idx / D, b = idx - a * D;
        M[idx] = In.cor[(size_t)ids[a] * p + ids[b]];
      }
      __syncthreads();
      const double r_crit = In.r_crit;
      const bool fast = FAST && In.fast;
      // NaN inputs keep the edge (cpuInd.cpp:23-25,31-33); checked only when
      // the instance's correlation matrix may hold NaN
      const bool chk = In.flags & kCorNaN;
      const int nn = deg - 1;
      const unsigned long long total = binom(bt, LV + 1, nn, LV);
      const double *Mi = M + (size_t)deg * D;  // row i
      for (long long t = t0 + wave; t < seg_end; t += kWavesPerBlock) {
        const long long local = t - off;
        const int y = (int)(local / nch);
        const long long e = base + y;
        const unsigned long long r_begin = (unsigned long long)(local % nch) * (unsigned long long)rpt;
        if (__atomic_load_n(&fpu[e], __ATOMIC_RELAXED) < r_begin) continue;
        const unsigned long long r_end = min(total, r_begin + (unsigned long long)rpt);
        const double a = Mi[y];
        const double *Mj = M + (size_t)y * D;  // row j
        // lane t tests the kr consecutive ranks rb = r0 + kr * t ..: one
        // unrank, then lexicographic successors; the chunk's lowest passing
        // rank is the lowest passing lane's first passing step.  kr = kRun,
        // or fewer when the task's range is shorter than a chunk, so its
        // ranks spread over all 64 lanes
        const unsigned long long span = r_end > r_begin ? r_end - r_begin : 0ull;
        const int kr = span >= 64ull * kRun ? kRun : max(1, (int)((span + 63ull) / 64ull));
        for (unsigned long long r0 = r_begin; r0 < r_end; r0 += 64ull * kr) {
          const unsigned long long rb = r0 + (unsigned long long)lane * kr;
          int first = kRun;       // this lane's first certain pass
          unsigned unc_bits = 0;  // its uncertain steps
          int q[LV];
#pragma unroll
          for (int c = 0; c < LV; ++c) q[c] = c;
          if (rb < r_end) unrank_lds<LV>((const lds_u64 *)bt, LV + 1, rb, nn, q);
#pragma unroll
          for (int st = 0; st < kRun; ++st) {
            if (st < kr && rb + st < r_end && first == kRun) {
              int pos[LV];
#pragma unroll
              for (int c = 0; c < LV; ++c) pos[c] = q[c] < y ? q[c] : q[c] + 1;
              double x[LV], yv[LV], C[LV][LV];
#pragma unroll
              for (int c = 0; c < LV; ++c) {
                x[c] = Mi[pos[c]];
                yv[c] = Mj[pos[c]];
              }
#pragma unroll
              for (int c = 0; c < LV; ++c)
#pragma unroll
                for (int r = c + 1; r < LV; ++r) C[r][c] = M[pos[c] * D + pos[r]];
              bool anynan = false;
              if (chk) {
                anynan = __builtin_isnan(a);
#pragma unroll
                for (int c = 0; c < LV; ++c)
                  anynan |= __builtin_isnan(x[c]) || __builtin_isnan(yv[c]);
#pragma unroll
                for (int c = 0; c < LV; ++c)
#pragma unroll
                  for (int r = c + 1; r < LV; ++r) anynan |= __builtin_isnan(C[r][c]);
              }
              if (!anynan) {
                bool fb = false, unc = false;
                double r;
                // the fast instantiation runs one straight-line path for
                // every instance (a branch on In.fast kept both paths'
                // registers live and halved occupancy); exact instances of a
                // mixed launch get the same statistic and ignore the bound
                if constexpr (LV == 1) {
                  r = pcor_l1(a, x[0], yv[0]);
                } else if constexpr (FAST) {
                  double sens;
                  r = pcor_canon_sens<LV>(a, x, yv, C, fb, sens);
                  unc = fast && (fb || !fast_certain(r, r_crit, sens, In.eps));
                } else {
                  r = pcor_canon<LV>(a, x, yv, C, fb);
                }
                if (removes(r, r_crit) && !unc) first = st;
                if (unc) unc_bits |= 1u << st;
                F += fb ? 1 : 0;
              }
            }
            if (st + 1 < kr) next_comb<LV>(q, nn);
          }
          const unsigned long long m = __ballot(first < kRun);
          unsigned long long best = ~0ull;
          if (m) {
            const int t = __ffsll((long long)m) - 1;
            best = r0 + (unsigned long long)t * kr + (unsigned long long)__shfl(first, t, 64);
          }
          if (fast) {
            // uncertain tests below the chunk's lowest certain pass
#pragma unroll
            for (int st = 0; st < kRun; ++st)
              queue_rank(L, ((unc_bits >> st) & 1u) && rb + st < best, e, rb + st, lane);
          }
          if (m) {
            if (lane == 0) atomicMin(&fpu[e], best);
            break;
          }
        }
      }
      t0 = seg_end;
    }
    // the block is finished on this side (other waves may still be on their
    // last tasks; the host then skips a block the device completes anyway)
    if (Q.done && tid == 0)
      __hip_atomic_store(&Q.done[blk], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) F += __shfl_xor(F, s, 64);
  if (lane == 0 && F) atomicAdd(&L.ctr->pinv_fallbacks, F);
}

// Level LV >= 2 of exact instances, by conditioning set: the same tests and
// results as search_ln_lds_kernel, organised around S instead of the edge.
// Every test (i, j | S) of row i has S an LV-subset of N(i) not holding j, so
// the row's C(deg, LV) sets S, each against every j of N(i) \ S, are exactly
// its tests.  The part of a test that depends on i and S only (the LDL^T
// factorisation of S's block, U^-1 x and s00; pcor_prep) is computed once per
// S and shared by its deg - LV tests (pcor_finish does the rest, with the same
// operations as a per-test call, so every statistic has the same bits).  A
// passing (j, S) is turned into its rank among the LV-subsets of N(i) \ {j} in
// lexicographic order of positions (the edge's own enumeration) and lowers the
// edge's first passing rank with atomicMin; a lane walks its sets in
// lexicographic order, in which each edge's ranks rise, so it skips an edge
// for the rest of its run once the edge passed.  The row's tasks of the plan
// share its sets evenly: the segment [t0, t1) of the row's U tasks takes sets
// [t0 U^-1 NS, t1 U^-1 NS) (floors, NS = C(deg, LV)), spread over the
// workgroup's lanes in runs of consecutive sets.  Rows longer than the LDS
// block are left to search_ln_kernel, as for search_ln_lds_kernel.
__device__ __forceinline__ unsigned long long mul_div(unsigned long long a, unsigned long long b,
                                                      unsigned long long c) {
  return (unsigned long long)(((unsigned __int128)a * b) / c);
}

// The pinv form of one test of the by-set search (S or the 2 x 2 Schur
// block numerically singular), its operands read again from the staged
// block: called rarely, and out of line, so that none of the per-set values
// it needs stays live through the fast path.  M is the staged block of side
// D with row i at Mi, s the members of S as positions, y the tested one.
#ifndef HPC_LS_PAIR
#define HPC_LS_PAIR 1
#endif
// two tests in flight per lane in the by-set search
constexpr bool kLsPair = HPC_LS_PAIR != 0;
// the highest level whose pinv form runs out of line
constexpr int kLsOutOfLine = 6;

template <int LV>
struct Members {
  int v[LV];
};

template <int LV>
__device__ __noinline__ bool ls_pinv_removes(const double *M, int D, const double *Mi, Members<LV> s,
                                             int y, double r_crit) {
  double x[LV], yv[LV], C[LV][LV];
#pragma unroll
  for (int c = 0; c < LV; ++c) {
    x[c] = Mi[s.v[c]];
    yv[c] = M[y * D + s.v[c]];
  }
#pragma unroll
  for (int c = 0; c < LV; ++c)
#pragma unroll
    for (int r = c + 1; r < LV; ++r) C[r][c] = M[s.v[c] * D + s.v[r]];
  bool fb;
  return removes(pcor_canon<LV>(Mi[y], x, yv, C, fb), r_crit);
}

template <int LV>
__global__ __launch_bounds__(64 * kWavesPerBlock) void search_ls_kernel(
    Layout L, const long long *__restrict__ task_off, const long long *__restrict__ row_tasks,
    int tpb, int dmax, TaskQueue Q) {
  static_assert(LV >= 2, "levels >= 2");
  extern __shared__ unsigned long long smem[];
  unsigned long long *bt = smem;  // binomial rows 0..dmax-1 x (LV+1)
  double *M = reinterpret_cast<double *>(smem + (size_t)dmax * (LV + 1));  // dmax x dmax
  int *ids = reinterpret_cast<int *>(M + (size_t)dmax * dmax);
  const lds_u64 *tb = (const lds_u64 *)bt;
  const int tid = threadIdx.x;
  const int nthreads = 64 * kWavesPerBlock;
  for (int t = tid; t < dmax * (LV + 1); t += nthreads) {
    const int m = t / (LV + 1), k = t % (LV + 1);
    bt[t] = L.binom[(size_t)m * kBinomK + k];
  }
  const long long n_tasks = (long long)L.ctr->n_tasks;
  // balanced: one block per workgroup, its equal contiguous share of tasks
  const long long n_blocks = Q.balanced ? gridDim.x : (n_tasks + tpb - 1) / tpb;
  unsigned long long F = 0;
  unsigned long long *fpu = reinterpret_cast<unsigned long long *>(L.fp);
  __shared__ long long s_blk;
  for (long long blk = Q.head ? claim_block(Q, n_blocks, s_blk) : blockIdx.x; blk < n_blocks;
       blk = Q.head ? claim_block(Q, n_blocks, s_blk) : blk + gridDim.x) {
    long long t0, t1;
    if (Q.balanced) {
      t0 = (long long)mul_div((unsigned long long)n_tasks, (unsigned long long)blk, gridDim.x);
      t1 = (long long)mul_div((unsigned long long)n_tasks, (unsigned long long)blk + 1, gridDim.x);
    } else {
      t0 = blk * tpb;
      t1 = min(n_tasks, t0 + tpb);
    }
    int prev = 0;  // the row of the block's previous segment
    while (t0 < t1) {
      const int i = task_row_wave(task_off, L.p_total, prev, t0);
      prev = i;
      const long long off = task_off[i];
      const long long U = row_tasks[i];
      const long long seg_end = min(t1, off + U);
      if (seg_end <= t0) break;  // (a consistent plan always advances)
      const int base = L.rowptr[i];
      const int deg = L.rowptr[i + 1] - base;
      const int D = deg + 1;
      if (D > dmax) {  // search_ln_kernel's row
        t0 = seg_end;
        continue;
      }
      const Inst &In = L.inst[L.row_inst[i]];
      const int p = In.p, s0 = In.start;
      __syncthreads();  // the previous segment is done with M (and bt is staged)
      for (int a = tid; a < D; a += nthreads) ids[a] = (a < deg ? L.nbr[base + a] : i) - s0;
      __syncthreads();
      for (int idx = tid; idx < D * D; idx += nthreads) {
        const int a = idx / D, b = idx - a * D;
        M[idx] = In.cor[(size_t)ids[a] * p + ids[b]];
      }
      __syncthreads();
      const double r_crit = In.r_crit, r2_lo = In.r2_lo, r2_hi = In.r2_hi;
      const bool chk = In.flags & kCorNaN;
      const double *Mi = M + (size_t)deg * D;  // row i
      const unsigned long long NS = tb[deg * (LV + 1) + LV];
      const unsigned long long s_lo = mul_div((unsigned long long)(t0 - off), NS, (unsigned long long)U);
      const unsigned long long s_hi = mul_div((unsigned long long)(seg_end - off), NS, (unsigned long long)U);
      const unsigned long long cnt = s_hi - s_lo;
      const unsigned long long kr = (cnt + nthreads - 1) / nthreads;
      const unsigned long long sb = s_lo + (unsigned long long)tid * kr;
      const unsigned long long se = min(s_hi, sb + kr);
      const int nn = deg - 1;  // an edge's enumeration: LV-subsets of N(i) \ {j}
      if (sb < se) {
        int q[LV];  // S as positions in N(i), ascending
        unrank_lds<LV>(tb, LV + 1, sb, deg, q);
        unsigned long long done0 = 0, done1 = 0;  // edges (positions) passed in this run
        const unsigned long long all0 = deg >= 64 ? ~0ull : (1ull << deg) - 1;
        const unsigned long long all1 = deg > 64 ? (1ull << (deg - 64)) - 1 : 0ull;
        for (unsigned long long sr = sb; sr < se; ++sr) {
          double x[LV], C[LV][LV];
          unsigned long long in0 = 0, in1 = 0;  // S's positions
#pragma unroll
          for (int c = 0; c < LV; ++c) {
            x[c] = Mi[q[c]];
            if (q[c] < 64) in0 |= 1ull << q[c]; else in1 |= 1ull << (q[c] - 64);
          }
#pragma unroll
          for (int c = 0; c < LV; ++c)
#pragma unroll
            for (int r = c + 1; r < LV; ++r) C[r][c] = M[q[c] * D + q[r]];
          bool nanS = false;
          if (chk) {
#pragma unroll
            for (int c = 0; c < LV; ++c) nanS |= __builtin_isnan(x[c]);
#pragma unroll
            for (int c = 0; c < LV; ++c)
#pragma unroll
              for (int r = c + 1; r < LV; ++r) nanS |= __builtin_isnan(C[r][c]);
          }
          // NaN inputs keep the edge (cpuInd.cpp:23-25,31-33): no test of S
          if (!nanS) {
            PcorPrep<LV> P;
            pcor_prep<LV>(x, C, P);
            // the edges to test: positions of N(i) outside S not yet passed,
            // two at a time (two independent statistics in flight per lane)
            unsigned long long c0 = ~(in0 | done0) & all0, c1 = ~(in1 | done1) & all1;
            auto pop = [&]() -> int {
              if (c0) {
                const int y = __builtin_ctzll(c0);
                c0 &= c0 - 1;
                return y;
              }
              const int y = 64 + __builtin_ctzll(c1);
              c1 &= c1 - 1;
              return y;
            };
            while (c0 | c1) {
              const int ya = pop();
              const bool has_b = kLsPair && (c0 | c1) != 0;
              const int yb = has_b ? pop() : ya;
              const double aa = Mi[ya], ab = Mi[yb];
              double va[LV], vb[LV];
#pragma unroll
              for (int c = 0; c < LV; ++c) {
                va[c] = M[ya * D + q[c]];
                vb[c] = M[yb * D + q[c]];
              }
              double s11, s01a, s01b, qa, qb, wv[LV];
              const bool sa = pcor_schur<LV>(P, aa, va, s11, s01a, qa, wv);
              const bool sb2 = pcor_schur<LV>(P, ab, vb, s11, s01b, qb, wv);
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                if (h == 1 && !has_b) break;
                const int y = h ? yb : ya;
                const double a = h ? ab : aa;
                if (chk) {
                  bool nan = __builtin_isnan(a);
#pragma unroll
                  for (int c = 0; c < LV; ++c) nan |= __builtin_isnan(h ? vb[c] : va[c]);
                  if (nan) continue;
                }
                bool rm;
                if (h ? sb2 : sa) {
                  // the pinv form, as per test: out of line at the levels
                  // where that frees registers of the fast path (at higher
                  // levels the call's argument passing costs more)
                  if constexpr (LV <= kLsOutOfLine) {
                    Members<LV> S;
#pragma unroll
                    for (int c = 0; c < LV; ++c) S.v[c] = q[c];
                    rm = ls_pinv_removes<LV>(M, D, Mi, S, y, r_crit);
                  } else {
                    bool fb;
                    rm = removes(pcor_canon<LV>(a, x, h ? vb : va, C, fb), r_crit);
                  }
                  ++F;
                } else {
                  rm = removes_sq(h ? s01b : s01a, h ? qb : qa, r_crit, r2_lo, r2_hi);
                }
                if (!rm) continue;
                // S's rank among the LV-subsets of N(i) \ {j}: positions past
                // y move down by one; rank = sum over members of
                // C(nn - prev - 1, LV - c) - C(nn - q'_c, LV - c)
                unsigned long long rk = 0;
                int prev = -1;
#pragma unroll
                for (int c = 0; c < LV; ++c) {
                  const int qc = q[c] - (q[c] > y ? 1 : 0);
                  rk += tb[(nn - prev - 1) * (LV + 1) + (LV - c)] - tb[(nn - qc) * (LV + 1) + (LV - c)];
                  prev = qc;
                }
                const long long e = base + y;
                if (rk < __atomic_load_n(&fpu[e], __ATOMIC_RELAXED)) atomicMin(&fpu[e], rk);
                if (y < 64) done0 |= 1ull << y; else done1 |= 1ull << (y - 64);
              }
            }
          }
          if (sr + 1 < se) next_comb<LV>(q, deg);
        }
      }
      t0 = seg_end;
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) F += __shfl_xor(F, s, 64);
  if ((tid & 63) == 0 && F) atomicAdd(&L.ctr->pinv_fallbacks, F);
}

// Tasks per row: deg * ceil(C(deg - 1, level) / rpt) (row_task_count) for rows
// of instances at this level (and in the row mask).
__global__ void task_count_kernel(Layout L, int level, int rpt, const int32_t *__restrict__ row_mask,
                                  long long *__restrict__ row_tasks) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= L.p_total) return;
  const Inst &In = L.inst[L.row_inst[g]];
  const int deg = L.rowptr[g + 1] - L.rowptr[g];
  long long n = 0;
  if ((In.flags & kActive) && In.level == level && (!row_mask || row_mask[g]) && deg > 0) {
    // a row that cannot be planned gets no tasks; the host fails its level
    n = max(0ll, row_task_count(binom(L.binom, kBinomK, deg - 1, level), deg, rpt));
  }
  row_tasks[g] = n;
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
__device__ __forceinline__ int find_pos(const int32_t *row, int len, int node) {
  int lo = 0, hi = len;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (row[mid] < node) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// The next level's counts from the row lengths the removals left (one
// workgroup, the removal kernel's last): per row the task count of nc.level,
// per compaction workgroup the sums of row lengths and tasks, replaced by
// their exclusive prefixes, and the edge total in rowptr[p_total].  The same
// values compact_count_kernel derives from the adjacency.  Returns (on every
// thread) the totals and the longest row, which the caller publishes; what
// is written here is read by the next kernel only, so no fence is needed.
// One row per thread and pass: a compaction workgroup's rows are an aligned
// group of lanes, summed with shuffles (the code stays small: this runs once
// per launch, from a cold instruction cache).
__device__ void next_count(const Layout &L, const NextCount &nc, const unsigned long long *tab,
                           int tab_rows, unsigned long long &te, unsigned long long &tt,
                           int &tmax) {
  __shared__ unsigned long long s_we[kCommitWaves], s_wt[kCommitWaves];
  __shared__ int s_wm[kCommitWaves];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
  const int nblk = (L.p_total + kCompRows - 1) / kCompRows;
  int mx = 0;
  for (int g0 = 0; g0 < L.p_total; g0 += nt) {
    const int g = g0 + tid;
    int c = 0;
    long long tk = 0;
    if (g < L.p_total) {
      c = __hip_atomic_load(&L.deg[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (nc.level >= 2 && c > 0) {
        // a row that cannot be planned gets no tasks; the host fails its level
        const unsigned long long *tb = c - 1 < tab_rows ? tab : L.binom;
        tk = max(0ll, row_task_count(binom(tb, kBinomK, c - 1, nc.level), c, nc.rpt));
      }
      nc.row_tasks[g] = tk;
    }
    mx = max(mx, c);
    unsigned long long e = (unsigned long long)c, t = (unsigned long long)tk;
#pragma unroll
    for (int sh = 1; sh < kCompRows; sh <<= 1) {
      e += __shfl_xor(e, sh, 64);
      t += __shfl_xor(t, sh, 64);
    }
    if (tid % kCompRows == 0 && g < L.p_total) {
      const size_t b = (size_t)(g / kCompRows);
      nc.blk[2 * b] = e;
      nc.blk[2 * b + 1] = t;
    }
  }
  __syncthreads();
  // exclusive prefixes over the threads' block ranges (thread order = block
  // order): a scan within each wave, then the waves' totals
  const int per = (nblk + nt - 1) / nt;
  const int lo = min(nblk, tid * per), hi = min(nblk, lo + per);
  unsigned long long se = 0, st = 0;
  for (int b = lo; b < hi; ++b) {
    se += nc.blk[2 * (size_t)b];
    st += nc.blk[2 * (size_t)b + 1];
  }
  unsigned long long ie = se, it = st;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const unsigned long long ve = __shfl_up(ie, s, 64), vt = __shfl_up(it, s, 64);
    if (lane >= s) {
      ie += ve;
      it += vt;
    }
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) mx = max(mx, __shfl_xor(mx, s, 64));
  if (lane == 63) {
    s_we[wv] = ie;
    s_wt[wv] = it;
  }
  if (lane == 0) s_wm[wv] = mx;
  __syncthreads();
  unsigned long long oe = 0, ot = 0, ae = 0, at = 0;
  int M = 0;
  for (int w = 0; w < (nt >> 6); ++w) {
    if (w < wv) {
      oe += s_we[w];
      ot += s_wt[w];
    }
    ae += s_we[w];
    at += s_wt[w];
    M = max(M, s_wm[w]);
  }
  unsigned long long re = oe + ie - se, rt = ot + it - st;
  for (int b = lo; b < hi; ++b) {
    const unsigned long long e = nc.blk[2 * (size_t)b], t = nc.blk[2 * (size_t)b + 1];
    nc.blk[2 * (size_t)b] = re;
    nc.blk[2 * (size_t)b + 1] = rt;
    re += e;
    rt += t;
  }
  if (tid == 0) L.rowptr[L.p_total] = (int32_t)ae;
  te = ae;
  tt = at;
  tmax = M;
}

// A published counter word with the next level's sizes in place (the device
// counters get them after the clear that follows the publication).
__device__ __forceinline__ unsigned long long with_next(int w, unsigned long long v,
                                                        unsigned long long te, unsigned long long tt,
                                                        int tmax) {
  auto put32 = [](unsigned long long word, size_t byte, int x) {
    const int sh = (int)(byte & 4) * 8;
    return (word & ~(0xffffffffull << sh)) | ((unsigned long long)(unsigned)x << sh);
  };
  constexpr size_t o_md = offsetof(Counters, max_deg);
  constexpr size_t o_imd = kIctrOffset + offsetof(InstCounters, max_deg);
  if (w == (int)(offsetof(Counters, n_edges) / 8)) return te;
  if (w == (int)(offsetof(Counters, n_tasks) / 8)) return tt;
  if (w == (int)(o_md / 8)) v = put32(v, o_md, tmax);
  if (w == (int)(o_imd / 8)) v = put32(v, o_imd, tmax);
  return v;
}

// Removal and sepset recording (cpuUtil.cpp:23-48).  Each removed edge at
// level l appends a record of 3 + l words to its instance's record buffer:
// i, j (local ids, i < j), l, then the members in ascending order.  A
// wavefront reserves the space of all its records with one atomic when it
// holds one instance; the record order is scheduling-dependent, the records
// are not.
//
// The rows 0..tab_rows-1 of the binomial table (every row a test of the
// level can reach: tab_rows = longest row + 1) are staged in LDS, so the
// unranking of a record's set does its dependent lookups there; tab_rows = 0
// reads the global table.
__global__ __launch_bounds__(64 * kCommitWaves) void commit_kernel(Layout L, long long n_edges,
                                                                 Publish pub, int tab_rows,
                                                                 NextCount nc) {
  extern __shared__ unsigned long long s_tab[];
  for (int t = threadIdx.x; t < tab_rows * kBinomK; t += blockDim.x) s_tab[t] = L.binom[t];
  // an overflowed recheck queue leaves the level undecided: the host redoes
  // it (the workgroups still take part in the publication below)
  const bool live = !L.ctr->overflow;
  // the snapshot's own size bounds the count a launch was sized with
  n_edges = min(n_edges, (long long)L.rowptr[L.p_total]);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int k = -1, level = 0;
  unsigned long long T = 0, U = 0, E = 0, R = 0;
  bool rm = false, passed = false;
  int i = 0, j = 0, y = 0, base = 0, nn = 0, nni = 0;
  unsigned long long rank = 0;
  if (live && e < n_edges) {
    i = L.erow[e];
    k = L.K == 1 ? 0 : L.row_inst[i];
    level = L.inst[k].level;
    // levels the binomial table covers (the host never packs a higher one)
    if (level >= 1 && level <= kMaxLevel) {
      j = L.nbr[e];
      const int bi = L.rowptr[i];
      nni = L.rowptr[i + 1] - bi - 1;
      const long long fi = L.fp[e];
      E = 1;  // until the total (read after the staging barrier) says otherwise
      if (fi >= 0) T = (unsigned long long)fi + 1;
      // the lower-id row's lowest passing set wins, else the higher-id
      // row's: row i records when it passed and either it is the lower id or
      // row j did not pass (the only case that needs the reverse edge)
      if (fi >= 0) {
        bool mine = true;
        if (j < i) {
          // the reverse edge from the level-1 search, else a search of row j
          const int rv = L.erev ? L.erev[e] : -1;
          long long fj;
          if (rv >= 0) {
            fj = L.fp[rv];
          } else {
            const int bj = L.rowptr[j];
            const int lenj = L.rowptr[j + 1] - bj;
            const int yj = find_pos(L.nbr + bj, lenj, i);
            // the adjacency is symmetric, so i is in N(j); guard anyway
            fj = (yj < lenj && L.nbr[bj + yj] == i) ? L.fp[bj + yj] : -1;
          }
          mine = fj < 0;
        }
        if (mine) {
          y = (int)(e - bi);
          base = bi;
          rank = (unsigned long long)fi;
          nn = nni;
          passed = true;
        }
      }
    }
  }
  // the wave's instance, if it holds only one (-1: no edge), and the
  // workgroup's, if all its waves hold the same one
  int kk = k;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) kk = max(kk, __shfl_xor(kk, s, 64));
  const bool uniform = __all(k == kk || k < 0);
  // sk: each wave's instance for the record reservation (written once, before
  // the barrier that every wave passes); sc: for the counts
  __shared__ int sk[kCommitWaves], sc[kCommitWaves];
  __shared__ unsigned long long sT[kCommitWaves], sU[kCommitWaves], sE[kCommitWaves],
      sR[kCommitWaves], sW[kCommitWaves];
  __shared__ unsigned long long s_base;
  if (lane == 0) sk[wv] = uniform ? kk : -2;
  __syncthreads();  // also: the staged binomial rows are in place
  // the staged rows when they cover this row's sets (they cover every row of
  // a launch sized from its own counts), else the global table
  const unsigned long long *tab = (tab_rows > 0 && nni < tab_rows) ? s_tab : L.binom;
  if (E) {
    const unsigned long long total = binom(tab, kBinomK, nni, level);
    if (total > 0) {
      if (!T) T = total;
      U = total;
    } else {
      T = 0;
      E = 0;
    }
    // a rank valid for this level and row (so the unranking below ends)
    rm = passed && rank < total;
  }
  int K0 = -1;
  bool bu = true;
#pragma unroll
  for (int w = 0; w < kCommitWaves; ++w) {
    const int v = sk[w];
    if (v == -2 || (v >= 0 && K0 >= 0 && v != K0)) bu = false;
    if (v >= 0 && K0 < 0) K0 = v;
  }
  // record space: one reservation per workgroup when it holds one instance,
  // else per wave or per edge
  const unsigned long long need = rm ? (unsigned long long)(3 + level) : 0ull;
  unsigned long long incl = need;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const unsigned long long v = __shfl_up(incl, s, 64);
    if (lane >= s) incl += v;
  }
  const unsigned long long tot = __shfl(incl, 63, 64);
  unsigned long long off = 0;
  if (bu) {
    if (lane == 63) sW[wv] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long sum = 0;
      for (int w = 0; w < kCommitWaves; ++w) {
        const unsigned long long t = sW[w];
        sW[w] = sum;  // exclusive prefix over the waves
        sum += t;
      }
      s_base = (sum && K0 >= 0) ? atomicAdd(&L.ictr[K0].rec_words, sum) : 0ull;
    }
    __syncthreads();
    off = s_base + sW[wv] + incl - need;
  } else if (uniform) {
    unsigned long long wbase = 0;
    if (lane == 63 && tot) wbase = atomicAdd(&L.ictr[kk].rec_words, tot);
    wbase = __shfl(wbase, 63, 64);
    off = wbase + incl - need;
  } else if (rm) {
    off = atomicAdd(&L.ictr[k].rec_words, need);
  }
  // a record that would pass the buffer's end is dropped (only a launch sized
  // for other data can get there: its results are discarded)
  if (rm && L.inst[k].rec_base + off + need > L.inst[k].rec_cap) rm = false;
  if (rm) {
    const Inst &In = L.inst[k];
    const int W = In.W, s0 = In.start;
    const int li = i - s0, lj = j - s0;
    atomicAnd(&In.adj[(size_t)li * W + (lj >> 6)], ~(1ull << (lj & 63)));
    atomicAnd(&In.adj[(size_t)lj * W + (li >> 6)], ~(1ull << (li & 63)));
    if (nc.on) {
      atomicSub(&L.deg[i], 1);
      atomicSub(&L.deg[j], 1);
    }
    int32_t *out = In.rec + In.rec_base + off;
    out[0] = min(li, lj);
    out[1] = max(li, lj);
    out[2] = level;
    // the members of the rank-th level-subset, as unrank() finds them (a
    // binary search per member over the binomials; the last member is the
    // remaining rank past the previous one)
    int v0 = 0;
    for (int m = 0; m < level; ++m) {
      const int km = level - m;
      int v;
      if (km == 1) {
        v = v0 + (int)rank;
      } else {
        const unsigned long long top = tab[(size_t)(nn - v0) * kBinomK + km];
        const unsigned long long X = top - rank;
        int lo = km - 1, hi = nn - 1 - v0;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (tab[(size_t)mid * kBinomK + km] < X) lo = mid; else hi = mid - 1;
        }
        v = nn - 1 - lo;
        rank = rank + tab[(size_t)(lo + 1) * kBinomK + km] - top;
      }
      out[3 + m] = L.nbr[base + (v < y ? v : v + 1)] - s0;
      v0 = v + 1;
    }
    R = 1;
  }
  // per-instance counts: reduced per wave when the wave holds one instance,
  // then merged across the workgroup, so a one-instance launch issues one
  // atomic per counter and workgroup
  if (uniform) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      T += __shfl_xor(T, s, 64);
      U += __shfl_xor(U, s, 64);
      E += __shfl_xor(E, s, 64);
      R += __shfl_xor(R, s, 64);
    }
    if (lane == 0) {
      sc[wv] = (E | R) ? kk : -1;
      sT[wv] = T; sU[wv] = U; sE[wv] = E; sR[wv] = R;
    }
  } else {
    if (E | R) {
      atomicAdd(&L.ictr[k].tests, T);
      atomicAdd(&L.ictr[k].tests_max, U);
      atomicAdd(&L.ictr[k].edges, E);
      if (R) atomicAdd(&L.ictr[k].removed, R);
    }
    if (lane == 0) sc[wv] = -1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // runs of waves with the same instance share one set of atomics
    int cur = -1;
    unsigned long long t = 0, u = 0, ed = 0, rmv = 0;
    for (int w = 0; w <= kCommitWaves; ++w) {
      const int kw = w < kCommitWaves ? sc[w] : -2;
      if (kw == -1) continue;
      if (kw != cur && cur >= 0) {
        atomicAdd(&L.ictr[cur].tests, t);
        atomicAdd(&L.ictr[cur].tests_max, u);
        atomicAdd(&L.ictr[cur].edges, ed);
        if (rmv) atomicAdd(&L.ictr[cur].removed, rmv);
        t = u = ed = rmv = 0;
      }
      if (kw < 0) break;
      cur = kw;
      t += sT[w]; u += sU[w]; ed += sE[w]; rmv += sR[w];
    }
  }
  if (!pub.flag) return;
  // the last workgroup to finish publishes the counters
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(&L.ctr->done2, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    __threadfence();
    unsigned long long nte = 0, ntt = 0;
    int nmax = 0;
    if (nc.on) next_count(L, nc, tab_rows > 0 ? s_tab : L.binom, tab_rows > 0 ? tab_rows : 1 << 30,
                          nte, ntt, nmax);
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(L.ctr);
    unsigned long long words_rec = 0;
    if (pub.advance && L.K == 1 && threadIdx.x == 0)
      words_rec = __hip_atomic_load(&L.ictr[0].rec_words, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int w = threadIdx.x; w < pub.words; w += blockDim.x) {
      unsigned long long v = __hip_atomic_load(&src[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (nc.on) v = with_next(w, v, nte, ntt, nmax);
      pub.dst[w] = v;
    }
    if (pub.sys) {
      __threadfence_system();
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_store(pub.flag, pub.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
      // (the words reach device memory before the kernel's end)
      __syncthreads();
    }
    if (pub.advance && L.K == 1) {
      // the next level of the (single) instance starts from clear counters
      unsigned long long *blk = reinterpret_cast<unsigned long long *>(L.ctr);
      for (int w = threadIdx.x; w < pub.words; w += blockDim.x) blk[w] = 0ull;
      __syncthreads();
      if (threadIdx.x == 0) {
        Inst *In = const_cast<Inst *>(L.inst);
        In[0].level += 1;
        In[0].rec_base += words_rec;
        if (nc.on) {
          // the next level's sizes stay for its search
          L.ctr->n_edges = nte;
          L.ctr->n_tasks = ntt;
          L.ctr->max_deg = nmax;
          L.ictr[0].max_deg = nmax;
        }
      }
    }
  }
}

__global__ void scatter_kernel(const int64_t *__restrict__ idx, const int64_t *__restrict__ val,
                               long long count, int64_t *__restrict__ fp) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) fp[idx[t]] = val[t];
}

// Lowers first passing ranks (-1 = none is the largest unsigned value).
__global__ void scatter_min_kernel(const int64_t *__restrict__ idx, const int64_t *__restrict__ val,
                                   long long count, int64_t *__restrict__ fp) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count)
    atomicMin(reinterpret_cast<unsigned long long *>(fp) + idx[t], (unsigned long long)val[t]);
}

// LDS bytes of search_ln_lds_kernel for blocks of side d (row length + 1):
// binomial rows 0..d-1, the d x d neighbourhood block and its ids.
size_t lds_bytes(int d_in, int lv) {
  const size_t d = (size_t)d_in;
  return d * (lv + 1) * 8 + d * d * 8 + ((d * 4 + 7) & ~(size_t)7);
}

// Largest block side whose staging fits `limit` bytes of LDS.
int lds_side(size_t limit, int lv) {
  int d = 0;
  while (lds_bytes(d + 1, lv) <= limit) ++d;
  return d;
}

// Resident workgroups of the grid-stride search kernels (the task count is
// only known on the device): enough to fill all 256 CUs several times.
constexpr int kSearchBlocks = 256 * 8;

// Ordered edges per workgroup block of the level-1 gather kernel.
constexpr int kL1GatherEpb = 64;

// Workgroups of 64 kWavesPerBlock threads and `lds` bytes of dynamic LDS that
// the device holds at once (the grid of a balanced by-set search).
int resident_blocks(const void *kernel, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 64 * kWavesPerBlock, lds) !=
          hipSuccess ||
      cus <= 0 || per <= 0)
    return kSearchBlocks;
  return cus * per;
}

// Rows whose block fits `limit` bytes of LDS go to the staged kernel, longer
// ones to the gather kernel (launched only when such a row exists).  Both
// read the same task plan; each skips the other's rows.
template <int LV, bool FAST>
void launch_ln_as(hipStream_t s, const Layout &L, int rpt, const long long *row_tasks,
                  const long long *task_off, int tpb, size_t limit, const TaskQueue &Q) {
  const int dmax = std::min(L.max_deg + 1, lds_side(limit, LV));
  // a testable row has deg - 1 >= LV, so a block side of at least LV + 2
  if (dmax >= LV + 2) {
    // exact levels >= 2 by conditioning set, unless host threads share the
    // queue (their side runs the per-edge tasks)
    if constexpr (!FAST && LV >= 2) {
      if (!Q.done) {
        int grid = kSearchBlocks;
        if (Q.balanced) {
          // the occupancy query once per LDS size and host thread
          thread_local size_t cached_lds = 0;
          thread_local int cached_grid = 0;
          const size_t lds = lds_bytes(dmax, LV);
          if (lds != cached_lds || cached_grid <= 0) {
            cached_grid = resident_blocks((const void *)search_ls_kernel<LV>, lds);
            cached_lds = lds;
          }
          grid = cached_grid;
        }
        hipLaunchKernelGGL((search_ls_kernel<LV>), dim3(grid), dim3(64 * kWavesPerBlock),
                           lds_bytes(dmax, LV), s, L, task_off, row_tasks, tpb, dmax, Q);
      } else {
        hipLaunchKernelGGL((search_ln_lds_kernel<LV, FAST>), dim3(kSearchBlocks),
                           dim3(64 * kWavesPerBlock), lds_bytes(dmax, LV), s, L, rpt, task_off,
                           row_tasks, tpb, dmax, Q);
      }
    } else {
      hipLaunchKernelGGL((search_ln_lds_kernel<LV, FAST>), dim3(kSearchBlocks),
                         dim3(64 * kWavesPerBlock), lds_bytes(dmax, LV), s, L, rpt, task_off,
                         row_tasks, tpb, dmax, Q);
    }
  }
  if (L.max_deg + 1 <= dmax) return;
  const size_t lds = (size_t)(L.max_deg + 1) * (LV + 1) * sizeof(unsigned long long);
  const int use = lds <= 48 * 1024;
  hipLaunchKernelGGL((search_ln_kernel<LV, FAST>), dim3(kSearchBlocks), dim3(64 * kWavesPerBlock),
                     use ? lds : 0, s, L, rpt, task_off, row_tasks, tpb, dmax, use);
}

// The fast-correlation variant (uncertainty bounds, recheck queue) is only
// instantiated into launches that hold a fast instance.
template <int LV>
void launch_ln(hipStream_t s, const Layout &L, int rpt, const long long *row_tasks,
               const long long *task_off, bool fast, int tpb, size_t limit, const TaskQueue &Q) {
  if (fast) launch_ln_as<LV, true>(s, L, rpt, row_tasks, task_off, tpb, limit, Q);
  else launch_ln_as<LV, false>(s, L, rpt, row_tasks, task_off, tpb, limit, Q);
}

// Level 1 through the task-plan kernels (the shared queue): exact only.
template <>
void launch_ln<1>(hipStream_t s, const Layout &L, int rpt, const long long *row_tasks,
                  const long long *task_off, bool, int tpb, size_t limit, const TaskQueue &Q) {
  launch_ln_as<1, false>(s, L, rpt, row_tasks, task_off, tpb, limit, Q);
}

}  // namespace

size_t search_lds_limit(size_t requested) {
  return requested == 0 || requested > kSearchLdsLimit ? kSearchLdsLimit : requested;
}

void launch_search_l1(hipStream_t s, const Layout &L, int64_t n_edges, const int32_t *row_mask,
                      bool fast, int epb, size_t limit) {
  if (n_edges <= 0) return;
  // a correlation row of every instance fits the LDS budget: by pivot row
  const size_t row_bytes = (size_t)L.pmax * sizeof(double);
  if (L.ccor && L.pmax > 0 && row_bytes <= limit) {
    const long long blocks = (n_edges + epb - 1) / epb;
    const long long per_xcd = (blocks + kL1Classes - 1) / kL1Classes;
    const dim3 grid((unsigned)(per_xcd * kL1Classes));
    if (fast)
      hipLaunchKernelGGL(search_l1_pivot_kernel<true>, grid, dim3(64 * kWavesPerBlock), row_bytes,
                         s, L, (long long)n_edges, row_mask, epb, per_xcd);
    else
      hipLaunchKernelGGL(search_l1_pivot_kernel<false>, grid, dim3(64 * kWavesPerBlock), row_bytes,
                         s, L, (long long)n_edges, row_mask, epb, per_xcd);
    return;
  }
  // rows of up to `cap` neighbours are staged (16 bytes per neighbour)
  const int cap = (int)std::min<size_t>((size_t)L.max_deg, limit / 16);
  if (cap >= 2) {
    const long long span = (long long)kL1Classes * epb;
    const long long blocks = kL1Classes * ((n_edges + span - 1) / span);
    const size_t staged = (size_t)cap * 16;
    if (fast)
      hipLaunchKernelGGL(search_l1_lds_kernel<true>, dim3((unsigned)blocks),
                         dim3(64 * kWavesPerBlock), staged, s, L, (long long)n_edges, row_mask,
                         epb, cap);
    else
      hipLaunchKernelGGL(search_l1_lds_kernel<false>, dim3((unsigned)blocks),
                         dim3(64 * kWavesPerBlock), staged, s, L, (long long)n_edges, row_mask,
                         epb, cap);
  }
  if (L.max_deg <= cap) return;
  long long blocks = (n_edges + kL1GatherEpb - 1) / kL1GatherEpb;
  if (blocks > 16384) blocks = 16384;
  if (fast)
    hipLaunchKernelGGL(search_l1_kernel<true>, dim3((unsigned)blocks), dim3(64 * kWavesPerBlock), 0,
                       s, L, (long long)n_edges, row_mask, cap, kL1GatherEpb);
  else
    hipLaunchKernelGGL(search_l1_kernel<false>, dim3((unsigned)blocks), dim3(64 * kWavesPerBlock),
                       0, s, L, (long long)n_edges, row_mask, cap, kL1GatherEpb);
}

void launch_task_plan(hipStream_t s, int level, const Layout &L, int rpt,
                      const int32_t *row_mask, long long *row_tasks, long long *task_off) {
  hipLaunchKernelGGL(task_count_kernel, dim3((L.p_total + 255) / 256), dim3(256), 0, s, L, level,
                     rpt, row_mask, row_tasks);
  hipLaunchKernelGGL(task_scan_kernel, dim3(1), dim3(kScan), 0, s, row_tasks, L.p_total, task_off,
                     L.ctr);
}

void launch_search_ln(hipStream_t s, int level, const Layout &L, int rpt,
                      const long long *row_tasks, const long long *task_off, bool fast, int tpb,
                      size_t limit, const TaskQueue &Q) {
  switch (level) {
    case 1: launch_ln<1>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 2: launch_ln<2>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 3: launch_ln<3>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 4: launch_ln<4>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 5: launch_ln<5>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 6: launch_ln<6>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 7: launch_ln<7>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 8: launch_ln<8>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 9: launch_ln<9>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 10: launch_ln<10>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 11: launch_ln<11>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 12: launch_ln<12>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 13: launch_ln<13>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 14: launch_ln<14>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 15: launch_ln<15>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    case 16: launch_ln<16>(s, L, rpt, row_tasks, task_off, fast, tpb, limit, Q); break;
    default: break;  // rejected by the caller (kMaxLevel)
  }
}

void launch_commit(hipStream_t s, const Layout &L, int64_t n_edges, const Publish &pub,
                   const NextCount &next) {
  if (n_edges <= 0) return;
  const int tpb = 64 * kCommitWaves;
  // binomial rows 0..max_deg in LDS when they fit in 48 KB
  int tab_rows = L.max_deg + 1;
  if ((size_t)tab_rows * kBinomK * sizeof(unsigned long long) > 48 * 1024) tab_rows = 0;
  hipLaunchKernelGGL(commit_kernel, dim3((unsigned)((n_edges + tpb - 1) / tpb)), dim3(tpb),
                     (size_t)tab_rows * kBinomK * sizeof(unsigned long long), s, L,
                     (long long)n_edges, pub, tab_rows, next);
}

void launch_scatter_fp(hipStream_t s, const int64_t *idx, const int64_t *val, long long count,
                       int64_t *fp) {
  if (count <= 0) return;
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, idx,
                     val, count, fp);
}

void launch_scatter_min_fp(hipStream_t s, const int64_t *idx, const int64_t *val, long long count,
                           int64_t *fp) {
  if (count <= 0) return;
  hipLaunchKernelGGL(scatter_min_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                     idx, val, count, fp);
}

}  // namespace hpc
