// search.hip -- width search, normalisation, threshold and ordered
// compaction over dedispersed planes (gfx950).
//
// Replaces hella::smooth (src/hella/src/smooth.cpp:21-127) and the
// threshold half of hella::find_peaks (src/hella/src/peaks.cu:31-117).
//
// Two width-trial families:
//  * BOXCAR (production): for width w in minW, 2minW, ..., maxW the sum of
//    w dedispersed samples centred at output sample s + maxW/2.  The
//    workgroup of one (row, 4096-sample block) keeps a u32 prefix sum of the
//    row segment in LDS, so every width costs two LDS reads per sample and
//    the sums stay exact integers (wrap-around u32 differences).
//  * HELLA (reference kernel): the 3 x (2sm+1) polynomial-Gaussian kernel
//    of smooth.cpp:45-58 applied as a direct FIR with replicate border.
// Statistics follow calculate_stddev (normalization.cu:125-162) over the
// first 256*floor(nout/256) samples with the canonical blocked reduction
// (4096-sample blocks, 256 lanes, pairwise lane tree, ascending blocks).
// Planes are never stored in BOXCAR mode: each pass recomputes the boxes
// from D, which is cheaper than writing and re-reading nw planes.
#include "reduce.hpp"
#include "search.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace hamd {

namespace {

constexpr int NT = 256;
constexpr int SB = SEARCH_BLOCK;
constexpr int MAXBOX = 1024;

enum Pass { P_STATS = 0, P_Y0 = 1, P_COUNT = 2, P_WRITE = 3 };

// exclusive scan of one u32 per thread over the 256-thread workgroup: wave
// scans, then the four wave totals through LDS (u32 sums: any order gives the
// same words)
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *total)
{
  const int l = threadIdx.x, lane = l & 63, wv = l >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) sh[wv] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const uint32_t c = sh[w];
    tot += c;
    before += w < wv ? c : 0u;
  }
  *total = tot;
  __syncthreads();
  return before + incl - v;
}

// P[1..L] holds a block's dedispersed samples (P[0] = 0); afterwards P[i] is
// the u32 (wrap-around) sum of the first i.  Thread l scans a contiguous
// segment; segment totals are scanned within each wave by DPP steps and across
// the four waves through wtot.
// MAXSEG: the largest segment; EXACT: L == MAXSEG * NT, so every thread's
// segment is full and no access is predicated.
template <int MAXSEG = (SB + MAXBOX + NT - 1) / NT, bool EXACT = false>
__device__ void block_prefix(uint32_t *P, int L, uint32_t *wtot)
{
  // thread l owns the seg consecutive words from 1 + l * seg; the loops run
  // over the largest segment with the tail predicated off, so every LDS read
  // of a segment is issued before the first addition (u32 sums wrap, and any
  // order of the additions gives the same words)
  const int l = threadIdx.x, lane = l & 63, wv = l >> 6;
  const int seg = EXACT ? MAXSEG : (L + NT - 1) / NT;
  const int i0 = 1 + l * seg, n = EXACT ? MAXSEG : max(0, min(seg, L + 1 - i0));
  uint32_t v[MAXSEG];
#pragma unroll
  for (int i = 0; i < MAXSEG; ++i) v[i] = (EXACT || i < n) ? P[i0 + i] : 0u;
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < MAXSEG; ++i) run += v[i];
  const uint32_t incl = wave_incl_scan(run);
  if (lane == 63) wtot[wv] = incl;
  __syncthreads();
  uint32_t acc = incl - run;
  for (int w = 0; w < wv; ++w) acc += wtot[w];
#pragma unroll
  for (int i = 0; i < MAXSEG; ++i) {
    acc += v[i];
    if (EXACT || i < n) P[i0 + i] = acc;
  }
  __syncthreads();
}

// The normalisation scalars, read through the constant address space: every
// field a kernel reads at a wave-uniform index is a scalar load (the search
// kernels that read them never write them; k_finish, which writes them, runs
// before, in another launch).
typedef __attribute__((address_space(4))) const SScal SScalK;

template <class S>
__device__ __forceinline__ float normed(double raw, const S &sc, int k)
{
  return sc.sigf[k] > 0.f ? (float)((raw - sc.mu[k]) * sc.inv[k]) : (float)raw;
}

constexpr int MAXL = SB + MAXBOX;        // samples a block's boxes read
constexpr int PF = (MAXL + NT - 1) / NT; // prefetch registers per thread

// A load through the global address space: it counts on vmcnt alone, so the
// waits of LDS accesses (lgkmcnt) never wait for it, as they wait for a flat
// load.  Plane loads are issued unconditionally at clamped addresses and the
// values outside the row are zeroed where they are used: a load that sits in
// a branch of its own is waited for inside that branch, one at a time.
template <class T>
__device__ __forceinline__ T gld(const T *p)
{
  typedef __attribute__((address_space(1))) const T gt;
  return *((gt *)p);
}
constexpr int PN = MAXL + 1;
constexpr int TREE_WORDS = 2 * SEARCH_MAXW * 128 * 2;
constexpr int LDSW = PN > TREE_WORDS ? PN : TREE_WORDS;

// Statistics passes over a run of bpw blocks of one row.
//  P_STATS: per width, the block sums of x and x*x (canonical lane order) and
//           the block's largest raw value (for the count pass)
//  P_Y0:    sums of y and y*y of the normalised first width
// BOXCAR builds the block's prefix sums in LDS; the next block's samples are
// loaded into registers while the current block is processed.
template <bool HELLA, int NW>
__global__ __launch_bounds__(NT) void k_stats(const SearchBeamDev *__restrict__ beams,
                                              SearchCommon C, int pass, int bpw)
{
  const SearchBeamDev B = beams[blockIdx.z];
  const int r = blockIdx.y;
  if (r >= B.rows) return;
  const int b0 = blockIdx.x * bpw, b1 = min(C.nblk, b0 + bpw);
  if (b0 >= b1) return;
  const int l = threadIdx.x;
  __shared__ __align__(16) uint32_t lds_u[LDSW];
  __shared__ uint32_t wtot[NT / 64];
  uint32_t *P = lds_u;
  double *shd = reinterpret_cast<double *>(lds_u);
  const SScalK &sc = *(SScalK *)B.scal;
  const uint32_t *drow = B.D + (int64_t)(r + B.row0) * B.ds;
  const int c0 = C.maxW / 2;
  const bool y0 = pass == P_Y0;  // launched with NW = 1
  const bool direct = !HELLA && y0 && C.minW == 1;
  const int L = SB + C.maxW;

  uint32_t nx[PF];
  auto fetch = [&](int b) {
    const int base = b * SB;
#pragma unroll
    for (int t = 0; t < PF; ++t) nx[t] = gld(drow + min(base + l + NT * t, C.ntime_dd - 1));
  };
  if (!HELLA) fetch(b0);

  for (int b = b0; b < b1; ++b) {
    const int base = b * SB;
    if (!HELLA) {
#pragma unroll
      for (int t = 0; t < PF; ++t) {
        const int i = l + NT * t;
        if (i < L) P[1 + i] = base + i < C.ntime_dd ? nx[t] : 0u;
      }
      if (l == 0) P[0] = 0u;
      if (b + 1 < b1) fetch(b + 1);
      __syncthreads();
      if (!direct) block_prefix(P, L, wtot);
    }
    auto raw = [&](int k, int j) -> double {
      if (HELLA) return (double)gld(B.planes + ((int64_t)k * B.rows + r) * B.ps + base + j);
      if (direct) return (double)P[1 + j + c0];
      const int w = C.minW << k;
      const int a = j + c0 - w / 2;
      return (double)(uint32_t)(P[a + w] - P[a]);
    };
    double acc[2 * NW];
    uint32_t mx[NW];
#pragma unroll
    for (int k = 0; k < 2 * NW; ++k) acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) mx[k] = 0u;
    for (int kk = 0; kk < SB / NT; ++kk) {
      const int j = l + NT * kk;
      if (base + j >= C.nout) break;
      const bool in_stats = base + j < C.wp;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        double v = raw(k, j);
        if (!y0) {
          const uint32_t bits = HELLA ? __float_as_uint((float)v) : (uint32_t)v;
          mx[k] = max(mx[k], bits);
        }
        if (y0) v = (double)normed(v, sc, 0);
        if (in_stats) {
          acc[k] += v;
          acc[NW + k] += v * v;
        }
      }
    }
    if (!y0) {
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const uint32_t m = wave_max(mx[k]);
        if ((l & 63) == 0)
          B.bmax[(((int64_t)k * B.rows + r) * C.nblk + b) * 4 + (l >> 6)] = m;
      }
    }
    __syncthreads();  // P is dead; its LDS becomes the tree scratch
    tree256_multi(acc, shd);
    if (l == 0) {
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const int64_t slot = ((int64_t)k * C.nblk + b) * B.rows + r;  // [k][block][row]
        B.part[slot] = acc[k];
        B.part2[slot] = acc[NW + k];
      }
    }
    __syncthreads();
  }
}

// P_Y0 when the first width is the raw dedispersed sample (BOXCAR, minW 1):
// the statistics of y = normed(D[c0 + j]) over a run of bpw blocks of one
// row, read straight from the plane (lane l takes samples l + 256k of a
// block, as k_stats does), the next block's samples loaded while the
// current block's are reduced; only the tree uses LDS.
__global__ __launch_bounds__(NT) void k_y0_direct(const SearchBeamDev *__restrict__ beams,
                                                  SearchCommon C, int bpw)
{
  const SearchBeamDev B = beams[blockIdx.z];
  const int r = blockIdx.y;
  if (r >= B.rows) return;
  const int b0 = blockIdx.x * bpw, b1 = min(C.nblk, b0 + bpw);
  if (b0 >= b1) return;
  const int l = threadIdx.x;
  __shared__ double shd[2 * 128];
  const SScalK &sc = *(SScalK *)B.scal;
  const uint32_t *drow = B.D + (int64_t)(r + B.row0) * B.ds + C.maxW / 2;
  constexpr int K = SB / NT;
  uint32_t nx[K];
  auto fetch = [&](int b) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = b * SB + l + NT * k;
      nx[k] = gld(drow + min(j, C.nout - 1));  // used only where j < nout
    }
  };
  fetch(b0);
  for (int b = b0; b < b1; ++b) {
    uint32_t cur[K];
#pragma unroll
    for (int k = 0; k < K; ++k) cur[k] = nx[k];
    if (b + 1 < b1) fetch(b + 1);
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = b * SB + l + NT * k;
      if (j >= C.nout || j >= C.wp) break;
      const double v = (double)normed((double)cur[k], sc, 0);
      acc[0] += v;
      acc[1] += v * v;
    }
    tree256_multi(acc, shd);
    if (l == 0) {
      const int64_t slot = (int64_t)b * B.rows + r;  // [block][row]
      B.part[slot] = acc[0];
      B.part2[slot] = acc[1];
    }
  }
}

__device__ __forceinline__ void add128(uint64_t &lo, uint64_t &hi, uint64_t blo, uint64_t bhi)
{
  const uint64_t t = lo + blo;
  hi += bhi + (t < lo ? 1ull : 0ull);
  lo = t;
}

// BOXCAR statistics over a run of bpw blocks of one row: per width the exact
// integer sums S1 = sum box (u64) and S2 = sum box^2 (u128) over the
// statistics columns, so any summation order gives the same totals; and per
// (block, wave) the largest box sum for the count pass.  The next block's
// samples are loaded into registers while the current block is processed.
// Every box sum is below 2^30 (checked on the host), so a thread's 16
// squares of one block fit a u64 before they are folded into the u128.
// W1: widths 1, 2, 4, ... (WIDTH_MIN 1), compile-time LDS offsets.
// NARROW: every box sum is so small that a thread's squares over all its
// blocks fit a u64 (box_max^2 x 16 x bpw < 2^64) and its 16 box sums of one
// block fit a u32 (box_max < 2^28), both checked on the host; the thread then
// keeps one u64 of squares per width, adds a block's boxes in a u32, and only
// the reduction widens to u128, which frees the registers of the per-thread
// u128 sums.
// The row is read as a buffer of ntime_dd words, so a load past its end
// returns 0 without a clamp or a select.
template <int NW, bool W1, bool NARROW>
__global__ __launch_bounds__(NT) void k_stats_box(const SearchBeamDev *__restrict__ beams,
                                                  SearchCommon C, int bpw)
{
  const SearchBeamDev B = beams[blockIdx.z];
  const int r = blockIdx.y;
  if (r >= B.rows) return;
  const int g = blockIdx.x;
  const int ng = (C.nblk + bpw - 1) / bpw;
  const int b0 = g * bpw, b1 = min(C.nblk, b0 + bpw);
  if (b0 >= b1) return;
  const int l = threadIdx.x, lane = l & 63, wv = l >> 6;
  __shared__ __align__(16) uint32_t P[PN];
  __shared__ uint32_t wtot[NT / 64];
  __shared__ uint64_t red[NT / 64][NW][3];
  const uint32_t *drow = B.D + (int64_t)(r + B.row0) * B.ds;
  const int c0 = C.maxW / 2;
  const int L = SB + C.maxW;
  uint64_t s1[NW], q[NW], s2lo[NW], s2hi[NW];
  uint32_t mx[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) s1[k] = s2lo[k] = s2hi[k] = q[k] = 0ull;
  // W1: the widths are 1 .. 2^(NW-1), so L = SB + 2^(NW-1) is known here;
  // when it is a multiple of NT every load, store and prefix segment is full
  constexpr int LW = SB + (1 << (NW - 1));
  constexpr int PFW = W1 ? (LW + NT - 1) / NT : PF;
  constexpr bool EX = W1 && LW % NT == 0;
  uint32_t nx[PFW];
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)drow, (short)0, C.ntime_dd * 4, 0x00020000);
  auto fetch = [&](int b) {
    const int vo = (b * SB + l) * 4;  // the whole offset in the range-checked part
#pragma unroll
    for (int t = 0; t < PFW; ++t) nx[t] = __builtin_amdgcn_raw_buffer_load_b32(rs, vo + NT * 4 * t, 0, 0);
  };
  uint32_t s1b[NW];  // NARROW: this block's box sums
  auto body = [&](int j, auto in_stats) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int w = W1 ? (1 << k) : (C.minW << k);
      const int a = j + c0 - w / 2;
      const uint32_t box = P[a + w] - P[a];
      mx[k] = max(mx[k], box);
      if constexpr (decltype(in_stats)::value) {
        if (NARROW)
          s1b[k] += box;
        else
          s1[k] += box;
        q[k] += (uint64_t)box * box;
      }
    }
  };
  fetch(b0);
  for (int b = b0; b < b1; ++b) {
    const int base = b * SB;
#pragma unroll
    for (int t = 0; t < PFW; ++t) {
      const int i = l + NT * t;
      if (EX || i < L) P[1 + i] = nx[t];
    }
    if (l == 0) P[0] = 0u;
    if (b + 1 < b1) fetch(b + 1);
    __syncthreads();
    if constexpr (EX)
      block_prefix<PFW, true>(P, L, wtot);
    else
      block_prefix(P, L, wtot);
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      if (!NARROW) q[k] = 0ull;
      s1b[k] = 0u;
      mx[k] = 0u;
    }
    // iterations whose 256 samples are all searched / all in the statistics
    // columns are uniform over the workgroup (wp is a multiple of 256)
    const int nrem = C.nout - base;
    const int nfull = min(SB / NT, nrem / NT);
    const int nst = min(nfull, max(0, (C.wp - base) / NT));
    int kk = 0;
    for (; kk < nst; ++kk) body(l + NT * kk, std::true_type{});
    for (; kk < nfull; ++kk) body(l + NT * kk, std::false_type{});
    if (kk < SB / NT) {
      const int j = l + NT * kk;
      if (j < nrem) {
        if (base + j < C.wp)
          body(j, std::true_type{});
        else
          body(j, std::false_type{});
      }
    }
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      if (NARROW) s1[k] += s1b[k];
      if (!NARROW) add128(s2lo[k], s2hi[k], q[k], 0ull);
      const uint32_t m = wave_max(mx[k]);
      if (lane == 0) B.bmax[(((int64_t)k * B.rows + r) * C.nblk + b) * 4 + wv] = m;
    }
    __syncthreads();  // P is rewritten by the next block
  }
  // integer totals: wave sums by shuffles, then the four waves
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    uint64_t a = s1[k], lo = NARROW ? q[k] : s2lo[k], hi = NARROW ? 0ull : s2hi[k];
#pragma unroll
    for (int sft = 32; sft >= 1; sft >>= 1) {
      a += __shfl_down(a, sft, 64);
      const uint64_t olo = __shfl_down(lo, sft, 64), ohi = __shfl_down(hi, sft, 64);
      add128(lo, hi, olo, ohi);
    }
    if (lane == 0) {
      red[wv][k][0] = a;
      red[wv][k][1] = lo;
      red[wv][k][2] = hi;
    }
  }
  __syncthreads();
  if (l < NW) {
    const int k = l;
    uint64_t a = 0, lo = 0, hi = 0;
    for (int w = 0; w < NT / 64; ++w) {
      a += red[w][k][0];
      add128(lo, hi, red[w][k][1], red[w][k][2]);
    }
    const int64_t slot = ((int64_t)k * ng + g) * B.rows + r;  // [k][group][row]
    reinterpret_cast<uint64_t *>(B.part)[slot] = a;
    reinterpret_cast<uint64_t *>(B.part2)[2 * slot] = lo;
    reinterpret_cast<uint64_t *>(B.part2)[2 * slot + 1] = hi;
  }
}

// Threshold passes.  A workgroup owns HIT_IPW consecutive (row, block) items
// (item = row * nblk + block, so a width's slots of the items are
// consecutive words):
//  P_COUNT: crossings per width; widths whose block maximum cannot cross the
//           threshold are 0 without reading D
//  P_WRITE: candidates of the block, ordered by (width, row, sample)
// One thread per item first finds the widths that can have crossings (the
// block maxima, or the counts of the count pass); the workgroup then walks
// only the items that have one.  Around a few bright pulses most blocks have
// none, and they cost one load round per HIT_IPW items instead of a
// workgroup each.
// LX: 0, or the block's span SB + maxW when it is a multiple of NT (every
// load, store and prefix segment full, nothing predicated)
constexpr int HIT_IPW = 16;
static_assert(HIT_IPW <= 64, "one wave selects the items");

template <bool HELLA, int LX = 0>
__global__ __launch_bounds__(NT) void k_hits(const SearchBeamDev *__restrict__ beams,
                                             SearchCommon C, int pass)
{
  const SearchBeamDev B = beams[blockIdx.z];
  const int64_t plane = (int64_t)B.rows * C.nblk;  // slots of one width
  const int64_t it0 = (int64_t)blockIdx.x * HIT_IPW;
  if (it0 >= plane) return;
  const int l = threadIdx.x;
  __shared__ __align__(16) uint32_t P[PN];
  __shared__ uint32_t wtot[NT / 64];
  __shared__ uint32_t shu[NT];
  __shared__ uint32_t shc[SEARCH_MAXW];
  __shared__ uint32_t s_it[HIT_IPW], s_can[HIT_IPW], s_n;
  const SScalK &sc = *(SScalK *)B.scal;
  const int c0 = C.maxW / 2;
  const float thr = sc.thr;

  // the items' widths that can have crossings (every width's words are
  // requested before the first is used); the items with any are listed in
  // item order
  if (l < 64) {
    const int64_t it = it0 + l;
    const bool mine = l < HIT_IPW && it < plane;
    const int64_t itc = mine ? it : it0;  // loads stay inside the slots
    uint32_t can = 0;
    if (pass == P_COUNT) {
      if (!sc.gated) {
        // each item's four wave maxima of a width in one 16-byte load
        typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
        u32x4v bm[SEARCH_MAXW];
#pragma unroll
        for (int k = 0; k < SEARCH_MAXW; ++k)
          bm[k] = gld(reinterpret_cast<const u32x4v *>(B.bmax) + min(k, C.nw - 1) * plane + itc);
#pragma unroll
        for (int k = 0; k < SEARCH_MAXW; ++k) {
          if (k < C.nw) {
            const uint32_t m = max(max(bm[k].x, bm[k].y), max(bm[k].z, bm[k].w));
            const bool c = HELLA ? normed((double)__uint_as_float(m), sc, k) > thr
                                 : (uint64_t)m >= sc.rthr[k];
            can |= c ? (1u << k) : 0u;
          }
        }
      }
      if (mine && !can)
        for (int k = 0; k < C.nw; ++k) B.counts[k * plane + it] = 0u;
    } else {
      uint32_t cn[SEARCH_MAXW];
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k) cn[k] = gld(B.counts + min(k, C.nw - 1) * plane + itc);
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k)
        if (k < C.nw) can |= cn[k] ? (1u << k) : 0u;
    }
    can = mine ? can : 0u;
    const uint64_t act = __ballot(can != 0u);
    if (can) {
      const int q = __popcll(act & ((1ull << l) - 1ull));
      s_it[q] = (uint32_t)l;
      s_can[q] = can;
    }
    if (l == 0) s_n = (uint32_t)__popcll(act);
  }
  __syncthreads();
  const int nact = (int)__builtin_amdgcn_readfirstlane(s_n);

  for (int q = 0; q < nact; ++q) {
    const int64_t it = it0 + (int64_t)__builtin_amdgcn_readfirstlane(s_it[q]);
    const uint32_t can = __builtin_amdgcn_readfirstlane(s_can[q]);
    const int r = (int)(it / C.nblk);
    const int b = (int)(it - (int64_t)r * C.nblk);
    const int base = b * SB;
    const uint32_t *drow = B.D + (int64_t)(r + B.row0) * B.ds;
    if (!HELLA) {
      // the block and its halo: every load issued before the first LDS store
      const int L = SB + C.maxW;
      constexpr int PFX = LX ? LX / NT : PF;
      uint32_t v[PFX];
#pragma unroll
      for (int t = 0; t < PFX; ++t) v[t] = gld(drow + min(base + l + NT * t, C.ntime_dd - 1));
#pragma unroll
      for (int t = 0; t < PFX; ++t) {
        const int i = l + NT * t;
        if (LX || i < L) P[i + 1] = base + i < C.ntime_dd ? v[t] : 0u;
      }
      if (l == 0) P[0] = 0u;
      __syncthreads();
      if constexpr (LX != 0)
        block_prefix<PFX, true>(P, L, wtot);
      else
        block_prefix(P, L, wtot);
    }
    auto raw = [&](int k, int j) -> double {
      if (HELLA) return (double)gld(B.planes + ((int64_t)k * B.rows + r) * B.ps + base + j);
      const int w = C.minW << k;
      const int a = j + c0 - w / 2;
      return (double)(uint32_t)(P[a + w] - P[a]);
    };

    if (pass == P_COUNT) {
      if (l < SEARCH_MAXW) shc[l] = 0u;
      __syncthreads();
      uint32_t cnt[SEARCH_MAXW];
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k) cnt[k] = 0;
      for (int kk = 0; kk < SB / NT; ++kk) {
        const int j = l + NT * kk;
        if (base + j >= C.nout) break;
#pragma unroll
        for (int k = 0; k < SEARCH_MAXW; ++k)
          if (k < C.nw && ((can >> k) & 1u)) {
            if (HELLA)
              cnt[k] += normed(raw(k, j), sc, k) > thr ? 1u : 0u;
            else
              cnt[k] += (uint64_t)raw(k, j) >= sc.rthr[k] ? 1u : 0u;
          }
      }
      // integer totals are order free: wave sums, then LDS atomics
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k) {
        if (k < C.nw) {
          const uint32_t x = wave_sum(cnt[k]);
          if ((l & 63) == 0 && x) atomicAdd(&shc[k], x);
        }
      }
      __syncthreads();
      if (l < C.nw) B.counts[l * plane + it] = shc[l];
    } else {
      // P_WRITE: ordered by (width, row, sample).  Per width, every wave
      // keeps the ballot of each 256-sample step and publishes its hit
      // count; after one barrier a hit's rank is the hits of the earlier
      // steps, of the earlier waves of its step, and of the lower lanes of
      // its wave.
      constexpr int KS = SB / NT;
      const int lane = l & 63, wv = l >> 6;
      const uint64_t below = (1ull << lane) - 1ull;
      for (int k = 0; k < C.nw; ++k) {
        if (!((can >> k) & 1u)) continue;
        uint64_t m[KS];
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
          const int j = l + NT * kk;
          bool hit = false;
          if (base + j < C.nout) {
            const double v = raw(k, j);
            hit = HELLA ? normed(v, sc, k) > thr : (uint64_t)v >= sc.rthr[k];
          }
          m[kk] = __ballot(hit);
          if (lane == 0) shu[kk * (NT / 64) + wv] = (uint32_t)__popcll(m[kk]);
        }
        __syncthreads();
        uint32_t pos = *(const __attribute__((address_space(4))) uint32_t *)(B.offs + k * plane + it);
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
          uint32_t tot = 0, before = 0;
#pragma unroll
          for (int w = 0; w < NT / 64; ++w) {
            const uint32_t cw = shu[kk * (NT / 64) + w];
            tot += cw;
            before += w < wv ? cw : 0u;
          }
          if ((m[kk] >> lane) & 1ull) {
            const uint32_t rank = before + (uint32_t)__popcll(m[kk] & below);
            if (pos + rank < (uint32_t)C.cap) {
              const int j = l + NT * kk;
              Cand c;
              c.snr = normed(raw(k, j), sc, k);
              c.samp = base + j;
              c.width = k;
              c.dm_idx = r + B.row0;
              B.cands[pos + rank] = c;
            }
          }
          pos += tot;
        }
        __syncthreads();  // shu is rewritten by the next width
      }
    }
    __syncthreads();  // P and shc are rewritten by the next item
  }
}

// Statistics from the block sums (calculate_stddev, normalization.cu:125-162):
// S1_r, S2_r = the row's sums of x and x*x; mu = canonical sum over rows of
// S1_r / n, divided by the row count; the row variance about mu is
// ((S2_r - 2 mu S1_r) + n mu mu) / n, combined the same way.
// Row sums: BOXCAR statistics hold exact integer partials per group of blocks
// (S1 u64, S2 u128 as lo/hi), summed exactly and rounded to double once
// (S2 = hi * 2^64 + lo); otherwise the double block partials are added in
// ascending block order.
__device__ __forceinline__ void row_sums(const SearchBeamDev &B, const SearchCommon &C, int k, int r,
                                         bool exact, int ngroups, bool want2, double &s1, double &s2)
{
  // the partials are loaded RS at a time (clamped, unconditional) before any
  // is added, and added in ascending order: one memory latency per RS
  // partials instead of one per partial
  constexpr int RS = 8;
  if (exact) {
    const uint64_t *p1 = reinterpret_cast<const uint64_t *>(B.part);
    const uint64_t *p2 = reinterpret_cast<const uint64_t *>(B.part2);
    // [k][group][row]: the lanes (consecutive rows) read consecutive words
    const int64_t base = (int64_t)k * ngroups * B.rows + r;
    uint64_t a = 0, lo = 0, hi = 0;
    for (int g0 = 0; g0 < ngroups; g0 += RS) {
      uint64_t x1[RS], xl[RS], xh[RS];
#pragma unroll
      for (int j = 0; j < RS; ++j) {
        const int64_t g = base + (int64_t)min(g0 + j, ngroups - 1) * B.rows;
        x1[j] = gld(p1 + g);
        xl[j] = want2 ? gld(p2 + 2 * g) : 0ull;
        xh[j] = want2 ? gld(p2 + 2 * g + 1) : 0ull;
      }
#pragma unroll
      for (int j = 0; j < RS; ++j) {
        if (g0 + j < ngroups) {
          a += x1[j];
          add128(lo, hi, xl[j], xh[j]);
        }
      }
    }
    s1 = (double)a;
    s2 = want2 ? (double)hi * 18446744073709551616.0 + (double)lo : 0.0;
    return;
  }
  // [k][block][row]: the lanes (consecutive rows) read consecutive words
  const int64_t base = (int64_t)k * C.nblk * B.rows + r;
  s1 = 0.0;
  s2 = 0.0;
  for (int b0 = 0; b0 < C.nblk; b0 += RS) {
    double x1[RS], x2[RS];
#pragma unroll
    for (int j = 0; j < RS; ++j) {
      const int64_t b = base + (int64_t)min(b0 + j, C.nblk - 1) * B.rows;
      x1[j] = gld(B.part + b);
      x2[j] = want2 ? gld(B.part2 + b) : 0.0;
    }
#pragma unroll
    for (int j = 0; j < RS; ++j)
      if (b0 + j < C.nblk) s1 += x1[j];
    if (want2) {
#pragma unroll
      for (int j = 0; j < RS; ++j)
        if (b0 + j < C.nblk) s2 += x2[j];
    }
  }
}

// The row values of a block of SB rows are formed by all FIN_NT threads at
// once (rowv); lane l < 256 then adds rows l + 256k in ascending k and the
// lanes are folded by the canonical tree.
constexpr int FIN_NT = 1024;

// the canonical sum of rowv[0 .. nr) (one chunk of at most SB rows): lane
// l < 256 adds rows l + 256k in ascending k, then the lanes are folded by the
// canonical tree; every thread returns it
__device__ double fold_rows(const double *rowv, int nr, double *sh)
{
  const int l = threadIdx.x;
  __syncthreads();
  double acc[1] = {0.0};
  if (l < NT) {
    for (int kk = 0; kk < SB / NT; ++kk) {
      if (l + NT * kk >= nr) break;
      acc[0] += rowv[l + NT * kk];
    }
  }
  tree256_multi(acc, sh);
  if (l == 0) sh[256] = acc[0];
  __syncthreads();
  const double t = sh[256];
  __syncthreads();
  return t;
}

__device__ double rows_csum(const SearchBeamDev &B, const SearchCommon &C, int k, double mu,
                            bool var, bool exact, int ngroups, double *rowv, double *sh)
{
  const int l = threadIdx.x;
  const double n = (double)C.wp;
  double total = 0.0;
  for (int r0 = 0; r0 < B.rows; r0 += SB) {
    for (int i = l; i < SB; i += FIN_NT) {
      const int r = r0 + i;
      double val = 0.0;
      if (r < B.rows) {
        double s1, s2;
        row_sums(B, C, k, r, exact, ngroups, var, s1, s2);
        val = !var ? s1 / n : ((s2 - (2.0 * mu) * s1) + (n * mu) * mu) / n;
      }
      rowv[i] = val;
    }
    total += fold_rows(rowv, min(SB, B.rows - r0), sh);
  }
  return total / (double)B.rows;
}

// mu and the variance about it (calculate_stddev's two passes).  When every
// row fits one chunk, each thread loads its rows' S1 and S2 once and keeps
// them for the second pass; the row values and the folds are those of
// rows_csum, so the results are the same.
__device__ void rows_mu_var(const SearchBeamDev &B, const SearchCommon &C, int k, bool exact,
                            int ngroups, double *rowv, double *sh, double &mu, double &var)
{
  if (B.rows > SB) {
    mu = rows_csum(B, C, k, 0.0, false, exact, ngroups, rowv, sh);
    var = rows_csum(B, C, k, mu, true, exact, ngroups, rowv, sh);
    return;
  }
  constexpr int RPT = SB / FIN_NT;  // rows per thread
  const int l = threadIdx.x;
  const double n = (double)C.wp;
  double s1[RPT], s2[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    s1[q] = s2[q] = 0.0;
    if (l + FIN_NT * q < B.rows) row_sums(B, C, k, l + FIN_NT * q, exact, ngroups, true, s1[q], s2[q]);
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) rowv[l + FIN_NT * q] = l + FIN_NT * q < B.rows ? s1[q] / n : 0.0;
  mu = (0.0 + fold_rows(rowv, B.rows, sh)) / (double)B.rows;  // rows_csum's total, from 0.0
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    rowv[l + FIN_NT * q] =
        l + FIN_NT * q < B.rows ? ((s2[q] - (2.0 * mu) * s1[q]) + (n * mu) * mu) / n : 0.0;
  var = (0.0 + fold_rows(rowv, B.rows, sh)) / (double)B.rows;
}

__global__ __launch_bounds__(FIN_NT) void k_finish(const SearchBeamDev *__restrict__ beams,
                                                   SearchCommon C, int which, int ngroups)
{
  const SearchBeamDev B = beams[blockIdx.y];
  const int k = blockIdx.x;
  __shared__ double rowv[SB];
  __shared__ double sh[260];  // 128 tree slots + the broadcast slot 256
  const bool exact = which == P_STATS && C.mode == SEARCH_BOXCAR;
  double mu, var;
  rows_mu_var(B, C, k, exact, ngroups, rowv, sh, mu, var);
  SScal &sc = *B.scal;
  const float sf = (float)sqrt(var);
  if (which == P_STATS) {
    if (threadIdx.x != 0) return;
    sc.mu[k] = mu;
    sc.sigf[k] = sf;
    sc.inv[k] = 1.0 / (double)sf;
    return;
  }
  const float thr = C.snr * sf;
  const int kw = threadIdx.x;
  if (kw < C.nw && C.mode == SEARCH_BOXCAR) {
    // binary search for the first integer sum that crosses the threshold; the
    // width's scalars are read once, not in each of the 33 steps (the steps
    // evaluate normed() on these same values)
    struct {
      double mu[1], inv[1];
      float sigf[1];
    } one = {{sc.mu[kw]}, {sc.inv[kw]}, {sc.sigf[kw]}};
    uint64_t lo = 0, hi = 1ull << 32;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (normed((double)mid, one, 0) > thr) hi = mid;
      else lo = mid + 1;
    }
    sc.rthr[kw] = lo;
  }
  if (threadIdx.x != 0) return;
  sc.mu0 = mu;
  sc.sig0 = sf;
  // peaks.cu:61, definitions.h:69-72: search only when 0.1 < sigma < 5
  sc.gated = (sf < 5.f && sf > 0.1f) ? 0 : 1;
  sc.thr = thr;
}

// ---- ordered compaction offsets: exclusive scan of counts per beam (tile
// totals, then each tile's scan from the totals before it)
__device__ __forceinline__ int64_t scan_len(const SearchBeamDev &B, const SearchCommon &C)
{
  return (int64_t)C.nw * B.rows * C.nblk;
}

__global__ __launch_bounds__(NT) void k_scan_tiles(const SearchBeamDev *__restrict__ beams,
                                                   SearchCommon C)
{
  const SearchBeamDev B = beams[blockIdx.y];
  const int64_t n = scan_len(B, C);
  const int64_t t0 = (int64_t)blockIdx.x * SB;
  if (t0 >= n) return;
  __shared__ uint32_t sh[NT];
  uint32_t s = 0;
  for (int i = threadIdx.x; i < SB; i += NT)
    if (t0 + i < n) s += B.counts[t0 + i];
  uint32_t tot;
  (void)block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) B.tiles[blockIdx.x] = tot;
}

__global__ __launch_bounds__(NT) void k_scan_down(const SearchBeamDev *__restrict__ beams,
                                                  SearchCommon C)
{
  const SearchBeamDev B = beams[blockIdx.y];
  const int64_t n = scan_len(B, C);
  const int64_t t0 = (int64_t)blockIdx.x * SB;
  if (t0 >= n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *B.ncand = 0u;  // a beam with nothing to scan
    return;
  }
  __shared__ uint32_t sh[NT];
  __shared__ uint32_t red[2][NT / 64];
  // this tile's offset is the sum of the earlier tiles' totals (k_scan_tiles);
  // the first tile's workgroup also stores the beam's crossings (u32 sums:
  // any order gives the same words)
  const int nt = (int)((n + SB - 1) / SB);
  uint32_t pre = 0, all = 0;
  for (int i = threadIdx.x; i < nt; i += NT) {
    const uint32_t v = B.tiles[i];
    all += v;
    pre += i < (int)blockIdx.x ? v : 0u;
  }
  pre = wave_sum(pre);
  all = wave_sum(all);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = pre;
    red[1][threadIdx.x >> 6] = all;
  }
  __syncthreads();
  uint32_t tile_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    tile_off += red[0][w];
    total += red[1][w];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.ncand = total;
  constexpr int PER = SB / NT;
  const int64_t i0 = t0 + (int64_t)threadIdx.x * PER;
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    v[i] = (i0 + i < n) ? B.counts[i0 + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, sh, &tot) + tile_off;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    if (i0 + i < n) B.offs[i0 + i] = run;
    run += v[i];
  }
}

// hella smoothing (smooth.cpp:42-87): plane[k][r][s] =
//   sum_{i<3} sum_{j<=2sm} K[i][j] * D[r+row0+i-1][min(s + maxW/2 + j, ntime_dd-1)]
__global__ __launch_bounds__(NT) void k_hella_fir(const SearchBeamDev *__restrict__ beams,
                                                  SearchCommon C, int k, int sm,
                                                  const float *__restrict__ kern)
{
  const SearchBeamDev B = beams[blockIdx.z];
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (r >= B.rows || s >= C.nout) return;
  const int n = 2 * sm + 1;
  const int c0 = C.maxW / 2;
  float acc = 0.f;
  for (int i = 0; i < 3; ++i) {
    const uint32_t *row = B.D + (int64_t)(r + B.row0 + i - 1) * B.ds;
    for (int j = 0; j < n; ++j) {
      int x = s + c0 + j;
      if (x > C.ntime_dd - 1) x = C.ntime_dd - 1;
      const float t = kern[i * n + j] * (float)row[x];
      acc = acc + t;
    }
  }
  B.planes[((int64_t)k * B.rows + r) * B.ps + s] = acc;
}

}  // namespace

int search_nwidth(int mode, int minW, int maxW)
{
  int n = 0;
  if (minW <= 0) return 0;
  if (mode == SEARCH_HELLA)
    for (int sm = minW; sm < maxW; sm *= 2) ++n;  // casm:186-191
  else
    for (int w = minW; w <= maxW; w *= 2) ++n;
  return n;
}

// smooth.cpp:45-58 (the polynomial approximation of a squared Gaussian,
// DM weights 0.3 / 1 / 0.3, normalised to unit sum)
void hella_kernel_host(int sm, std::vector<float> &k)
{
  const float w[3] = {0.3f, 1.f, 0.3f};
  const int n = 2 * sm + 1;
  k.assign((size_t)3 * n, 0.f);
  float filtSum = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < n; ++j) {
      const double u = (j - sm) / (sm / 2.355);
      const float v = (float)(1. - 0.5 * u * u + 0.25 * u * u * 0.25 * u * u -
                              0.083 * u * u * 0.083 * u * u * 0.083 * u * u);
      k[(size_t)i * n + j] = w[i] * v * v;
      filtSum += w[i] * v * v;
    }
  for (auto &x : k) x /= filtSum;
}

template <bool HELLA>
static void launch_stats(const SearchBeamDev *d_beams, const SearchCommon &C, int pass, int bpw,
                         dim3 grid, hipStream_t s)
{
  const int nw = pass == P_Y0 ? 1 : C.nw;
  switch (nw) {
#define HAMD_STATS_CASE(N)                                                                 \
  case N:                                                                                  \
    hipLaunchKernelGGL((k_stats<HELLA, N>), grid, dim3(NT), 0, s, d_beams, C, pass, bpw); \
    break;
    HAMD_STATS_CASE(1) HAMD_STATS_CASE(2) HAMD_STATS_CASE(3) HAMD_STATS_CASE(4)
    HAMD_STATS_CASE(5) HAMD_STATS_CASE(6) HAMD_STATS_CASE(7) HAMD_STATS_CASE(8)
    HAMD_STATS_CASE(9) HAMD_STATS_CASE(10) HAMD_STATS_CASE(11) HAMD_STATS_CASE(12)
#undef HAMD_STATS_CASE
    default:
      throw HipError{hipErrorInvalidValue, "width count", __LINE__};
  }
}

static void launch_stats_box(const SearchBeamDev *d_beams, const SearchCommon &C, int bpw,
                             dim3 grid, hipStream_t s)
{
  // a thread adds at most 16 box sums and 16 squares per block, over bpw blocks
  const unsigned __int128 bm = (unsigned __int128)(uint64_t)std::max<int64_t>(C.box_max, 0);
  const bool narrow = C.box_max > 0 && C.box_max < (1ll << 28) &&
                      bm * bm * 16u * (unsigned)bpw < ((unsigned __int128)1 << 64);
  switch (C.nw) {
#define HAMD_BOX_CASE(N)                                                                        \
  case N:                                                                                       \
    if (C.minW == 1 && narrow)                                                                  \
      hipLaunchKernelGGL((k_stats_box<N, true, true>), grid, dim3(NT), 0, s, d_beams, C, bpw);   \
    else if (C.minW == 1)                                                                       \
      hipLaunchKernelGGL((k_stats_box<N, true, false>), grid, dim3(NT), 0, s, d_beams, C, bpw);  \
    else if (narrow)                                                                            \
      hipLaunchKernelGGL((k_stats_box<N, false, true>), grid, dim3(NT), 0, s, d_beams, C, bpw);  \
    else                                                                                        \
      hipLaunchKernelGGL((k_stats_box<N, false, false>), grid, dim3(NT), 0, s, d_beams, C, bpw); \
    break;
    HAMD_BOX_CASE(1) HAMD_BOX_CASE(2) HAMD_BOX_CASE(3) HAMD_BOX_CASE(4)
    HAMD_BOX_CASE(5) HAMD_BOX_CASE(6) HAMD_BOX_CASE(7) HAMD_BOX_CASE(8)
    HAMD_BOX_CASE(9) HAMD_BOX_CASE(10) HAMD_BOX_CASE(11) HAMD_BOX_CASE(12)
#undef HAMD_BOX_CASE
    default:
      throw HipError{hipErrorInvalidValue, "width count", __LINE__};
  }
}

// blocks per workgroup of the statistics passes: enough workgroups for a few
// per CU, the rest of a row's blocks walked with register prefetch
static int stats_bpw(const SearchCommon &C, int max_rows, int nbeam)
{
  const long long wgs_wanted = 256LL * 24;
  const long long per = (long long)C.nblk * max_rows * nbeam / wgs_wanted;
  return (int)std::max(1LL, std::min((long long)C.nblk, per));
}

void SearchRunner::run(const SearchBeamDev *d_beams, int nbeam, int max_rows,
                       const SearchCommon &C, const float *d_kern, const std::vector<int> &kofs,
                       hipStream_t s)
{
  smooth(d_beams, nbeam, max_rows, C, d_kern, kofs, s);
  peaks(d_beams, nbeam, max_rows, C, s);
}

void SearchRunner::smooth(const SearchBeamDev *d_beams, int nbeam, int max_rows,
                          const SearchCommon &C, const float *d_kern, const std::vector<int> &kofs,
                          hipStream_t s)
{
  if (nbeam <= 0 || max_rows <= 0 || C.nout <= 0) return;
  if (C.nw > SEARCH_MAXW || C.maxW > MAXBOX) throw HipError{hipErrorInvalidValue, "widths", __LINE__};
  const bool hella = C.mode == SEARCH_HELLA;
  const int bpw = stats_bpw(C, max_rows, nbeam);
  const dim3 gp((unsigned)((C.nblk + bpw - 1) / bpw), (unsigned)max_rows, (unsigned)nbeam);
  if (hella) {
    for (int k = 0; k < C.nw; ++k) {
      const int sm = C.minW << k;
      hipLaunchKernelGGL(k_hella_fir, dim3((unsigned)((C.nout + NT - 1) / NT), (unsigned)max_rows,
                                           (unsigned)nbeam),
                         dim3(NT), 0, s, d_beams, C, k, sm, d_kern + kofs[(size_t)k]);
    }
  }
  if (hella)
    launch_stats<true>(d_beams, C, (int)P_STATS, bpw, gp, s);
  else
    launch_stats_box(d_beams, C, bpw, gp, s);
  hipLaunchKernelGGL(k_finish, dim3((unsigned)C.nw, (unsigned)nbeam), dim3(FIN_NT), 0, s, d_beams, C,
                     (int)P_STATS, (int)gp.x);
  HAMD_CHECK(hipGetLastError());
}

void SearchRunner::peaks(const SearchBeamDev *d_beams, int nbeam, int max_rows,
                         const SearchCommon &C, hipStream_t s)
{
  if (nbeam <= 0 || max_rows <= 0 || C.nout <= 0) return;
  const bool hella = C.mode == SEARCH_HELLA;
  const int bpw = stats_bpw(C, max_rows, nbeam);
  const dim3 gs((unsigned)((C.nblk + bpw - 1) / bpw), (unsigned)max_rows, (unsigned)nbeam);
  const dim3 gp((unsigned)C.nblk, (unsigned)max_rows, (unsigned)nbeam);
  // threshold passes: HIT_IPW (row, block) items per workgroup
  const int64_t nitem = (int64_t)max_rows * C.nblk;
  const dim3 gh((unsigned)((nitem + HIT_IPW - 1) / HIT_IPW), 1u, (unsigned)nbeam);
  auto pass = [&](int p) {
    if (p == P_Y0) {
      if (hella)
        launch_stats<true>(d_beams, C, p, bpw, gs, s);
      else if (C.minW == 1)
        // one block per workgroup: a plain stream over the plane with every
        // block's loads in flight at once (its partials are per block)
        hipLaunchKernelGGL(k_y0_direct, gp, dim3(NT), 0, s, d_beams, C, 1);
      else
        launch_stats<false>(d_beams, C, p, bpw, gs, s);
    } else if (hella) {
      hipLaunchKernelGGL(k_hits<true>, gh, dim3(NT), 0, s, d_beams, C, p);
    } else {
      // the usual widest boxes get the kernel with every block span full
      const int lx = SB + C.maxW;
      if (lx == SB + 256)
        hipLaunchKernelGGL((k_hits<false, SB + 256>), gh, dim3(NT), 0, s, d_beams, C, p);
      else if (lx == SB + 512)
        hipLaunchKernelGGL((k_hits<false, SB + 512>), gh, dim3(NT), 0, s, d_beams, C, p);
      else if (lx == SB + 1024)
        hipLaunchKernelGGL((k_hits<false, SB + 1024>), gh, dim3(NT), 0, s, d_beams, C, p);
      else
        hipLaunchKernelGGL(k_hits<false>, gh, dim3(NT), 0, s, d_beams, C, p);
    }
  };
  auto fin = [&](int p) {
    hipLaunchKernelGGL(k_finish, dim3(1u, (unsigned)nbeam), dim3(FIN_NT), 0, s, d_beams, C, p, (int)gs.x);
  };
  pass(P_Y0);
  fin(P_Y0);
  pass(P_COUNT);
  const int64_t n = (int64_t)C.nw * max_rows * C.nblk;
  const unsigned ntiles = (unsigned)((n + SB - 1) / SB);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C);
  hipLaunchKernelGGL(k_scan_down, dim3(ntiles, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C);
  pass(P_WRITE);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
