// Candidate clustering on the device: the crossings of a beam's gulp fall into
// cells of CLUSTER_DT samples x CLUSTER_DDM DM trials; 8-connected occupied
// cells form one cluster, represented by its highest-S/N member (lowest list
// index on ties); clusters are listed by representative index.  One
// workgroup per beam:
//   1. every candidate marks its cell with the lowest candidate index in it
//      (atomic minimum into the beam's cell grid, issued only by the first
//      lane of each run of a wave's lanes that share a cell);
//   2. the first candidate of each cell joins the cells around it (union-find
//      over candidate indices, the larger root hooked under the smaller by
//      compare-and-swap, so a component's root is its lowest cell index);
//   3. every candidate folds itself into its root: the (S/N, -index) key by
//      atomic maximum, the count, and the sample and DM bounds; lanes of a
//      wave that share a root (neighbouring candidates mostly do) first
//      combine their values by DPP reductions over the wave, so a large
//      cluster does not serialise one atomic per candidate on its root;
//   4. the candidates that represent their cluster are compacted in index
//      order (wave ballots, workgroup scan).
// The candidates taken are those the host keeps: the first min(crossings,
// CAND_CAP, MAX_GIANTS) of the beam, and within a gulp's list no more than
// MAX_GIANTS in all (the beams before this one counted in order).
// The union-find parents live in LDS for lists of up to PCAP candidates, the
// other per-candidate words (keys, counts, bounds) for up to WCAP; beyond,
// the work set's buffers hold them; the same generic pointers serve both.
// A beam's words and cell grid are touched by its workgroup alone during the
// launch, so every atomic, and every read of a word other waves change by
// atomics, is a workgroup-scope atomic (a device-scope one would have to be
// made coherent across the chip's L2s), and the phases are ordered by the
// workgroup barriers alone.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "reduce.hpp"
#include "search.hpp"

namespace hamd {
namespace {

constexpr int CT = 1024;   // threads per workgroup
constexpr int PCAP = 16384;  // candidates whose parents fit in LDS (4 B each)
constexpr int WCAP = 3072;   // candidates whose other words fit in LDS (28 B each)
constexpr size_t CLUSTER_LDS = (size_t)WCAP * 8 + (size_t)PCAP * 4 + (size_t)5 * WCAP * 4;
constexpr uint32_t NONE = 0xffffffffu;

constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;

__device__ __forceinline__ uint32_t ld(const uint32_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, WG);
}

__device__ __forceinline__ unsigned long long ld64(const unsigned long long *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, WG);
}

template <class T>
__device__ __forceinline__ void amin(T *p, T v)
{
  (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, WG);
}

template <class T>
__device__ __forceinline__ void amax(T *p, T v)
{
  (void)__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, WG);
}

template <class T>
__device__ __forceinline__ void aadd(T *p, T v)
{
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, WG);
}

// compare-and-swap; the word's value before the operation
__device__ __forceinline__ uint32_t acas(uint32_t *p, uint32_t expect, uint32_t want)
{
  __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED, WG);
  return expect;
}

// root of x: parents only ever point to lower indices
__device__ __forceinline__ uint32_t find(uint32_t *par, uint32_t x)
{
  uint32_t p = ld(par + x);
  while (p != x) {
    x = p;
    p = ld(par + x);
  }
  return x;
}

__device__ __forceinline__ void unite(uint32_t *par, uint32_t a, uint32_t b)
{
  while (true) {
    a = find(par, a);
    b = find(par, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // hook the larger root under the smaller; retry if a was hooked meanwhile
    if (acas(par + a, a, b) == a) return;
  }
}

// p, known to the compiler as a global-memory address (a generic pointer read
// from a descriptor is otherwise accessed by flat instructions, which also
// count on the LDS counter and serialise with the LDS traffic around them)
template <class T>
__device__ __forceinline__ T *as_global(T *p)
{
  // through an integer: a cast of the generic pointer itself would be folded away
  typedef __attribute__((address_space(1))) T gT;
  return (T *)(gT *)(uintptr_t)p;
}

// float -> u32 with the same order (every S/N here is finite)
__device__ __forceinline__ uint32_t ordered(float f)
{
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ uint32_t scan_excl(uint32_t v, uint32_t *sh, uint32_t *total)
{
  const int l = threadIdx.x, lane = l & 63, wv = l >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) sh[wv] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < CT / 64; ++w) {
    const uint32_t c = sh[w];
    tot += c;
    before += w < wv ? c : 0u;
  }
  *total = tot;
  __syncthreads();
  return before + incl - v;
}

// The four phases and the grid clean-up.  RES: the list has at most
// CT x RES_PT candidates, so each thread keeps its candidates (S/N, sample,
// DM row), their cells and, after phase 1, their cells' heads in registers,
// loaded once: the later phases read no candidate or grid word from memory,
// and every round's loads are issued together.  Otherwise each round reads
// them from memory.
constexpr int RES_PT = 10;

template <bool RES>
__device__ __forceinline__ void cluster_phases(const SearchBeamDev &B, const ClusterArgs &A, int n,
                                               uint32_t *par, unsigned long long *key,
                                               uint32_t *cnt, uint32_t *smin, uint32_t *smax,
                                               uint32_t *dmin, uint32_t *dmax, uint32_t *sh)
{
  const int l = threadIdx.x, lane = l & 63;
  const int nround = (n + CT - 1) / CT;
  constexpr int UNR = RES ? RES_PT : 1;  // rounds unrolled (all of them when resident)
  auto cell_of = [&](int samp, int dm) { return (dm / A.cd) * B.cl_cols + samp / A.ct; };
  float rsnr[RES ? RES_PT : 1];
  int rsamp[RES ? RES_PT : 1], rdm[RES ? RES_PT : 1], rcell[RES ? RES_PT : 1];
  uint32_t rhead[RES ? RES_PT : 1];
  if constexpr (RES) {
#pragma unroll
    for (int k = 0; k < RES_PT; ++k) {
      const int i = l + CT * k;
      if (i < n) {
        const Cand c = B.cands[i];
        rsnr[k] = c.snr;
        rsamp[k] = c.samp;
        rdm[k] = c.dm_idx;
        rcell[k] = cell_of(c.samp, c.dm_idx);
      }
    }
  }
  // candidate i = l + CT * k of this thread, and its cell
  auto cand = [&](int k, float &snr, int &samp, int &dm, int &cl) {
    if constexpr (RES) {
      snr = rsnr[k];
      samp = rsamp[k];
      dm = rdm[k];
      cl = rcell[k];
    } else {
      const Cand c = B.cands[l + CT * k];
      snr = c.snr;
      samp = c.samp;
      dm = c.dm_idx;
      cl = cell_of(samp, dm);
    }
  };

  // 1. cells, and every candidate's words (only the cell heads' are used)
#pragma unroll UNR
  for (int k = 0; k < (RES ? RES_PT : nround); ++k) {
    const int i = l + CT * k;
    if (i >= n) break;
    float snr;
    int samp, dm, cl;
    cand(k, snr, samp, dm, cl);
    // the lanes of a wave hold consecutive candidates, and the list is
    // ordered by (width, DM row, sample), so neighbours mostly share a cell:
    // a lane whose predecessor is in the same cell leaves the cell's minimum
    // to it (one atomic per run of a cell instead of one per candidate; the
    // atomics of a cell's run all serialise on one grid word)
    const int prev = __builtin_amdgcn_update_dpp(0, cl, 0x138, 0xF, 0xF, false);  // wave_shr:1
    if (lane == 0 || prev != cl) amin(B.cl_grid + cl, (uint32_t)i);
    par[i] = (uint32_t)i;
    key[i] = 0ull;
    cnt[i] = 0u;
    smin[i] = NONE;
    smax[i] = 0u;
    dmin[i] = NONE;
    dmax[i] = 0u;
  }
  __syncthreads();
  // the head of every candidate's cell (fixed from here on)
  auto head = [&](int k, int cl) -> uint32_t {
    if constexpr (RES)
      return rhead[k];
    else
      return ld(B.cl_grid + cl);
  };
  if constexpr (RES) {
#pragma unroll
    for (int k = 0; k < RES_PT; ++k)
      if (l + CT * k < n) rhead[k] = ld(B.cl_grid + rcell[k]);
  }
  // 2. cell heads join the occupied cells around them
#pragma unroll UNR
  for (int k = 0; k < (RES ? RES_PT : nround); ++k) {
    const int i = l + CT * k;
    if (i >= n) break;
    float snr;
    int samp, dm, cl;
    cand(k, snr, samp, dm, cl);
    if (head(k, cl) != (uint32_t)i) continue;
    const int r = dm / A.cd, t = samp / A.ct;
    uint32_t g[8];
    int q = 0;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
      for (int dt = -1; dt <= 1; ++dt) {
        if (!dr && !dt) continue;
        const int rr = r + dr, tt = t + dt;
        g[q++] = (rr < 0 || rr >= B.cl_rows || tt < 0 || tt >= B.cl_cols)
                     ? NONE
                     : ld(B.cl_grid + rr * B.cl_cols + tt);
      }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (g[j] != NONE) unite(par, (uint32_t)i, g[j]);
  }
  __syncthreads();
  // 3. every candidate into its root (all lanes take part in every round)
#pragma unroll UNR
  for (int k = 0; k < (RES ? RES_PT : nround); ++k) {
    if (CT * k >= n) break;
    const int i = l + CT * k;
    bool act = i < n;
    uint32_t root = 0, sm = 0, dmv = 0;
    unsigned long long kv = 0ull;
    if (act) {
      float snr;
      int samp, dm, cl;
      cand(k, snr, samp, dm, cl);
      root = find(par, head(k, cl));
      kv = ((unsigned long long)ordered(snr) << 32) | (unsigned long long)(NONE - (uint32_t)i);
      sm = (uint32_t)samp;
      dmv = (uint32_t)dm;
    }
    // up to four groups of lanes with a common root, each combined over the
    // wave by DPP reductions (identities from the lanes outside the group);
    // the rest add directly
    for (int g = 0; g < 4; ++g) {
      const unsigned long long pend = __ballot(act);
      if (!pend) break;
      const int lead = __ffsll((long long)pend) - 1;
      const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)root, lead);
      const bool mine = act && root == r;
      const unsigned long long kk = wave_max_u64(mine ? kv : 0ull);
      const uint32_t cc = (uint32_t)__popcll(__ballot(mine));
      const uint32_t s0 = wave_min(mine ? sm : NONE), s1 = wave_max(mine ? sm : 0u);
      const uint32_t d0 = wave_min(mine ? dmv : NONE), d1 = wave_max(mine ? dmv : 0u);
      if (lane == lead) {
        amax(key + r, kk);
        aadd(cnt + r, cc);
        amin(smin + r, s0);
        amax(smax + r, s1);
        amin(dmin + r, d0);
        amax(dmax + r, d1);
      }
      act = act && !mine;
    }
    if (act) {
      amax(key + root, kv);
      aadd(cnt + root, 1u);
      amin(smin + root, sm);
      amax(smax + root, sm);
      amin(dmin + root, dmv);
      amax(dmax + root, dmv);
    }
  }
  __syncthreads();
  // 4. representatives in index order
  uint32_t run = 0;
#pragma unroll UNR
  for (int k = 0; k < (RES ? RES_PT : nround); ++k) {
    if (CT * k >= n) break;
    const int i = l + CT * k;
    uint32_t root = 0;
    bool rep = false;
    if (i < n) {
      float snr;
      int samp, dm, cl;
      cand(k, snr, samp, dm, cl);
      root = find(par, head(k, cl));
      rep = NONE - (uint32_t)(ld64(key + root) & 0xffffffffull) == (uint32_t)i;
    }
    uint32_t tot;
    const uint32_t slot = run + scan_excl(rep ? 1u : 0u, sh, &tot);
    if (rep) {
      ClusterRec o;
      o.rep = i;
      o.count = (int)ld(cnt + root);
      o.samp_min = (int)ld(smin + root);
      o.samp_max = (int)ld(smax + root);
      o.dm_min = (int)ld(dmin + root);
      o.dm_max = (int)ld(dmax + root);
      o.beam = 0;
      B.clusters[slot] = o;
    }
    run += tot;
  }
  if (l == 0) B.cl_n[0] = run;
  // the grid is left empty for the next gulp
  __syncthreads();
#pragma unroll UNR
  for (int k = 0; k < (RES ? RES_PT : nround); ++k) {
    const int i = l + CT * k;
    if (i >= n) break;
    float snr;
    int samp, dm, cl;
    cand(k, snr, samp, dm, cl);
    B.cl_grid[cl] = NONE;
  }
}

__global__ __launch_bounds__(CT) void k_cluster(const SearchBeamDev *__restrict__ beams,
                                                ClusterArgs A)
{
  const int b = blockIdx.x;
  const SearchBeamDev B = beams[b];
  const int l = threadIdx.x;
  extern __shared__ __align__(16) unsigned long long lds_key[];  // [WCAP] keys, [PCAP] parents, 5 x [WCAP]
  __shared__ uint32_t sh[CT / 64];
  // candidates taken for this beam
  auto taken = [&](const SearchBeamDev &X) {
    return min(min(*X.ncand, (uint32_t)A.cand_cap), (uint32_t)A.max_giants);
  };
  uint32_t before = (uint32_t)A.prefix;
  if (A.cumulative)
    for (int q = 0; q < b; ++q) {
      const uint32_t t = taken(beams[q]);
      const uint32_t left = before < (uint32_t)A.max_giants ? (uint32_t)A.max_giants - before : 0u;
      before += min(t, left);
    }
  const uint32_t left = before < (uint32_t)A.max_giants ? (uint32_t)A.max_giants - before : 0u;
  const int n = (int)min(taken(B), left);
  if (l == 0) B.cl_n[1] = (uint32_t)n;
  if (n == 0) {
    if (l == 0) B.cl_n[0] = 0u;
    return;
  }
  // every placement of the words is a call site of its own, so each pointer
  // has one known address space (LDS or global) in the code that uses it
  SearchBeamDev G = B;
  G.cands = as_global(B.cands);
  G.cl_grid = as_global(B.cl_grid);
  G.clusters = as_global(B.clusters);
  uint32_t *const lds_par = reinterpret_cast<uint32_t *>(lds_key + WCAP);
  uint32_t *const gwork = as_global(B.cl_work);
  unsigned long long *const gkey = as_global(B.cl_key);
  auto run = [&](auto res, uint32_t *par, unsigned long long *key, uint32_t *cnt, int cap)
                 __attribute__((always_inline)) {
    uint32_t *smin = cnt + cap;
    uint32_t *smax = smin + cap;
    uint32_t *dmin = smax + cap;
    uint32_t *dmax = dmin + cap;
    cluster_phases<decltype(res)::value>(G, A, n, par, key, cnt, smin, smax, dmin, dmax, sh);
  };
  // WCAP <= CT * RES_PT < PCAP: a list whose words fit in LDS is held in registers
  if (n <= WCAP)
    run(std::true_type{}, lds_par, lds_key, lds_par + PCAP, WCAP);
  else if (n <= CT * RES_PT)
    run(std::true_type{}, lds_par, gkey, gwork + B.cl_cap, B.cl_cap);
  else if (n <= PCAP)
    run(std::false_type{}, lds_par, gkey, gwork + B.cl_cap, B.cl_cap);
  else
    run(std::false_type{}, gwork, gkey, gwork + B.cl_cap, B.cl_cap);
}

}  // namespace

void cluster_kernel_setup()
{
  HAMD_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_cluster),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)CLUSTER_LDS));
}

namespace {

// A beam's results into pinned host memory, HV_WG workgroups striding over
// the lists: only what exists is written (one pass over PCIe, no copy-engine
// transfers)
constexpr int HV_WG = 8;

__global__ __launch_bounds__(256) void k_harvest(const SearchBeamDev *__restrict__ beam,
                                                 uint32_t *h_n, Cand *h_c, int cap_c,
                                                 ClusterRec *h_k, int cap_k)
{
  const SearchBeamDev B = *beam;
  const uint32_t nc = B.ncand[0], nk = B.cl_n[0];
  const int tc = (int)min(nc, (uint32_t)cap_c), tk = (int)min(nk, (uint32_t)cap_k);
  const int g = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  if (g < 3) h_n[g] = B.ncand[g];
  for (int j = g; j < tc; j += stride) h_c[j] = B.cands[j];
  for (int j = g; j < tk; j += stride) h_k[j] = B.clusters[j];
}

}  // namespace

void SearchRunner::harvest(const SearchBeamDev *d_beam, uint32_t *h_n, Cand *h_c, int cap_c,
                           ClusterRec *h_k, int cap_k, hipStream_t s)
{
  hipLaunchKernelGGL(k_harvest, dim3(HV_WG), dim3(256), 0, s, d_beam, h_n, h_c, cap_c, h_k, cap_k);
  HAMD_CHECK(hipGetLastError());
}

void SearchRunner::cluster(const SearchBeamDev *d_beams, int nbeam, const ClusterArgs &A,
                           hipStream_t s)
{
  if (nbeam <= 0) return;
  // the 148 KiB LDS limit was raised by kernels_setup() at set-up
  const size_t lds = CLUSTER_LDS;
  hipLaunchKernelGGL(k_cluster, dim3((unsigned)nbeam), dim3(CT), lds, s, d_beams, A);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
