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

#include <cmath>

namespace hamd {

namespace {

constexpr int NT = 256;
constexpr int SB = SEARCH_BLOCK;
constexpr int MAXBOX = 1024;

enum Pass { P_STATS = 0, P_Y0 = 1, P_COUNT = 2, P_WRITE = 3 };

// exclusive scan of one u32 per thread over the workgroup
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *total)
{
  const int l = threadIdx.x;
  sh[l] = v;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    const uint32_t t = (l >= off) ? sh[l - off] : 0u;
    __syncthreads();
    sh[l] += t;
    __syncthreads();
  }
  const uint32_t incl = sh[l];
  *total = sh[NT - 1];
  __syncthreads();
  return incl - v;
}

__device__ __forceinline__ float normed(double raw, const SScal &sc, int k)
{
  return sc.sigf[k] > 0.f ? (float)((raw - sc.mu[k]) * sc.inv[k]) : (float)raw;
}

template <bool HELLA>
__global__ __launch_bounds__(NT) void k_pass(const SearchBeamDev *__restrict__ beams,
                                             SearchCommon C, int pass)
{
  const SearchBeamDev &B = beams[blockIdx.z];
  const int r = blockIdx.y;
  if (r >= B.rows) return;
  const int b = blockIdx.x;
  const int base = b * SB;
  const int l = threadIdx.x;
  // P (prefix sums) and the reduction scratch share the LDS
  constexpr int PN = SB + MAXBOX + 1;
  constexpr int TREE_WORDS = 2 * SEARCH_MAXW * 128 * 2;
  __shared__ __align__(16) uint32_t lds_u[(PN > TREE_WORDS ? PN : TREE_WORDS)];
  __shared__ uint32_t shu[NT];
  __shared__ uint32_t shc[SEARCH_MAXW];
  uint32_t *P = lds_u;
  double *shd = reinterpret_cast<double *>(lds_u);
  const SScal &sc = *B.scal;
  const uint32_t *drow = B.D + (int64_t)(r + B.row0) * B.ds;
  const int c0 = C.maxW / 2;

  if (pass == P_COUNT && sc.gated) {
    for (int k = l; k < C.nw; k += NT) B.counts[((int64_t)k * B.rows + r) * C.nblk + b] = 0u;
    return;
  }
  if (pass == P_WRITE) {
    uint32_t any = 0;
    for (int k = 0; k < C.nw; ++k) any |= B.counts[((int64_t)k * B.rows + r) * C.nblk + b];
    if (!any) return;
  }
  // width 1 needs no prefix sums: its box is the dedispersed sample itself
  const bool direct = !HELLA && pass == P_Y0 && C.minW == 1;
  if (!HELLA && !direct) {
    // prefix sums of D[row][base .. base + L) in LDS (u32, wrap-around)
    const int L = SB + C.maxW;
    for (int i = l; i < L; i += NT) {
      const int col = base + i;
      P[i + 1] = col < C.ntime_dd ? drow[col] : 0u;
    }
    if (l == 0) P[0] = 0u;
    __syncthreads();
    const int seg = (L + NT - 1) / NT;
    const int i0 = 1 + l * seg, i1 = min(1 + (l + 1) * seg, L + 1);
    uint32_t run = 0;
    for (int i = i0; i < i1; ++i) run += P[i];
    uint32_t tot;
    const uint32_t off = block_excl_scan(run, shu, &tot);
    uint32_t acc = off;
    for (int i = i0; i < i1; ++i) {
      acc += P[i];
      P[i] = acc;
    }
    __syncthreads();
  }

  auto raw = [&](int k, int j) -> double {
    if (HELLA) return (double)B.planes[((int64_t)k * B.rows + r) * B.ps + base + j];
    if (direct) return (double)drow[base + j + c0];
    const int w = C.minW << k;
    const int a = j + c0 - w / 2;
    return (double)(uint32_t)(P[a + w] - P[a]);
  };

  if (pass == P_STATS || pass == P_Y0) {
    const int nwp = pass == P_Y0 ? 1 : C.nw;
    double acc[2 * SEARCH_MAXW];
#pragma unroll
    for (int k = 0; k < 2 * SEARCH_MAXW; ++k) acc[k] = 0.0;
    for (int kk = 0; kk < SB / NT; ++kk) {
      const int j = l + NT * kk;
      if (base + j >= C.wp) break;
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k) {
        if (k < nwp) {
          double v = raw(k, j);
          if (pass == P_Y0) v = (double)normed(v, sc, 0);
          acc[k] += v;
          acc[SEARCH_MAXW + k] += v * v;
        }
      }
    }
    __syncthreads();  // P is dead; its LDS becomes the tree scratch
    if (pass == P_Y0) {
      double two[2] = {acc[0], acc[SEARCH_MAXW]};
      tree256_multi(two, shd);
      acc[0] = two[0];
      acc[SEARCH_MAXW] = two[1];
    } else {
      tree256_multi(acc, shd);
    }
    if (l == 0) {
#pragma unroll
      for (int k = 0; k < SEARCH_MAXW; ++k) {
        if (k < nwp) {
          const int64_t slot = ((int64_t)k * B.rows + r) * C.nblk + b;
          B.part[slot] = acc[k];
          B.part2[slot] = acc[SEARCH_MAXW + k];
        }
      }
    }
    return;
  }

  const float thr = sc.thr;
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
        if (k < C.nw) {
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
        uint32_t x = cnt[k];
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) x += __shfl_down(x, s, 64);
        if ((l & 63) == 0 && x) atomicAdd(&shc[k], x);
      }
    }
    __syncthreads();
    if (l < C.nw) B.counts[((int64_t)l * B.rows + r) * C.nblk + b] = shc[l];
    return;
  }

  // P_WRITE: ordered by (width, row, sample)
  for (int k = 0; k < C.nw; ++k) {
    const int64_t slot = ((int64_t)k * B.rows + r) * C.nblk + b;
    if (B.counts[slot] == 0u) continue;
    uint32_t pos = B.offs[slot];
    for (int kk = 0; kk < SB / NT; ++kk) {
      const int j = l + NT * kk;
      const bool in = base + j < C.nout;
      float y = 0.f;
      bool hit = false;
      if (in) {
        const double v = raw(k, j);
        hit = HELLA ? normed(v, sc, k) > thr : (uint64_t)v >= sc.rthr[k];
        if (hit) y = normed(v, sc, k);
      }
      uint32_t tot;
      const uint32_t rank = block_excl_scan(hit ? 1u : 0u, shu, &tot);
      if (hit && pos + rank < (uint32_t)C.cap) {
        Cand c;
        c.snr = y;
        c.samp = base + j;
        c.width = k;
        c.dm_idx = r + B.row0;
        B.cands[pos + rank] = c;
      }
      pos += tot;
    }
  }
}

// Statistics from the block sums (calculate_stddev, normalization.cu:125-162):
// S1_r, S2_r = ascending sums of the row's block partials; mu = canonical
// sum over rows of S1_r / n, divided by the row count; the row variance about
// mu is ((S2_r - 2 mu S1_r) + n mu mu) / n, combined the same way.
__device__ double rows_csum(const SearchBeamDev &B, const SearchCommon &C, int k, double mu,
                            bool var, double *sh)
{
  const int l = threadIdx.x;
  const double n = (double)C.wp;
  double total = 0.0;
  for (int r0 = 0; r0 < B.rows; r0 += SB) {
    double acc[1] = {0.0};
    for (int kk = 0; kk < SB / NT; ++kk) {
      const int r = r0 + l + NT * kk;
      if (r >= B.rows) break;
      const int64_t base = ((int64_t)k * B.rows + r) * C.nblk;
      double s1 = 0.0;
      for (int b = 0; b < C.nblk; ++b) s1 += B.part[base + b];
      if (!var) {
        acc[0] += s1 / n;
      } else {
        double s2 = 0.0;
        for (int b = 0; b < C.nblk; ++b) s2 += B.part2[base + b];
        acc[0] += ((s2 - (2.0 * mu) * s1) + (n * mu) * mu) / n;
      }
    }
    tree256_multi(acc, sh);
    if (l == 0) sh[256] = acc[0];
    __syncthreads();
    total += sh[256];
    __syncthreads();
  }
  return total / (double)B.rows;
}

__global__ __launch_bounds__(NT) void k_finish(const SearchBeamDev *__restrict__ beams,
                                               SearchCommon C, int which)
{
  const SearchBeamDev &B = beams[blockIdx.y];
  const int k = blockIdx.x;
  __shared__ double sh[260];  // 128 tree slots + the broadcast slot 256
  const double mu = rows_csum(B, C, k, 0.0, false, sh);
  const double var = rows_csum(B, C, k, mu, true, sh);
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
    // binary search for the first integer sum that crosses the threshold
    uint64_t lo = 0, hi = 1ull << 32;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (normed((double)mid, sc, kw) > thr) hi = mid;
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

// ---- ordered compaction offsets: exclusive scan of counts per beam
__device__ __forceinline__ int64_t scan_len(const SearchBeamDev &B, const SearchCommon &C)
{
  return (int64_t)C.nw * B.rows * C.nblk;
}

__global__ __launch_bounds__(NT) void k_scan_tiles(const SearchBeamDev *__restrict__ beams,
                                                   SearchCommon C)
{
  const SearchBeamDev &B = beams[blockIdx.y];
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

__global__ __launch_bounds__(NT) void k_scan_top(const SearchBeamDev *__restrict__ beams,
                                                 SearchCommon C)
{
  const SearchBeamDev &B = beams[blockIdx.x];
  const int64_t n = scan_len(B, C);
  const int nt = (int)((n + SB - 1) / SB);
  __shared__ uint32_t sh[NT];
  const int seg = (nt + NT - 1) / NT;
  const int i0 = threadIdx.x * seg, i1 = min(i0 + seg, nt);
  uint32_t s = 0;
  for (int i = i0; i < i1; ++i) s += B.tiles[i];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, sh, &tot);
  for (int i = i0; i < i1; ++i) {
    const uint32_t v = B.tiles[i];
    B.tiles[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) *B.ncand = tot;
}

__global__ __launch_bounds__(NT) void k_scan_down(const SearchBeamDev *__restrict__ beams,
                                                  SearchCommon C)
{
  const SearchBeamDev &B = beams[blockIdx.y];
  const int64_t n = scan_len(B, C);
  const int64_t t0 = (int64_t)blockIdx.x * SB;
  if (t0 >= n) return;
  __shared__ uint32_t sh[NT];
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
  uint32_t run = block_excl_scan(s, sh, &tot) + B.tiles[blockIdx.x];
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
  const SearchBeamDev &B = beams[blockIdx.z];
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
  const dim3 gp((unsigned)C.nblk, (unsigned)max_rows, (unsigned)nbeam);
  if (hella) {
    for (int k = 0; k < C.nw; ++k) {
      const int sm = C.minW << k;
      hipLaunchKernelGGL(k_hella_fir, dim3((unsigned)((C.nout + NT - 1) / NT), (unsigned)max_rows,
                                           (unsigned)nbeam),
                         dim3(NT), 0, s, d_beams, C, k, sm, d_kern + kofs[(size_t)k]);
    }
  }
  if (hella)
    hipLaunchKernelGGL(k_pass<true>, gp, dim3(NT), 0, s, d_beams, C, (int)P_STATS);
  else
    hipLaunchKernelGGL(k_pass<false>, gp, dim3(NT), 0, s, d_beams, C, (int)P_STATS);
  hipLaunchKernelGGL(k_finish, dim3((unsigned)C.nw, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C,
                     (int)P_STATS);
  HAMD_CHECK(hipGetLastError());
}

void SearchRunner::peaks(const SearchBeamDev *d_beams, int nbeam, int max_rows,
                         const SearchCommon &C, hipStream_t s)
{
  if (nbeam <= 0 || max_rows <= 0 || C.nout <= 0) return;
  const bool hella = C.mode == SEARCH_HELLA;
  const dim3 gp((unsigned)C.nblk, (unsigned)max_rows, (unsigned)nbeam);
  auto pass = [&](int p) {
    if (hella)
      hipLaunchKernelGGL(k_pass<true>, gp, dim3(NT), 0, s, d_beams, C, p);
    else
      hipLaunchKernelGGL(k_pass<false>, gp, dim3(NT), 0, s, d_beams, C, p);
  };
  auto fin = [&](int p) {
    hipLaunchKernelGGL(k_finish, dim3(1u, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C, p);
  };
  pass(P_Y0);
  fin(P_Y0);
  pass(P_COUNT);
  const int64_t n = (int64_t)C.nw * max_rows * C.nblk;
  const unsigned ntiles = (unsigned)((n + SB - 1) / SB);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C);
  hipLaunchKernelGGL(k_scan_top, dim3((unsigned)nbeam), dim3(NT), 0, s, d_beams, C);
  hipLaunchKernelGGL(k_scan_down, dim3(ntiles, (unsigned)nbeam), dim3(NT), 0, s, d_beams, C);
  pass(P_WRITE);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
