// dedisp_kernel.hip -- brute-force integer dedispersion for gfx950.
//
// Replaces the kernel of the external dedisp library that casm-hella calls at
// src/hella/src/dedisperse.cpp:32-38:
//   out[d][t] = sum_{live c} in[c][t + delay(d,c)]      (exact integer sum)
//
// Layout: the input is channel-major ("FT") u8: row c holds the channel's
// samples; column x lives at (start + x) mod ring so a per-beam circular
// store can be read without shifting it.
//
// Work decomposition: one workgroup of 256 threads owns a block of ND
// consecutive DM trials x a tile of T output samples.  It walks the live
// channels in chunks of CC: each channel's window
//   [t0 + delay(d_first,c), t0 + T + delay(d_last,c) + 20)
// is staged in LDS with 16-byte loads, then every thread (one DM, V = 36
// consecutive samples) adds its bytes.  Four samples of one channel come
// from two LDS dwords; v_perm_b32 extracts the even and the odd samples into
// two u16x2 lanes at the thread's byte alignment, so one 32-bit add performs
// two sample additions.  u16 lanes hold at most 257*255, so they are widened
// into u32 totals every 256 channels.  V/4 = 9 is odd, which spreads the
// threads of one DM over all 32 LDS banks.
#include <type_traits>

#include "internal.hpp"

namespace hamd {

namespace {

constexpr int NT = 256;
constexpr int V = 36;          // samples per thread
constexpr int NW = V / 4 + 1;  // LDS dwords read per channel per thread

__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
  return __builtin_amdgcn_perm(hi, lo, sel);
}

template <int G>
__global__ __launch_bounds__(NT) void k_dedisp(const DDBeamDev *__restrict__ beams,
                                               const DDTileDesc *__restrict__ tiles, int chunk,
                                               int lrow)
{
  constexpr int ND = NT / G;
  constexpr int T = G * V;
  extern __shared__ __align__(16) uint8_t lds[];
  const DDTileDesc td = tiles[blockIdx.x];
  if (td.beam < 0) return;
  const DDBeamDev B = beams[td.beam];
  uint8_t *sdat = lds;                                                      // [chunk][lrow]
  int32_t *sdly = reinterpret_cast<int32_t *>(lds + (size_t)chunk * lrow);  // [chunk][ND]
  int32_t *soff = sdly + chunk * ND;                                        // [chunk]

  const int tid = threadIdx.x;
  const int dl = tid / G;
  const int g = tid % G;
  const int d0 = td.d0;
  const int t0 = td.t0;
  const int wave = tid >> 6, lane = tid & 63;
  const int nlive = B.nlive, ndm = B.ndm;

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;

  int since_flush = 0;
  for (int k0 = 0; k0 < nlive; k0 += chunk) {
    const int nc = min(chunk, nlive - k0);
    __syncthreads();
    for (int i = tid; i < nc * ND; i += NT) {
      const int cc = i / ND, dd = i - (i / ND) * ND;
      const int dm = min(d0 + dd, ndm - 1);
      sdly[cc * ND + dd] = B.delta[(int64_t)dm * nlive + k0 + cc];
    }
    __syncthreads();
    // stage windows: one wave per channel, 16-byte units
    for (int cc = wave; cc < nc; cc += NT / 64) {
      const int base_d = sdly[cc * ND];
      const int spread = sdly[cc * ND + ND - 1] - base_d;
      const int64_t first = (B.start + t0 + base_d) % B.ring;
      const int64_t al = first & ~(int64_t)15;
      const int off = (int)(first - al);
      const int n16 = (off + T + spread + 4 + 15) >> 4;
      const uint8_t *row = B.base + (int64_t)B.live[k0 + cc] * B.row_stride;
      uint8_t *dst = sdat + (size_t)cc * lrow;
      for (int u = lane; u < n16; u += 64) {
        int64_t col = al + 16 * (int64_t)u;
        if (col >= B.ring) col %= B.ring;
        *reinterpret_cast<uint4 *>(dst + 16 * u) = *reinterpret_cast<const uint4 *>(row + col);
      }
      if (lane == 0) soff[cc] = off;
    }
    __syncthreads();
    for (int cc = 0; cc < nc; ++cc) {
      const int p = g * V + soff[cc] + sdly[cc * ND + dl] - sdly[cc * ND];
      const uint32_t *wrow = reinterpret_cast<const uint32_t *>(sdat + (size_t)cc * lrow) + (p >> 2);
      const uint32_t s = (uint32_t)(p & 3);
      const uint32_t sel_e = s | 0x0c00u | ((s + 2u) << 16) | 0x0c000000u;
      const uint32_t sel_o = (s + 1u) | 0x0c00u | ((s + 3u) << 16) | 0x0c000000u;
      uint32_t w[NW];
#pragma unroll
      for (int i = 0; i < NW; ++i) w[i] = wrow[i];
#pragma unroll
      for (int k = 0; k < V / 4; ++k) {
        acc[2 * k] += perm(w[k + 1], w[k], sel_e);
        acc[2 * k + 1] += perm(w[k + 1], w[k], sel_o);
      }
    }
    since_flush += nc;
    if (since_flush >= 256 || k0 + chunk >= nlive) {
#pragma unroll
      for (int k = 0; k < V / 4; ++k) {
        tot[4 * k + 0] += acc[2 * k] & 0xffffu;
        tot[4 * k + 2] += acc[2 * k] >> 16;
        tot[4 * k + 1] += acc[2 * k + 1] & 0xffffu;
        tot[4 * k + 3] += acc[2 * k + 1] >> 16;
        acc[2 * k] = 0;
        acc[2 * k + 1] = 0;
      }
      since_flush = 0;
    }
  }

  const int d = d0 + dl;
  if (d >= ndm) return;
  const int tb = t0 + g * V;
  const int nout = B.nout;
  if (B.out_f32) {
    float *o = static_cast<float *>(B.out) + (int64_t)d * B.out_stride + tb;
    if (tb + V <= nout && ((reinterpret_cast<uintptr_t>(o) & 15) == 0)) {
#pragma unroll
      for (int k = 0; k < V / 4; ++k)
        reinterpret_cast<float4 *>(o)[k] = make_float4((float)tot[4 * k], (float)tot[4 * k + 1],
                                                       (float)tot[4 * k + 2], (float)tot[4 * k + 3]);
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j)
        if (tb + j < nout) o[j] = (float)tot[j];
    }
  } else {
    uint32_t *o = static_cast<uint32_t *>(B.out) + (int64_t)d * B.out_stride + tb;
    if (tb + V <= nout && ((reinterpret_cast<uintptr_t>(o) & 15) == 0)) {
#pragma unroll
      for (int k = 0; k < V / 4; ++k)
        reinterpret_cast<uint4 *>(o)[k] =
            make_uint4(tot[4 * k], tot[4 * k + 1], tot[4 * k + 2], tot[4 * k + 3]);
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j)
        if (tb + j < nout) o[j] = tot[j];
    }
  }
}

template <int G>
void launch_g(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, int chunk,
              int lrow, hipStream_t s)
{
  constexpr int ND = NT / G;
  const size_t lds = (size_t)chunk * lrow + (size_t)chunk * ND * 4 + (size_t)chunk * 4;
  hipLaunchKernelGGL(k_dedisp<G>, dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams, d_tiles,
                     chunk, lrow);
  HAMD_CHECK(hipGetLastError());
}

__global__ void k_fill_rows(uint32_t *out, int64_t stride, int n)
{
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n) out[(int64_t)blockIdx.y * stride + t] = 0u;  // 0u is also 0.0f
}

// u8 [nsamps][in_stride] (time-major) -> [nch][row_stride] (channel-major)
__global__ __launch_bounds__(256) void k_transpose_tf_ft(const uint8_t *__restrict__ in,
                                                         int64_t nsamps, int64_t in_stride,
                                                         int nch, uint8_t *__restrict__ out,
                                                         int64_t row_stride, int64_t ring)
{
  __shared__ uint8_t tile[64][65];
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  for (int r = ty; r < 64; r += 4) {
    const int64_t t = t0 + r;
    const int c = c0 + tx;
    tile[r][tx] = (t < nsamps && c < nch) ? in[t * in_stride + c] : (uint8_t)0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int c = c0 + r;
    const int64_t t = t0 + tx;
    if (c < nch && t < ring) out[(int64_t)c * row_stride + t] = tile[tx][r];
  }
}

}  // namespace

int dd_samples_per_thread() { return V; }

// Largest delay spread over any nd consecutive DM trials (any block start,
// so sub-ranges of the DM list are covered too).  Delays grow with DM.
int dd_max_spread(const int32_t *delta, int ndm, int nlive, int nd)
{
  int best = 0;
  for (int d0 = 0; d0 < ndm; ++d0) {
    int d1 = d0 + nd - 1 < ndm - 1 ? d0 + nd - 1 : ndm - 1;
    for (int i = 0; i < nlive; ++i) {
      int sp = delta[(int64_t)d1 * nlive + i] - delta[(int64_t)d0 * nlive + i];
      if (sp > best) best = sp;
    }
  }
  return best;
}

static int lrow_for(const DDTile &t, int max_spread)
{
  const int T = t.lanes_per_dm * V;
  return (int)round_up((size_t)(T + max_spread + 16 + 4 + 16), 16);
}

size_t dd_lds_bytes(const DDTile &t, int max_spread)
{
  const int ND = NT / t.lanes_per_dm;
  return (size_t)t.chunk * lrow_for(t, max_spread) + (size_t)t.chunk * ND * 4 + (size_t)t.chunk * 4;
}

void dd_build_tiles(const std::vector<DDBeamDev> &beams, const DDTile &t,
                    std::vector<DDTileDesc> &tiles, int order)
{
  const int T = dd_tile_T(t), ND = dd_tile_ND(t);
  std::vector<DDTileDesc> list;
  for (size_t b = 0; b < beams.size(); ++b) {
    const DDBeamDev &B = beams[b];
    if (B.nout <= 0 || B.ndm <= 0 || B.nlive <= 0) continue;
    if (order == 0) {
      for (int d0 = 0; d0 < B.ndm; d0 += ND)
        for (int t0 = 0; t0 < B.nout; t0 += T) list.push_back(DDTileDesc{(int)b, d0, t0, 0});
    } else {
      for (int t0 = 0; t0 < B.nout; t0 += T)
        for (int d0 = 0; d0 < B.ndm; d0 += ND) list.push_back(DDTileDesc{(int)b, d0, t0, 0});
    }
  }
  // the dispatcher hands workgroup i to XCD i % 8: give each XCD a contiguous
  // run of the list so neighbouring time tiles (overlapping windows) share
  // that XCD's L2
  const size_t nx = 8;
  const size_t q = (list.size() + nx - 1) / nx;
  tiles.assign(q * nx, DDTileDesc{-1, 0, 0, 0});
  for (size_t k = 0; k < list.size(); ++k) tiles[(k % q) * nx + k / q] = list[k];
}

void dd_launch(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, const DDTile &t,
               int max_spread, hipStream_t s)
{
  if (ntiles <= 0) return;
  const int lrow = lrow_for(t, max_spread);
  switch (t.lanes_per_dm) {
    case 8: launch_g<8>(d_beams, d_tiles, ntiles, t.chunk, lrow, s); break;
    case 16: launch_g<16>(d_beams, d_tiles, ntiles, t.chunk, lrow, s); break;
    case 32: launch_g<32>(d_beams, d_tiles, ntiles, t.chunk, lrow, s); break;
    case 64: launch_g<64>(d_beams, d_tiles, ntiles, t.chunk, lrow, s); break;
    default: throw HipError{hipErrorInvalidValue, "lanes_per_dm", __LINE__};
  }
}

void fill_zero_rows(void *out, int64_t stride, int n, int rows, hipStream_t s)
{
  if (n <= 0 || rows <= 0) return;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)rows);
  hipLaunchKernelGGL(k_fill_rows, grid, dim3(256), 0, s, static_cast<uint32_t *>(out), stride, n);
  HAMD_CHECK(hipGetLastError());
}

void transpose_tf_to_ft(const uint8_t *in, int64_t nsamps, int64_t in_stride, int nch,
                        uint8_t *out, int64_t row_stride, int64_t ring, hipStream_t s)
{
  dim3 grid((unsigned)((ring + 63) / 64), (unsigned)((nch + 63) / 64));
  hipLaunchKernelGGL(k_transpose_tf_ft, grid, dim3(256), 0, s, in, nsamps, in_stride, nch, out,
                     row_stride, ring);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
