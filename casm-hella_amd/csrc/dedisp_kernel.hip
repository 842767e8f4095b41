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
// is staged in LDS with 16-byte loads, then every thread (one DM, V = 28
// consecutive samples) adds its bytes.  Four samples of one channel come
// from two LDS dwords; v_perm_b32 extracts the even and the odd samples into
// two u16x2 lanes at the thread's byte alignment, so one 32-bit add performs
// two sample additions.  u16 lanes hold at most 257*255, so they are widened
// into u32 totals every 256 channels.  V/4 = 7 is odd, which spreads the
// threads of one DM over all 32 LDS banks.
#include <mutex>
#include <type_traits>
#include <vector>

#include "internal.hpp"

namespace hamd {

namespace {

constexpr int NT = 256;
constexpr int V = 28;          // samples per thread
constexpr int NW = V / 4 + 1;  // LDS dwords read per channel per thread

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Read-only table entry at a wave-uniform index through the scalar cache (a
// constant-address-space load): it waits on lgkmcnt, so it does not force the
// window loads in flight on vmcnt to complete first.
template <class T>
__device__ __forceinline__ T sload(const T *p, int64_t i)
{
  typedef __attribute__((address_space(4))) const T ct;
  return ((ct *)p)[i];
}

// load through a global-address-space pointer: a global (not flat) load,
// tracked by the vector memory counter only
template <class T>
__device__ __forceinline__ T gload(const T *p)
{
  typedef __attribute__((address_space(1))) const T gt;
  return *((gt *)p);
}

__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
  return __builtin_amdgcn_perm(hi, lo, sel);
}

// write one thread's V sums of DM row d starting at output sample tb
__device__ __forceinline__ void store_out(const DDBeamDev &B, int d, int tb, const uint32_t *tot)
{
  if (d >= B.ndm) return;
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

// One DM trial's place in a staged chunk: the byte address of the LDS dword
// holding the trial's first sample for lane group 0 (row byte offset + q
// rounded down to a dword; q = the trial's offset into the channel window)
// and the v_perm selector that extracts the even samples at that byte
// alignment (the odd-sample selector is one more in each half).  Built once
// per (channel, trial) while the chunk is staged.
__device__ __forceinline__ uint2 tab_entry(int row_byte, int q)
{
  const uint32_t s = (uint32_t)(q & 3);
  return make_uint2((uint32_t)(row_byte + (q & ~3)), s | 0x0c00u | ((s + 2u) << 16) | 0x0c000000u);
}

// add the nc staged channels of a chunk to a thread's V sums (u16x2 lanes).
// Channels go in pairs; the next pair's table entries are read while the
// current pair is summed, so only the window reads wait on LDS.  tab holds
// at least cap entries per trial column (reads past nc are never used).
template <int ND>
__device__ __forceinline__ void sum_chunk(uint32_t (&acc)[V / 2], const uint8_t *sdat,
                                          const uint2 *tab, int nc, int cap, int dl, int g)
{
  auto one = [&](uint2 e) {
    const uint32_t *wrow = reinterpret_cast<const uint32_t *>(sdat + e.x) + g * (V / 4);
    const uint32_t sel_o = e.y + 0x00010001u;
    uint32_t w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = wrow[i];
#pragma unroll
    for (int k = 0; k < V / 4; ++k) {
      acc[2 * k] += perm(w[k + 1], w[k], e.y);
      acc[2 * k + 1] += perm(w[k + 1], w[k], sel_o);
    }
  };
  uint2 e0 = tab[dl], e1 = tab[min(1, cap - 1) * ND + dl];
  int cc = 0;
  for (; cc + 1 < nc; cc += 2) {
    const uint2 n0 = tab[min(cc + 2, cap - 1) * ND + dl];
    const uint2 n1 = tab[min(cc + 3, cap - 1) * ND + dl];
    one(e0);
    one(e1);
    e0 = n0;
    e1 = n1;
  }
  if (cc < nc) one(e0);
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
  uint2 *tab = reinterpret_cast<uint2 *>(sdly + chunk * ND + chunk);        // [chunk][ND]

  const int tid = threadIdx.x;
  const int dl = tid / G;
  const int g = tid % G;
  const int d0 = td.d0;
  const int t0 = td.t0;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int nlive = B.nlive, ndm = B.ndm;

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;

  // delays of the first chunk; later chunks are prefetched into registers
  // during the previous chunk's arithmetic (one (channel, DM) pair per thread)
  auto dly_load = [&](int k0, int i) -> int32_t {
    const int nc = min(chunk, nlive - k0);
    if (i >= nc * ND) return 0;
    const int cc = i / ND, dd = i - (i / ND) * ND;
    const int dm = min(d0 + dd, ndm - 1);
    return gload(B.delta + (int64_t)dm * nlive + k0 + cc);
  };
  int32_t pre = dly_load(0, tid);
  const int ring = (int)B.ring;
  int since_flush = 0;
  for (int k0 = 0; k0 < nlive; k0 += chunk) {
    const int nc = min(chunk, nlive - k0);
    __syncthreads();
    if (tid < nc * ND) sdly[tid] = pre;
    for (int i = tid + NT; i < nc * ND; i += NT) sdly[i] = dly_load(k0, i);
    __syncthreads();
    if (k0 + chunk < nlive) pre = dly_load(k0 + chunk, tid);
    // stage windows: one wave per channel, 16-byte units
    for (int cc = wave; cc < nc; cc += NT / 64) {
      const int base_d = sdly[cc * ND];
      const int spread = sdly[cc * ND + ND - 1] - base_d;
      int first = (int)B.start + t0 + base_d;
      while (first >= ring) first -= ring;
      const int al = first & ~15;
      const int off = first - al;
      const int n16 = (off + T + spread + 4 + 15) >> 4;
      const uint8_t *row = B.base + (int64_t)sload(B.live, k0 + cc) * B.row_stride;
      uint8_t *dst = sdat + (size_t)cc * lrow;
      for (int u = lane; u < n16; u += 64) {
        int col = al + 16 * u;
        while (col >= ring) col -= ring;
        *reinterpret_cast<u32x4 *>(dst + 16 * u) = gload(reinterpret_cast<const u32x4 *>(row + col));
      }
      if (lane < ND) tab[cc * ND + lane] = tab_entry(cc * lrow, off + sdly[cc * ND + lane] - base_d);
    }
    __syncthreads();
    sum_chunk<ND>(acc, sdat, tab, nc, chunk, dl, g);
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

  store_out(B, d0 + dl, t0 + g * V, tot);
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void glb_void_t;

// Same arithmetic as k_dedisp; the channel windows of the next chunk are
// copied global -> LDS by the DMA path of global_load_lds_dwordx4 (each lane
// moves 16 bytes to LDS base + 16 * lane) while the current chunk is summed,
// so the loads never wait in registers.  Two LDS buffers alternate.
template <int G, int CC>
__global__ __launch_bounds__(NT) void k_dedisp_dma(const DDBeamDev *__restrict__ beams,
                                                   const DDTileDesc *__restrict__ tiles,
                                                   int lrow)
{
  constexpr int chunk = CC;
  constexpr int ND = NT / G;
  constexpr int T = G * V;
  extern __shared__ __align__(16) uint8_t lds[];
  const DDTileDesc td = tiles[blockIdx.x];
  if (td.beam < 0) return;
  const DDBeamDev B = beams[td.beam];
  // layout: data[2][chunk][lrow] | tab[2][chunk][ND]
  uint8_t *sdat = lds;
  uint2 *tab = reinterpret_cast<uint2 *>(lds + (size_t)2 * chunk * lrow);

  const int tid = threadIdx.x;
  const int dl = tid / G;
  const int g = tid % G;
  const int d0 = td.d0;
  const int t0 = td.t0;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int nlive = B.nlive, ndm = B.ndm;
  const int dlast = min(d0 + ND - 1, ndm - 1);

  // Each wave stages channels wave, wave+4, ... of a chunk.  Their delays
  // (first DM, last DM, this lane's DM) are loaded one chunk ahead into
  // registers so issuing the DMA copies never waits on memory.
  constexpr int MAXC = CC / (NT / 64);
  const int ring = (int)B.ring;
  int32_t pb[MAXC], pl[MAXC], pr[MAXC];
  const int nlast = dlast - d0;
  const int mylane = min(lane, nlast);
  auto prefetch = [&](int k0) {
    const int nc = min(chunk, nlive - k0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int cc = min(wave + 4 * i, nc - 1);
      const int32_t *dt = B.delta_t + (int64_t)(k0 + cc) * B.dt_stride + d0;
      pb[i] = sload(dt, 0);
      pl[i] = sload(dt, nlast);
      pr[i] = gload(dt + mylane);
    }
  };
  // ring >= lrow + 16 (checked by the launcher): one subtraction wraps a column
  auto stage = [&](int k0, int buf) {
    const int nc = min(chunk, nlive - k0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int cc = wave + 4 * i;
      if (cc >= nc) break;
      const int base_d = pb[i];
      const int spread = pl[i] - base_d;
      int first = (int)B.start + t0 + base_d;
      first -= first >= ring ? ring : 0;
      const int al = first & ~15;
      const int off = first - al;
      const int n16 = (off + T + spread + 4 + 15) >> 4;
      const uint8_t *row = B.base + (int64_t)sload(B.live, k0 + cc) * B.row_stride;
      uint8_t *dst = sdat + ((size_t)buf * chunk + cc) * lrow;
      for (int j = 0; j * 64 < n16; ++j) {
        const int u = j * 64 + lane;
        int col = al + 16 * u;
        col -= col >= ring ? ring : 0;
        if (u < n16)
          __builtin_amdgcn_global_load_lds((glb_void_t *)(row + col), (lds_void_t *)(dst + 1024 * j),
                                           16, 0, 0);
      }
      if (lane < ND)
        tab[((size_t)buf * chunk + cc) * ND + lane] =
            tab_entry((buf * chunk + cc) * lrow, off + pr[i] - base_d);
    }
  };

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;

  prefetch(0);
  stage(0, 0);
  if (chunk < nlive) prefetch(chunk);
  int since_flush = 0, buf = 0;
  for (int k0 = 0; k0 < nlive; k0 += chunk, buf ^= 1) {
    const int nc = min(chunk, nlive - k0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (k0 + chunk < nlive) {
      stage(k0 + chunk, buf ^ 1);
      if (k0 + 2 * chunk < nlive) prefetch(k0 + 2 * chunk);
    }
    sum_chunk<ND>(acc, sdat, tab + (size_t)buf * chunk * ND, nc, chunk, dl, g);
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
  store_out(B, d0 + dl, t0 + g * V, tot);
}

// Register-pipelined variant: while chunk k is summed from LDS, the 16-byte
// units of chunk k+1's windows are already loading into registers (UPL units
// per lane per staged channel, through global loads that only the vector
// memory counter tracks); after the sums they are written to the single LDS
// buffer.  The first and last trial delays of a staged channel are scalar
// loads from the channel-major table two chunks ahead; each lane's own trial
// delay travels with the window units.  Every load is unconditional (clamped
// indices), so none has to complete where a branch joins.  Units beyond
// UPL*64 of a window (very wide delay spreads) are copied directly.
template <int G, int CC, int UPL>
__global__ __launch_bounds__(NT) void k_dedisp_pf(const DDBeamDev *__restrict__ beams,
                                                  const DDTileDesc *__restrict__ tiles, int lrow)
{
  constexpr int ND = NT / G;
  constexpr int T = G * V;
  constexpr int MAXC = CC / (NT / 64);
  extern __shared__ __align__(16) uint8_t lds[];
  const DDTileDesc td = tiles[blockIdx.x];
  if (td.beam < 0) return;
  const DDBeamDev B = beams[td.beam];
  uint8_t *sdat = lds;                                             // [CC][lrow]
  uint2 *tab = reinterpret_cast<uint2 *>(lds + (size_t)CC * lrow);  // [CC][ND]

  const int tid = threadIdx.x;
  const int dl = tid / G;
  const int g = tid % G;
  const int d0 = td.d0;
  const int t0 = td.t0;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int nlive = B.nlive;
  const int nlast = min(ND - 1, B.ndm - 1 - d0);  // last trial of the block (relative)
  const int mylane = min(lane, nlast);
  const int ring = (int)B.ring;

  int32_t pb[MAXC], pl[MAXC];  // scalar: first / last trial delay of each slot
  int32_t pr[MAXC];            // this lane's trial delay
  u32x4 ureg[MAXC * UPL];
  int ubias[MAXC], un16[MAXC], ualgn[MAXC];
  int64_t urow[MAXC];

  auto dly = [&](int k0) {
    const int nc = min(CC, nlive - k0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int cc = min(wave + 4 * i, nc - 1);
      const int32_t *dt = B.delta_t + (int64_t)(k0 + cc) * B.dt_stride + d0;
      pb[i] = sload(dt, 0);
      pl[i] = sload(dt, nlast);
    }
  };
  auto load = [&](int k0) {
    const int nc = min(CC, nlive - k0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int cc = min(wave + 4 * i, nc - 1);
      pr[i] = gload(B.delta_t + (int64_t)(k0 + cc) * B.dt_stride + d0 + mylane);
      int first = (int)B.start + t0 + pb[i];
      first -= first >= ring ? ring : 0;
      const int al = first & ~15;
      ubias[i] = (first - al) - pb[i];
      ualgn[i] = al;
      un16[i] = ((first - al) + T + (pl[i] - pb[i]) + 4 + 15) >> 4;
      urow[i] = (int64_t)sload(B.live, k0 + cc) * B.row_stride;
#pragma unroll
      for (int j = 0; j < UPL; ++j) {
        const int u = min(j * 64 + lane, un16[i] - 1);
        int col = al + 16 * u;
        col -= col >= ring ? ring : 0;
        ureg[i * UPL + j] = gload(reinterpret_cast<const u32x4 *>(B.base + urow[i] + col));
      }
    }
  };
  auto store = [&](int k0) {
    const int nc = min(CC, nlive - k0);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
      const int cc = wave + 4 * i;
      if (cc < nc) {
        uint8_t *dst = sdat + (size_t)cc * lrow;
#pragma unroll
        for (int j = 0; j < UPL; ++j) {
          const int u = j * 64 + lane;
          if (u < un16[i]) *reinterpret_cast<u32x4 *>(dst + 16 * u) = ureg[i * UPL + j];
        }
        for (int u = UPL * 64 + lane; u < un16[i]; u += 64) {
          int col = ualgn[i] + 16 * u;
          col -= col >= ring ? ring : 0;
          *reinterpret_cast<u32x4 *>(dst + 16 * u) =
              gload(reinterpret_cast<const u32x4 *>(B.base + urow[i] + col));
        }
        if (lane < ND) tab[cc * ND + lane] = tab_entry(cc * lrow, ubias[i] + pr[i]);
      }
    }
  };

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;

  dly(0);
  load(0);
  if (CC < nlive) dly(CC);
  int since_flush = 0;
  for (int k0 = 0; k0 < nlive; k0 += CC) {
    const int nc = min(CC, nlive - k0);
    __syncthreads();  // everyone is done with the previous chunk
    store(k0);
    __syncthreads();
    if (k0 + CC < nlive) {
      load(k0 + CC);  // in flight during the sums below
      if (k0 + 2 * CC < nlive) dly(k0 + 2 * CC);
    }
    sum_chunk<ND>(acc, sdat, tab, nc, CC, dl, g);
    since_flush += nc;
    if (since_flush >= 256 || k0 + CC >= nlive) {
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
  store_out(B, d0 + dl, t0 + g * V, tot);
}

template <int G, int CC>
void launch_pf(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, int lrow,
               hipStream_t s)
{
  constexpr int ND = NT / G;
  const size_t lds = (size_t)CC * lrow + (size_t)CC * ND * sizeof(uint2);
  const int upl = (lrow + 1023) / 1024;
  if (upl <= 1)
    hipLaunchKernelGGL((k_dedisp_pf<G, CC, 1>), dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams,
                       d_tiles, lrow);
  else if (upl == 2)
    hipLaunchKernelGGL((k_dedisp_pf<G, CC, 2>), dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams,
                       d_tiles, lrow);
  else
    hipLaunchKernelGGL((k_dedisp_pf<G, CC, 3>), dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams,
                       d_tiles, lrow);
}

// One DM trial per wave.  A workgroup of NDW waves owns NDW consecutive
// trials x T = 64 * V output samples.  Every delay a wave needs is uniform
// across the wave, so the scalar unit loads the delays (one scalar load per
// chunk from the DM-major table) and derives each channel's LDS offset and
// v_perm selectors; the vector ALU does little besides the additions.  Wave
// w stages channel k0 + w of every NDW-channel chunk: NU 16-byte units per
// lane (NU * 1024 >= lrow) are read into registers one chunk ahead through a
// buffer resource spanning the channel's ring row, and written to the other
// of two LDS buffers once the current chunk is summed, so one barrier per
// chunk suffices.  Channels past the last live one stage zeros, so every
// chunk is a full NDW channels and the sums are straight-line code.
// The LDS entry of a (trial, channel) comes from the beam's precomputed
// table (dd_entries) when it has one for blocks of NDW trials and the tile
// starts at a ring column that is a multiple of 16; otherwise the scalar
// unit derives it from the delays.
// Requirements: ring % 16 == 0 (every beam's ring is a multiple of 32 or
// 64 columns) and the DM-major delay table readable for DD_DELTA_PAD
// entries past its end.
// s_waitcnt immediate (gfx9 encoding): at most 8 LDS/scalar requests
// outstanding, vector memory and export counters not waited on
constexpr int WAIT_LGKM_8 = 15 | (7 << 4) | (8 << 8) | (3 << 14);

template <int NDW, int NU, bool PIPE>
__global__ __launch_bounds__(64 * NDW) void k_dedisp_w(const DDBeamDev *__restrict__ beams,
                                                     const DDTileDesc *__restrict__ tiles,
                                                     int lrow)
{
  constexpr int CC = NDW;
  typedef int32_t dvec __attribute__((ext_vector_type(CC), aligned(4)));
  typedef __attribute__((address_space(4))) const dvec cdvec;
  extern __shared__ __align__(16) uint8_t lds[];  // [2][CC][lrow]
  const DDTileDesc td = tiles[blockIdx.x];
  if (td.beam < 0) return;
  const DDBeamDev B = beams[td.beam];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int d0 = td.d0, t0 = td.t0;
  const int ndm = B.ndm, nlive = B.nlive;
  const int ring = (int)B.ring;
  // ring column of output sample t0 (window starts stay below 2 * ring)
  const int S = (int)B.start + t0;
  const int32_t *own_row = B.delta + (int64_t)min(d0 + wave, ndm - 1) * nlive;
  const int32_t *first_row = B.delta + (int64_t)d0 * nlive;
  const int32_t *last_row = B.delta + (int64_t)min(d0 + NDW - 1, ndm - 1) * nlive;

  // staging: scalar parameters of this wave's channel of a chunk, then its
  // window in registers
  int sb = 0, sl = 0, sch = 0;
  auto stage_params = [&](int k) {
    const int kk = min(k, nlive - 1);
    sb = sload(first_row, kk);
    sl = sload(last_row, kk);
    sch = k < nlive ? sload(B.live, kk) : -1;
  };
  u32x4 P[NU];
  const int lane16 = lane * 16;
  // bytes of the staged row that the block's trials read: a trial starts at
  // the dword holding its first sample, (first & 15) + (its delay - sb) <=
  // (first & 15) + (sl - sb) bytes in, and reads 64 * V + 4 bytes, so the
  // rest of the row (sized for the widest block of the beam) is neither
  // loaded nor stored.  A channel past the last live one is staged whole:
  // its readers take their offsets from padding entries.
  int need = lrow;
  // One load site for every case, so the loads stay in flight until the
  // store after the sums (loads issued on several paths make the compiler
  // merge their results with register moves, which wait for the data).  A
  // channel past the last live one gets a resource of zero records: every
  // load of it is out of range and returns zeros.
  auto fetch = [&]() {
    const bool live = sch >= 0;
    int first = S + sb;
    first -= first >= ring ? ring : 0;
    const int al = first & ~15;
    need = live ? (first & 15) + (sl - sb) + 64 * V + 4 : lrow;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(B.base + (live ? (int64_t)sch * B.row_stride : 0)), (short)0, live ? ring : 0,
        0x00020000);
    // no wrap: one lane offset, the units 1024 bytes apart by the scalar
    // offset (no per-unit vector additions)
    const bool nowrap = al + NU * 1024 <= ring;
    int col[NU];
    if (nowrap) {
#pragma unroll
      for (int u = 0; u < NU; ++u) col[u] = al + lane16;
    } else {
      // the staged units run past the ring end (short rings may wrap more
      // than once)
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        int c = al + lane16 + u * 1024;
        while (c >= ring) c -= ring;
        col[u] = c;
      }
    }
    // unit 0 is always needed (need >= 64 * V + 4 > 1024)
#pragma unroll
    for (int u = 0; u < NU; ++u)
      if (u == 0 || lane16 + u * 1024 < need)
        P[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, col[u], nowrap ? u * 1024 : 0, 0);
  };
  auto store = [&](int buf) {
    uint8_t *dst = lds + (size_t)(buf * CC + wave) * lrow + lane16;
#pragma unroll
    for (int u = 0; u < NU; ++u)
      if (u == 0 || lane16 + u * 1024 < need) *reinterpret_cast<u32x4 *>(dst + u * 1024) = P[u];
  };

  // LDS entry of this wave's trial for the CC channels of a chunk (scalar
  // registers): the trial's dword offset in the staged row and its v_perm
  // selector of the even samples.  From the precomputed table when the
  // beam has one for this block size (one scalar load per chunk), else from
  // this trial's and the block's first trial's delays; past the last live
  // channel the offset is that of the zero row's start.
  typedef uint32_t evec __attribute__((ext_vector_type(2 * CC), aligned(4)));
  typedef __attribute__((address_space(4))) const evec cevec;
  const bool tab = B.ent != nullptr && B.ent_nd == NDW && (S & 15) == 0;
  const uint32_t *ent_row = tab ? B.ent + (int64_t)min(d0 + wave, ndm - 1) * B.ent_stride * 2 : nullptr;
  uint32_t dw[CC], se[CC];
  auto entries = [&](int k0) {
    if (tab) {
      const evec e = *(cevec *)(ent_row + 2 * k0);
#pragma unroll
      for (int cc = 0; cc < CC; ++cc) {
        dw[cc] = e[2 * cc];
        se[cc] = e[2 * cc + 1];
      }
      return;
    }
    const dvec o = *(cdvec *)(own_row + k0);
    const dvec f = *(cdvec *)(first_row + k0);
#pragma unroll
    for (int cc = 0; cc < CC; ++cc) {
      const int bas = f[cc];
      const int own = k0 + cc < nlive ? o[cc] : bas;
      // the row starts at the 16-byte aligned column below the first trial's start
      const int q = ((S + bas) & 15) + own - bas;
      const uint32_t s = (uint32_t)(q & 3);
      dw[cc] = (uint32_t)(q & ~3);
      se[cc] = s | 0x0c00u | ((s + 2u) << 16) | 0x0c000000u;
    }
  };

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;

  const int lane_off = lane * V;  // byte offset of the lane's first sample
  auto one = [&](int buf, int cc) {
    const uint32_t sel_e = se[cc];
    const uint32_t sel_o = sel_e + 0x00010001u;
    const uint32_t *wrow = reinterpret_cast<const uint32_t *>(
        lds + (buf * CC + cc) * lrow + dw[cc] + lane_off);
    uint32_t w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = wrow[i];
#pragma unroll
    for (int k = 0; k < V / 4; ++k) {
      acc[2 * k] += perm(w[k + 1], w[k], sel_e);
      acc[2 * k + 1] += perm(w[k + 1], w[k], sel_o);
    }
  };

  const int nchunks = (nlive + CC - 1) / CC;
  stage_params(wave);
  fetch();
  entries(0);
  store(0);
  if (CC < nlive) stage_params(CC + wave);
  __syncthreads();
  int since_flush = 0;
  for (int ci = 0; ci < nchunks; ++ci) {
    const int buf = ci & 1;
    const bool more = ci + 1 < nchunks;
    if (more) fetch();  // the next chunk's window, in flight during the sums
    if (PIPE) {
      // every channel's LDS words of the chunk are requested before the
      // first v_perm, so the reads of later channels overlap the sums
      uint32_t w[CC][NW], so[CC];
      auto rd = [&](int cc) {
        so[cc] = se[cc] + 0x00010001u;
        const uint32_t *wrow = reinterpret_cast<const uint32_t *>(
            lds + (buf * CC + cc) * lrow + dw[cc] + lane_off);
#pragma unroll
        for (int i = 0; i < NW; ++i) w[cc][i] = wrow[i];
      };
      // two channels per accumulator update (three-operand additions)
      auto sum2 = [&](int cc) {
#pragma unroll
        for (int k = 0; k < V / 4; ++k) {
          acc[2 * k] += perm(w[cc][k + 1], w[cc][k], se[cc]) +
                        perm(w[cc + 1][k + 1], w[cc + 1][k], se[cc + 1]);
          acc[2 * k + 1] += perm(w[cc][k + 1], w[cc][k], so[cc]) +
                            perm(w[cc + 1][k + 1], w[cc + 1][k], so[cc + 1]);
        }
      };
      if constexpr (CC == 4) {
        // three channels' reads in flight (12 of the 15 LDS requests the
        // counter tracks); the fourth is issued once the first has landed,
        // so the sums of each pair overlap the reads of the next
        rd(0);
        rd(1);
        rd(2);
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM_8);
        __builtin_amdgcn_sched_barrier(0);
        rd(3);
        __builtin_amdgcn_sched_barrier(0);
        sum2(0);
        sum2(2);
      } else if constexpr (CC == 8) {
        // the same rolling window over eight channels: at most three
        // channels' reads in flight, the next one issued as soon as the
        // oldest has landed, each pair summed once both are in
        rd(0);
        rd(1);
        rd(2);
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM_8);
        __builtin_amdgcn_sched_barrier(0);
        rd(3);
        __builtin_amdgcn_sched_barrier(0);
        sum2(0);
        __builtin_amdgcn_sched_barrier(0);
        rd(4);
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM_8);
        __builtin_amdgcn_sched_barrier(0);
        rd(5);
        __builtin_amdgcn_sched_barrier(0);
        sum2(2);
        __builtin_amdgcn_sched_barrier(0);
        rd(6);
        __builtin_amdgcn_s_waitcnt(WAIT_LGKM_8);
        __builtin_amdgcn_sched_barrier(0);
        rd(7);
        __builtin_amdgcn_sched_barrier(0);
        sum2(4);
        sum2(6);
      } else {
#pragma unroll
        for (int cc = 0; cc < CC; ++cc) rd(cc);
#pragma unroll
        for (int cc = 0; cc < CC; cc += 2) sum2(cc);
      }
    } else {
#pragma unroll
      for (int cc = 0; cc < CC; cc += 2) {
        one(buf, cc);
        one(buf, cc + 1);
      }
    }
    since_flush += CC;
    if (since_flush >= 256 || !more) {
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
    if (more) {
      entries((ci + 1) * CC);
      if (ci + 2 < nchunks) stage_params((ci + 2) * CC + wave);
      store(buf ^ 1);
    }
    __syncthreads();
  }
  store_out(B, d0 + wave, t0 + lane * V, tot);
}

template <int NDW, bool PIPE>
void launch_w(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, int lrow,
              hipStream_t s)
{
  const size_t lds = (size_t)2 * NDW * lrow;
  const int nu = (lrow + 1023) / 1024;
  // the LDS limit above 64 KiB was raised by kernels_setup() at set-up, so
  // nothing here is a non-stream runtime call (launches may be captured)
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(64 * NDW), lds, s, d_beams, d_tiles,
                       lrow);
  };
  switch (nu < 2 ? 2 : nu) {
    case 2: go(k_dedisp_w<NDW, 2, PIPE>); break;
    case 3: go(k_dedisp_w<NDW, 3, PIPE>); break;
    case 4: go(k_dedisp_w<NDW, 4, PIPE>); break;
    case 5: go(k_dedisp_w<NDW, 5, PIPE>); break;
    case 6: go(k_dedisp_w<NDW, 6, PIPE>); break;
    default: throw HipError{hipErrorInvalidValue, "dedispersion window too wide for k_dedisp_w", __LINE__};
  }
}

// Zeros for the windows of channels past the last live one (DMA source).
__device__ __align__(16) uint8_t dd_zero_page[6 * 1024];

// Workgroup barrier that waits for this wave's LDS reads but not for its
// LDS-DMA copies in flight (the fence of __syncthreads() would wait for
// those too); the memory clobber keeps the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void barrier_keep_dma()
{
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// k_dedisp_w with the windows copied global -> LDS by the DMA path of
// global_load_lds_dwordx4 two chunks ahead, into three statically separate
// LDS buffers: the sums of chunk i read one buffer while the copies of chunk
// i + 2 land in another, and the compiler, seeing distinct objects, does not
// make the reads wait for the copies.  No staging registers.  The chunk loop
// is unrolled by three so every code site names its buffers statically.
// Every wave issues exactly NU copies per chunk (channels past the last live
// one copy zeros), so "at most NU copies outstanding" means the previous
// chunk's have landed.  Rows are NU * 1024 bytes.  NDW = 4 trials per
// workgroup; the entries come from the beam's table or the delays as in
// k_dedisp_w.
template <int NU>
__global__ __launch_bounds__(256) void k_dedisp_wd(const DDBeamDev *__restrict__ beams,
                                                 const DDTileDesc *__restrict__ tiles)
{
  constexpr int NDW = 4, CC = 4;
  constexpr int LROW = NU * 1024;
  typedef int32_t dvec __attribute__((ext_vector_type(CC), aligned(4)));
  typedef __attribute__((address_space(4))) const dvec cdvec;
  typedef uint32_t evec __attribute__((ext_vector_type(2 * CC), aligned(4)));
  typedef __attribute__((address_space(4))) const evec cevec;
  __shared__ __align__(16) uint8_t buf0[CC * LROW];
  __shared__ __align__(16) uint8_t buf1[CC * LROW];
  __shared__ __align__(16) uint8_t buf2[CC * LROW];
  const DDTileDesc td = tiles[blockIdx.x];
  if (td.beam < 0) return;
  const DDBeamDev B = beams[td.beam];
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int d0 = td.d0, t0 = td.t0;
  const int ndm = B.ndm, nlive = B.nlive;
  const int ring = (int)B.ring;
  const int S = (int)B.start + t0;
  const int32_t *own_row = B.delta + (int64_t)min(d0 + wave, ndm - 1) * nlive;
  const int32_t *first_row = B.delta + (int64_t)d0 * nlive;
  const bool tab = B.ent != nullptr && B.ent_nd == NDW && (S & 15) == 0;
  const uint32_t *ent_row = tab ? B.ent + (int64_t)min(d0 + wave, ndm - 1) * B.ent_stride * 2 : nullptr;
  const int nchunks = (nlive + CC - 1) / CC;

  // copy this wave's channel of chunk ci into the wave's row of dst; the
  // copies are issued from one place (after the source addresses are
  // chosen), so the compiler keeps track of their LDS destination
  auto stage = [&](int ci, uint8_t *dst) {
    const int k = ci * CC + wave;
    const uint8_t *src[NU];
    if (k < nlive) {
      const int bas = sload(first_row, k);
      int first = S + bas;
      first -= first >= ring ? ring : 0;
      const int al = first & ~15;
      const uint8_t *row = B.base + (int64_t)sload(B.live, k) * B.row_stride;
      const bool wraps = al + LROW > ring;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        int col = al + 16 * (u * 64 + lane);
        if (wraps)
          while (col >= ring) col -= ring;
        src[u] = row + col;
      }
    } else {
#pragma unroll
      for (int u = 0; u < NU; ++u) src[u] = dd_zero_page + 16 * (u * 64 + lane);
    }
    uint8_t *row_lds = dst + wave * LROW;
#pragma unroll
    for (int u = 0; u < NU; ++u)
      __builtin_amdgcn_global_load_lds((glb_void_t *)src[u], (lds_void_t *)(row_lds + 1024 * u), 16, 0, 0);
  };

  uint32_t dw[CC], se[CC];
  auto entries = [&](int k0) {
    if (tab) {
      const evec e = *(cevec *)(ent_row + 2 * k0);
#pragma unroll
      for (int cc = 0; cc < CC; ++cc) {
        dw[cc] = e[2 * cc];
        se[cc] = e[2 * cc + 1];
      }
      return;
    }
    const dvec o = *(cdvec *)(own_row + k0);
    const dvec f = *(cdvec *)(first_row + k0);
#pragma unroll
    for (int cc = 0; cc < CC; ++cc) {
      const int bas = f[cc];
      const int own = k0 + cc < nlive ? o[cc] : bas;
      const int q = ((S + bas) & 15) + own - bas;
      const uint32_t s = (uint32_t)(q & 3);
      dw[cc] = (uint32_t)(q & ~3);
      se[cc] = s | 0x0c00u | ((s + 2u) << 16) | 0x0c000000u;
    }
  };

  uint32_t acc[V / 2];
  uint32_t tot[V];
#pragma unroll
  for (int i = 0; i < V / 2; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < V; ++i) tot[i] = 0;
  const int lane_off = lane * V;
  int since_flush = 0;

  // waitcnt immediates: vmcnt(n), no wait on the other counters
  constexpr int WAIT_VM_NU = (NU & 15) | (7 << 4) | (15 << 8) | ((NU >> 4) << 14);
  constexpr int WAIT_VM_0 = (7 << 4) | (15 << 8);

  auto step = [&](int ci, const uint8_t *cur, uint8_t *ahead) {
    const bool copy = ci + 2 < nchunks;
#pragma unroll
    for (int cc = 0; cc < CC; cc += 2) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int c = cc + h;
        const uint32_t sel_e = se[c], sel_o = se[c] + 0x00010001u;
        const uint32_t *wrow = reinterpret_cast<const uint32_t *>(cur + c * LROW + dw[c] + lane_off);
        uint32_t w[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = wrow[i];
#pragma unroll
        for (int k = 0; k < V / 4; ++k) {
          acc[2 * k] += perm(w[k + 1], w[k], sel_e);
          acc[2 * k + 1] += perm(w[k + 1], w[k], sel_o);
        }
      }
    }
    since_flush += CC;
    if (since_flush >= 256 || ci + 1 >= nchunks) {
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
    if (ci + 1 < nchunks) entries((ci + 1) * CC);
    // chunk ci + 2 goes to the buffer every wave finished reading before the
    // previous barrier
    if (copy) stage(ci + 2, ahead);
    // the next chunk's copies (issued one step earlier) have landed when at
    // most this step's NU copies are outstanding
    if (copy) __builtin_amdgcn_s_waitcnt(WAIT_VM_NU);
    else __builtin_amdgcn_s_waitcnt(WAIT_VM_0);
    barrier_keep_dma();
  };

  stage(0, buf0);
  if (nchunks > 1) stage(1, buf1);
  entries(0);
  if (nchunks > 1) __builtin_amdgcn_s_waitcnt(WAIT_VM_NU);
  else __builtin_amdgcn_s_waitcnt(WAIT_VM_0);
  barrier_keep_dma();
  for (int ci = 0; ci < nchunks; ci += 3) {
    step(ci, buf0, buf2);
    if (ci + 1 >= nchunks) break;
    step(ci + 1, buf1, buf0);
    if (ci + 2 >= nchunks) break;
    step(ci + 2, buf2, buf1);
  }
  store_out(B, d0 + wave, t0 + lane * V, tot);
}

void launch_wd(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, int nu,
               hipStream_t s)
{
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(256), 0, s, d_beams, d_tiles);
  };
  switch (nu < 2 ? 2 : nu) {
    case 2: go(k_dedisp_wd<2>); break;
    case 3: go(k_dedisp_wd<3>); break;
    case 4: go(k_dedisp_wd<4>); break;
    case 5: go(k_dedisp_wd<5>); break;
    default: throw HipError{hipErrorInvalidValue, "dedispersion window too wide for k_dedisp_wd", __LINE__};
  }
}

template <int G>
void launch_g(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, int chunk,
              int lrow, int dma, hipStream_t s)
{
  constexpr int ND = NT / G;
  if (dma == 2) {
    if (chunk == 8) launch_pf<G, 8>(d_beams, d_tiles, ntiles, lrow, s);
    else launch_pf<G, 16>(d_beams, d_tiles, ntiles, lrow, s);
  } else if (dma == 1) {
    const size_t lds = (size_t)2 * chunk * lrow + (size_t)2 * chunk * ND * sizeof(uint2);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams, d_tiles, lrow);
    };
    if (chunk == 8) go(k_dedisp_dma<G, 8>);
    else if (chunk == 16) go(k_dedisp_dma<G, 16>);
    else go(k_dedisp_dma<G, 32>);
  } else {
    const size_t lds = (size_t)chunk * lrow + (size_t)chunk * ND * 4 + (size_t)chunk * 4 +
                       (size_t)chunk * ND * sizeof(uint2);
    hipLaunchKernelGGL(k_dedisp<G>, dim3((unsigned)ntiles), dim3(NT), lds, s, d_beams, d_tiles,
                       chunk, lrow);
  }
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

template <int NDW, bool PIPE>
void raise_lds_limit_w(int bytes)
{
  const void *k[] = {reinterpret_cast<const void *>(k_dedisp_w<NDW, 2, PIPE>),
                     reinterpret_cast<const void *>(k_dedisp_w<NDW, 3, PIPE>),
                     reinterpret_cast<const void *>(k_dedisp_w<NDW, 4, PIPE>),
                     reinterpret_cast<const void *>(k_dedisp_w<NDW, 5, PIPE>),
                     reinterpret_cast<const void *>(k_dedisp_w<NDW, 6, PIPE>)};
  for (const void *f : k)
    HAMD_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

}  // namespace

void dd_kernel_setup()
{
  // every one-trial-per-wave instance may take the device's whole
  // per-workgroup LDS (launches above 64 KiB need the raised limit)
  const int bytes = (int)dd_lds_budget(DDTile{64, 4, 3});
  raise_lds_limit_w<4, false>(bytes);
  raise_lds_limit_w<8, false>(bytes);
  raise_lds_limit_w<4, true>(bytes);
  raise_lds_limit_w<8, true>(bytes);
}

void kernels_setup()
{
  // once per device: kernel attributes are per device, and setting them is
  // not a stream operation, so it must not happen inside a stream capture
  // (a fixed table: after the first call on a device nothing is allocated)
  constexpr int MAXDEV = 64;
  static std::mutex mu;
  static bool done[MAXDEV] = {};
  int dev = 0;
  HAMD_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (dev >= 0 && dev < MAXDEV && done[dev]) return;
  dd_kernel_setup();
  cluster_kernel_setup();
  if (dev >= 0 && dev < MAXDEV) done[dev] = true;
}

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
  const int l = (int)round_up((size_t)(T + max_spread + 16 + 4 + 16), 16);
  return t.dma == 5 ? (int)round_up((size_t)l, 1024) : l;  // DMA rows: whole 1024-byte units
}

size_t dd_lds_bytes(const DDTile &t, int max_spread)
{
  if (t.dma == 5) return (size_t)3 * t.chunk * lrow_for(t, max_spread);
  if (t.dma >= 3) return (size_t)2 * t.chunk * lrow_for(t, max_spread);
  const int ND = NT / t.lanes_per_dm;
  // window rows, the chunk's delays (register loader), a pad word per
  // channel and the (address, selector) table of every (channel, trial)
  const size_t one = (size_t)t.chunk * lrow_for(t, max_spread) + (size_t)t.chunk * ND * 4 +
                     (size_t)t.chunk * 4 + (size_t)t.chunk * ND * 8;
  return t.dma == 1 ? 2 * one : one;
}

void dd_build_tiles(const std::vector<DDBeamDev> &beams, const DDTile &t,
                    std::vector<DDTileDesc> &tiles, int order)
{
  const int T = dd_tile_T(t), ND = dd_tile_ND(t);
  // the dispatcher hands workgroup i to XCD i % 8.  Every beam's tile list is
  // cut into 8 contiguous runs, one per XCD, so neighbouring time tiles
  // (overlapping windows) share that XCD's L2, and every XCD gets the same
  // share of every beam: beams of different channel counts cost different
  // amounts per tile, and a split of the whole list by count would leave the
  // XCDs holding the widest beams with most of the work.  Each beam's runs
  // are padded to equal length with empty tiles (at most 7 per beam).
  const size_t nx = 8;
  tiles.clear();
  std::vector<DDTileDesc> list;
  for (size_t b = 0; b < beams.size(); ++b) {
    const DDBeamDev &B = beams[b];
    if (B.nout <= 0 || B.ndm <= 0 || B.nlive <= 0) continue;
    list.clear();
    if (order == 0) {
      for (int d0 = 0; d0 < B.ndm; d0 += ND)
        for (int t0 = 0; t0 < B.nout; t0 += T) list.push_back(DDTileDesc{(int)b, d0, t0, 0});
    } else {
      for (int t0 = 0; t0 < B.nout; t0 += T)
        for (int d0 = 0; d0 < B.ndm; d0 += ND) list.push_back(DDTileDesc{(int)b, d0, t0, 0});
    }
    const size_t q = (list.size() + nx - 1) / nx, base = tiles.size();
    tiles.resize(base + q * nx, DDTileDesc{-1, 0, 0, 0});
    for (size_t k = 0; k < list.size(); ++k) tiles[base + (k % q) * nx + k / q] = list[k];
  }
}

size_t dd_lds_budget(const DDTile &t)
{
  if (t.dma < 3) return DD_LDS_BUDGET;
  // the one-trial-per-wave kernel may use the device's whole per-workgroup LDS
  static size_t limit = 0;
  if (limit == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      limit = (size_t)v;
    else
      limit = DD_LDS_BUDGET;
  }
  return limit;
}

bool dd_tile_fits(const DDTile &t, int max_spread, int64_t min_ring)
{
  if (t.dma < 3) return dd_lds_bytes(t, max_spread) <= dd_lds_budget(t);
  const int lrow = lrow_for(t, max_spread);
  (void)min_ring;  // windows wrap any number of times
  if (t.dma == 5)  // three static buffers of 4 rows of up to 5 KB
    return t.lanes_per_dm == 64 && t.chunk == 4 && lrow <= 5 * 1024;
  return t.lanes_per_dm == 64 && (t.chunk == 4 || t.chunk == 8) &&
         dd_lds_bytes(t, max_spread) <= dd_lds_budget(t) && lrow <= 6 * 1024;
}

void dd_launch(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, const DDTile &t_in,
               int max_spread, hipStream_t s, int64_t min_ring)
{
  if (ntiles <= 0) return;
  if (t_in.dma >= 3) {
    // the DM block of the tile table is the workgroup's wave count: no
    // shape can be changed here, so a shape that does not fit is an error
    if (!dd_tile_fits(t_in, max_spread, min_ring))
      throw HipError{hipErrorInvalidValue, "dedispersion tile (one trial per wave) does not fit",
                     __LINE__};
    const int lrow = lrow_for(t_in, max_spread);
    if (t_in.dma == 5) {
      launch_wd(d_beams, d_tiles, ntiles, lrow / 1024, s);
    } else if (t_in.dma == 4) {
      if (t_in.chunk == 4) launch_w<4, true>(d_beams, d_tiles, ntiles, lrow, s);
      else launch_w<8, true>(d_beams, d_tiles, ntiles, lrow, s);
    } else {
      if (t_in.chunk == 4) launch_w<4, false>(d_beams, d_tiles, ntiles, lrow, s);
      else launch_w<8, false>(d_beams, d_tiles, ntiles, lrow, s);
    }
    HAMD_CHECK(hipGetLastError());
    return;
  }
  DDTile t = t_in;
  // shrink the channel chunk (then drop the second DMA buffer) until the
  // workgroup's LDS fits the launch limit
  while (dd_lds_bytes(t, max_spread) > DD_LDS_BUDGET && t.chunk > 8) t.chunk /= 2;
  if (dd_lds_bytes(t, max_spread) > DD_LDS_BUDGET && t.dma == 1) t.dma = 0;
  if (dd_lds_bytes(t, max_spread) > DD_LDS_BUDGET)
    throw HipError{hipErrorInvalidValue, "dedispersion window exceeds LDS (delay spread too wide)",
                   __LINE__};
  const int lrow = lrow_for(t, max_spread);
  // the prefetch and DMA loaders wrap columns with a single subtraction
  if (t.dma >= 1 && min_ring < (int64_t)lrow + 16) t.dma = 0;
  if (t.dma == 2 && t.chunk > 16) t.dma = 0;
  switch (t.lanes_per_dm) {
    case 8: launch_g<8>(d_beams, d_tiles, ntiles, t.chunk, lrow, t.dma, s); break;
    case 16: launch_g<16>(d_beams, d_tiles, ntiles, t.chunk, lrow, t.dma, s); break;
    case 32: launch_g<32>(d_beams, d_tiles, ntiles, t.chunk, lrow, t.dma, s); break;
    case 64: launch_g<64>(d_beams, d_tiles, ntiles, t.chunk, lrow, t.dma, s); break;
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
