// reduce.hpp -- the canonical 256-lane pairwise tree and the wave-wide DPP
// scans and reductions shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hamd {

// ---- wave-wide integer scans and reductions, all in the vector ALU (DPP)

// inclusive prefix sum over the wave (u32, wrapping): a Hillis-Steele scan
// within each 16-lane row by DPP row shifts (a lane with no source lane adds
// 0), then the row totals carried forward by the row broadcasts of lane 15
// (into rows 1 and 3) and of lane 31 (into rows 2 and 3).  (Shuffles through
// ds_bpermute would wait on LDS latency six times.)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

// op over the wave, in every lane: four DPP steps within each 16-lane row
// (lane pairs, quads, half rows, rows: every lane has a source), then the four
// row results read into scalar registers.  Every lane of the wave must be
// active.
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t x, Op op)
{
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));  // row_half_mirror
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));  // row_mirror
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)x, 0);
  const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)x, 16);
  const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)x, 32);
  const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
  return op(op(a, b), op(c, d));
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x)
{
  return wave_reduce(x, [](uint32_t p, uint32_t q) { return max(p, q); });
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x)
{
  return wave_reduce(x, [](uint32_t p, uint32_t q) { return min(p, q); });
}

// u32 sum, wrapping
__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
  return wave_reduce(x, [](uint32_t p, uint32_t q) { return p + q; });
}

// one DPP step of wave_max_u64: both halves moved by the same control
template <int CTRL>
__device__ __forceinline__ unsigned long long max_dpp_u64(unsigned long long x)
{
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, 0xF, 0xF, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, false);
  const unsigned long long y = ((unsigned long long)hi << 32) | lo;
  return y > x ? y : x;
}

// largest u64 of the wave, in every lane (the steps of wave_reduce)
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x)
{
  x = max_dpp_u64<0xB1>(x);
  x = max_dpp_u64<0x4E>(x);
  x = max_dpp_u64<0x141>(x);
  x = max_dpp_u64<0x140>(x);
  unsigned long long m = 0;
#pragma unroll
  for (int r = 0; r < 64; r += 16) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, r);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), r);
    const unsigned long long y = ((unsigned long long)hi << 32) | lo;
    m = y > m ? y : m;
  }
  return m;
}

// x of lane l + s for lanes l < s (the lanes the tree keeps): s = 16 through
// the swizzle crossbar (lane id xor 16 within 32), s = 8, 4, 2, 1 by DPP row
// shifts (within 16 lanes), s = 32 by a shuffle; other lanes get values the
// tree never reads.
template <int S>
__device__ __forceinline__ double lane_down(double x)
{
  if constexpr (S == 32) {
    return __shfl_down(x, 32, 64);
  } else {
    const long long b = __double_as_longlong(x);
    int lo = (int)b, hi = (int)(b >> 32);
    if constexpr (S == 16) {
      lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
      hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
    } else {
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x100 + S, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x100 + S, 0xF, 0xF, false);
    }
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  }
}

// Lane l of a 256-thread workgroup holds v[k]; afterwards lane 0 holds
//   tree(v) with steps p[l] += p[l + s] for s = 128, 64, ..., 1
// for every k.  The two cross-wave steps go through LDS (sh: N x 128
// doubles), the six in-wave steps move lanes by a shuffle, a swizzle and DPP
// row shifts; the additions and their order are those of the sequential
// pairwise tree.  In a larger workgroup the
// threads from 256 on only take part in the barriers.
template <int N>
__device__ void tree256_multi(double (&v)[N], double *sh)
{
  const int l = threadIdx.x;
  if (l >= 128 && l < 256) {
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * 128 + (l - 128)] = v[k];
  }
  __syncthreads();
  if (l < 128) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += sh[k * 128 + l];
  }
  __syncthreads();
  if (l >= 64 && l < 128) {
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * 128 + (l - 64)] = v[k];
  }
  __syncthreads();
  if (l < 64) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double x = v[k] + sh[k * 128 + l];
      x += lane_down<32>(x);
      x += lane_down<16>(x);
      x += lane_down<8>(x);
      x += lane_down<4>(x);
      x += lane_down<2>(x);
      x += lane_down<1>(x);
      v[k] = x;
    }
  }
  __syncthreads();
}

}  // namespace hamd
