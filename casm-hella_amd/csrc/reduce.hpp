// reduce.hpp -- the canonical 256-lane pairwise tree shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace hamd {

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
