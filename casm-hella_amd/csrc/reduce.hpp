// reduce.hpp -- the canonical 256-lane pairwise tree shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace hamd {

// Lane l of a 256-thread workgroup holds v[k]; afterwards lane 0 holds
//   tree(v) with steps p[l] += p[l + s] for s = 128, 64, ..., 1
// for every k.  The two cross-wave steps go through LDS (sh: N x 128
// doubles), the six in-wave steps use shuffles; the additions and their order
// are those of the sequential pairwise tree.  In a larger workgroup the
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
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) x += __shfl_down(x, s, 64);
      v[k] = x;
    }
  }
  __syncthreads();
}

}  // namespace hamd
