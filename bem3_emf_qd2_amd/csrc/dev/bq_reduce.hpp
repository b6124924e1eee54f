// bq_reduce.hpp -- fixed-order reductions (deterministic: the same inputs give the same
// bits regardless of what else runs on the device).
#pragma once
#include <hip/hip_runtime.h>

namespace bq {

// butterfly sum over aligned groups of W lanes (W power of two <= 64); every lane of the
// group ends with the group total
template <int W, int K>
__device__ __forceinline__ void group_sum(double (&v)[K])
{
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, W);
  }
}

// sum over a 256-thread block; result valid in every thread.  lds: K*4 doubles.
template <int K>
__device__ __forceinline__ void block_sum256(double (&v)[K], double* lds)
{
  group_sum<64, K>(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * 4 + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (lds[k * 4 + 0] + lds[k * 4 + 1]) + (lds[k * 4 + 2] + lds[k * 4 + 3]);
  __syncthreads();
}

}  // namespace bq
