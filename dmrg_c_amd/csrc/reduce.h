// Deterministic block reductions shared by the single-solve and batched kernels.  A solve
// is bit-identical in both only because both sum with these functions: per wavefront by
// shuffles, per workgroup over the wavefronts in order, and over a launch's blocks by one
// finish block in block-index order.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace dmrgx {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the workgroup; the result is valid in thread 0.  Uses LDS, so callers that
// reduce twice in one kernel must separate the calls with __syncthreads().
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double lds[kBlock / 64];
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) t += lds[i];
  }
  return t;
}

// This block's share of a reduction, stored with a plain store at *slot.
__device__ __forceinline__ void block_partial_to(double v, double *slot) {
  const double t = block_sum(v);
  if (threadIdx.x == 0) *slot = t;
}
__device__ __forceinline__ void block_partial(double v, double *partial) { block_partial_to(v, partial + blockIdx.x); }

// Sum of partial[0..nblk) by one workgroup (valid in thread 0): thread t adds
// partial[t], partial[t + kBlock], ... in index order, with eight loads in flight ahead of
// the additions, then the workgroup sum.
__device__ __forceinline__ double finish_sum(const double *__restrict__ partial, int nblk) {
  double s = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * kBlock < nblk; i += 8 * kBlock) {
    double p[8];
#pragma unroll
    for (int u = 0; u < 8; u++) p[u] = partial[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; u++) s += p[u];
  }
  for (; i < nblk; i += kBlock) s += partial[i];
  return block_sum(s);
}

}  // namespace dmrgx
