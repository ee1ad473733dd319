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
// partial[t], partial[t + kBlock], ... in index order, then the workgroup sum.  The loads
// go out in batches of D per thread (the partials were just written by other workgroups,
// often on other XCDs, so each batch costs a full memory round trip); the additions stay
// in index order, so the batch size does not change the result.  load() issues the first
// batch, sum() adds it (and any further batches) and reduces, so that a kernel can overlap
// the first round trip with other work.
template <int D>
struct PartialSum {
  double p[D];
  __device__ __forceinline__ void load(const double *__restrict__ partial, int nblk) {
    const int i = threadIdx.x;
#pragma unroll
    for (int u = 0; u < D; u++) p[u] = (i + u * kBlock < nblk) ? partial[i + u * kBlock] : 0.0;
  }
  __device__ __forceinline__ double sum(const double *__restrict__ partial, int nblk) {
    double s = 0.0;
    int i = threadIdx.x;
#pragma unroll
    for (int u = 0; u < D; u++)
      if (i + u * kBlock < nblk) s += p[u];
    for (i += D * kBlock; i < nblk; i += D * kBlock) {
      load_at(partial, nblk, i);
#pragma unroll
      for (int u = 0; u < D; u++)
        if (i + u * kBlock < nblk) s += p[u];
    }
    return block_sum(s);
  }

 private:
  __device__ __forceinline__ void load_at(const double *__restrict__ partial, int nblk, int i) {
#pragma unroll
    for (int u = 0; u < D; u++) p[u] = (i + u * kBlock < nblk) ? partial[i + u * kBlock] : 0.0;
  }
};

// finish_sum's result computed by each wavefront on its own (no LDS, no barrier): lane l
// plays the finishing workgroup's threads l, l + 64, ... in turn, each adding its partials
// in index order, and their wavefront sums are added in wavefront order -- the same
// operations in the same order as finish_sum.  load() issues D partials per played thread.
template <int D>
struct WavePartialSum {
  static constexpr int W = kBlock / 64;
  double p[W][D];
  __device__ __forceinline__ void load(const double *__restrict__ partial, int nblk) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int w = 0; w < W; w++)
#pragma unroll
      for (int u = 0; u < D; u++) {
        const int i = 64 * w + l + u * kBlock;
        p[w][u] = (i < nblk) ? partial[i] : 0.0;
      }
  }
  __device__ __forceinline__ double sum(const double *__restrict__ partial, int nblk) const {
    const int l = threadIdx.x & 63;
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      const int t = 64 * w + l;
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < D; u++)
        if (t + u * kBlock < nblk) s += p[w][u];
      for (int i = t + D * kBlock; i < nblk; i += kBlock) s += partial[i];
      tot += __shfl(wave_sum(s), 0, 64);
    }
    return tot;
  }
};

__device__ __forceinline__ double finish_sum(const double *__restrict__ partial, int nblk) {
  PartialSum<16> ps;
  ps.load(partial, nblk);
  return ps.sum(partial, nblk);
}

}  // namespace dmrgx
