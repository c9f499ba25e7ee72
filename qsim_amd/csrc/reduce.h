// reduce.h -- wave64 / workgroup reductions and scans in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace qhip {

__device__ inline double WaveSum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a workgroup of BS threads; the result is valid in thread 0.
// `part` is LDS scratch of BS / 64 doubles.
template <int BS>
__device__ inline double BlockSum(double v, double* part) {
  static_assert(BS % 64 == 0 && BS <= 1024, "block size");
  v = WaveSum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) part[w] = v;
  __syncthreads();
  if (w == 0) {
    v = l < BS / 64 ? part[l] : 0.0;
    v = WaveSum(v);
  }
  __syncthreads();
  return v;
}

template <int BS>
__device__ inline double BlockSum(double v) {
  __shared__ double part[BS / 64];
  return BlockSum<BS>(v, part);
}

template <int BS>
__device__ inline Dbl2 BlockSum(Dbl2 v) {
  static_assert(BS % 64 == 0 && BS <= 1024, "block size");
  __shared__ double px[BS / 64], py[BS / 64];
  v.x = WaveSum(v.x);
  v.y = WaveSum(v.y);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    px[w] = v.x;
    py[w] = v.y;
  }
  __syncthreads();
  if (w == 0) {
    v.x = l < BS / 64 ? px[l] : 0.0;
    v.y = l < BS / 64 ? py[l] : 0.0;
    v.x = WaveSum(v.x);
    v.y = WaveSum(v.y);
  }
  __syncthreads();
  return v;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ inline double WaveInclusiveScan(double v) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    double u = __shfl_up(v, o, 64);
    if (l >= o) v += u;
  }
  return v;
}

// Sums `count` partials in a fixed order with one workgroup of 256 threads;
// writes the total to out[0].
template <typename Acc>
__global__ void __launch_bounds__(256)
FinalSumKernel(const Acc* __restrict__ partials, unsigned count,
               Acc* __restrict__ out) {
  Acc acc{};
  for (unsigned i = threadIdx.x; i < count; i += 256) acc += partials[i];
  acc = BlockSum<256>(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

}  // namespace qhip
