// Shared helpers for the gfx950 kernels of the randomized-NLA library.
//
// Conventions used by every kernel and entry point in this library:
//  * matrices are row-major (numpy C order) with an explicit leading
//    dimension, i.e. element (i, j) of an r x c matrix X lives at X[i*ld + j];
//  * all device pointers are caller-owned; entry points never allocate;
//  * every entry point launches on the hipStream_t it is given and returns an
//    int status (0 ok, <0 bad argument -index, >0 numerical event).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/parla_amd.h"

#define PAMD_WAVE 64

// Return -idx from the calling entry point when cond is false.
#define PAMD_ARG_CHECK(cond, idx) \
  do {                            \
    if (!(cond)) return -(idx);   \
  } while (0)

#define PAMD_HIP_CHECK(expr)                            \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return PAMD_ERR_HIP;          \
  } while (0)

#define PAMD_LAUNCH_CHECK()                             \
  do {                                                  \
    hipError_t e_ = hipGetLastError();                  \
    if (e_ != hipSuccess) return PAMD_ERR_HIP;          \
  } while (0)

__host__ __device__ static inline int64_t pamd_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ static inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
__host__ __device__ static inline int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// wavefront / block reductions (64-wide wavefronts)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, PAMD_WAVE);
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, PAMD_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, PAMD_WAVE));
  return v;
}

// Sum over the whole block; every thread receives the total.  `red` must hold
// at least blockDim.x / 64 doubles of shared memory.  Contains barriers, so
// every thread of the block must call it.  The summation order is fixed
// (lane tree, then waves in index order), so results are run-to-run identical.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (PAMD_WAVE - 1);
  const int wid = threadIdx.x / PAMD_WAVE;
  const int nw = blockDim.x / PAMD_WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}
