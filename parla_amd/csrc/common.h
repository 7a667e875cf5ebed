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

// Order LDS accesses between lanes of one wavefront (no block barrier): a
// compiler barrier plus wavefront-scope release/acquire fences.
#define PAMD_WAVE_SYNC()                                   \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
  } while (0)

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

// Library-internal entry shared between sources (not part of the C ABI):
// C = alpha op(A) op(B) + beta Cin, the input C read from Cin (blas.hip).
int pamd_gemm_cin_f64(int transA, int transB, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                      int64_t lda, const double* B, int64_t ldb, double beta, const double* Cin, int64_t ldcin,
                      double* C, int64_t ldc, void* workspace, int64_t workspace_bytes, void* stream);

__host__ __device__ static inline int64_t pamd_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ static inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
__host__ __device__ static inline int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// wavefront / block reductions (64-wide wavefronts)
//
// Cross-lane traffic goes through DPP row operations (VALU, a few cycles) and
// v_readlane, not ds_bpermute shuffles: on gfx950 a 64-lane FP64 shuffle tree
// costs ~200 ns per reduction, the DPP sequence a small fraction of that.
// ---------------------------------------------------------------------------

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// MI355X dispatches workgroup w to XCD w % 8, each XCD with its own L2.
// Returns the tile that workgroup w should process so that XCD j handles the
// contiguous tile range [start_j, start_j + count_j): neighbouring tiles,
// which share operand rows, then meet in the same L2.  A bijection on
// [0, total) for any total.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t w, uint32_t total) {
  constexpr uint32_t NXCD = 8;
  const uint32_t q = total / NXCD, r = total % NXCD;
  const uint32_t xcd = w % NXCD, local = w / NXCD;
  return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Returns x unchanged but opaque to the optimiser, at the point where it is
// called: values derived from it are not hoisted out of the enclosing loop.
// Used where loop-invariant lane masks would otherwise be precomputed for
// every unrolled step and spilled.
__device__ __forceinline__ int pamd_opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// value of v in lane `lane` (uniform index), for every lane
__device__ __forceinline__ double wave_bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ float wave_bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Full-wave sum, returned to every lane.  quad_perm(1,0,3,2), quad_perm(2,3,0,1),
// row_ror:4, row_ror:8 give each 16-lane row its total; row_bcast:15 (rows 1, 3)
// and row_bcast:31 (rows 2, 3) fold the rows into lane 63.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xb1>(v);
  v += dpp_f64<0x4e>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return wave_bcast(v, 63);
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xb1>(v);
  v += dpp_f32<0x4e>(v);
  v += dpp_f32<0x124>(v);
  v += dpp_f32<0x128>(v);
  v += dpp_f32<0x142, 0xa>(v);
  v += dpp_f32<0x143, 0xc>(v);
  return wave_bcast(v, 63);
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
