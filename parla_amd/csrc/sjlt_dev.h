// SJLT pieces shared by the single-problem and the batched sketch kernels:
// the per-column draw and the offsets pass of the CSC -> CSR transpose.
#pragma once

#include "common.h"
#include "rng.h"

namespace {

// Column j of a wide SJLT with n_short rows: k rows drawn one at a time,
// each redrawn on a repeat of an earlier row of the column (up to 64
// attempts, then the first unused row); with replacement when n_short < k.
// Bit l of `signs` is set when entry l is negative.  KMAX >= k is a compile-
// time bound so the picked rows stay in registers.
template <int KMAX>
__device__ __forceinline__ void sjlt_column(int64_t n_short, int k, uint64_t seed, int64_t j, int32_t (&picked)[KMAX],
                                            uint64_t& signs) {
  const bool with_repl = n_short < k;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  signs = 0;
#pragma unroll
  for (int l = 0; l < KMAX; ++l) {
    if (l >= k) continue;  // constant trip count keeps `picked` in registers
    int32_t r = 0;
    uint32_t sign = 0;
    bool placed = false;
    for (uint32_t attempt = 0; attempt < 64 && !placed; ++attempt) {
      const U4 u = philox4x32_10(U4{(uint32_t)j, (uint32_t)(j >> 32), (uint32_t)l, 0x510E527Fu + attempt}, k0, k1);
      r = (int32_t)(((uint64_t)u.x * (uint64_t)n_short) >> 32);
      sign = u.y & 1u;
      bool dup = false;
      if (!with_repl) {
#pragma unroll
        for (int q = 0; q < KMAX; ++q)
          if (q < l) dup |= (picked[q] == r);
      }
      placed = !dup;
    }
    if (!placed) {  // 64 straight repeats (practically unreachable): first unused row
      for (r = 0; r < (int32_t)n_short; ++r) {
        bool dup = false;
#pragma unroll
        for (int q = 0; q < KMAX; ++q)
          if (q < l) dup |= (picked[q] == r);
        if (!dup) break;
      }
    }
    picked[l] = r;
    signs |= (uint64_t)sign << l;
  }
}

// CSC -> CSR offsets by one 1024-thread workgroup: rowptr = exclusive scan of
// the row counts summed over the nblk column blocks (cnt[b][r], row stride
// n_short), and cnt[b][r] <- the first slot of block b in row r, so that
// blocks fill every row in column order.
__device__ __forceinline__ void csr_scan_body(int64_t n_short, int nblk, int32_t* __restrict__ cnt,
                                              int64_t* __restrict__ rowptr) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t base = 0; base < n_short; base += blockDim.x) {
    const int64_t r = base + threadIdx.x;
    int64_t c = 0;
    if (r < n_short)
      for (int b = 0; b < nblk; ++b) c += cnt[(int64_t)b * n_short + r];
    // inclusive scan inside the wave
    int64_t v = c;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t t = __shfl_up(v, off, 64);
      if (lane >= off) v += t;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    int64_t wpre = 0;
    for (int w = 0; w < wave; ++w) wpre += wsum[w];
    const int64_t excl = carry + wpre + v - c;
    if (r < n_short) {
      rowptr[r] = excl;
      int64_t run = excl;
      for (int b = 0; b < nblk; ++b) {
        const int32_t t = cnt[(int64_t)b * n_short + r];
        cnt[(int64_t)b * n_short + r] = (int32_t)run;
        run += t;
      }
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry = excl + c;
    __syncthreads();
  }
  if (threadIdx.x == 0) rowptr[n_short] = carry;
}

}  // namespace
