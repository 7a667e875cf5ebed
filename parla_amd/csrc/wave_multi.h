// Wavefront reductions of several values at once (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

// Wave-wide sums of N values at once (N a power of two, <= 64): in round k
// each lane swaps half of its remaining values with lane ^ 2^k and keeps the
// half selected by bit k of its lane, so log2 N rounds leave one value per
// lane, and the remaining rounds add it across the wave.  Returns, in lane l,
// the wave total of v[idx], idx = sum_k bit_k(l) N / 2^(k+1) (the low log2 N
// bits of l reversed); every lane with the same low bits holds the same
// total.  About 2N - 1 + 6 - log2 N shuffles instead of 6 N.
template <int N>
__device__ __forceinline__ double wave_sum_multi(double (&v)[N], int lane) {
  constexpr int LOG2N = (N >= 64) ? 6 : (N >= 32) ? 5 : (N >= 16) ? 4 : (N >= 8) ? 3 : (N >= 4) ? 2 : (N >= 2) ? 1 : 0;
  static_assert((1 << LOG2N) == N && N <= 64, "N a power of two, at most 64");
#pragma unroll
  for (int round = 0; round < LOG2N; ++round) {
    const bool hi = (lane >> round) & 1;
#pragma unroll
    for (int i = 0; i < (N >> (round + 1)); ++i) {
      const int h = N >> (round + 1);
      // both values in registers before the selects (a select of the two
      // addresses would keep v in scratch memory)
      double lo = v[i], up = v[i + h];
      asm volatile("" : "+v"(lo), "+v"(up));
      const double send = hi ? lo : up;
      const double keep = hi ? up : lo;
      v[i] = keep + __shfl_xor(send, 1 << round, 64);
    }
  }
  double t = v[0];
#pragma unroll
  for (int m = N; m < 64; m <<= 1) t += __shfl_xor(t, m, 64);
  return t;
}

template <int N>
__device__ __forceinline__ int wave_sum_multi_index(int lane) {
  int idx = 0;
#pragma unroll
  for (int k = 0; (N >> k) > 1; ++k) idx += ((lane >> k) & 1) * (N >> (k + 1));
  return idx;
}
