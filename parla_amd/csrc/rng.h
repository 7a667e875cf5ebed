// Counter-based random numbers shared by the sketch kernels.
#pragma once

#include "common.h"

namespace {

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in (0, 1]
__device__ __forceinline__ double u53_open0(uint32_t hi, uint32_t lo) {
  const uint64_t b = (((uint64_t)hi << 32) | lo) >> 11;
  return ((double)b + 1.0) * (1.0 / 9007199254740992.0);
}

// Standard normal pair for counter `pair` (one Philox call -> two normals).
__device__ __forceinline__ void normal_pair(uint64_t seed, uint64_t pair, double& z0, double& z1) {
  const U4 r = philox4x32_10(U4{(uint32_t)pair, (uint32_t)(pair >> 32), 0x6A09E667u, 0u}, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  const double u1 = u53_open0(r.x, r.y);
  const double u2 = u53_open0(r.z, r.w);
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  z0 = rad * c;
  z1 = rad * s;
}

}  // namespace
