// Blocked BF16 operand layout shared by the split-precision GEMM and the exact
// BF16 sketch, and the direct global -> LDS load they stream it with.
//
// A K-contiguous operand (rows x kp BF16, kp a multiple of 64) is stored as
// 1 KiB blocks of 16 rows x 32 K: element (i, k) sits at
//   ((i / 16) (kp / 32) + k / 32) 512 + (i % 16) 32 + 8 (c ^ ((i % 16) / 4)) + k % 8,
// c = (k % 32) / 8 the 16-byte chunk of the row.  A GEMM tile's K step is a
// set of whole blocks, each moved by one global_load_lds_dwordx4 per
// wavefront (64 lanes x 16 bytes, contiguous in LDS), and the chunk swizzle
// makes the MFMA operand reads (16 rows, one chunk) conflict-free in LDS
// without padding.
#pragma once

#include "common.h"

typedef __attribute__((address_space(3))) void blk_lds_void;
typedef const __attribute__((address_space(1))) void blk_glb_void;

namespace {

// offset (in BF16 elements) of element (i, k) of a blocked array with kp columns
__device__ __forceinline__ int64_t blk_off(int64_t i, int64_t k, int64_t kp) {
  const int r = (int)(i & 15), c = (int)((k & 31) >> 3);
  return ((i >> 4) * (kp >> 5) + (k >> 5)) * 512 + r * 32 + ((c ^ (r >> 2)) << 3) + (k & 7);
}

// this lane's 16-byte chunk (in uint4 units) within a block for an MFMA
// operand read: row lane % 16, logical chunk lane / 16
__device__ __forceinline__ int blk_frag(int lane) {
  const int lr = lane & 15, lk = lane >> 4;
  return lr * 4 + (lk ^ (lr >> 2));
}

__device__ __forceinline__ uint4 blk_pack(const uint16_t (&v)[8]) {
  uint4 w;
  w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
  w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
  w.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
  w.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
  return w;
}

// one wavefront moves the 1 KiB block at src (lane l: bytes 16 l .. 16 l + 15)
// to LDS at dst; completion is counted by vmcnt
__device__ __forceinline__ void blk_load_lds(const uint4* src_block, uint4* dst_block, int lane) {
  __builtin_amdgcn_global_load_lds((blk_glb_void*)(src_block + lane), (blk_lds_void*)dst_block, 16, 0, 0);
}

}  // namespace
