// Right triangular solve X = Y R^{-1} (R upper k x k, k <= 512) on FP64
// MFMA, shared by the single-problem and the batched entry points.
#pragma once

#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

// ---------------------------------------------------------------------------
// X = Y R^{-1} for tall Y (m x k, k <= 512) on FP64 MFMA: one workgroup per
// 16-row strip, held in LDS and solved in place left to right in 16-column
// blocks J:  X_J = (Y_J - X_{<J} R_{<J,J}) D_J  with D_J the inverse of the
// diagonal block R_JJ (from trsm_dinv_body).  The update's K range is split
// over the four wavefronts and their 16 x 16 partials are summed in wave
// order, so the result does not depend on timing.
// ---------------------------------------------------------------------------
constexpr int TS_B = 16;

// D_J = R_JJ^{-1} (16 x 16, identity beyond k) by one 64-thread workgroup
__device__ __forceinline__ void trsm_dinv_body(int J, int k, const double* __restrict__ R, int64_t ldr,
                                               double* __restrict__ Dinv) {
  __shared__ double Rb[TS_B][TS_B + 1];
  const int j0 = J * TS_B, t = threadIdx.x;
  for (int e = t; e < TS_B * TS_B; e += 64) {
    const int i = e / TS_B, j = e % TS_B;
    Rb[i][j] = (j0 + i < k && j0 + j < k) ? R[(int64_t)(j0 + i) * ldr + j0 + j] : ((i == j) ? 1.0 : 0.0);
  }
  __syncthreads();
  if (t < TS_B) {  // column t of the inverse: backward substitution on e_t
    double x[TS_B];
#pragma unroll
    for (int i = TS_B - 1; i >= 0; --i) {
      double s = (i == t) ? 1.0 : 0.0;
#pragma unroll
      for (int l = 0; l < TS_B; ++l)
        if (l > i) s -= Rb[i][l] * x[l];
      x[i] = (i <= t) ? s / Rb[i][i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < TS_B; ++i) Dinv[(int64_t)J * TS_B * TS_B + i * TS_B + t] = x[i];
  }
}

// strip `strip` (rows 16 strip .. +15) by one 256-thread workgroup; dynamic
// LDS: trsm_strip_lds(k) bytes
__device__ __forceinline__ void trsm_strip_body(int64_t strip, int64_t m, int k, const double* __restrict__ Y,
                                                int64_t ldy, const double* __restrict__ R, int64_t ldr,
                                                const double* __restrict__ Dinv, double* __restrict__ X,
                                                int64_t ldx) {
  extern __shared__ double smem[];
  const int kp = (k + TS_B - 1) / TS_B * TS_B;
  const int lds = kp + 1;
  double* S = smem;                  // TS_B x lds: the strip, Y overwritten by X block by block
  double* P = S + TS_B * lds;        // 4 x 16 x 17 wave partials
  double* U = P + 4 * TS_B * 17;     // 16 x 17 updated block
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = strip * TS_B;
  for (int e = tid; e < TS_B * kp; e += 256) {
    const int i = e / kp, j = e % kp;
    S[i * lds + j] = (r0 + i < m && j < k) ? Y[(r0 + i) * ldy + j] : 0.0;
  }
  __syncthreads();
  const int nb = kp / TS_B;
  for (int J = 0; J < nb; ++J) {
    const int j0 = J * TS_B;
    const int jc = j0 + lr;  // this lane's column of R / of the block
    // partial of X_{<J} R_{<J,J}: wave w takes the k-steps kk = 4 (w + 4 q),
    // in order; the R values of four steps are loaded before their MFMAs
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int kk = 4 * wave; kk < j0; kk += 64) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k2 = kk + 16 * u;
        b[u] = (k2 < j0 && jc < k) ? R[(int64_t)(k2 + lk) * ldr + jc] : 0.0;
        a[u] = (k2 < j0) ? S[lr * lds + k2 + lk] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (kk + 16 * u < j0) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) P[(wave * TS_B + lk + 4 * r) * 17 + lr] = acc[r];
    __syncthreads();
    if (wave == 0) {
      // U = Y_J - sum_w P_w, then X_J = U D_J on MFMA (K = 16: four steps)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = lk + 4 * r;
        double u = S[i * lds + j0 + lr];
#pragma unroll
        for (int w = 0; w < 4; ++w) u -= P[(w * TS_B + i) * 17 + lr];
        U[i * 17 + lr] = u;
      }
      PAMD_WAVE_SYNC();
      d4 x = {0.0, 0.0, 0.0, 0.0};
      const double* D = Dinv + (int64_t)J * TS_B * TS_B;
#pragma unroll
      for (int q = 0; q < TS_B; q += 4) {
        const double a = U[lr * 17 + q + lk];
        const double b = D[(q + lk) * TS_B + lr];
        x = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, x, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) S[(lk + 4 * r) * lds + j0 + lr] = x[r];
    }
    __syncthreads();
  }
  for (int e = tid; e < TS_B * k; e += 256) {
    const int i = e / k, j = e % k;
    if (r0 + i < m) X[(r0 + i) * ldx + j] = S[i * lds + j];
  }
}

inline size_t trsm_strip_lds(int k) {
  const int kp = (k + TS_B - 1) / TS_B * TS_B;
  return ((size_t)TS_B * (kp + 1) + 4 * TS_B * 17 + TS_B * 17) * sizeof(double);
}

}  // namespace
