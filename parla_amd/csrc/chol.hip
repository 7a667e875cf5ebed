// Cholesky factorisation and right triangular solve for CholeskyQR.
//
// CholeskyQR2 orthonormalises a tall-skinny Y (m x k) with two rounds of
//   G = Y^T Y (MFMA GEMM),  G = R^T R (potrf),  Y <- Y R^{-1} (trsm_right),
// so the O(m k^2) work is GEMM-shaped and the only sequential parts act on
// the k x k Gram matrix.  It replaces the Householder QR inside the power
// iteration stabilizer (utils/linalg_wrappers.py:6-7) where Y is well enough
// conditioned (the caller falls back to Householder when potrf fails).
//
// potrf: one 1024-thread workgroup, blocked right-looking with 32-wide
//   panels (see the kernel).  trtri: R^{-1} one column per workgroup, so that
//   Y R^{-1} is a plain MFMA GEMM.  trsm_right: the in-place alternative for
//   callers that cannot afford the k x k inverse; one workgroup per tile of
//   rows, staged in LDS, 16 lanes of one wavefront per row.
#include "common.h"
#include "trsv_dev.h"

namespace {

constexpr int CB = 32;
constexpr int POTRF_MAX_N = 512;  // block row of R staged in LDS: 32 x 512 doubles

// Blocked right-looking upper Cholesky, one 512-thread workgroup, 32-wide
// blocks.  Per block:
//   1. one wavefront factors the 32 x 32 diagonal block with its columns in
//      registers (lane c holds column c; pivots and row entries move by
//      v_readlane) and forms the block's inverse the same way;
//   2. the block row R[jb:jb+32, c0:n] = D^{-T} G[jb:jb+32, c0:n] is a product
//      with that inverse, spread over all threads (no sequential substitution);
//   3. the trailing upper triangle gets the rank-32 update from LDS in 4 x 4
//      register tiles.
// The diagonal-block inverse has forward error ~ kappa(D) u; CholeskyQR2's
// second round (kappa ~ 1) removes its effect on the final Q.
__global__ __launch_bounds__(512) void potrf_kernel(int64_t n, double* __restrict__ A, int64_t lda,
                                                    int* __restrict__ info) {
  __shared__ double D[CB][CB + 1];
  __shared__ double Di[CB][CB + 1];
  __shared__ double Rs[CB][POTRF_MAX_N];
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nthr = blockDim.x;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int64_t jb = 0; jb < n; jb += CB) {
    const int nb = (int)imin64(CB, n - jb);
    const int64_t c0 = jb + nb;
    const int nt = (int)(n - c0);
    for (int e = tid; e < CB * CB; e += nthr) {
      const int i = e / CB, j = e % CB;
      D[i][j] = (i < nb && j < nb) ? A[(jb + i) * lda + jb + j] : ((i == j) ? 1.0 : 0.0);
    }
    for (int e = tid; e < nb * nt; e += nthr) {
      const int i = e / nt, c = e % nt;
      Rs[i][c] = A[(jb + i) * lda + c0 + c];
    }
    __syncthreads();
    if (wave == 0) {
      const int c = lane & (CB - 1);
      double col[CB];
#pragma unroll
      for (int i = 0; i < CB; ++i) col[i] = D[i][c];
      int fail = 0;
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const double djj = wave_bcast(col[j], j);
        if (!fail && !(djj > 0.0 && djj < INFINITY)) fail = j + 1;
        if (fail) continue;  // wave-uniform: djj is a broadcast
        const double r = sqrt(djj);
        if (c == j)
          col[j] = r;
        else if (c > j)
          col[j] = col[j] / r;
        const double rjc = col[j];
#pragma unroll
        for (int i = j + 1; i < CB; ++i) {
          const double rji = wave_bcast(col[j], i);
          if (c >= i) col[i] -= rji * rjc;
        }
      }
      if (fail) {
        if (lane == 0 && fail <= nb) bad = (int)(jb + fail);
      } else {
        // D^{-1}, column c: backward substitution on e_c; D[i][l] is lane l's col[i]
        double x[CB];
#pragma unroll
        for (int i = CB - 1; i >= 0; --i) {
          const double rii = 1.0 / wave_bcast(col[i], i);
          double sacc = 0.0;
#pragma unroll
          for (int l = i + 1; l < CB; ++l) {
            const double dil = wave_bcast(col[i], l);
            if (l <= c) sacc += dil * x[l];
          }
          x[i] = (i == c) ? rii : ((i < c) ? -sacc * rii : 0.0);
        }
        if (lane < CB) {
#pragma unroll
          for (int i = 0; i < CB; ++i) {
            D[i][c] = (i <= c) ? col[i] : 0.0;
            Di[i][c] = x[i];
          }
        }
      }
    }
    __syncthreads();
    if (bad) break;
    for (int e = tid; e < nb * nb; e += nthr) {
      const int i = e / nb, j = e % nb;
      if (j >= i) A[(jb + i) * lda + jb + j] = D[i][j];
    }
    // block row R[i][c] = sum_{l <= i} Di[l][i] G[l][c].  Row i needs G rows
    // 0..i only, so rows are produced bottom-up in groups of 8 and written over
    // the G rows they no longer need.
    for (int ig = ((nb - 1) / 8) * 8; ig >= 0; ig -= 8) {
      const int nrow = imin64(8, nb - ig);
      const int nout = nrow * nt;
      constexpr int MAXO = (8 * POTRF_MAX_N + 511) / 512;
      double o[MAXO];
#pragma unroll
      for (int q = 0; q < MAXO; ++q) {
        const int e = tid + q * nthr;
        double acc = 0.0;
        if (e < nout) {
          const int i = ig + e / nt, cc = e % nt;
          for (int l = 0; l <= i; ++l) acc += Di[l][i] * Rs[l][cc];
        }
        o[q] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < MAXO; ++q) {
        const int e = tid + q * nthr;
        if (e < nout) {
          const int i = ig + e / nt, cc = e % nt;
          Rs[i][cc] = o[q];
          A[(jb + i) * lda + c0 + cc] = o[q];
        }
      }
      __syncthreads();
    }
    // trailing update on 4 x 4 tiles (ti <= tc) of the upper triangle
    const int nT = (nt + 3) / 4;
    const int ntiles = nT * (nT + 1) / 2;
    for (int t = tid; t < ntiles; t += nthr) {
      int ti = 0, rem = t;
      while (rem >= nT - ti) {
        rem -= nT - ti;
        ++ti;
      }
      const int tc = ti + rem;
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
      for (int l = 0; l < nb; ++l) {
        double ri[4], rc[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int i = ti * 4 + a, cc = tc * 4 + a;
          ri[a] = (i < nt) ? Rs[l][i] : 0.0;
          rc[a] = (cc < nt) ? Rs[l][cc] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] += ri[a] * rc[b];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int i = ti * 4 + a, cc = tc * 4 + b;
          if (i < nt && cc < nt && cc >= i) A[(c0 + i) * lda + c0 + cc] -= acc[a][b];
        }
    }
    __syncthreads();
  }
  if (tid == 0) info[0] = bad;
}

// Rinv = R^{-1} (upper; zeros below): workgroup j solves R x = e_j on rows 0..j.
__global__ __launch_bounds__(256) void trtri_kernel(int64_t n, const double* __restrict__ R, int64_t ldr,
                                                    double* __restrict__ Ri, int64_t ldi) {
  extern __shared__ double smem[];
  const int64_t j = blockIdx.x;
  const int sz = (int)j + 1;
  double* xs = smem;
  for (int i = threadIdx.x; i < sz; i += blockDim.x) xs[i] = (i == sz - 1) ? 1.0 : 0.0;
  __syncthreads();
  trsv_upper_block(0, sz, R, ldr, xs, smem + sz);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Ri[i * ldi + j] = (i <= j) ? xs[i] : 0.0;
}

// rows of B per workgroup: 16 lanes per row, tile of rows x n doubles <= 64 KiB
inline int trsm_rows(int64_t n) {
  int64_t r = 65536 / (8 * imax64(n, 1));
  r = imin64(64, r) / 4 * 4;
  return (int)imax64(4, r);
}

// B (m x n) <- B R^{-1}, R upper n x n.  LDS holds the `rows`-row tile of B.
__global__ __launch_bounds__(1024) void trsm_right_kernel(int64_t m, int64_t n, int rows,
                                                          const double* __restrict__ R, int64_t ldr,
                                                          double* __restrict__ B, int64_t ldb) {
  extern __shared__ double T[];  // rows x n
  const int64_t r0 = (int64_t)blockIdx.x * rows;
  const int nr = (int)imin64(rows, m - r0);
  for (int64_t e = threadIdx.x; e < (int64_t)rows * n; e += blockDim.x) {
    const int64_t rr = e / n, c = e % n;
    T[e] = (rr < nr) ? B[(r0 + rr) * ldb + c] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane >> 4, q = lane & 15;  // 4 rows per wave, 16 lanes per row
  const int rr = wave * 4 + sub;
  double* row = T + (int64_t)rr * n;
  for (int64_t c = 0; c < n; ++c) {
    double s = 0.0;
    for (int64_t l = q; l < c; l += 16) s += row[l] * R[l * ldr + c];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    s += __shfl_xor(s, 8, 64);
    const double v = (row[c] - s) / R[c * ldr + c];
    PAMD_WAVE_SYNC();
    if (q == 0) row[c] = v;
    PAMD_WAVE_SYNC();
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)nr * n; e += blockDim.x) {
    const int64_t r = e / n, c = e % n;
    B[(r0 + r) * ldb + c] = T[r * n + c];
  }
}

}  // namespace

extern "C" int pamd_potrf_upper_f64(int64_t n, double* A, int64_t lda, int* info, void* stream) {
  PAMD_ARG_CHECK(n >= 0 && n <= POTRF_MAX_N, 1);
  PAMD_ARG_CHECK(A != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 3);
  PAMD_ARG_CHECK(info != nullptr, 4);
  potrf_kernel<<<1, 512, 0, (hipStream_t)stream>>>(n, A, lda, info);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

#define PAMD_TRSM_MAX_N 2048

extern "C" int pamd_trsm_right_upper_f64(int64_t m, int64_t n, const double* R, int64_t ldr, double* B, int64_t ldb,
                                         void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= PAMD_TRSM_MAX_N, 2);
  PAMD_ARG_CHECK(R != nullptr || n == 0, 3);
  PAMD_ARG_CHECK(ldr >= n && ldr >= 1, 4);
  PAMD_ARG_CHECK(B != nullptr || m * n == 0, 5);
  PAMD_ARG_CHECK(ldb >= n && ldb >= 1, 6);
  if (m == 0 || n == 0) return PAMD_OK;
  const int rows = trsm_rows(n);
  const size_t shmem = (size_t)rows * n * sizeof(double);
  trsm_right_kernel<<<(unsigned)pamd_cdiv(m, rows), rows * 16, shmem, (hipStream_t)stream>>>(m, n, rows, R, ldr, B,
                                                                                             ldb);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_trtri_upper_f64(int64_t n, const double* R, int64_t ldr, double* Rinv, int64_t ldi,
                                    void* stream) {
  PAMD_ARG_CHECK(n >= 0 && n <= PAMD_TRSV_MAX_N, 1);
  PAMD_ARG_CHECK(R != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(ldr >= n && ldr >= 1, 3);
  PAMD_ARG_CHECK(Rinv != nullptr || n == 0, 4);
  PAMD_ARG_CHECK(ldi >= n && ldi >= 1, 5);
  if (n == 0) return PAMD_OK;
  const size_t shmem = (size_t)(n + TRSV_SCRATCH_DOUBLES) * sizeof(double);
  trtri_kernel<<<(unsigned)n, 256, shmem, (hipStream_t)stream>>>(n, R, ldr, Rinv, ldi);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
