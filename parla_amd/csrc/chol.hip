// Cholesky factorisation and right triangular solve for CholeskyQR.
//
// CholeskyQR2 orthonormalises a tall-skinny Y (m x k) with two rounds of
//   G = Y^T Y (MFMA GEMM),  G = R^T R (potrf),  Y <- Y R^{-1} (trsm_right),
// so the O(m k^2) work is GEMM-shaped and the only sequential parts act on
// the k x k Gram matrix.  It replaces the Householder QR inside the power
// iteration stabilizer (utils/linalg_wrappers.py:6-7) where Y is well enough
// conditioned (the caller falls back to Householder when potrf fails).
//
// potrf: one 512-thread workgroup, blocked right-looking with 32-wide
//   panels (see the kernel).  trtri: R^{-1} one column per workgroup, so that
//   Y R^{-1} is a plain MFMA GEMM.  trsm_right: the in-place alternative for
//   callers that cannot afford the k x k inverse; one workgroup per tile of
//   rows, staged in LDS, 16 lanes of one wavefront per row.
#include "common.h"
#include "trsm_dev.h"
#include "trsv_dev.h"

namespace {

constexpr int CB = 32;
constexpr int POTRF_MAX_N = 512;  // block row of R staged in LDS: 32 x 512 doubles
constexpr int POTRF_THREADS = 512;
constexpr int SYRK_G = 4;  // trailing-update tiles per wavefront whose loads are issued together


// Wavefront 0 factors the 32 x 32 diagonal block D = R11^T R11 held in Rd:
// lane c holds column c in registers; at step j the finished row j of R11
// goes to LDS (Rd row j, 1 / R_jj to rinv) and every lane reads it back with
// same-address (broadcast) reads.  Returns 0, or j + 1 when pivot j is not
// positive and finite (the wavefront stops there).
__device__ __forceinline__ int potrf_factor32(double (*Rd)[CB + 1], double* rinv, int lane) {
  const int c = pamd_opaque(lane & (CB - 1));
  double col[CB];
#pragma unroll
  for (int i = 0; i < CB; ++i) col[i] = Rd[i][c];
  PAMD_WAVE_SYNC();  // Rd is overwritten row by row below
  int fail = 0;
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    const double djj = wave_bcast(col[j], j);
    if (!fail && !(djj > 0.0 && djj < INFINITY)) fail = j + 1;
    if (fail) continue;  // wave-uniform: djj is a broadcast
    // 1/sqrt(djj): hardware estimate + two Newton steps y <- y (3 - d y^2) / 2
    // (each squares the relative error), then sqrt(djj) = djj * y
    double y = __builtin_amdgcn_rsq(djj);
    const double h = 0.5 * djj;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    const double r = djj * y;
    // R11[j][c] for c >= j; lanes c < j produce entries below the diagonal
    // that no later step reads (lanes 32..63 repeat lanes 0..31)
    const double v = (c == j) ? r : col[j] * y;
    Rd[j][c] = v;
    if (lane == 0) rinv[j] = y;
    PAMD_WAVE_SYNC();
    double rowj[CB];
#pragma unroll
    for (int i = 0; i < CB; ++i) rowj[i] = (i > j) ? Rd[j][i] : 0.0;
    __builtin_amdgcn_sched_barrier(0);  // issue the row's reads together, then wait once
#pragma unroll
    for (int i = 0; i < CB; ++i)  // constant trip count: unrolls before the j loop does
      if (i > j) col[i] -= rowj[i] * v;
  }
  return fail;
}

// Blocked right-looking upper Cholesky in one 512-thread workgroup, 32-wide
// blocks.  Per block (G = the not yet factored part of A):
//   1. wavefront 0 factors the 32 x 32 diagonal block D = R11^T R11: lane c
//      holds column c in registers; at step j the finished row j of R11 goes
//      to LDS and every lane reads it back with same-address (broadcast)
//      reads, so the only cross-lane move per step is the pivot's readlane;
//   2. the block row R12 = R11^{-T} G12 is a forward substitution, one thread
//      per column of G12, in axpy order (row l, once known, updates all later
//      rows), with R11 broadcast from LDS and multiplications by 1 / R_ll;
//   3. the trailing upper triangle G22 -= R12^T R12 is a rank-32 SYRK on
//      v_mfma_f64_16x16x4f64, 16 x 16 upper-triangle tiles spread over the
//      wavefronts.
// Where a step needs a row of R11 or an operand tile, all of its LDS reads
// are issued before the arithmetic (sched_barrier), so the step waits for
// LDS once instead of once per read.
// Rows nb..31 of the staged block are zero and the diagonal block is padded
// with the identity, so a last block narrower than 32 needs no special case.
__device__ __forceinline__ void potrf_block(int64_t n, double* __restrict__ A, int64_t lda, int* __restrict__ info) {
  __shared__ double Rd[CB][CB + 1];  // the diagonal block; R11 row j replaces its row j at step j
  __shared__ double rinv[CB];        // 1 / R_jj
  __shared__ double Rs[CB][POTRF_MAX_N];
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nthr = POTRF_THREADS, nwaves = POTRF_THREADS / 64;
  const int lr = lane & 15, lk = lane >> 4;  // MFMA operand coordinates
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int64_t jb = 0; jb < n; jb += CB) {
    const int nb = (int)imin64(CB, n - jb);
    const int64_t c0 = jb + nb;
    const int nt = (int)(n - c0);
    const int ntp = (nt + 15) & ~15;  // staged width, padded to whole tiles
    for (int e = tid; e < CB * CB; e += nthr) {
      const int i = e / CB, j = e % CB;
      Rd[i][j] = (i < nb && j < nb) ? A[(jb + i) * lda + jb + j] : ((i == j) ? 1.0 : 0.0);
    }
    for (int e = tid; e < CB * ntp; e += nthr) {
      const int i = e / ntp, c = e % ntp;
      Rs[i][c] = (i < nb && c < nt) ? A[(jb + i) * lda + c0 + c] : 0.0;
    }
    __syncthreads();
    if (wave == 0) {
      const int fail = potrf_factor32(Rd, rinv, lane);
      if (fail && lane == 0 && fail <= nb) bad = (int)(jb + fail);
    }
    __syncthreads();
    if (bad) break;
    for (int e = tid; e < nb * nb; e += nthr) {
      const int i = e / nb, j = e % nb;
      if (j >= i) A[(jb + i) * lda + jb + j] = Rd[i][j];
    }
    if (nt == 0) break;
    // 2. block row, thread per column (nt < POTRF_THREADS, a single pass)
    static_assert(POTRF_MAX_N - CB <= POTRF_THREADS, "one thread per trailing column");
    if (tid < nt) {  // block row R12 = R11^{-T} G12, column c = tid: forward substitution in
       // axpy order (row l, once known, updates all later rows), R11 broadcast
       // from LDS, multiplications by 1 / R_ll
      const int c = tid;
      double t[CB];
    #pragma unroll
      for (int l = 0; l < CB; ++l) t[l] = Rs[l][c];
    #pragma unroll
      for (int l = 0; l < CB; ++l) {
        double rowl[CB];
    #pragma unroll
        for (int i = 0; i < CB; ++i) rowl[i] = (i > l) ? Rd[l][i] : 0.0;
        const double ril = rinv[l];
        __builtin_amdgcn_sched_barrier(0);
        const double y = t[l] * ril;
        t[l] = y;
    #pragma unroll
        for (int i = 0; i < CB; ++i)
          if (i > l) t[i] -= rowl[i] * y;
      }
    #pragma unroll
      for (int l = 0; l < CB; ++l) Rs[l][c] = t[l];  // rows >= nb stay zero
    }
    __syncthreads();
    for (int e = tid; e < nb * nt; e += nthr) {
      const int i = e / nt, c = e % nt;
      A[(jb + i) * lda + c0 + c] = Rs[i][c];
    }
    // 3. trailing SYRK on upper-triangle tiles (ti <= tc); tile index t
    //    enumerates row ti's tiles tc = ti .. T-1 in order.  A wavefront
    //    takes its tiles SYRK_G at a time and issues the group's global loads
    //    together, so it waits for memory once per group, not once per tile.
    const int T = ntp >> 4;
    const int ntiles = T * (T + 1) / 2;
    for (int t0 = wave; t0 < ntiles; t0 += SYRK_G * nwaves) {
      int tis[SYRK_G], tcs[SYRK_G];
      double cv[SYRK_G][4];
#pragma unroll
      for (int g = 0; g < SYRK_G; ++g) {
        const int t = t0 + g * nwaves;
        int ti = 0, rem = (t < ntiles) ? t : 0;
        while (rem >= T - ti) {
          rem -= T - ti;
          ++ti;
        }
        tis[g] = (t < ntiles) ? ti : -1;
        tcs[g] = ti + rem;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti * 16 + lk + 4 * r, cc = tcs[g] * 16 + lr;
          cv[g][r] = (t < ntiles && i < nt && cc < nt && cc >= i) ? A[(c0 + i) * lda + c0 + cc] : 0.0;
        }
      }
#pragma unroll
      for (int g = 0; g < SYRK_G; ++g) {
        if (tis[g] < 0) continue;  // wave-uniform
        const int ti = tis[g], tc = tcs[g];
        double av[CB / 4], bv[CB / 4];
#pragma unroll
        for (int q = 0; q < CB / 4; ++q) {
          av[q] = Rs[4 * q + lk][ti * 16 + lr];
          bv[q] = Rs[4 * q + lk][tc * 16 + lr];
        }
        __builtin_amdgcn_sched_barrier(0);
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < CB / 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti * 16 + lk + 4 * r, cc = tc * 16 + lr;
          if (i < nt && cc < nt && cc >= i) A[(c0 + i) * lda + c0 + cc] = cv[g][r] - acc[r];
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) info[0] = bad;
}

// Look-ahead variant for n <= LA_MAXN (the CholeskyQR Gram matrices of the
// power iteration): the block row of the current block and of the next one
// live in two LDS buffers, so the next block is formed straight into LDS and
// its factorisation overlaps the rest of the current trailing update.  Per
// block, with R12 of the current block in Rs[cur]:
//   (a) all wavefronts: the next block's diagonal block and block row
//       (trailing tile rows 0 and 1) = A - R12^T R12, into Rd and Rs[cur ^ 1];
//   (b) wavefront 0 factors that diagonal block while wavefronts 1..7 update
//       the rest of the trailing upper triangle in global memory;
//   (c) all: the next block row R12' = R11'^{-T} (...), written to A with R11'.
constexpr int LA_MAXN = 288;
constexpr int LA_W = LA_MAXN - CB;  // staged trailing width

__device__ __forceinline__ void potrf_la(int64_t n, double* __restrict__ A, int64_t lda, int* __restrict__ info) {
  __shared__ double Rd[CB][CB + 1];
  __shared__ double rinv[CB];
  __shared__ double Rs[2][CB][LA_W];
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int nthr = POTRF_THREADS, nwaves = POTRF_THREADS / 64;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid == 0) bad = 0;
  // block 0: stage, factor, block row
  {
    const int nb = (int)imin64(CB, n);
    const int nt = (int)(n - nb);
    const int ntp = (nt + 15) & ~15;
    for (int e = tid; e < CB * CB; e += nthr) {
      const int i = e / CB, j = e % CB;
      Rd[i][j] = (i < nb && j < nb) ? A[i * lda + j] : ((i == j) ? 1.0 : 0.0);
    }
    for (int e = tid; e < CB * ntp; e += nthr) {
      const int i = e / ntp, c = e % ntp;
      Rs[0][i][c] = (i < nb && c < nt) ? A[i * lda + nb + c] : 0.0;
    }
    __syncthreads();
    if (wave == 0) {
      const int fail = potrf_factor32(Rd, rinv, lane);
      if (fail && lane == 0 && fail <= nb) bad = fail;
    }
    __syncthreads();
    if (bad) {
      if (tid == 0) info[0] = bad;
      return;
    }
    for (int e = tid; e < nb * nb; e += nthr) {
      const int i = e / nb, j = e % nb;
      if (j >= i) A[i * lda + j] = Rd[i][j];
    }
    if (tid < nt) {
      double(*R0)[LA_W] = Rs[0];
      {  // block row R12 = R11^{-T} G12, column c = tid: forward substitution in
         // axpy order (row l, once known, updates all later rows), R11 broadcast
         // from LDS, multiplications by 1 / R_ll
        const int c = tid;
        double t[CB];
      #pragma unroll
        for (int l = 0; l < CB; ++l) t[l] = R0[l][c];
      #pragma unroll
        for (int l = 0; l < CB; ++l) {
          double rowl[CB];
      #pragma unroll
          for (int i = 0; i < CB; ++i) rowl[i] = (i > l) ? Rd[l][i] : 0.0;
          const double ril = rinv[l];
          __builtin_amdgcn_sched_barrier(0);
          const double y = t[l] * ril;
          t[l] = y;
      #pragma unroll
          for (int i = 0; i < CB; ++i)
            if (i > l) t[i] -= rowl[i] * y;
        }
      #pragma unroll
        for (int l = 0; l < CB; ++l) R0[l][c] = t[l];  // rows >= nb stay zero
      }
      for (int i = 0; i < nb; ++i) A[i * lda + nb + tid] = Rs[0][i][tid];
    }
    __syncthreads();
  }
  int cur = 0;
  for (int64_t jb = 0; jb < n; jb += CB) {
    const int nb = (int)imin64(CB, n - jb);
    const int64_t c0 = jb + nb;  // first row / column of the trailing matrix
    const int nt = (int)(n - c0);
    if (nt <= 0) break;
    const int T = ((nt + 15) & ~15) >> 4;
    const int nb2 = nt < CB ? nt : CB;  // the next block
    const int nt2 = nt - nb2;
    const int ntp2 = (nt2 + 15) & ~15;
    double(*Rn)[LA_W] = Rs[cur ^ 1];
    const double(*Rc)[LA_W] = Rs[cur];
    // (a) trailing tile rows 0 and 1 (the next block's rows) into LDS; padding
    //     of the next diagonal block (identity) and block row (zeros) first
    for (int e = tid; e < CB * CB; e += nthr) {
      const int i = e / CB, j = e % CB;
      if (i >= nb2 || j >= nb2) Rd[i][j] = (i == j) ? 1.0 : 0.0;
    }
    for (int e = tid; e < CB * ntp2; e += nthr) {
      const int i = e / ntp2, c = e % ntp2;
      if (i >= nb2 || c >= nt2) Rn[i][c] = 0.0;
    }
    {
      const int nta = (T >= 2) ? (2 * T - 1) : T;  // tiles (0, 0..T-1) and (1, 1..T-1)
      for (int t = wave; t < nta; t += nwaves) {
        const int ti = (t < T) ? 0 : 1, tc = (t < T) ? t : t - T + 1;
        double av[CB / 4], bv[CB / 4], cv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti * 16 + lk + 4 * r, cc = tc * 16 + lr;
          cv[r] = (i < nt && cc < nt && cc >= i) ? A[(c0 + i) * lda + c0 + cc] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < CB / 4; ++q) {
          av[q] = Rc[4 * q + lk][ti * 16 + lr];
          bv[q] = Rc[4 * q + lk][tc * 16 + lr];
        }
        __builtin_amdgcn_sched_barrier(0);
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < CB / 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti * 16 + lk + 4 * r, cc = tc * 16 + lr;
          if (i < nb2 && cc < nt && cc >= i) {
            const double v = cv[r] - acc[r];
            if (cc < nb2)
              Rd[i][cc] = v;
            else
              Rn[i][cc - nb2] = v;
          }
        }
      }
    }
    __syncthreads();
    // (b) factor the next diagonal block || the rest of the trailing update
    if (wave == 0) {
      const int fail = potrf_factor32(Rd, rinv, lane);
      if (fail && lane == 0 && fail <= nb2) bad = (int)(c0 + fail);
    } else if (T > 2) {
      const int ntb = (T - 2) * (T - 1) / 2;  // tiles ti = 2 .. T-1, tc = ti .. T-1
      for (int t0 = wave - 1; t0 < ntb; t0 += SYRK_G * (nwaves - 1)) {
        int tis[SYRK_G], tcs[SYRK_G];
        double cv[SYRK_G][4];
#pragma unroll
        for (int g = 0; g < SYRK_G; ++g) {
          const int t = t0 + g * (nwaves - 1);
          int ti = 2, rem = (t < ntb) ? t : 0;
          while (rem >= T - ti) {
            rem -= T - ti;
            ++ti;
          }
          tis[g] = (t < ntb) ? ti : -1;
          tcs[g] = ti + rem;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = ti * 16 + lk + 4 * r, cc = tcs[g] * 16 + lr;
            cv[g][r] = (t < ntb && i < nt && cc < nt && cc >= i) ? A[(c0 + i) * lda + c0 + cc] : 0.0;
          }
        }
#pragma unroll
        for (int g = 0; g < SYRK_G; ++g) {
          if (tis[g] < 0) continue;  // wave-uniform
          const int ti = tis[g], tc = tcs[g];
          double av[CB / 4], bv[CB / 4];
#pragma unroll
          for (int q = 0; q < CB / 4; ++q) {
            av[q] = Rc[4 * q + lk][ti * 16 + lr];
            bv[q] = Rc[4 * q + lk][tc * 16 + lr];
          }
          __builtin_amdgcn_sched_barrier(0);
          d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int q = 0; q < CB / 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc, 0, 0, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = ti * 16 + lk + 4 * r, cc = tc * 16 + lr;
            if (i < nt && cc < nt && cc >= i) A[(c0 + i) * lda + c0 + cc] = cv[g][r] - acc[r];
          }
        }
      }
    }
    __syncthreads();
    if (bad) break;
    // (c) the next block's R11 and R12 to A
    for (int e = tid; e < nb2 * nb2; e += nthr) {
      const int i = e / nb2, j = e % nb2;
      if (j >= i) A[(c0 + i) * lda + c0 + j] = Rd[i][j];
    }
    if (tid < nt2) {
      {  // block row R12 = R11^{-T} G12, column c = tid: forward substitution in
         // axpy order (row l, once known, updates all later rows), R11 broadcast
         // from LDS, multiplications by 1 / R_ll
        const int c = tid;
        double t[CB];
      #pragma unroll
        for (int l = 0; l < CB; ++l) t[l] = Rn[l][c];
      #pragma unroll
        for (int l = 0; l < CB; ++l) {
          double rowl[CB];
      #pragma unroll
          for (int i = 0; i < CB; ++i) rowl[i] = (i > l) ? Rd[l][i] : 0.0;
          const double ril = rinv[l];
          __builtin_amdgcn_sched_barrier(0);
          const double y = t[l] * ril;
          t[l] = y;
      #pragma unroll
          for (int i = 0; i < CB; ++i)
            if (i > l) t[i] -= rowl[i] * y;
        }
      #pragma unroll
        for (int l = 0; l < CB; ++l) Rn[l][c] = t[l];  // rows >= nb stay zero
      }
      for (int i = 0; i < nb2; ++i) A[(c0 + i) * lda + c0 + nb2 + tid] = Rn[i][tid];
    }
    __syncthreads();
    cur ^= 1;
  }
  if (tid == 0) info[0] = bad;
}

__global__ __launch_bounds__(POTRF_THREADS) void potrf_kernel(int64_t n, double* __restrict__ A, int64_t lda,
                                                              int* __restrict__ info) {
  potrf_block(n, A, lda, info);
}

__global__ __launch_bounds__(POTRF_THREADS) void potrf_la_kernel(int64_t n, double* __restrict__ A, int64_t lda,
                                                                 int* __restrict__ info) {
  potrf_la(n, A, lda, info);
}

// problem i: the n[i] x n[i] matrix at pool + off[i] (ld n[i]); one workgroup each
__global__ __launch_bounds__(POTRF_THREADS) void potrf_batched_kernel(const int64_t* __restrict__ nv,
                                                                      const int64_t* __restrict__ off,
                                                                      double* __restrict__ pool,
                                                                      int* __restrict__ info) {
  const int64_t n = nv[blockIdx.x];
  potrf_block(n, pool + off[blockIdx.x], imax64(n, 1), info + blockIdx.x);
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
  trsv_upper_block<8>(0, sz, R, ldr, xs, smem + sz);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Ri[i * ldi + j] = (i <= j) ? xs[i] : 0.0;
}

// batched R^{-1}: grid (max n, problems); column j of problem i when j < n[i]
__global__ __launch_bounds__(256) void trtri_batched_kernel(const int64_t* __restrict__ nv,
                                                            const int64_t* __restrict__ off,
                                                            const double* __restrict__ Rp, double* __restrict__ Rip,
                                                            double* __restrict__ RiTp) {
  extern __shared__ double smem[];
  const int64_t n = nv[blockIdx.y];
  const int64_t j = blockIdx.x;
  if (j >= n) return;
  const double* R = Rp + off[blockIdx.y];
  double* Ri = Rip + off[blockIdx.y];
  const int sz = (int)j + 1;
  double* xs = smem;
  for (int i = threadIdx.x; i < sz; i += blockDim.x) xs[i] = (i == sz - 1) ? 1.0 : 0.0;
  __syncthreads();
  trsv_upper_block<8>(0, sz, R, n, xs, smem + sz);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Ri[i * n + j] = (i <= j) ? xs[i] : 0.0;
  if (RiTp != nullptr) {  // row j of the transpose: contiguous
    double* RiT = RiTp + off[blockIdx.y];
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) RiT[j * n + i] = (i <= j) ? xs[i] : 0.0;
  }
}

__global__ __launch_bounds__(64) void trsm_dinv_kernel(int k, const double* __restrict__ R, int64_t ldr,
                                                       double* __restrict__ Dinv) {
  trsm_dinv_body(blockIdx.x, k, R, ldr, Dinv);
}

__global__ __launch_bounds__(256) void trsm_strip_kernel(int64_t m, int k, const double* __restrict__ Y, int64_t ldy,
                                                         const double* __restrict__ R, int64_t ldr,
                                                         const double* __restrict__ Dinv, double* __restrict__ X,
                                                         int64_t ldx) {
  trsm_strip_body(blockIdx.x, m, k, Y, ldy, R, ldr, Dinv, X, ldx);
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

// ---------------------------------------------------------------------------
// Small k x k helpers of CholeskyQR2, one workgroup each (k <= 512):
//   combine: out = A + alpha B + sigma I, or H(A + alpha B + sigma I) where H
//            keeps the upper triangle and halves the diagonal; optionally
//            ||A + alpha B + sigma I||_F (rows summed in order by each thread,
//            then the waves in order);
//   stats:   min |G_ii|, max |G_ii|, the norm above and the potrf info, packed
//            into four doubles for one read-back.
// ---------------------------------------------------------------------------
constexpr int SC_THREADS = 1024;

__global__ __launch_bounds__(SC_THREADS) void small_combine_kernel(int k, const double* __restrict__ A, int64_t lda,
                                                                  const double* __restrict__ B, int64_t ldb,
                                                                  double alpha, double sigma, int upper_half,
                                                                  double* __restrict__ out, int64_t ldo,
                                                                  double* __restrict__ nrm) {
  constexpr int U = 8;  // elements per thread whose loads are in flight together
  __shared__ double red[SC_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double ss = 0.0;
  const int total = k * k;  // k <= 512
  // thread tid takes the elements tid, tid + SC_THREADS, ... in that order
  for (int base = 0; base < total; base += SC_THREADS * U) {
    double va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + tid + SC_THREADS * u;
      const int i = e / k, j = e - (e / k) * k;
      va[u] = (e < total) ? A[(int64_t)i * lda + j] : 0.0;
      vb[u] = (e < total && B != nullptr) ? B[(int64_t)i * ldb + j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + tid + SC_THREADS * u;
      if (e < total) {
        const int i = e / k, j = e - (e / k) * k;
        double v = va[u];
        if (B != nullptr) v += alpha * vb[u];
        if (i == j) v += sigma;
        ss += v * v;
        if (upper_half) v = (j > i) ? v : ((j == i) ? 0.5 * v : 0.0);
        out[(int64_t)i * ldo + j] = v;
      }
    }
  }
  if (nrm != nullptr) {
    ss = wave_sum(ss);
    if (lane == 0) red[wave] = ss;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < SC_THREADS / 64; ++w) t += red[w];
      nrm[0] = sqrt(t);
    }
  }
}

// the same without the norm, over many workgroups: 8 elements per thread
constexpr int SCW_PER_BLOCK = 256 * 8;
__global__ __launch_bounds__(256) void small_combine_wide_kernel(int k, const double* __restrict__ A, int64_t lda,
                                                                 const double* __restrict__ B, int64_t ldb,
                                                                 double alpha, double sigma, int upper_half,
                                                                 double* __restrict__ out, int64_t ldo) {
  constexpr int U = SCW_PER_BLOCK / 256;
  const int total = k * k;  // k <= 512
  const int base = (int)blockIdx.x * SCW_PER_BLOCK + (int)threadIdx.x;
  double va[U], vb[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = base + 256 * u;
    const int i = e / k, j = e - (e / k) * k;
    va[u] = (e < total) ? A[(int64_t)i * lda + j] : 0.0;
    vb[u] = (e < total && B != nullptr) ? B[(int64_t)i * ldb + j] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = base + 256 * u;
    if (e < total) {
      const int i = e / k, j = e - (e / k) * k;
      double v = va[u];
      if (B != nullptr) v += alpha * vb[u];
      if (i == j) v += sigma;
      if (upper_half) v = (j > i) ? v : ((j == i) ? 0.5 * v : 0.0);
      out[(int64_t)i * ldo + j] = v;
    }
  }
}

__global__ __launch_bounds__(256) void cholqr_stats_kernel(int k, const double* __restrict__ G, int64_t ldg,
                                                          const int* __restrict__ info,
                                                          const double* __restrict__ enorm,
                                                          double* __restrict__ out) {
  __shared__ double smin[4], smax[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double mn = INFINITY, mx = 0.0;
  for (int i = tid; i < k; i += 256) {
    const double d = fabs(G[(int64_t)i * ldg + i]);
    mn = fmin(mn, d);
    mx = fmax(mx, d);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, off, 64));
    mx = fmax(mx, __shfl_xor(mx, off, 64));
  }
  if (lane == 0) {
    smin[wave] = mn;
    smax[wave] = mx;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; ++w) {
      mn = fmin(mn, smin[w]);
      mx = fmax(mx, smax[w]);
    }
    out[0] = mn;
    out[1] = mx;
    out[2] = (enorm != nullptr) ? enorm[0] : 0.0;
    out[3] = (double)info[0];
  }
}

}  // namespace

extern "C" int pamd_potrf_upper_f64(int64_t n, double* A, int64_t lda, int* info, void* stream) {
  PAMD_ARG_CHECK(n >= 0 && n <= POTRF_MAX_N, 1);
  PAMD_ARG_CHECK(A != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 3);
  PAMD_ARG_CHECK(info != nullptr, 4);
  if (n >= 1 && n <= LA_MAXN)
    potrf_la_kernel<<<1, POTRF_THREADS, 0, (hipStream_t)stream>>>(n, A, lda, info);
  else
    potrf_kernel<<<1, POTRF_THREADS, 0, (hipStream_t)stream>>>(n, A, lda, info);
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

extern "C" int pamd_potrf_batched_f64(int64_t nprob, const int64_t* n, const int64_t* off, double* pool, int* info,
                                      void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(n != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(off != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(pool != nullptr || nprob == 0, 4);
  PAMD_ARG_CHECK(info != nullptr || nprob == 0, 5);
  if (nprob == 0) return PAMD_OK;
  potrf_batched_kernel<<<(unsigned)nprob, POTRF_THREADS, 0, (hipStream_t)stream>>>(n, off, pool, info);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_trtri_batched_f64(int64_t nprob, int64_t nmax, const int64_t* n, const int64_t* off,
                                      const double* R, double* Rinv, double* RinvT, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(nmax >= 0 && nmax <= PAMD_TRSV_MAX_N, 2);
  PAMD_ARG_CHECK(n != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(off != nullptr || nprob == 0, 4);
  PAMD_ARG_CHECK(R != nullptr || nprob == 0, 5);
  PAMD_ARG_CHECK(Rinv != nullptr || nprob == 0, 6);
  if (nprob == 0 || nmax == 0) return PAMD_OK;
  const size_t shmem = (size_t)(nmax + TRSV_SCRATCH_DOUBLES) * sizeof(double);
  dim3 grid((unsigned)nmax, (unsigned)nprob);
  trtri_batched_kernel<<<grid, 256, shmem, (hipStream_t)stream>>>(n, off, R, Rinv, RinvT);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

#define PAMD_TRSM_MFMA_MAX_K 512

extern "C" int pamd_trsm_mfma_workspace(int64_t k, int64_t* bytes) {
  PAMD_ARG_CHECK(k >= 0 && k <= PAMD_TRSM_MFMA_MAX_K, 1);
  PAMD_ARG_CHECK(bytes != nullptr, 2);
  *bytes = imax64(1, pamd_cdiv(k, TS_B)) * TS_B * TS_B * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_trsm_right_upper_mfma_f64(int64_t m, int64_t k, const double* R, int64_t ldr, const double* Y,
                                              int64_t ldy, double* X, int64_t ldx, void* workspace, void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(k >= 0 && k <= PAMD_TRSM_MFMA_MAX_K, 2);
  PAMD_ARG_CHECK(R != nullptr || k == 0, 3);
  PAMD_ARG_CHECK(ldr >= k && ldr >= 1, 4);
  PAMD_ARG_CHECK(Y != nullptr || m * k == 0, 5);
  PAMD_ARG_CHECK(ldy >= k && ldy >= 1, 6);
  PAMD_ARG_CHECK(X != nullptr || m * k == 0, 7);
  PAMD_ARG_CHECK(ldx >= k && ldx >= 1, 8);
  PAMD_ARG_CHECK(workspace != nullptr || k == 0, 9);
  if (m == 0 || k == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)pamd_cdiv(k, TS_B);
  trsm_dinv_kernel<<<nb, 64, 0, st>>>((int)k, R, ldr, (double*)workspace);
  PAMD_LAUNCH_CHECK();
  trsm_strip_kernel<<<(unsigned)pamd_cdiv(m, TS_B), 256, trsm_strip_lds((int)k), st>>>(m, (int)k, Y, ldy, R, ldr,
                                                                     (const double*)workspace, X, ldx);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_small_combine_f64(int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                                      double alpha, double sigma, int upper_half, double* out, int64_t ldo,
                                      double* nrm, void* stream) {
  PAMD_ARG_CHECK(k >= 0 && k <= 512, 1);
  PAMD_ARG_CHECK(A != nullptr || k == 0, 2);
  PAMD_ARG_CHECK(lda >= k && lda >= 1, 3);
  PAMD_ARG_CHECK(B == nullptr || ldb >= k, 5);
  PAMD_ARG_CHECK(upper_half == 0 || upper_half == 1, 8);
  PAMD_ARG_CHECK(out != nullptr || k == 0, 9);
  PAMD_ARG_CHECK(ldo >= k && ldo >= 1, 10);
  if (k == 0 && nrm == nullptr) return PAMD_OK;
  if (nrm == nullptr) {  // no reduction: spread over the chip
    small_combine_wide_kernel<<<(unsigned)pamd_cdiv(k * k, SCW_PER_BLOCK), 256, 0, (hipStream_t)stream>>>(
        (int)k, A, lda, B, ldb, alpha, sigma, upper_half, out, ldo);
  } else {
    small_combine_kernel<<<1, SC_THREADS, 0, (hipStream_t)stream>>>((int)k, A, lda, B, ldb, alpha, sigma, upper_half,
                                                                   out, ldo, nrm);
  }
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_cholqr_stats_f64(int64_t k, const double* G, int64_t ldg, const int* info, const double* enorm,
                                     double* out, void* stream) {
  PAMD_ARG_CHECK(k >= 1 && k <= 512, 1);
  PAMD_ARG_CHECK(G != nullptr, 2);
  PAMD_ARG_CHECK(ldg >= k, 3);
  PAMD_ARG_CHECK(info != nullptr, 4);
  PAMD_ARG_CHECK(out != nullptr, 6);
  cholqr_stats_kernel<<<1, 256, 0, (hipStream_t)stream>>>((int)k, G, ldg, info, enorm, out);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
