// Block-level upper-triangular solve used by the standalone TRSV entry point
// and by the LSQR iteration kernel.
//
// The solution vector lives in LDS (global memory for n beyond what LDS
// holds).  The triangle is processed in 64-row
// blocks: the off-diagonal part of each block is a GEMV spread over all the
// block's wavefronts (row-wise for R x = b, column-wise for R^T x = b, both
// reading contiguous 512-byte row segments of the row-major R), then the
// 64 x 64 diagonal block -- staged in LDS -- is solved by one wavefront with
// one FMA per lane per step (the pivot is broadcast with v_readlane and the
// division replaced by a multiplication with the reciprocal of the diagonal).
#pragma once

#include "common.h"

#define PAMD_TRSV_MAX_N 2048
constexpr int TRSV_SCRATCH_DOUBLES = 64 * 65 + 16 * 64 + 64;

// The 64 x 64 diagonal block at rows / columns r0.. of R (bs = its size): a
// thread's share held in registers, element e = threadIdx.x + blockDim.x * k
// for k < DPT (row e / 64, column e % 64); zero outside the block.
template <int DPT>
__device__ __forceinline__ void trsv_load_diag(const double* __restrict__ R, int64_t ldr, int r0, int bs,
                                               double (&d)[DPT]) {
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const int e = threadIdx.x + blockDim.x * k;
    const int i = e >> 6, j = e & 63;
    d[k] = (e < 64 * 64 && i < bs && j < bs) ? R[(int64_t)(r0 + i) * ldr + r0 + j] : 0.0;
  }
}

// dinv[q] (8 x 8, row-major) = the inverse of R's 8 x 8 diagonal sub-block
// at rows / columns 8 q .. (identity where it passes n), for q < ceil(n / 8):
// one thread per (sub-block, column) by back substitution.  All threads of
// the block call this; the caller synchronises before dinv is read.
__device__ inline void trsv_dinv8(int n, const double* __restrict__ R, int64_t ldr, double* __restrict__ dinv) {
  const int nq = (n + 7) / 8;
  for (int t = threadIdx.x; t < nq * 8; t += blockDim.x) {
    const int q = t >> 3, j = t & 7, b0 = 8 * q;
    double z[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (i == j) ? 1.0 : 0.0;
    if (b0 + j < n) {  // a column of the triangle: every row i <= j is inside it
#pragma unroll
      for (int i = 7; i >= 0; --i) {
        if (i > j) continue;
        const double* row = R + (int64_t)(b0 + i) * ldr + b0;
        double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = i + 1; k < 8; ++k)
          if (k <= j) s -= row[k] * z[k];
        z[i] = s / row[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) dinv[(int64_t)q * 64 + i * 8 + j] = z[i];
  }
}

// All threads of the block (blockDim.x a multiple of 64, at most 1024) must
// call this.  trans = 0 solves R x = b, trans = 1 solves R^T x = b; xs holds b
// on entry and x on exit, in LDS or (for n beyond what LDS holds) in global
// memory that only this workgroup touches.  scratch: TRSV_SCRATCH_DOUBLES
// doubles of LDS.  DPT: diagonal-block elements a thread keeps in registers;
// 4096 / blockDim.x covers the whole block, a smaller value is still correct
// (the rest is read from R when it is stored to LDS).  dinv (optional, from
// trsv_dinv8): the 64 x 64 diagonal blocks are then solved in eight steps of
// 8 rows, each a product with the sub-block's inverse, instead of 64 steps of
// one row; the inverses' application error is of order u kappa(8 x 8 block).
template <int DPT = 4>
__device__ inline void trsv_upper_block(int trans, int n, const double* __restrict__ R, int64_t ldr,
                                        double* xs, double* scratch, const double* __restrict__ dinv = nullptr) {
  double(*Rd)[65] = reinterpret_cast<double(*)[65]>(scratch);
  double(*red)[64] = reinterpret_cast<double(*)[64]>(scratch + 64 * 65);
  double* rdiag = scratch + 64 * 65 + 16 * 64;  // 1 / R_ii of the current block
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = blockDim.x >> 6;
  const int nb = (n + 63) / 64;
  // The diagonal block of the next step is loaded into registers while the
  // current one is solved, and stored to LDS after the next step's
  // off-diagonal product: its memory latency is hidden twice over.
  double dnext[DPT];
  {
    const int blk = trans ? 0 : (nb - 1);
    trsv_load_diag(R, ldr, blk * 64, min(64, n - blk * 64), dnext);
  }
  for (int bi = 0; bi < nb; ++bi) {
    const int blk = trans ? bi : (nb - 1 - bi);
    const int r0 = blk * 64;
    const int bs = min(64, n - r0);
    const int r1 = r0 + bs;
    double dcur[DPT];
#pragma unroll
    for (int k = 0; k < DPT; ++k) dcur[k] = dnext[k];
    if (!trans) {
      // xs[r0:r1] -= R[r0:r1, r1:n] @ xs[r1:n]; a lane's terms in order, the
      // row values of four of them loaded before they are used
      if (r1 < n) {
        for (int i = wave; i < bs; i += nwaves) {
          const double* row = R + (int64_t)(r0 + i) * ldr;
          double s = 0.0;
          for (int j = r1 + lane; j < n; j += 256) {
            double rv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) rv[u] = (j + 64 * u < n) ? row[j + 64 * u] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (j + 64 * u < n) s += rv[u] * xs[j + 64 * u];
          }
          s = wave_sum(s);
          if (lane == 0) xs[r0 + i] -= s;
        }
      }
    } else {
      // xs[r0:r1] -= R[0:r0, r0:r1]^T @ xs[0:r0]; rows i = wave, + nwaves, ...
      // in order, four rows' values loaded before they are used
      double s = 0.0;
      if (lane < bs)
        for (int i = wave; i < r0; i += 4 * nwaves) {
          double rv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int ii = i + u * nwaves;
            rv[u] = (ii < r0) ? R[(int64_t)ii * ldr + r0 + lane] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int ii = i + u * nwaves;
            if (ii < r0) s += rv[u] * xs[ii];
          }
        }
      red[wave][lane] = s;
    }
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const int e = tid + blockDim.x * k;
      if (e < 64 * 64) {
        const int i = e >> 6, j = e & 63;
        Rd[i][j] = (i < bs && j < bs) ? dcur[k] : ((i == j) ? 1.0 : 0.0);
        if (i == j) rdiag[i] = (i < bs) ? 1.0 / dcur[k] : 0.0;
      }
    }
    for (int e = tid + blockDim.x * DPT; e < 64 * 64; e += blockDim.x) {
      const int i = e >> 6, j = e & 63;
      const double v = (i < bs && j < bs) ? R[(int64_t)(r0 + i) * ldr + r0 + j] : 0.0;
      Rd[i][j] = (i < bs && j < bs) ? v : ((i == j) ? 1.0 : 0.0);
      if (i == j) rdiag[i] = (i < bs) ? 1.0 / v : 0.0;
    }
    __syncthreads();
    if (bi + 1 < nb) {
      const int nblk = trans ? bi + 1 : (nb - 2 - bi);
      trsv_load_diag(R, ldr, nblk * 64, min(64, n - nblk * 64), dnext);
    }
    if (wave == 0 && dinv) {
      double y = (lane < bs) ? xs[r0 + lane] : 0.0;
      if (trans) {
        double t = 0.0;
        for (int w = 0; w < nwaves; ++w) t += red[w][lane];
        y -= t;
      }
      // this lane's row (trans: column) of its sub-block's inverse (zero past
      // the triangle, whose sub-blocks dinv does not hold)
      const double* dsb = dinv + (int64_t)((r0 + (lane & ~7)) >> 3) * 64;
      double di[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        di[k] = (lane < bs) ? (trans ? dsb[k * 8 + (lane & 7)] : dsb[(lane & 7) * 8 + k]) : 0.0;
      const int nsb = (bs + 7) / 8;
      for (int q = 0; q < nsb; ++q) {
        const int i0 = 8 * (trans ? q : (nsb - 1 - q));
        double r[8], yk[8], xk[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = trans ? Rd[i0 + k][lane] : Rd[lane][i0 + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) yk[k] = wave_bcast(y, i0 + k);
        // both 8-term sums as four pairs added in a tree: a dependent chain of
        // four operations instead of eight
        double px[4], pu[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) px[k] = fma(di[2 * k + 1], yk[2 * k + 1], di[2 * k] * yk[2 * k]);
        const double xl = (px[0] + px[1]) + (px[2] + px[3]);
#pragma unroll
        for (int k = 0; k < 8; ++k) xk[k] = wave_bcast(xl, i0 + k);
        // every lane forms the update (so R's LDS reads stay at the top of the
        // step, ahead of the broadcasts), the lanes it applies to keep it
#pragma unroll
        for (int k = 0; k < 4; ++k) pu[k] = fma(r[2 * k + 1], xk[2 * k + 1], r[2 * k] * xk[2 * k]);
        const double yu = y - ((pu[0] + pu[1]) + (pu[2] + pu[3]));
        const bool below = trans ? (lane >= i0 + 8) : (lane < i0);
        y = (lane >= i0 && lane < i0 + 8) ? xl : (below ? yu : y);
      }
      if (lane < bs) xs[r0 + lane] = y;
    } else if (wave == 0) {
      double y = (lane < bs) ? xs[r0 + lane] : 0.0;
      const double rd = rdiag[lane];  // lane i holds 1 / R_ii
      if (trans) {
        double t = 0.0;
        for (int w = 0; w < nwaves; ++w) t += red[w][lane];
        y -= t;
        // forward: lane l needs Rd[i][l] at step i; fetch 8 steps at a time
        for (int i0 = 0; i0 < bs; i0 += 8) {
          double r[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) r[k] = Rd[i0 + k][lane];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int i = i0 + k;
            if (i < bs) {
              const double xi = wave_bcast(y, i) * wave_bcast(rd, i);
              if (lane == i)
                y = xi;
              else if (lane > i)
                y -= r[k] * xi;
            }
          }
        }
      } else {
        // backward: lane l needs Rd[l][i] at step i
        for (int i0 = ((bs - 1) / 8) * 8; i0 >= 0; i0 -= 8) {
          double r[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) r[k] = Rd[lane][i0 + k];
#pragma unroll
          for (int k = 7; k >= 0; --k) {
            const int i = i0 + k;
            if (i < bs) {
              const double xi = wave_bcast(y, i) * wave_bcast(rd, i);
              if (lane == i)
                y = xi;
              else if (lane < i)
                y -= r[k] * xi;
            }
          }
        }
      }
      if (lane < bs) xs[r0 + lane] = y;
    }
    __syncthreads();
  }
}
