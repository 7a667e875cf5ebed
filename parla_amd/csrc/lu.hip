// LU with partial pivoting returning P @ L (scipy.linalg.lu(..., permute_l=True)[0]),
// the alternative stabilizer of power iteration (utils/linalg_wrappers.py:10-12).
//
// One 1024-thread workgroup runs the unblocked right-looking factorisation on
// a workspace copy of the (tall, narrow) input: per column a block argmax
// (ties -> smallest row, as LAPACK's idamax), a row swap, a column scale and a
// rank-1 update.  Meant for the n x k matrices of power iteration (k in the
// hundreds), where it is a small fraction of the surrounding GEMMs.
#include "common.h"

namespace {

__global__ __launch_bounds__(1024) void lu_kernel(int64_t m, int64_t n, double* __restrict__ W,
                                                  int32_t* __restrict__ perm) {
  __shared__ double vbest[16];
  __shared__ int64_t ibest[16];
  __shared__ int64_t piv;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  for (int64_t i = tid; i < m; i += blockDim.x) perm[i] = (int32_t)i;
  __syncthreads();
  const int64_t k = imin64(m, n);
  for (int64_t j = 0; j < k; ++j) {
    // argmax |W[i][j]| over i >= j
    double bv = -1.0;
    int64_t bi = m;
    for (int64_t i = j + tid; i < m; i += blockDim.x) {
      const double v = fabs(W[i * n + j]);
      if (v > bv) {
        bv = v;
        bi = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int64_t oi = __shfl_xor(bi, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      vbest[wave] = bv;
      ibest[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double v = vbest[0];
      int64_t ii = ibest[0];
      for (int w = 1; w < nw; ++w)
        if (vbest[w] > v || (vbest[w] == v && ibest[w] < ii)) {
          v = vbest[w];
          ii = ibest[w];
        }
      piv = (ii < m) ? ii : j;
    }
    __syncthreads();
    const int64_t p = piv;
    if (p != j) {
      for (int64_t c = tid; c < n; c += blockDim.x) {
        const double t = W[j * n + c];
        W[j * n + c] = W[p * n + c];
        W[p * n + c] = t;
      }
      if (tid == 0) {
        const int32_t t = perm[j];
        perm[j] = perm[p];
        perm[p] = t;
      }
    }
    __syncthreads();
    const double d = W[j * n + j];
    if (d != 0.0) {
      const double inv = 1.0 / d;
      for (int64_t i = j + 1 + tid; i < m; i += blockDim.x) {
        double* row = W + i * n;
        const double l = row[j] * inv;
        row[j] = l;
        const double* prow = W + j * n;
        for (int64_t c = j + 1; c < n; ++c) row[c] -= l * prow[c];
      }
    }
    __syncthreads();
  }
}

// PL[perm[i]][c] = L[i][c] (unit diagonal, zeros above) for c < k
__global__ void lu_pl_kernel(int64_t m, int64_t n, int64_t k, const double* __restrict__ W,
                             const int32_t* __restrict__ perm, double* __restrict__ PL, int64_t ldpl) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * k) return;
  const int64_t i = idx / k, c = idx % k;
  const double v = (c < i) ? W[i * n + c] : ((c == i) ? 1.0 : 0.0);
  PL[(int64_t)perm[i] * ldpl + c] = v;
}

__global__ void copy_kernel(int64_t m, int64_t n, const double* __restrict__ A, int64_t lda, double* __restrict__ W) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  W[idx] = A[(idx / n) * lda + idx % n];
}

}  // namespace

extern "C" int pamd_lu_workspace(int64_t m, int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 3);
  *bytes = m * n * (int64_t)sizeof(double) + (m + 64) * (int64_t)sizeof(int32_t) + 256;
  return PAMD_OK;
}

extern "C" int pamd_lu_pl_f64(int64_t m, int64_t n, const double* A, int64_t lda, double* PL, int64_t ldpl,
                              void* workspace, void* stream) {
  PAMD_ARG_CHECK(m >= 0 && m < (1LL << 31), 1);
  PAMD_ARG_CHECK(n >= 0, 2);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 3);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 4);
  const int64_t k = imin64(m, n);
  PAMD_ARG_CHECK(PL != nullptr || m * k == 0, 5);
  PAMD_ARG_CHECK(ldpl >= k && ldpl >= 1, 6);
  PAMD_ARG_CHECK(workspace != nullptr, 7);
  if (m == 0 || n == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  double* W = (double*)workspace;
  int32_t* perm = (int32_t*)(W + m * n);
  copy_kernel<<<(unsigned)pamd_cdiv(m * n, 256), 256, 0, st>>>(m, n, A, lda, W);
  PAMD_LAUNCH_CHECK();
  lu_kernel<<<1, 1024, 0, st>>>(m, n, W, perm);
  PAMD_LAUNCH_CHECK();
  lu_pl_kernel<<<(unsigned)pamd_cdiv(m * k, 256), 256, 0, st>>>(m, n, k, W, perm, PL, ldpl);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
