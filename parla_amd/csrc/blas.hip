// Dense FP64 kernels: MFMA GEMM, GEMV, 2-norm, upper-triangular solves.
//
// GEMM: 64 or 96 x 64 or 96 output tile per 256-thread workgroup (four
// wavefronts, each owning a quadrant of 16x16 v_mfma_f64_16x16x4f64 blocks;
// the tile edge per dimension is the one that pads that dimension less), K staged
// through LDS 16 at a time with the next K tile prefetched into registers
// while the current one is multiplied.  Optional split-K writes per-slice
// partials that a second kernel sums in slice order (deterministic).
#include "common.h"
#include "trsv_dev.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int GBK = 16;

// op(A)[m][k] for the BM x 16 tile (m0.., k0..) -> registers, two consecutive
// doubles per slot, BM / 32 slots per thread
template <int TA, int BM>
__device__ __forceinline__ void load_a(const double* __restrict__ A, int64_t lda, int64_t M, int64_t K,
                                       int64_t m0, int64_t k0, double (&r)[BM / 16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < BM / 32; ++i) {
    const int sl = t + 256 * i;
    if (TA == 0) {  // A is M x K: rows of 16 contiguous k, 8 slots per row
      const int64_t m = m0 + sl / 8;
      const int64_t kq = k0 + (sl % 8) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) r[2 * i + e] = (m < M && kq + e < K) ? A[m * lda + kq + e] : 0.0;
    } else {  // A is K x M: rows of BM contiguous m
      const int64_t k = k0 + sl / (BM / 2);
      const int64_t mq = m0 + (sl % (BM / 2)) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) r[2 * i + e] = (k < K && mq + e < M) ? A[k * lda + mq + e] : 0.0;
    }
  }
}

template <int TA, int BM>
__device__ __forceinline__ void store_a(double (*As)[BM + 1], const double (&r)[BM / 16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < BM / 32; ++i) {
    const int sl = t + 256 * i;
    if (TA == 0) {
      const int m = sl / 8, kq = (sl % 8) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) As[kq + e][m] = r[2 * i + e];
    } else {
      const int k = sl / (BM / 2), mq = (sl % (BM / 2)) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) As[k][mq + e] = r[2 * i + e];
    }
  }
}

// op(B)[k][n] for the 16 x BN tile (k0.., n0..)
template <int TB, int BN>
__device__ __forceinline__ void load_b(const double* __restrict__ B, int64_t ldb, int64_t N, int64_t K,
                                       int64_t n0, int64_t k0, double (&r)[BN / 16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < BN / 32; ++i) {
    const int sl = t + 256 * i;
    if (TB == 0) {  // B is K x N: rows of BN contiguous n
      const int64_t k = k0 + sl / (BN / 2);
      const int64_t nq = n0 + (sl % (BN / 2)) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) r[2 * i + e] = (k < K && nq + e < N) ? B[k * ldb + nq + e] : 0.0;
    } else {  // B is N x K: rows of 16 contiguous k
      const int64_t n = n0 + sl / 8;
      const int64_t kq = k0 + (sl % 8) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) r[2 * i + e] = (n < N && kq + e < K) ? B[n * ldb + kq + e] : 0.0;
    }
  }
}

template <int TB, int BN>
__device__ __forceinline__ void store_b(double (*Bs)[BN + 1], const double (&r)[BN / 16]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < BN / 32; ++i) {
    const int sl = t + 256 * i;
    if (TB == 0) {
      const int k = sl / (BN / 2), nq = (sl % (BN / 2)) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) Bs[k][nq + e] = r[2 * i + e];
    } else {
      const int n = sl / 8, kq = (sl % 8) * 2;
#pragma unroll
      for (int e = 0; e < 2; ++e) Bs[kq + e][n] = r[2 * i + e];
    }
  }
}

// BM x BN output tile per workgroup (BM, BN in {64, 96}); four wavefronts in
// 2 x 2, each owning a (BM/2) x (BN/2) quadrant of 16 x 16 MFMA blocks.
// grid: (ceil(N/BN), ceil(M/BM), splits).  Slice z covers K range
// [z*kchunk, min(K, (z+1)*kchunk)).  With splits > 1 the raw product of each
// slice goes to part[z] (M x N, ld N); otherwise C = alpha*prod + beta*Cin
// (Cin, ldcin: the input C, which may be C itself).
template <int TA, int TB, int BM, int BN>
__global__ __launch_bounds__(256) void gemm_f64_kernel(int64_t M, int64_t N, int64_t K, int64_t kchunk,
                                                       double alpha, const double* __restrict__ A, int64_t lda,
                                                       const double* __restrict__ B, int64_t ldb, double beta,
                                                       const double* Cin, int64_t ldcin, double* C, int64_t ldc,
                                                       double* __restrict__ part) {
  constexpr int IM = BM / 32, JN = BN / 32;  // MFMA blocks per wavefront
  __shared__ double As[GBK][BM + 1];
  __shared__ double Bs[GBK][BN + 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // XCD-aware tile order: workgroup w runs on XCD w % 8, so tile t is given to
  // the workgroup that keeps runs of consecutive tiles on one XCD and its L2.
  // Consecutive tiles step along the dimension with fewer tiles: those tiles
  // share the panel of the larger operand (a row panel of op(A) when N has
  // fewer tiles, a column panel of op(B) when M has).
  const uint32_t gx = gridDim.x, gy = gridDim.y;
  const uint32_t tot = gx * gy * gridDim.z;
  const uint32_t wid = (blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x;
  const uint32_t tile = xcd_tile(wid, tot);
  uint32_t tx, ty;
  const uint32_t tz = tile / (gx * gy), t2 = tile % (gx * gy);
  if (gx <= gy) {
    tx = t2 % gx;
    ty = t2 / gx;
  } else {
    ty = t2 % gy;
    tx = t2 / gy;
  }
  const int64_t m0 = (int64_t)ty * BM;
  const int64_t n0 = (int64_t)tx * BN;
  const int64_t kbeg = (int64_t)tz * kchunk;
  const int64_t kend = min(K, kbeg + kchunk);

  d4 acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[BM / 16], rb[BN / 16];
  if (kbeg < kend) {
    load_a<TA, BM>(A, lda, M, kend, m0, kbeg, ra);
    load_b<TB, BN>(B, ldb, N, kend, n0, kbeg, rb);
  }
  for (int64_t k0 = kbeg; k0 < kend; k0 += GBK) {
    store_a<TA, BM>(As, ra);
    store_b<TB, BN>(Bs, rb);
    __syncthreads();
    if (k0 + GBK < kend) {  // prefetch the next K tile while this one is used
      load_a<TA, BM>(A, lda, M, kend, m0, k0 + GBK, ra);
      load_b<TB, BN>(B, ldb, N, kend, n0, k0 + GBK, rb);
    }
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[IM], b[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) a[i] = As[kr][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < JN; ++j) b[j] = Bs[kr][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }

  // D layout of v_mfma_f64_16x16x4f64: lane l, item r -> row l/16 + 4*r, col l%16
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * (BM / 2) + i * 16 + (lane >> 4) + 4 * r;
        const int64_t col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
        if (row < M && col < N) {
          const double v = acc[i][j][r];
          if (part) {
            part[((int64_t)tz * M + row) * N + col] = v;
          } else {
            C[row * ldc + col] = (beta == 0.0) ? alpha * v : alpha * v + beta * Cin[row * ldcin + col];
          }
        }
      }
}

// Tile edge for a dimension of size d: 96 when that pads at least 5% less
// than 64 (the power iteration's k = 266 pads to 288 instead of 320).
inline int gemm_edge(int64_t d) {
  const int64_t p64 = pamd_cdiv(d, 64) * 64, p96 = pamd_cdiv(d, 96) * 96;
  return (p96 * 20 <= p64 * 19) ? 96 : 64;
}

// C = alpha * sum_z part[z] + beta * Cin, summed in slice order.
__global__ void splitk_reduce_kernel(int64_t M, int64_t N, int splits, double alpha, const double* __restrict__ part,
                                     double beta, const double* Cin, int64_t ldcin, double* C, int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t row = idx / N, col = idx % N;
  double s = 0.0;
  for (int z = 0; z < splits; ++z) s += part[z * M * N + idx];
  C[row * ldc + col] = (beta == 0.0) ? alpha * s : alpha * s + beta * Cin[row * ldcin + col];
}

int gemm_splits(int64_t M, int64_t N, int64_t K) {
  const int64_t tiles = pamd_cdiv(M, gemm_edge(M)) * pamd_cdiv(N, gemm_edge(N));
  if (K < 4 * 256) {
    // a short K: split only an output of few tiles (the k x k products of
    // CholeskyQR at k = 266: 9 tiles, 36 -> 18 us with 64 K per slice); a
    // larger output fills the chip without, and its partials would cost more
    // than they save (8192 x 266 x 266: 43 us unsplit, 49 us in 3 slices)
    if (tiles >= 128) return 1;
    const int64_t s = imin64(pamd_cdiv(256, tiles), K / 64);
    return (int)imax64(1, imin64(s, 64));
  }
  // a large output with a long K (the power iteration's A S, A^T Y and Q^T A
  // at 8192 x 4096, k = 266): split K until there are about 1536 workgroups,
  // six per CU.  Kernel plus reduction of the partials, interleaved medians:
  // 1024 / 1536 / 2048 / 3072 workgroups took 436 / 398 / 439 / 415 us
  // (A^T Y), 432 / 400 / 412 / 415 (A S) and 535 / 427 / 484 / 445 (Q^T A)
  if (M * N >= (1LL << 20) && K >= 2048) {
    if (tiles >= 1536) return 1;
    const int64_t s = imin64(pamd_cdiv(1536, tiles), K / 256);
    return (int)imax64(1, imin64(s, 64));
  }
  // small outputs: about 1024 workgroups, at least 256 K per slice
  const int64_t target = 1024;
  int64_t s = pamd_cdiv(target, tiles);
  s = min(s, K / 256);  // keep >= 256 K per slice
  return (int)imax64(1, imin64(s, 64));
}

// ---------------------------------------------------------------------------
// GEMV
// ---------------------------------------------------------------------------

// y[i] = alpha * A[i,:] x + beta * y[i]; one wavefront per row.
__global__ __launch_bounds__(256) void gemv_n_kernel(int64_t m, int64_t n, double alpha,
                                                     const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ x, double beta,
                                                     double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= m) return;
  const double* a = A + row * lda;
  double s = 0.0;
  for (int64_t j = lane; j < n; j += 64) s += a[j] * x[j];
  s = wave_sum(s);
  if (lane == 0) y[row] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[row];
}

// Column block of 64 (lanes) x row slice (grid.y); 4 waves split the rows.
// With part != NULL, slice sums go to part[slice * n + col].
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, int64_t rows_per_slice,
                                                     double alpha, const double* __restrict__ A, int64_t lda,
                                                     const double* __restrict__ x, double beta,
                                                     double* __restrict__ y, double* __restrict__ part) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t col = (int64_t)blockIdx.x * 64 + lane;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_slice;
  const int64_t r1 = min(m, r0 + rows_per_slice);
  double s = 0.0;
  if (col < n)
    for (int64_t i = r0 + wave; i < r1; i += 4) s += A[i * lda + col] * x[i];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && col < n) {
    const double t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    if (part)
      part[(int64_t)blockIdx.y * n + col] = t;
    else
      y[col] = (beta == 0.0) ? alpha * t : alpha * t + beta * y[col];
  }
}

__global__ void gemv_t_reduce_kernel(int64_t n, int slices, double alpha, const double* __restrict__ part,
                                     double beta, double* __restrict__ y) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  double s = 0.0;
  for (int z = 0; z < slices; ++z) s += part[(int64_t)z * n + col];
  y[col] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[col];
}

int gemv_t_slices(int64_t m, int64_t n) {
  const int64_t cb = pamd_cdiv(n, 64);
  int64_t s = pamd_cdiv(1024, cb);
  s = imin64(s, pamd_cdiv(m, 256));
  return (int)imax64(1, imin64(s, 1024));
}

__global__ __launch_bounds__(1024) void nrm2_kernel(int64_t n, const double* __restrict__ x,
                                                    double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += x[i] * x[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}

// sum of squares: one partial per workgroup, then an ordered final sum
__global__ __launch_bounds__(256) void sumsq_partial_kernel(int64_t n, const double* __restrict__ x,
                                                            double* __restrict__ part) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * x[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void sumsq_final_kernel(int64_t nparts, const double* __restrict__ part,
                                                           double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}

__global__ __launch_bounds__(1024) void trsv_kernel(int trans, int n, const double* __restrict__ R, int64_t ldr,
                                                    double* __restrict__ x) {
  extern __shared__ double smem[];
  double* xs = smem;  // n
  for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = x[i];
  __syncthreads();
  trsv_upper_block(trans, n, R, ldr, xs, smem + n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xs[i];
}

// n beyond what LDS holds: the solution vector stays in global memory (one
// workgroup, so its stores are visible to its own waves after each barrier);
// only the 64 x 64 diagonal block and the reductions use LDS.
__global__ __launch_bounds__(1024) void trsv_global_kernel(int trans, int n, const double* __restrict__ R,
                                                           int64_t ldr, double* x) {
  extern __shared__ double smem[];
  trsv_upper_block(trans, n, R, ldr, x, smem);
}

}  // namespace

extern "C" int pamd_gemm_workspace(int64_t M, int64_t N, int64_t K, int64_t* bytes) {
  PAMD_ARG_CHECK(bytes != nullptr, 4);
  const int s = gemm_splits(M, N, K);
  *bytes = (s > 1) ? (int64_t)s * M * N * (int64_t)sizeof(double) : 0;
  return PAMD_OK;
}

extern "C" int pamd_gemm_f64(int transA, int transB, int64_t M, int64_t N, int64_t K, double alpha,
                             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                             int64_t ldc, void* workspace, int64_t workspace_bytes, void* stream) {
  return pamd_gemm_cin_f64(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc, workspace,
                           workspace_bytes, stream);
}

// C = alpha op(A) op(B) + beta Cin with the input C read from Cin (which may
// be C): the library's own out-of-place updates (recursive TRSM) use it.
int pamd_gemm_cin_f64(int transA, int transB, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                      int64_t lda, const double* B, int64_t ldb, double beta, const double* Cin, int64_t ldcin,
                      double* C, int64_t ldc, void* workspace, int64_t workspace_bytes, void* stream) {
  PAMD_ARG_CHECK(transA == 0 || transA == 1, 1);
  PAMD_ARG_CHECK(transB == 0 || transB == 1, 2);
  PAMD_ARG_CHECK(M >= 0, 3);
  PAMD_ARG_CHECK(N >= 0, 4);
  PAMD_ARG_CHECK(K >= 0, 5);
  PAMD_ARG_CHECK(A != nullptr || M * K == 0, 7);
  PAMD_ARG_CHECK(lda >= (transA ? M : K) && lda >= 1, 8);
  PAMD_ARG_CHECK(B != nullptr || K * N == 0, 9);
  PAMD_ARG_CHECK(ldb >= (transB ? K : N) && ldb >= 1, 10);
  PAMD_ARG_CHECK(C != nullptr || M * N == 0, 12);
  PAMD_ARG_CHECK(ldc >= N && ldc >= 1, 13);
  PAMD_ARG_CHECK(Cin != nullptr || M * N == 0 || beta == 0.0, 12);
  PAMD_ARG_CHECK(ldcin >= N && ldcin >= 1, 13);
  if (M == 0 || N == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  int splits = gemm_splits(M, N, K);
  if (splits > 1 && (workspace == nullptr || workspace_bytes < (int64_t)splits * M * N * (int64_t)sizeof(double)))
    splits = 1;
  const int64_t kchunk = splits > 1 ? pamd_cdiv(pamd_cdiv(K, splits), GBK) * GBK : imax64(K, 1);
  if (splits > 1) splits = (int)pamd_cdiv(K, kchunk);
  double* part = splits > 1 ? (double*)workspace : nullptr;
  const int bm = gemm_edge(M), bn = gemm_edge(N);
  dim3 grid((unsigned)pamd_cdiv(N, bn), (unsigned)pamd_cdiv(M, bm), (unsigned)splits);
#define PAMD_GEMM(TA, TB, BM, BN) \
  gemm_f64_kernel<TA, TB, BM, BN><<<grid, 256, 0, st>>>(M, N, K, kchunk, alpha, A, lda, B, ldb, beta, Cin, ldcin, \
                                                        C, ldc, part)
#define PAMD_GEMM_T(BM, BN)                                  \
  do {                                                       \
    if (transA == 0 && transB == 0) PAMD_GEMM(0, 0, BM, BN); \
    else if (transA == 0) PAMD_GEMM(0, 1, BM, BN);           \
    else if (transB == 0) PAMD_GEMM(1, 0, BM, BN);           \
    else PAMD_GEMM(1, 1, BM, BN);                            \
  } while (0)
  if (bm == 64 && bn == 64) PAMD_GEMM_T(64, 64);
  else if (bm == 64) PAMD_GEMM_T(64, 96);
  else if (bn == 64) PAMD_GEMM_T(96, 64);
  else PAMD_GEMM_T(96, 96);
#undef PAMD_GEMM_T
#undef PAMD_GEMM
  PAMD_LAUNCH_CHECK();
  if (splits > 1) {
    const int64_t tot = M * N;
    splitk_reduce_kernel<<<(unsigned)pamd_cdiv(tot, 256), 256, 0, st>>>(M, N, splits, alpha, part, beta, Cin, ldcin,
                                                                        C, ldc);
    PAMD_LAUNCH_CHECK();
  }
  return PAMD_OK;
}

extern "C" int pamd_gemv_workspace(int trans, int64_t m, int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(bytes != nullptr, 4);
  *bytes = 0;
  if (trans == 1) {
    const int s = gemv_t_slices(m, n);
    if (s > 1) *bytes = (int64_t)s * n * (int64_t)sizeof(double);
  }
  return PAMD_OK;
}

extern "C" int pamd_gemv_f64(int trans, int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                             const double* x, double beta, double* y, void* workspace, int64_t workspace_bytes,
                             void* stream) {
  PAMD_ARG_CHECK(trans == 0 || trans == 1, 1);
  PAMD_ARG_CHECK(m >= 0, 2);
  PAMD_ARG_CHECK(n >= 0, 3);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 5);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 6);
  hipStream_t st = (hipStream_t)stream;
  if (trans == 0) {
    if (m == 0) return PAMD_OK;
    gemv_n_kernel<<<(unsigned)pamd_cdiv(m, 4), 256, 0, st>>>(m, n, alpha, A, lda, x, beta, y);
    PAMD_LAUNCH_CHECK();
    return PAMD_OK;
  }
  if (n == 0) return PAMD_OK;
  int slices = gemv_t_slices(m, n);
  if (slices > 1 && (workspace == nullptr || workspace_bytes < (int64_t)slices * n * (int64_t)sizeof(double)))
    slices = 1;
  const int64_t rps = imax64(1, pamd_cdiv(m, slices));
  slices = (int)imax64(1, pamd_cdiv(m, rps));
  double* part = slices > 1 ? (double*)workspace : nullptr;
  dim3 grid((unsigned)pamd_cdiv(n, 64), (unsigned)slices);
  gemv_t_kernel<<<grid, 256, 0, st>>>(m, n, rps, alpha, A, lda, x, beta, y, part);
  PAMD_LAUNCH_CHECK();
  if (part) {
    gemv_t_reduce_kernel<<<(unsigned)pamd_cdiv(n, 256), 256, 0, st>>>(n, slices, alpha, part, beta, y);
    PAMD_LAUNCH_CHECK();
  }
  return PAMD_OK;
}

extern "C" int pamd_nrm2_f64(int64_t n, const double* x, double* result, void* stream) {
  PAMD_ARG_CHECK(n >= 0, 1);
  PAMD_ARG_CHECK(x != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(result != nullptr, 3);
  nrm2_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(n, x, result);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

constexpr int64_t NRM2_MAX_PARTS = 1024;

extern "C" int pamd_nrm2_workspace(int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(n >= 0, 1);
  PAMD_ARG_CHECK(bytes != nullptr, 2);
  *bytes = NRM2_MAX_PARTS * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_nrm2_large_f64(int64_t n, const double* x, double* result, void* workspace, void* stream) {
  PAMD_ARG_CHECK(n >= 0, 1);
  PAMD_ARG_CHECK(x != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(result != nullptr, 3);
  PAMD_ARG_CHECK(workspace != nullptr, 4);
  hipStream_t st = (hipStream_t)stream;
  const int64_t parts = imax64(1, imin64(NRM2_MAX_PARTS, pamd_cdiv(n, 256 * 16)));
  sumsq_partial_kernel<<<(unsigned)parts, 256, 0, st>>>(n, x, (double*)workspace);
  PAMD_LAUNCH_CHECK();
  sumsq_final_kernel<<<1, 1024, 0, st>>>(parts, (const double*)workspace, result);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_trsv_upper_f64(int trans, int64_t n, const double* R, int64_t ldr, double* x, void* stream) {
  PAMD_ARG_CHECK(trans == 0 || trans == 1, 1);
  PAMD_ARG_CHECK(n >= 0 && n < (int64_t)1 << 30, 2);
  PAMD_ARG_CHECK(R != nullptr || n == 0, 3);
  PAMD_ARG_CHECK(ldr >= n && ldr >= 1, 4);
  PAMD_ARG_CHECK(x != nullptr || n == 0, 5);
  if (n == 0) return PAMD_OK;
  if (n <= PAMD_TRSV_MAX_N) {
    const size_t shmem = ((size_t)n + TRSV_SCRATCH_DOUBLES) * sizeof(double);
    trsv_kernel<<<1, 1024, shmem, (hipStream_t)stream>>>(trans, (int)n, R, ldr, x);
  } else {
    trsv_global_kernel<<<1, 1024, TRSV_SCRATCH_DOUBLES * sizeof(double), (hipStream_t)stream>>>(trans, (int)n, R,
                                                                                               ldr, x);
  }
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
