// Thin SVD of a small dense matrix by one-sided block Jacobi (Hestenes'
// method), and the two helpers of the symmetric eigendecomposition built on
// it (the eigenvalues of a symmetric C are the singular values of C with the
// signs of u_j . v_j).
//
// Replaces the small LAPACK decompositions of the low-rank drivers: la.svd(B)
// of SVD1 (drivers/svd.py:164), la.eigh(C) of EVD1 (drivers/evd.py:279), the
// SVD preconditioner (comps/preconditioning.py:74) and the SVD-based solve of
// sketch-and-solve (drivers/least_squares.py:186).
//
// The callers precondition as Drmac and Veselic do: the Jacobi runs on R^T of
// a Householder QR, whose columns are far closer to orthogonal than A's (9
// sweeps instead of 27 for a 266 x 266 matrix with Haar singular vectors).
//
// Method.  The columns of W (m x n, m >= n, initially A) are rotated pairwise
// until every pair is orthogonal to tol = m * eps relative to the pair's
// norms; the rotations accumulate in V.  Then A = W V^T, the column norms of W
// are the singular values and its normalised columns the left singular
// vectors.  Each rotation is computed from fresh inner products of the two
// columns, so the result has one-sided Jacobi's accuracy (backward stable,
// with small singular values determined to high relative accuracy for graded
// matrices), not the squared conditioning of a Gram-matrix method.
//
// Parallel order.  The n columns (padded with zero columns) form nb blocks
// of w columns, nb even.  A sweep is nb - 1 rounds of the round-robin
// ("circle") tournament over the blocks: in a round every block meets one
// other block, and each block pair is one workgroup.  The workgroup copies
// its 2w columns of W into LDS (column-major, so a column is contiguous and a
// wavefront reads it without bank conflicts), runs one full inner sweep over
// the 2w columns (2w - 1 steps of w disjoint column pairs, one wavefront per
// pair: three wave-reduced dot products, the rotation, the update in LDS),
// writes the columns back, and applies the accumulated 2w x 2w rotation J to
// its 2w columns of V (V <- V J, read and written once).
//
// Convergence without host synchronisation.  A workgroup that rotated any
// pair increments the sweep's counter.  Every round kernel of sweep s > 0
// returns at once when sweep s - 1 rotated nothing, so the host can enqueue
// max_sweeps sweeps without reading anything back.
//
// Layout in the workspace: W as np columns of m doubles (column j at
// W + j*m), V as np columns of np doubles, then the column norms, the sorted
// positions and the per-sweep counters.  w = 8: a round workgroup is eight
// wavefronts, and its 16 LDS columns take 64 KiB (m <= 512) or 128 KiB
// (m <= 1024).  Narrow blocks spread a sweep over more workgroups at once
// (17 per round at n = 266) with fewer wavefronts per SIMD; the rotation's
// reciprocal and square roots are hardware estimates refined by Newton steps.
#include <mutex>

#include "common.h"

namespace {

constexpr int JAC_BW = 8;                 // columns per block: 16 per pair, one wavefront per column pair
constexpr int JAC_THREADS = 64 * JAC_BW;  // one round workgroup
constexpr int JAC_MAX_SWEEPS = 64;
constexpr int64_t JAC_MAX_M = 1024;

__host__ __device__ inline int jac_block(int64_t) { return JAC_BW; }

__host__ __device__ inline int64_t jac_nblocks(int64_t n, int w) {
  int64_t nb = pamd_cdiv(n, w);
  nb += nb & 1;
  return nb < 2 ? 2 : nb;
}

struct JacLayout {
  int64_t W, V, sig, pos, cnt, total;  // offsets in doubles
};

__host__ __device__ inline JacLayout jac_layout(int64_t m, int64_t n) {
  const int w = jac_block(m);
  const int64_t np = jac_nblocks(n, w) * w;
  JacLayout L;
  L.W = 0;
  L.V = L.W + np * m;
  L.sig = L.V + np * np;
  L.pos = L.sig + np;                   // np int32 (np / 2 doubles, rounded up)
  L.cnt = L.pos + (np + 1) / 2 + 1;     // JAC_MAX_SWEEPS int32
  L.total = L.cnt + JAC_MAX_SWEEPS / 2 + 8;
  return L;
}

// the pair of participant k in round r of the circle tournament over L + 1
// players (L odd): player L sits still, the others rotate
__device__ __forceinline__ void circle_pair(int r, int k, int L, int& a, int& b) {
  if (k == 0) {
    a = L;
    b = r;
  } else {
    a = (r + k) % L;
    b = (r - k + L) % L;
  }
}

// W (column-major, ld m) <- A (row-major, lda) padded with zero columns to np;
// V <- I (np x np); sweep counters <- 0
__global__ void jac_init_kernel(int64_t m, int64_t n, int64_t np, const double* __restrict__ A, int64_t lda,
                                double* __restrict__ ws) {
  const JacLayout L = jac_layout(m, n);
  double* W = ws + L.W;
  double* V = ws + L.V;
  int* cnt = reinterpret_cast<int*>(ws + L.cnt);
  const int64_t j = blockIdx.y;  // column
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    W[j * m + i] = (j < n) ? A[i * lda + j] : 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x)
    V[j * np + i] = (i == j) ? 1.0 : 0.0;
  if (j == 0 && blockIdx.x == 0 && threadIdx.x < JAC_MAX_SWEEPS) cnt[threadIdx.x] = 0;
}

// 1 / x and 1 / sqrt(x) from the hardware estimates and two Newton steps
// each (the estimates' relative errors are squared twice: a few ulps)
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * fma(-x, y, 2.0);
  return y * fma(-x, y, 2.0);
}

__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// One round of sweep `sweep`: workgroup k orthogonalises block pair k.
// MMAX >= m rows of the pair's columns are staged in LDS.
template <int BW, int MMAX>
__global__ __launch_bounds__(64 * BW) void jac_round_kernel(int64_t m, int64_t n, int nb, int sweep, int round,
                                                            double tol, double* __restrict__ ws) {
  constexpr int P = 2 * BW;         // columns of the pair
  constexpr int E = MMAX / 64;      // rows per lane
  constexpr int NT = 64 * BW;       // threads
  __shared__ double Ws[P][MMAX];    // the pair's columns of W, column-major
  __shared__ double Js[P][P + 1];   // accumulated rotation, Js[col][row]
  const JacLayout L = jac_layout(m, n);
  const int* cnt = reinterpret_cast<const int*>(ws + L.cnt);
  if (sweep > 0 && cnt[sweep - 1] == 0) return;  // an earlier sweep rotated nothing: converged
  double* W = ws + L.W;
  double* V = ws + L.V;
  const int np = nb * BW;
  const int mi = (int)m;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  int bp, bq;
  circle_pair(round, blockIdx.x, nb - 1, bp, bq);
  // global column of pair column c
  auto gcol = [&](int c) -> int64_t { return (c < BW) ? (int64_t)bp * BW + c : (int64_t)bq * BW + (c - BW); };

  for (int idx = tid; idx < P * mi; idx += NT) {
    const int c = idx / mi, i = idx - c * mi;
    Ws[c][i] = W[gcol(c) * m + i];
  }
  for (int idx = tid; idx < P * P; idx += NT) Js[idx / P][idx % P] = (idx / P == idx % P) ? 1.0 : 0.0;
  __syncthreads();

  const int em = (mi + 63) / 64;
  const double tol2 = tol * tol;
  int rotated = 0;
  for (int t = 0; t < P - 1; ++t) {
    int a, b;
    circle_pair(t, wave, P - 1, a, b);
    double xa[E], xb[E];
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      xa[e] = 0.0;
      xb[e] = 0.0;
      const int i = lane + 64 * e;
      if (e < em && i < mi) {
        xa[e] = Ws[a][i];
        xb[e] = Ws[b][i];
      }
      al = fma(xa[e], xa[e], al);
      be = fma(xb[e], xb[e], be);
      ga = fma(xa[e], xb[e], ga);
    }
    al = wave_sum(al);
    be = wave_sum(be);
    ga = wave_sum(ga);
    // rotate unless |ga| <= tol sqrt(al be) (also skips zero columns)
    if (ga * ga > tol2 * (al * be)) {
      // a' = c a - s b, b' = s a + c b with a'.b' = 0: t = tan(theta) is the
      // smaller root of t^2 + 2 zeta t - 1 = 0, zeta = (be - al) / (2 ga)
      const double zeta = (be - al) * rcp_nr(2.0 * ga);
      const double az = fabs(zeta);
      double tt;
      if (az > 1e150) {
        tt = 0.5 * rcp_nr(zeta);
      } else {
        const double q = fma(zeta, zeta, 1.0);
        tt = copysign(rcp_nr(az + q * rsq_nr(q)), zeta);
      }
      const double c = rsq_nr(fma(tt, tt, 1.0));
      const double s = c * tt;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = lane + 64 * e;
        if (e < em && i < mi) {
          Ws[a][i] = c * xa[e] - s * xb[e];
          Ws[b][i] = s * xa[e] + c * xb[e];
        }
      }
      if (lane < P) {
        const double ja = Js[a][lane], jb = Js[b][lane];
        Js[a][lane] = c * ja - s * jb;
        Js[b][lane] = s * ja + c * jb;
      }
      rotated = 1;
    }
    __syncthreads();
  }
  const int any = __syncthreads_or(rotated);
  if (!any) return;  // the pair was already orthogonal: W and V unchanged

  for (int idx = tid; idx < P * mi; idx += NT) {
    const int c = idx / mi, i = idx - c * mi;
    W[gcol(c) * m + i] = Ws[c][i];
  }
  __syncthreads();
  // V(:, pair) <- V(:, pair) J: the pair's V columns (np <= MMAX rows) are
  // staged in the LDS the W columns used, then each thread forms one row's
  // 2w outputs, eight at a time
  for (int idx = tid; idx < P * np; idx += NT) {
    const int c = idx / np, i = idx - c * np;
    Ws[c][i] = V[gcol(c) * np + i];
  }
  __syncthreads();
  for (int i = tid; i < np; i += NT) {
#pragma unroll
    for (int j0 = 0; j0 < P; j0 += 8) {
      double acc[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) acc[jj] = 0.0;
#pragma unroll 4
      for (int c = 0; c < P; ++c) {
        const double v = Ws[c][i];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) acc[jj] = fma(v, Js[j0 + jj][c], acc[jj]);
      }
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) V[gcol(j0 + jj) * np + i] = acc[jj];
    }
  }
  if (tid == 0) atomicAdd(reinterpret_cast<int*>(ws + L.cnt) + sweep, 1);
}

// sigma_j = ||W(:, j)|| for the first n columns, one workgroup per column
__global__ __launch_bounds__(256) void jac_norms_kernel(int64_t m, int64_t n, double* __restrict__ ws) {
  __shared__ double red[4];
  const JacLayout L = jac_layout(m, n);
  const double* W = ws + L.W;
  const int64_t j = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
    const double v = W[j * m + i];
    acc = fma(v, v, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) ws[L.sig + j] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// position of column j in descending order of sigma (ties: lower index first),
// the sorted sigma, the convergence flag
__global__ __launch_bounds__(JAC_THREADS) void jac_sort_kernel(int64_t m, int64_t n, int max_sweeps,
                                                               double* __restrict__ s, double* __restrict__ ws,
                                                               int* __restrict__ info) {
  const JacLayout L = jac_layout(m, n);
  const double* sig = ws + L.sig;
  int* pos = reinterpret_cast<int*>(ws + L.pos);
  const int* cnt = reinterpret_cast<const int*>(ws + L.cnt);
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
    const double sj = sig[j];
    int p = 0;
    for (int64_t k = 0; k < n; ++k) {
      const double sk = sig[k];
      p += (sk > sj || (sk == sj && k < j)) ? 1 : 0;
    }
    pos[j] = p;
    s[p] = sj;
  }
  if (threadIdx.x == 0 && info) info[0] = (cnt[max_sweeps - 1] != 0) ? 1 : 0;
}

// U(:, pos_j) = W(:, j) / sigma_j (0 for sigma_j = 0); Vt(pos_j, :) = V(:, j)^T
__global__ void jac_output_kernel(int64_t m, int64_t n, int64_t np, double* __restrict__ U, int64_t ldu,
                                  double* __restrict__ Vt, int64_t ldvt, const double* __restrict__ ws) {
  const JacLayout L = jac_layout(m, n);
  const double* W = ws + L.W;
  const double* V = ws + L.V;
  const int* pos = reinterpret_cast<const int*>(ws + L.pos);
  const int64_t j = blockIdx.y;
  const int64_t p = pos[j];
  const double sj = ws[L.sig + j];
  const double inv = (sj > 0.0) ? 1.0 / sj : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    if (U) U[i * ldu + p] = W[j * m + i] * inv;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (Vt) Vt[p * ldvt + i] = V[j * np + i];
  }
}

// eigenvalues of the symmetric matrix from its singular triplets: lambda_j =
// sigma_j sign(u_j . v_j) - shift; info[0] = 2 when, for a singular value
// above 1e-10 sigma_max, |u_j . v_j| falls short of 1 by more than 1e-6 (a
// pair lambda, -lambda of nearly equal magnitude mixed the vectors; below
// that size the sign of a null-space value is noise, as it is for eigh)
__global__ __launch_bounds__(256) void jac_eig_sign_kernel(int64_t n, const double* __restrict__ U, int64_t ldu,
                                                           const double* __restrict__ Vt, int64_t ldvt,
                                                           const double* __restrict__ s, double shift,
                                                           double* __restrict__ lamb, int* __restrict__ info) {
  __shared__ double red[4];
  const int64_t j = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc = fma(U[i * ldu + j], Vt[j * ldvt + i], acc);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double d = red[0] + red[1] + red[2] + red[3];
    const double sj = s[j];
    lamb[j] = (d < 0.0 ? -sj : sj) - shift;
    if (sj > 1e-10 * s[0] && fabs(fabs(d) - 1.0) > 1e-6 && info) atomicMax(info, 2);
  }
}

int jac_launch_sweeps(int64_t m, int64_t n, int max_sweeps, double* ws, hipStream_t st) {
  const int w = jac_block(m);
  const int nb = (int)jac_nblocks(n, w);
  const double tol = (double)imax64(m, 16) * 2.220446049250313e-16;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    for (int r = 0; r < nb - 1; ++r) {
      if (m <= 512)
        jac_round_kernel<JAC_BW, 512><<<nb / 2, JAC_THREADS, 0, st>>>(m, n, nb, sw, r, tol, ws);
      else
        jac_round_kernel<JAC_BW, 1024><<<nb / 2, JAC_THREADS, 0, st>>>(m, n, nb, sw, r, tol, ws);
      PAMD_LAUNCH_CHECK();
    }
  }
  return PAMD_OK;
}

int jac_finish(int64_t m, int64_t n, int max_sweeps, double* U, int64_t ldu, double* s, double* Vt, int64_t ldvt,
               double* ws, int* info, hipStream_t st) {
  const int w = jac_block(m);
  const int64_t np = jac_nblocks(n, w) * w;
  jac_norms_kernel<<<(unsigned)n, 256, 0, st>>>(m, n, ws);
  PAMD_LAUNCH_CHECK();
  jac_sort_kernel<<<1, JAC_THREADS, 0, st>>>(m, n, max_sweeps, s, ws, info);
  PAMD_LAUNCH_CHECK();
  dim3 grid((unsigned)imin64(pamd_cdiv(imax64(m, n), 256), 4), (unsigned)n);
  jac_output_kernel<<<grid, 256, 0, st>>>(m, n, np, U, ldu, Vt, ldvt, ws);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

}  // namespace

extern "C" int pamd_svd_jacobi_workspace(int64_t m, int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 1 && m <= JAC_MAX_M, 1);
  PAMD_ARG_CHECK(n >= 1 && n <= m, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 3);
  *bytes = jac_layout(m, n).total * (int64_t)sizeof(double);
  return PAMD_OK;
}

namespace {

// the whole SVD on a caller-supplied or static workspace ws (arguments checked)
int jac_svd(int64_t m, int64_t n, const double* A, int64_t lda, double* U, int64_t ldu, double* s, double* Vt,
            int64_t ldvt, int max_sweeps, double* ws, int* info, hipStream_t st) {
  const int w = jac_block(m);
  const int64_t np = jac_nblocks(n, w) * w;
  dim3 grid((unsigned)imin64(pamd_cdiv(imax64(m, np), 256), 4), (unsigned)np);
  jac_init_kernel<<<grid, 256, 0, st>>>(m, n, np, A, lda, ws);
  PAMD_LAUNCH_CHECK();
  const int rc = jac_launch_sweeps(m, n, max_sweeps, ws, st);
  if (rc != PAMD_OK) return rc;
  return jac_finish(m, n, max_sweeps, U, ldu, s, Vt, ldvt, ws, info, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// The allocation-free variant: the same kernels on a file-scope device array
// of JAC_SCRATCH_DOUBLES doubles (16.8 MB, the workspace of the largest SVD
// this file does, m = n = 1024).  Each GPU holds its own copy of the array
// until the process ends; its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and used for every launch --
// the array's name is only a host-side stand-in.  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race.  Nothing in it
// needs clearing between calls: the init kernel writes W, V and the sweep
// counters, and the norms and positions are written before they are read.
// ---------------------------------------------------------------------------
constexpr int64_t JAC_SCRATCH_DOUBLES = 2 * JAC_MAX_M * JAC_MAX_M + 4 * JAC_MAX_M + 256;
static __device__ __align__(256) double scratch[JAC_SCRATCH_DOUBLES];

namespace {

constexpr int JAC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[JAC_MAX_DEVICES] = {};

// device address of the current device's copy of `scratch`
int jac_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= JAC_MAX_DEVICES) return PAMD_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return PAMD_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PAMD_OK;
}

}  // namespace

extern "C" int pamd_svd_jacobi_noalloc_f64(int64_t m, int64_t n, const double* A, int64_t lda, double* U,
                                           int64_t ldu, double* s, double* Vt, int64_t ldvt, int max_sweeps,
                                           int* info, void* stream) {
  PAMD_ARG_CHECK(m >= 1 && m <= JAC_MAX_M, 1);
  PAMD_ARG_CHECK(n >= 1 && n <= m, 2);
  PAMD_ARG_CHECK(A != nullptr, 3);
  PAMD_ARG_CHECK(lda >= n, 4);
  PAMD_ARG_CHECK(U == nullptr || ldu >= n, 6);
  PAMD_ARG_CHECK(s != nullptr, 7);
  PAMD_ARG_CHECK(Vt == nullptr || ldvt >= n, 9);
  PAMD_ARG_CHECK(max_sweeps >= 1 && max_sweeps <= JAC_MAX_SWEEPS, 10);
  if (jac_layout(m, n).total > JAC_SCRATCH_DOUBLES) return PAMD_ERR_UNSUPPORTED;
  double* ws = nullptr;
  const int rc = jac_scratch(&ws);
  if (rc != PAMD_OK) return rc;
  return jac_svd(m, n, A, lda, U, ldu, s, Vt, ldvt, max_sweeps, ws, info, (hipStream_t)stream);
}

extern "C" int pamd_svd_jacobi_f64(int64_t m, int64_t n, const double* A, int64_t lda, double* U, int64_t ldu,
                                   double* s, double* Vt, int64_t ldvt, int max_sweeps, void* workspace, int* info,
                                   void* stream) {
  PAMD_ARG_CHECK(m >= 1 && m <= JAC_MAX_M, 1);
  PAMD_ARG_CHECK(n >= 1 && n <= m, 2);
  PAMD_ARG_CHECK(A != nullptr, 3);
  PAMD_ARG_CHECK(lda >= n, 4);
  PAMD_ARG_CHECK(U == nullptr || ldu >= n, 6);
  PAMD_ARG_CHECK(s != nullptr, 7);
  PAMD_ARG_CHECK(Vt == nullptr || ldvt >= n, 9);
  PAMD_ARG_CHECK(max_sweeps >= 1 && max_sweeps <= JAC_MAX_SWEEPS, 10);
  PAMD_ARG_CHECK(workspace != nullptr, 11);
  return jac_svd(m, n, A, lda, U, ldu, s, Vt, ldvt, max_sweeps, (double*)workspace, info, (hipStream_t)stream);
}

namespace {

// out = the lower triangle of A mirrored, plus shift * I (n x n)
__global__ void sym_shift_kernel(int64_t n, const double* __restrict__ A, int64_t lda, double shift,
                                 double* __restrict__ out, int64_t ldo) {
  const int64_t i = blockIdx.y;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double a = (i >= j) ? A[i * lda + j] : A[j * lda + i];
    if (i == j) a += shift;
    out[i * ldo + j] = a;
  }
}

}  // namespace

extern "C" int pamd_sym_shift_f64(int64_t n, const double* A, int64_t lda, double shift, double* out,
                                  int64_t ldo, void* stream) {
  PAMD_ARG_CHECK(n >= 0 && n < (int64_t)1 << 31, 1);
  PAMD_ARG_CHECK(A != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 3);
  PAMD_ARG_CHECK(out != nullptr || n == 0, 5);
  PAMD_ARG_CHECK(ldo >= n && ldo >= 1, 6);
  if (n == 0) return PAMD_OK;
  dim3 grid((unsigned)imin64(pamd_cdiv(n, 256), 16), (unsigned)n);
  sym_shift_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(n, A, lda, shift, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_eig_from_svd_f64(int64_t n, const double* U, int64_t ldu, const double* Vt, int64_t ldvt,
                                     const double* s, double shift, double* lamb, int* info, void* stream) {
  PAMD_ARG_CHECK(n >= 0 && n < (int64_t)1 << 31, 1);
  PAMD_ARG_CHECK(U != nullptr || n == 0, 2);
  PAMD_ARG_CHECK(ldu >= n && ldu >= 1, 3);
  PAMD_ARG_CHECK(Vt != nullptr || n == 0, 4);
  PAMD_ARG_CHECK(ldvt >= n && ldvt >= 1, 5);
  PAMD_ARG_CHECK(s != nullptr || n == 0, 6);
  PAMD_ARG_CHECK(lamb != nullptr || n == 0, 8);
  PAMD_ARG_CHECK(lamb != s || n == 0, 8);
  if (n == 0) return PAMD_OK;
  jac_eig_sign_kernel<<<(unsigned)n, 256, 0, (hipStream_t)stream>>>(n, U, ldu, Vt, ldvt, s, shift, lamb, info);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
