// Blocked Householder QR (geqrf) and Q formation (orgqr), FP64.
//
// Right-looking, panel width QR_NB = 8.  Per panel:
//   1. qr_panel_kernel (one 1024-thread workgroup) factors the m_r x 8 panel
//      column by column.  Each column costs one fused pass -- sum of squares of
//      the sub-diagonal part and its dot products with the remaining panel
//      columns, reduced together -- and one update pass.  At the end it
//      writes V (explicit unit diagonal, zeros above) and the compact-WY
//      factor T (LAPACK dlarft, forward / columnwise) to the workspace.
//   2. trailing update C <- (I - V T^T V^T) C on the columns right of the
//      panel: W = V^T C (qr_vtc_kernel, row-chunked partials), W2 = T^T W
//      summed over chunks in order (qr_w2_kernel), C -= V W2
//      (qr_update_kernel).  Every reduction has a fixed order, so results are
//      run-to-run identical.
// orgqr applies the stored blocks in reverse to [I; 0] with W2 = T W.
#include "common.h"
#include "wave_multi.h"

namespace {

constexpr int QR_NB = 8;
constexpr int QR_CHUNK = 64;  // rows per partial in V^T C (16 per wavefront, all loads in flight)

struct QrLayout {
  int64_t npan;       // number of panels
  int64_t v_off;      // doubles: start of V blocks (panel p at v_off + vrow(p) * NB)
  int64_t t_off;      // doubles: T blocks (NB x NB each)
  int64_t tau_off;    // doubles: tau (npan * NB)
  int64_t wpart_off;  // doubles: V^T C partials (maxchunks * NB * ncols)
  int64_t w2_off;     // doubles: NB * ncols
  int64_t flag_off;   // doubles: one int flag (singular)
  int64_t total;      // doubles
};

// sum over panels q < p of (m - q*NB): row offset of panel p's V block
__host__ __device__ inline int64_t qr_vrow(int64_t m, int64_t p) { return p * m - QR_NB * (p * (p - 1) / 2); }

QrLayout qr_layout(int64_t m, int64_t n, int64_t ncols) {
  QrLayout L;
  L.npan = pamd_cdiv(n, QR_NB);
  L.v_off = 0;
  L.t_off = L.v_off + qr_vrow(m, L.npan) * QR_NB;
  L.tau_off = L.t_off + L.npan * QR_NB * QR_NB;
  L.wpart_off = L.tau_off + L.npan * QR_NB;
  const int64_t maxchunks = pamd_cdiv(m, QR_CHUNK);
  L.wpart_off = (L.wpart_off + 31) / 32 * 32;
  L.w2_off = L.wpart_off + maxchunks * QR_NB * imax64(ncols, 1);
  L.flag_off = L.w2_off + QR_NB * imax64(ncols, 1);
  L.total = L.flag_off + 4;
  return L;
}


constexpr int QR_NWAVES = 16;  // panel kernels run 1024 threads
constexpr int QR_NPAIR = QR_NB * (QR_NB - 1) / 2;

// The compact-WY factor T (LAPACK dlarft, forward / columnwise) from the
// Gram entries Gp of the panel's V columns (pairs a < b in row-major order):
// T[i, j] = -tau_j sum_{i <= l < j} T[i, l] G[l, j].  Row i of T needs only
// row i, so thread i forms it.  Ends with T and tau written.
__device__ __forceinline__ void qr_panel_T(const double* Gp, const double* taus, int nb, double* __restrict__ T,
                                           double* __restrict__ tau_out, int tid) {
  if (tid < QR_NB) {
    const int i = tid;
    double Trow[QR_NB];
#pragma unroll
    for (int b = 0; b < QR_NB; ++b) Trow[b] = 0.0;
    if (i < nb) Trow[i] = taus[i];
#pragma unroll
    for (int jj = 1; jj < QR_NB; ++jj) {
      if (jj < nb && jj > i) {
        double sum = 0.0;
#pragma unroll
        for (int l = 0; l < jj; ++l)
          if (l >= i) sum += Trow[l] * Gp[l * (2 * QR_NB - l - 1) / 2 + (jj - l - 1)];
        Trow[jj] = -taus[jj] * sum;
      }
    }
    tau_out[i] = (i < nb) ? taus[i] : 0.0;
#pragma unroll
    for (int b = 0; b < QR_NB; ++b) T[i * QR_NB + b] = Trow[b];
  }
}

// After red[w][q] holds wave w's partial of the Gram entry q (1024 threads):
// G by QR_NPAIR threads (waves summed in order), then T.
__device__ __forceinline__ void qr_panel_tail(double (*red)[QR_NPAIR + 1], double* Gp, const double* taus, int nb,
                                              double* __restrict__ T, double* __restrict__ tau_out, int tid) {
  if (tid < QR_NPAIR) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < QR_NWAVES; ++w) sum += red[w][tid];
    Gp[tid] = sum;
  }
  __syncthreads();
  qr_panel_T(Gp, taus, nb, T, tau_out, tid);
}

// Factor columns [c0, c0 + nb) of rows [c0, m) (nb <= QR_NB).
__global__ __launch_bounds__(1024) void qr_panel_kernel(int64_t m, int64_t c0, int nb, double* __restrict__ A,
                                                        int64_t lda, double* __restrict__ V,
                                                        double* __restrict__ T, double* __restrict__ tau_out,
                                                        int* __restrict__ singular) {
  __shared__ double red[QR_NWAVES][QR_NPAIR + 1];
  __shared__ double tots[QR_NB];
  __shared__ double taus[QR_NB];
  __shared__ double Gp[QR_NPAIR];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;

  for (int j = 0; j < nb; ++j) {
    const int64_t pj = c0 + j;
    // pass 1: part[0] = ||x||^2, part[c] = x . A[:, c0+c] for c > j (rows > pj)
    double part[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) part[c] = 0.0;
    for (int64_t i = pj + 1 + tid; i < m; i += blockDim.x) {
      const double* row = A + i * lda + c0;
      const double x = row[j];
      part[0] += x * x;
#pragma unroll
      for (int c = 1; c < QR_NB; ++c)
        if (c > j && c < nb) part[c] += x * row[c];
    }
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) {
      const double s = wave_sum(part[c]);
      if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (tid < QR_NB) {  // one lane per value sums the waves in order
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < QR_NWAVES; ++w) s += red[w][tid];
      tots[tid] = s;
    }
    __syncthreads();
    double tot[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) tot[c] = tots[c];
    double* prow = A + pj * lda + c0;
    const double alpha = prow[j];
    double beta, tau, scale;
    if (tot[0] == 0.0) {  // dlarfg: nothing below the diagonal -> H = I
      beta = alpha;
      tau = 0.0;
      scale = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + tot[0]), alpha);
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    double w[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) w[c] = (c > j && c < nb) ? tau * (prow[c] + scale * tot[c]) : 0.0;
    __syncthreads();  // every thread has read prow before it is rewritten
    // pass 2: v = x * scale; A[:, c] -= tau * w_c * v
    for (int64_t i = pj + 1 + tid; i < m; i += blockDim.x) {
      double* row = A + i * lda + c0;
      const double v = row[j] * scale;
      row[j] = v;
#pragma unroll
      for (int c = 1; c < QR_NB; ++c)
        if (c > j && c < nb) row[c] -= w[c] * v;
    }
    if (tid == 0) {
      prow[j] = beta;
#pragma unroll
      for (int c = 1; c < QR_NB; ++c)
        if (c > j && c < nb) prow[c] -= w[c];
      taus[j] = tau;
      if (beta == 0.0) singular[0] = 1;
    }
    __syncthreads();
  }

  // V block with explicit unit diagonal, zeros above and zero padding columns
  const int64_t mr = m - c0;
  for (int64_t e = tid; e < mr * QR_NB; e += blockDim.x) {
    const int64_t r = e / QR_NB;
    const int c = (int)(e % QR_NB);
    double v;
    if (c >= nb || c > r)
      v = 0.0;
    else if (c == r)
      v = 1.0;
    else
      v = A[(c0 + r) * lda + c0 + c];
    V[e] = v;
  }
  __syncthreads();
  // G = V^T V (upper part, pairs a < b), reduced in fixed order
  double g[QR_NPAIR];
#pragma unroll
  for (int q = 0; q < QR_NPAIR; ++q) g[q] = 0.0;
  for (int64_t r = tid; r < mr; r += blockDim.x) {
    double vr[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) vr[c] = V[r * QR_NB + c];
    int q = 0;
#pragma unroll
    for (int a = 0; a < QR_NB; ++a)
#pragma unroll
      for (int b = a + 1; b < QR_NB; ++b) g[q++] += vr[a] * vr[b];
  }
#pragma unroll
  for (int q = 0; q < QR_NPAIR; ++q) {
    const double s = wave_sum(g[q]);
    if (lane == 0) red[wave][q] = s;
  }
  __syncthreads();
  qr_panel_tail(red, Gp, taus, nb, T, tau_out, tid);
}


// Same factorisation with the panel held in registers: 512 threads, thread t
// owns the panel rows t + 512 q (q < RPT).  Each column costs one fused
// reduction (sum of squares + dot products, 8 values) and one barrier, after
// which every wavefront adds the waves' partials itself; the shared values of
// a column are double-buffered by its parity, so no barrier is needed between
// columns.  Global memory is touched only by the staged loads at the start and
// the staged stores at the end.
constexpr int QR_REG_THREADS = 512;

template <int RPT>
__global__ __launch_bounds__(QR_REG_THREADS) void qr_panel_reg_kernel(int64_t m, int64_t c0, int nb,
                                                                      double* __restrict__ A, int64_t lda,
                                                                      double* __restrict__ V,
                                                                      double* __restrict__ T,
                                                                      double* __restrict__ tau_out,
                                                                      int* __restrict__ singular) {
  constexpr int NT = QR_REG_THREADS, NW = NT / 64;
  constexpr int NG = 32;  // Gram pairs padded to a power of two
  static_assert(QR_NPAIR <= NG, "pairs fit the padded reduction");
  __shared__ double red[2][NW][QR_NB];
  __shared__ double prow[2][QR_NB];
  __shared__ double taus[QR_NB];
  __shared__ double gred[NW][NG];
  __shared__ double Gp[QR_NPAIR];
  // row slices of NT rows pass between global memory and the thread-per-row
  // registers through LDS: globally, eight consecutive lanes move one row's
  // QR_NB values, so a wavefront touches 8 rows per access instead of 64
  __shared__ double stg[2][NT][QR_NB + 1];
  static_assert(QR_NB == 8, "eight lanes per row in the staged copies");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t mr = m - c0;
  const int sc = tid & 7, sr = tid >> 3;  // staged copies: column, first row (rows sr + 64 i)
  double a[RPT][QR_NB];
  {
    double ld[RPT][QR_NB];  // all global loads in flight at once
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
      for (int i = 0; i < QR_NB; ++i) {
        const int64_t r = (int64_t)NT * q + sr + (NT / QR_NB) * i;
        ld[q][i] = (r < mr && sc < nb) ? A[(c0 + r) * lda + c0 + sc] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
#pragma unroll
      for (int i = 0; i < QR_NB; ++i) stg[0][sr + (NT / QR_NB) * i][sc] = ld[q][i];
      __syncthreads();
#pragma unroll
      for (int c = 0; c < QR_NB; ++c) a[q][c] = stg[0][tid][c];
      __syncthreads();
    }
  }
  const int idx8 = wave_sum_multi_index<QR_NB>(lane);
#pragma unroll
  for (int j = 0; j < QR_NB; ++j) {  // unrolled: every register index is static
    if (j >= nb) break;
    const int buf = j & 1;
    if (tid == j) {
#pragma unroll
      for (int c = 0; c < QR_NB; ++c) prow[buf][c] = a[0][c];
    }
    double part[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) part[c] = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int64_t r = tid + NT * q;
      if (r > j && r < mr) {
        const double x = a[q][j];
        part[0] += x * x;
#pragma unroll
        for (int c = 1; c < QR_NB; ++c)
          if (c > j) part[c] += x * a[q][c];
      }
    }
    const double wsum = wave_sum_multi<QR_NB>(part, lane);
    if (lane < QR_NB) red[buf][wave][idx8] = wsum;
    __syncthreads();  // the one barrier of the column: red and prow are double-buffered by parity
    // every wavefront forms the totals itself: lane l sums value l % QR_NB over
    // the waves in wave order, then the eight totals are broadcast in the wave
    double mine = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) mine += red[buf][w][lane & (QR_NB - 1)];
    double tot[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) tot[c] = wave_bcast(mine, c);
    const double alpha = prow[buf][j];
    double beta, tau, scale;
    if (tot[0] == 0.0) {
      beta = alpha;
      tau = 0.0;
      scale = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + tot[0]), alpha);
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    double w[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) w[c] = (c > j && c < nb) ? tau * (prow[buf][c] + scale * tot[c]) : 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int64_t r = tid + NT * q;
      if (r > j && r < mr) {
        const double v = a[q][j] * scale;
        a[q][j] = v;
#pragma unroll
        for (int c = 1; c < QR_NB; ++c)
          if (c > j) a[q][c] -= w[c] * v;
      }
    }
    if (tid == j) {
      a[0][j] = beta;
#pragma unroll
      for (int c = 1; c < QR_NB; ++c)
        if (c > j) a[0][c] -= w[c];
      taus[j] = tau;
      if (beta == 0.0) singular[0] = 1;
    }
  }
  // write back: R above the diagonal, v below; V with explicit unit diagonal;
  // Gram partials of V's columns (pairs x < y) on the way
  double g[NG];
#pragma unroll
  for (int p = 0; p < NG; ++p) g[p] = 0.0;
  // indices of the staged store, opaque here: addresses derived from them are
  // not precomputed before the column loop, where they would spill
  const int wr = pamd_opaque(sr), wc = pamd_opaque(sc);
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t r = tid + NT * q;
    double vr[QR_NB];
#pragma unroll
    for (int c = 0; c < QR_NB; ++c) {
      vr[c] = (c >= nb || c > r || r >= mr) ? 0.0 : ((c == r) ? 1.0 : a[q][c]);
      stg[0][tid][c] = a[q][c];
      stg[1][tid][c] = vr[c];
    }
    int p = 0;
#pragma unroll
    for (int x = 0; x < QR_NB; ++x)
#pragma unroll
      for (int y = x + 1; y < QR_NB; ++y) g[p++] += vr[x] * vr[y];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < QR_NB; ++i) {
      const int rl = wr + (NT / QR_NB) * i;
      const int64_t rs = (int64_t)NT * q + rl;
      if (rs < mr) {
        if (wc < nb) A[(c0 + rs) * lda + c0 + wc] = stg[0][rl][wc];
        V[rs * QR_NB + wc] = stg[1][rl][wc];
      }
    }
    __syncthreads();
  }
  const double gsum = wave_sum_multi<NG>(g, lane);
  if (lane < NG) gred[wave][wave_sum_multi_index<NG>(lane)] = gsum;
  __syncthreads();  // also: taus of every column are written
  if (tid < QR_NPAIR) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += gred[w][tid];
    Gp[tid] = sum;
  }
  __syncthreads();
  qr_panel_T(Gp, taus, nb, T, tau_out, tid);
}

// Wpart[chunk][r][c] = sum_{i in chunk} V[i][r] * C[i][c]; C = A[rows0.., cols0..].
// Wavefront w sums the chunk's rows w, w + 4, ... in order (its 16 loads of C
// issued together); the four wavefronts are added in order.
__global__ __launch_bounds__(256) void qr_vtc_kernel(int64_t mr, int64_t nc, const double* __restrict__ V,
                                                     const double* __restrict__ C, int64_t ldc,
                                                     double* __restrict__ Wpart) {
  constexpr int RW = QR_CHUNK / 4;  // rows per wavefront
  __shared__ double red[4][QR_NB][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t col = (int64_t)blockIdx.x * 64 + lane;
  const int64_t i0 = (int64_t)blockIdx.y * QR_CHUNK;
  double acc[QR_NB];
#pragma unroll
  for (int r = 0; r < QR_NB; ++r) acc[r] = 0.0;
  if (col < nc) {
    double c[RW];
#pragma unroll
    for (int u = 0; u < RW; ++u) {
      const int64_t i = i0 + wave + 4 * u;
      c[u] = (i < mr) ? C[i * ldc + col] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RW; ++u) {
      const int64_t i = i0 + wave + 4 * u;
      if (i < mr) {
        const double* v = V + i * QR_NB;
#pragma unroll
        for (int r = 0; r < QR_NB; ++r) acc[r] += v[r] * c[u];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < QR_NB; ++r) red[wave][r][lane] = acc[r];
  __syncthreads();
  if (wave == 0 && col < nc) {
    double* out = Wpart + (int64_t)blockIdx.y * QR_NB * nc;
#pragma unroll
    for (int r = 0; r < QR_NB; ++r)
      out[r * nc + col] = ((red[0][r][lane] + red[1][r][lane]) + red[2][r][lane]) + red[3][r][lane];
  }
}

// W2[r][c] = sum_r' op(T)[r][r'] * (sum_chunks Wpart[chunk][r'][c]); op = T^T (apply Q^T) or T (apply Q).
// Thread (r, c) of a 32-column block sums its chunk partials in chunk order
// (eight loads in flight), then the block applies op(T) through LDS.
__global__ __launch_bounds__(256) void qr_w2_kernel(int64_t nc, int nchunks, const double* __restrict__ Wpart,
                                                    const double* __restrict__ T, int transT,
                                                    double* __restrict__ W2) {
  __shared__ double sw[QR_NB][33];
  const int cl = threadIdx.x & 31, r = threadIdx.x >> 5;
  const int64_t col = (int64_t)blockIdx.x * 32 + cl;
  double w = 0.0;
  if (col < nc) {
    const double* p = Wpart + (int64_t)r * nc + col;
    const int64_t cs = (int64_t)QR_NB * nc;
    int ch = 0;
    for (; ch + 8 <= nchunks; ch += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(ch + u) * cs];
#pragma unroll
      for (int u = 0; u < 8; ++u) w += x[u];
    }
    for (; ch < nchunks; ++ch) w += p[(int64_t)ch * cs];
  }
  sw[r][cl] = w;
  __syncthreads();
  if (col < nc) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < QR_NB; ++q) s += (transT ? T[q * QR_NB + r] : T[r * QR_NB + q]) * sw[q][cl];
    W2[r * nc + col] = s;
  }
}

// C[i][c] -= sum_r V[i][r] * W2[r][c]
__global__ __launch_bounds__(256) void qr_update_kernel(int64_t mr, int64_t nc, const double* __restrict__ V,
                                                        const double* __restrict__ W2, double* __restrict__ C,
                                                        int64_t ldc) {
  const int64_t col = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t row0 = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
  if (col >= nc) return;
  double w[QR_NB];
#pragma unroll
  for (int r = 0; r < QR_NB; ++r) w[r] = W2[r * nc + col];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = row0 + q;
    if (i < mr) {
      const double* v = V + i * QR_NB;
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < QR_NB; ++r) s += v[r] * w[r];
      C[i * ldc + col] -= s;
    }
  }
}

__global__ void qr_eye_kernel(int64_t m, int64_t n, double* __restrict__ Q, int64_t ldq) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const int64_t i = idx / n, j = idx % n;
  Q[i * ldq + j] = (i == j) ? 1.0 : 0.0;
}

__global__ void qr_zero_flag_kernel(int* flag) { flag[0] = 0; }

int apply_block(int64_t mr, int64_t nc, const double* V, const double* T, int transT, double* C, int64_t ldc,
                double* Wpart, double* W2, hipStream_t st) {
  if (nc <= 0 || mr <= 0) return PAMD_OK;
  const int nchunks = (int)pamd_cdiv(mr, QR_CHUNK);
  dim3 g1((unsigned)pamd_cdiv(nc, 64), (unsigned)nchunks);
  qr_vtc_kernel<<<g1, 256, 0, st>>>(mr, nc, V, C, ldc, Wpart);
  PAMD_LAUNCH_CHECK();
  qr_w2_kernel<<<(unsigned)pamd_cdiv(nc, 32), 256, 0, st>>>(nc, nchunks, Wpart, T, transT, W2);
  PAMD_LAUNCH_CHECK();
  dim3 g3((unsigned)pamd_cdiv(nc, 64), (unsigned)pamd_cdiv(mr, 16));
  qr_update_kernel<<<g3, 256, 0, st>>>(mr, nc, V, W2, C, ldc);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

}  // namespace

extern "C" int pamd_geqrf_workspace(int64_t m, int64_t n, int64_t ncols, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= m, 2);
  PAMD_ARG_CHECK(ncols >= n, 3);
  PAMD_ARG_CHECK(bytes != nullptr, 4);
  *bytes = qr_layout(m, n, ncols).total * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_geqrf_f64(int64_t m, int64_t n, int64_t ncols, double* A, int64_t lda, void* workspace,
                              void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= m, 2);
  PAMD_ARG_CHECK(ncols >= n, 3);
  PAMD_ARG_CHECK(A != nullptr || m * ncols == 0, 4);
  PAMD_ARG_CHECK(lda >= ncols && lda >= 1, 5);
  PAMD_ARG_CHECK(workspace != nullptr, 6);
  if (n == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const QrLayout L = qr_layout(m, n, ncols);
  double* ws = (double*)workspace;
  int* flag = (int*)(ws + L.flag_off);
  qr_zero_flag_kernel<<<1, 1, 0, st>>>(flag);
  PAMD_LAUNCH_CHECK();
  for (int64_t p = 0; p < L.npan; ++p) {
    const int64_t c0 = p * QR_NB;
    const int nb = (int)imin64(QR_NB, n - c0);
    double* V = ws + L.v_off + qr_vrow(m, p) * QR_NB;
    double* T = ws + L.t_off + p * QR_NB * QR_NB;
    const int64_t mrows = m - c0;
    double* tau = ws + L.tau_off + p * QR_NB;
    constexpr int NT = QR_REG_THREADS;
    if (mrows <= NT)
      qr_panel_reg_kernel<1><<<1, NT, 0, st>>>(m, c0, nb, A, lda, V, T, tau, flag);
    else if (mrows <= 2 * NT)
      qr_panel_reg_kernel<2><<<1, NT, 0, st>>>(m, c0, nb, A, lda, V, T, tau, flag);
    else if (mrows <= 4 * NT)
      qr_panel_reg_kernel<4><<<1, NT, 0, st>>>(m, c0, nb, A, lda, V, T, tau, flag);
    else if (mrows <= 8 * NT)
      qr_panel_reg_kernel<8><<<1, NT, 0, st>>>(m, c0, nb, A, lda, V, T, tau, flag);
    else
      qr_panel_kernel<<<1, 1024, 0, st>>>(m, c0, nb, A, lda, V, T, tau, flag);
    PAMD_LAUNCH_CHECK();
    const int64_t mr = m - c0;
    const int64_t nc = ncols - (c0 + nb);
    int rc = apply_block(mr, nc, V, T, 1, A + c0 * lda + c0 + nb, lda, ws + L.wpart_off, ws + L.w2_off, st);
    if (rc != PAMD_OK) return rc;
  }
  return PAMD_OK;
}

extern "C" int pamd_orgqr_f64(int64_t m, int64_t n, int64_t ncols, const void* workspace, double* Q, int64_t ldq,
                              void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= m, 2);
  PAMD_ARG_CHECK(ncols >= n, 3);
  PAMD_ARG_CHECK(workspace != nullptr, 4);
  PAMD_ARG_CHECK(Q != nullptr || m * n == 0, 5);
  PAMD_ARG_CHECK(ldq >= n && ldq >= 1, 6);
  if (n == 0 || m == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const QrLayout L = qr_layout(m, n, ncols);
  double* ws = (double*)workspace;  // W partials are scratch; V and T are read only
  qr_eye_kernel<<<(unsigned)pamd_cdiv(m * n, 256), 256, 0, st>>>(m, n, Q, ldq);
  PAMD_LAUNCH_CHECK();
  for (int64_t p = L.npan - 1; p >= 0; --p) {
    const int64_t c0 = p * QR_NB;
    const double* V = ws + L.v_off + qr_vrow(m, p) * QR_NB;
    const double* T = ws + L.t_off + p * QR_NB * QR_NB;
    int rc = apply_block(m - c0, n - c0, V, T, 0, Q + c0 * ldq + c0, ldq, ws + L.wpart_off, ws + L.w2_off, st);
    if (rc != PAMD_OK) return rc;
  }
  return PAMD_OK;
}

extern "C" int pamd_qr_singular_flag(const void* workspace, int64_t m, int64_t n, int64_t ncols, int** flag_ptr) {
  PAMD_ARG_CHECK(workspace != nullptr, 1);
  PAMD_ARG_CHECK(flag_ptr != nullptr, 5);
  const QrLayout L = qr_layout(m, n, ncols);
  *flag_ptr = (int*)((double*)workspace + L.flag_off);
  return PAMD_OK;
}
