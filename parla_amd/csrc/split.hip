// Split-precision GEMM for the QB power iteration: an FP64 operand is held as
// P BF16 pieces x = x_0 + x_1 + ... (each piece the BF16 rounding of what the
// previous ones leave), and C = L R^T is formed on v_mfma_f32_16x16x32_bf16
// from the piece products L_p R_q^T with p + q < P, accumulated in FP32.
//   P = 2: 16 significant bits per element, three products per tile;
//   P = 3: 24 bits (the FP32 significand), six products.
// The dropped products are below 2^-(8P) of |L| |R| elementwise.
//
// Layout: both operands K-contiguous ("TN"): L_p is M x K row-major, R_q is
// N x K row-major (the transpose of the K x N right factor), K zero-padded
// to a multiple of 64 and R's rows zero-padded to a multiple of 288.  The
// split kernels produce either orientation from a row-major FP64 matrix, so
// A (m x n) serves A S (L = A) and A^T Y (L = A^T) and each tall k-column
// factor is split straight into its transpose.
//
// GEMM: one 256-thread workgroup per 128 x TN output tile and K range; the
// four wavefronts split the tile 2 x 2 (P = 2, TN = 288) or 4 x 1 (P = 3,
// TN = 144).  K advances 32 at a time through double-buffered LDS with the
// next step's global loads in registers.  Skinny products (N ~ 300) have few
// tiles, so K is split until there are at least 256 workgroups; the splits'
// FP32 partials are summed in split order in FP64 by a second kernel, which
// also writes C in either orientation.  Tiles are assigned XCD-aware: the
// workgroups of one XCD share their K range and column tile, so the right
// operand's slice is fetched into that XCD's L2 once.
#include "common.h"

typedef __bf16 sg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sg_f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int SG_TM = 128;
constexpr int SG_BK = 32;
constexpr int SG_LDR = SG_BK + 8;  // LDS row stride in BF16: 80 bytes, conflict-free 16-byte reads
constexpr int SG_THREADS = 256;
constexpr int SG_KALIGN = 64;
constexpr int SG_NALIGN = 288;

__device__ __forceinline__ uint16_t sg_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ double sg_bf16_val(uint16_t h) { return (double)__uint_as_float((uint32_t)h << 16); }

// pieces of x: h[p] = BF16(x - h[0] - ... - h[p-1]) (each difference exact in FP64)
template <int P>
__device__ __forceinline__ void sg_split(double x, uint16_t (&h)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    h[p] = sg_bf16_rne((float)x);
    x -= sg_bf16_val(h[p]);
  }
}

__device__ __forceinline__ uint4 sg_pack(const uint16_t (&v)[8]) {
  uint4 w;
  w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
  w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
  w.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
  w.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
  return w;
}

// pieces[p][i][k] of X (rows x cols, row-major) for i < rows_out, k < kp;
// zero outside X.  Thread: 8 consecutive k of one row.
template <int P>
__global__ __launch_bounds__(256) void sg_split_rows_kernel(int64_t rows, int64_t cols, const double* __restrict__ X,
                                                            int64_t ldx, int64_t rows_out, int64_t kp,
                                                            uint16_t* __restrict__ pieces, int64_t ldk,
                                                            int64_t pstride) {
  const int64_t k0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 8;
  if (k0 >= kp) return;
  for (int64_t i = blockIdx.y; i < rows_out; i += gridDim.y) {
    uint16_t h[P][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t j = k0 + e;
      const double x = (i < rows && j < cols) ? X[i * ldx + j] : 0.0;
      uint16_t t[P];
      sg_split<P>(x, t);
#pragma unroll
      for (int p = 0; p < P; ++p) h[p][e] = t[p];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) *reinterpret_cast<uint4*>(pieces + p * pstride + i * ldk + k0) = sg_pack(h[p]);
  }
}

// pieces[p][j][i] of X^T for j < rows_out, i < kp (kp a multiple of 64):
// 64 x 64 tiles transposed through LDS, coalesced on both sides
template <int P>
__global__ __launch_bounds__(256) void sg_split_t_kernel(int64_t rows, int64_t cols, const double* __restrict__ X,
                                                         int64_t ldx, int64_t rows_out, uint16_t* __restrict__ pieces,
                                                         int64_t ldk, int64_t pstride) {
  __shared__ double T[64][65];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int ii = tid / 64 + 4 * q, jj = tid % 64;
    const int64_t i = i0 + ii, j = j0 + jj;
    T[ii][jj] = (i < rows && j < cols) ? X[i * ldx + j] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int ch = tid + 256 * h;
    const int jj = ch / 8, ii = (ch % 8) * 8;
    const int64_t j = j0 + jj;
    if (j >= rows_out) continue;
    uint16_t v[P][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      uint16_t t[P];
      sg_split<P>(T[ii + e][jj], t);
#pragma unroll
      for (int p = 0; p < P; ++p) v[p][e] = t[p];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) *reinterpret_cast<uint4*>(pieces + p * pstride + j * ldk + i0 + ii) = sg_pack(v[p]);
  }
}

// ws[ks][i][j] (FP32) = sum over p + q < P of L_p[i, K_ks] R_q[j, K_ks]^T
template <int P, int WM, int WN, int MB, int NB>
__global__ __launch_bounds__(SG_THREADS) void sg_gemm_kernel(int64_t M, int64_t kp, int64_t kc, int mtiles,
                                                             int ntiles, int nsplit,
                                                             const uint16_t* __restrict__ L, int64_t ldl,
                                                             int64_t lps, const uint16_t* __restrict__ R,
                                                             int64_t ldr, int64_t rps, float* __restrict__ ws,
                                                             int64_t npad) {
  constexpr int TM = WM * MB * 16, TN = WN * NB * 16;
  static_assert(TM == SG_TM && WM * WN == 4, "four wavefronts on a 128-row tile");
  constexpr int LCH = P * TM * 4, RCH = P * TN * 4;  // 16-byte chunks per K step
  constexpr int LPT = (LCH + SG_THREADS - 1) / SG_THREADS, RPT = (RCH + SG_THREADS - 1) / SG_THREADS;
  __shared__ __attribute__((aligned(16))) uint16_t Ls[2][P][TM][SG_LDR];
  __shared__ __attribute__((aligned(16))) uint16_t Rs[2][P][TN][SG_LDR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int lr = lane & 15, lk = lane >> 4;
  const uint32_t total = (uint32_t)mtiles * ntiles * nsplit;
  const uint32_t t = xcd_tile(blockIdx.x, total);
  const int ks = (int)(t / ((uint32_t)mtiles * ntiles));
  const uint32_t rem = t % ((uint32_t)mtiles * ntiles);
  const int nt = (int)(rem / mtiles), mt = (int)(rem % mtiles);
  const int64_t i0 = (int64_t)mt * TM, j0 = (int64_t)nt * TN;
  const int64_t kbeg = (int64_t)ks * kc;
  const int64_t kend = imin64(kp, kbeg + kc);
  const int nsteps = (int)((kend - kbeg) / SG_BK);

  uint4 rl[LPT], rr[RPT];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int h = 0; h < LPT; ++h) {
      const int ch = tid + SG_THREADS * h;
      if (ch < LCH) {
        const int p = ch / (TM * 4), r = (ch / 4) % TM, c = ch % 4;
        const int64_t gi = i0 + r;
        rl[h] = (gi < M) ? *reinterpret_cast<const uint4*>(L + p * lps + gi * ldl + k0 + 8 * c) : uint4{0, 0, 0, 0};
      }
    }
#pragma unroll
    for (int h = 0; h < RPT; ++h) {
      const int ch = tid + SG_THREADS * h;
      if (ch < RCH) {
        const int q = ch / (TN * 4), r = (ch / 4) % TN, c = ch % 4;
        rr[h] = *reinterpret_cast<const uint4*>(R + q * rps + (j0 + r) * ldr + k0 + 8 * c);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int h = 0; h < LPT; ++h) {
      const int ch = tid + SG_THREADS * h;
      if (ch < LCH) {
        const int p = ch / (TM * 4), r = (ch / 4) % TM, c = ch % 4;
        *reinterpret_cast<uint4*>(&Ls[buf][p][r][8 * c]) = rl[h];
      }
    }
#pragma unroll
    for (int h = 0; h < RPT; ++h) {
      const int ch = tid + SG_THREADS * h;
      if (ch < RCH) {
        const int q = ch / (TN * 4), r = (ch / 4) % TN, c = ch % 4;
        *reinterpret_cast<uint4*>(&Rs[buf][q][r][8 * c]) = rr[h];
      }
    }
  };

  sg_f4 acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = sg_f4{0.f, 0.f, 0.f, 0.f};

  if (nsteps > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  if (nsteps > 1) load(kbeg + SG_BK);
  for (int s = 0; s < nsteps; ++s) {
    const int buf = s & 1;
    sg_bf16x8 a[P][MB];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        a[p][mb] = *reinterpret_cast<const sg_bf16x8*>(&Ls[buf][p][wm * MB * 16 + mb * 16 + lr][8 * lk]);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      sg_bf16x8 b[P];
#pragma unroll
      for (int q = 0; q < P; ++q)
        b[q] = *reinterpret_cast<const sg_bf16x8*>(&Rs[buf][q][wn * NB * 16 + nb * 16 + lr][8 * lk]);
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < P; ++q) {
          if (p + q >= P) continue;
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p][mb], b[q], acc[mb][nb], 0, 0, 0);
        }
    }
    if (s + 1 < nsteps) {
      store(buf ^ 1);
      if (s + 2 < nsteps) load(kbeg + (int64_t)(s + 2) * SG_BK);
    }
    __syncthreads();
  }
  // D layout of v_mfma_f32_16x16x32_bf16: lane l, item r -> row 4 (l / 16) + r, column l % 16
  float* w = ws + (int64_t)ks * M * npad;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wm * MB * 16 + mb * 16 + 4 * lk + r;
        const int64_t j = j0 + wn * NB * 16 + nb * 16 + lr;
        if (i < M) w[i * npad + j] = acc[mb][nb][r];
      }
}

// C[i * lci + j * lcj] = sum over splits (in split order, FP64) of ws[ks][i][j];
// 64 x 64 tiles through LDS so both the partial reads and the stores of
// either orientation are coalesced
__global__ __launch_bounds__(256) void sg_reduce_kernel(int64_t M, int64_t N, int64_t npad, int nsplit,
                                                        const float* __restrict__ ws, double* __restrict__ C,
                                                        int64_t lci, int64_t lcj) {
  __shared__ double T[64][65];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int ii = tid / 64 + 4 * q, jj = tid % 64;
    const int64_t i = i0 + ii, j = j0 + jj;
    double s = 0.0;
    if (i < M && j < N)
      for (int k = 0; k < nsplit; ++k) s += (double)ws[((int64_t)k * M + i) * npad + j];
    T[ii][jj] = s;
  }
  __syncthreads();
  const bool by_row = (lcj == 1);  // contiguous along j: threads walk j
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int a = tid / 64 + 4 * q, b = tid % 64;
    const int ii = by_row ? a : b, jj = by_row ? b : a;
    const int64_t i = i0 + ii, j = j0 + jj;
    if (i < M && j < N) C[i * lci + j * lcj] = T[ii][jj];
  }
}

struct SgShape {
  int mtiles, ntiles, nsplit;
  int64_t kp, kc, npad;
};

SgShape sg_shape(int P, int64_t M, int64_t N, int64_t K) {
  SgShape s;
  const int tn = (P == 2) ? 288 : 144;
  s.kp = pamd_cdiv(imax64(K, 1), SG_KALIGN) * SG_KALIGN;
  s.npad = pamd_cdiv(imax64(N, 1), SG_NALIGN) * SG_NALIGN;
  s.mtiles = (int)pamd_cdiv(imax64(M, 1), SG_TM);
  s.ntiles = (int)pamd_cdiv(imax64(N, 1), tn);
  const int64_t tiles = (int64_t)s.mtiles * s.ntiles;
  const int64_t ksteps = s.kp / SG_BK;
  int64_t split = imin64(ksteps, imax64(1, pamd_cdiv(256, tiles)));
  s.kc = pamd_cdiv(ksteps, split) * SG_BK;
  s.nsplit = (int)pamd_cdiv(s.kp, s.kc);
  return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int pamd_split_layout(int npieces, int64_t M, int64_t N, int64_t K, int64_t* kp, int64_t* npad,
                                 int64_t* workspace_bytes) {
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 1);
  PAMD_ARG_CHECK(M >= 0, 2);
  PAMD_ARG_CHECK(N >= 0, 3);
  PAMD_ARG_CHECK(K >= 0, 4);
  PAMD_ARG_CHECK(kp != nullptr, 5);
  PAMD_ARG_CHECK(npad != nullptr, 6);
  PAMD_ARG_CHECK(workspace_bytes != nullptr, 7);
  const SgShape s = sg_shape(npieces, M, N, K);
  *kp = s.kp;
  *npad = s.npad;
  *workspace_bytes = (int64_t)s.nsplit * imax64(M, 1) * s.npad * (int64_t)sizeof(float);
  return PAMD_OK;
}

extern "C" int pamd_split_bf16(int transpose, int npieces, int64_t rows, int64_t cols, const double* X, int64_t ldx,
                               int64_t rows_out, int64_t kp, uint16_t* pieces, int64_t ldk, int64_t pstride,
                               void* stream) {
  PAMD_ARG_CHECK(transpose == 0 || transpose == 1, 1);
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 2);
  PAMD_ARG_CHECK(rows >= 0, 3);
  PAMD_ARG_CHECK(cols >= 0, 4);
  PAMD_ARG_CHECK(X != nullptr || rows * cols == 0, 5);
  PAMD_ARG_CHECK(ldx >= cols && ldx >= 1, 6);
  // rows_out: rows of the piece array (X's rows, or its columns when transposed)
  PAMD_ARG_CHECK(rows_out >= (transpose ? cols : rows), 7);
  PAMD_ARG_CHECK(kp % SG_KALIGN == 0 && kp >= (transpose ? rows : cols), 8);
  PAMD_ARG_CHECK(pieces != nullptr, 9);
  PAMD_ARG_CHECK(ldk >= kp && ldk % 8 == 0, 10);
  PAMD_ARG_CHECK(pstride >= ldk * rows_out && pstride % 8 == 0, 11);
  if (rows_out == 0 || kp == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!transpose) {
    dim3 grid((unsigned)pamd_cdiv(kp / 8, 256), (unsigned)imin64(rows_out, 65535));
    if (npieces == 2)
      sg_split_rows_kernel<2><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, kp, pieces, ldk, pstride);
    else
      sg_split_rows_kernel<3><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, kp, pieces, ldk, pstride);
  } else {
    PAMD_ARG_CHECK(pamd_cdiv(rows_out, 64) <= 65535, 7);
    dim3 grid((unsigned)(kp / 64), (unsigned)pamd_cdiv(rows_out, 64));
    if (npieces == 2)
      sg_split_t_kernel<2><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, pieces, ldk, pstride);
    else
      sg_split_t_kernel<3><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, pieces, ldk, pstride);
  }
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_split_gemm_f64(int npieces, int64_t M, int64_t N, int64_t K, const uint16_t* L, int64_t ldl,
                                   int64_t lps, const uint16_t* R, int64_t ldr, int64_t rps, double* C, int64_t lci,
                                   int64_t lcj, void* workspace, void* stream) {
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 1);
  PAMD_ARG_CHECK(M >= 0, 2);
  PAMD_ARG_CHECK(N >= 0, 3);
  PAMD_ARG_CHECK(K >= 0, 4);
  if (M == 0 || N == 0) return PAMD_OK;
  const SgShape s = sg_shape(npieces, M, N, K);
  PAMD_ARG_CHECK(L != nullptr, 5);
  PAMD_ARG_CHECK(ldl >= s.kp && ldl % 8 == 0, 6);
  PAMD_ARG_CHECK(lps >= ldl * M && lps % 8 == 0, 7);
  PAMD_ARG_CHECK(R != nullptr, 8);
  PAMD_ARG_CHECK(ldr >= s.kp && ldr % 8 == 0, 9);
  PAMD_ARG_CHECK(rps >= ldr * s.npad && rps % 8 == 0, 10);
  PAMD_ARG_CHECK(C != nullptr, 11);
  PAMD_ARG_CHECK((lcj == 1 && lci >= N) || (lci == 1 && lcj >= M), 12);
  PAMD_ARG_CHECK(workspace != nullptr, 14);
  hipStream_t st = (hipStream_t)stream;
  const int64_t total = (int64_t)s.mtiles * s.ntiles * s.nsplit;
  PAMD_ARG_CHECK(total < (1LL << 31), 2);
  float* ws = (float*)workspace;
  if (npieces == 2)
    sg_gemm_kernel<2, 2, 2, 4, 9><<<(unsigned)total, SG_THREADS, 0, st>>>(
        M, s.kp, s.kc, s.mtiles, s.ntiles, s.nsplit, L, ldl, lps, R, ldr, rps, ws, s.npad);
  else
    sg_gemm_kernel<3, 4, 1, 2, 9><<<(unsigned)total, SG_THREADS, 0, st>>>(
        M, s.kp, s.kc, s.mtiles, s.ntiles, s.nsplit, L, ldl, lps, R, ldr, rps, ws, s.npad);
  PAMD_LAUNCH_CHECK();
  dim3 rgrid((unsigned)pamd_cdiv(M, 64), (unsigned)pamd_cdiv(N, 64));
  sg_reduce_kernel<<<rgrid, 256, 0, st>>>(M, N, s.npad, s.nsplit, ws, C, lci, lcj);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
