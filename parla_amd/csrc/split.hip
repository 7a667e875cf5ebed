// Split-precision GEMM for the QB power iteration: an FP64 operand is held as
// P BF16 pieces x = x_0 + x_1 + ... (each piece the BF16 rounding of what the
// previous ones leave), and C = L R^T is formed on v_mfma_f32_16x16x32_bf16
// from the piece products L_p R_q^T with p + q < P, accumulated in FP32.
//   P = 2: 16 significant bits per element, three products per tile;
//   P = 3: 24 bits (the FP32 significand), six products.
// The dropped products are below 2^-(8P) of |L| |R| elementwise.
//
// Piece layout ("blocked"): both operands K-contiguous ("TN"), L_p with M
// rows and R_q with N rows (the transpose of the K x N right factor), stored
// as 1 KiB blocks of 16 rows x 32 K: element (i, k) of a piece with kp
// (multiple of 64) columns sits at
//   ((i / 16) (kp / 32) + k / 32) 512 + (i % 16) 32 + 8 (c ^ ((i % 16) / 4)) + k % 8,
// c = (k % 32) / 8 the 16-byte chunk of the row.  A workgroup's K step is
// then whole 1 KiB blocks (full-burst loads), and the chunk swizzle makes the
// MFMA operand reads from LDS conflict-free without padding.  Rows beyond
// the matrix and columns beyond K are zero.  The split kernels write this
// layout from a row-major FP64 matrix, in either orientation, so A (m x n)
// serves A S (L = A) and A^T Y (L = A^T) and each tall k-column factor is
// split straight into its transpose.
//
// GEMM: one workgroup per 128 x TN output tile and K range; eight wavefronts
// split the tile 4 x 2 (P = 2, TN = 288) or four split it 4 x 1 (P = 3,
// TN = 144).  K advances 32 at a time through three LDS buffers filled by
// direct global -> LDS loads two steps ahead (sg_gemm_kernel).  Skinny products
// (N ~ 300) have few tiles, so K is split until there are at least 256
// workgroups; the splits' FP32 partials are summed in split order in FP64 by
// a second kernel, which also writes C in either orientation.  Tiles are
// assigned XCD-aware: the workgroups of one XCD share their K range and
// column tile, so the right operand's slice is fetched into that XCD's L2
// once.
#include "blocked.h"

typedef __bf16 sg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sg_f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int SG_TM = 128;
constexpr int SG_BK = 32;
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

// blocked pieces of X (rows x cols, row-major): one wavefront per 16 x 32
// block, lane l = row l / 4, chunk l % 4 (eight consecutive k)
template <int P>
__global__ __launch_bounds__(256) void sg_split_rows_kernel(int64_t rows, int64_t cols, const double* __restrict__ X,
                                                            int64_t ldx, int64_t nblk, int64_t kb_per_row,
                                                            uint16_t* __restrict__ pieces, int64_t pstride) {
  const int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nblk) return;
  const int lane = threadIdx.x & 63, r = lane >> 2, c = lane & 3;
  const int64_t i = (blk / kb_per_row) * 16 + r;
  const int64_t k0 = (blk % kb_per_row) * 32 + 8 * c;
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
  const int64_t off = blk * 512 + r * 32 + ((c ^ (r >> 2)) << 3);
#pragma unroll
  for (int p = 0; p < P; ++p) *reinterpret_cast<uint4*>(pieces + p * pstride + off) = blk_pack(h[p]);
}

// blocked pieces of X^T (row j of the piece array = column j of X, K = X's
// rows): a 64 x 64 tile of X transposed through LDS, then each wavefront
// writes two 16 x 32 blocks
template <int P>
__global__ __launch_bounds__(256) void sg_split_t_kernel(int64_t rows, int64_t cols, const double* __restrict__ X,
                                                         int64_t ldx, int64_t rows_out, int64_t kp,
                                                         uint16_t* __restrict__ pieces, int64_t pstride) {
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
  const int lane = tid & 63, wave = tid >> 6, r = lane >> 2, c = lane & 3;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int b = wave * 2 + h;  // block (jr, kr) of the tile: 4 row blocks x 2 K blocks
    const int jr = b >> 1, kr = b & 1;
    const int64_t j = j0 + jr * 16 + r;
    if (j >= rows_out) continue;
    uint16_t v[P][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      uint16_t t[P];
      sg_split<P>(T[kr * 32 + 8 * c + e][jr * 16 + r], t);
#pragma unroll
      for (int p = 0; p < P; ++p) v[p][e] = t[p];
    }
    const int64_t off = blk_off(j, i0 + kr * 32 + 8 * c, kp);
#pragma unroll
    for (int p = 0; p < P; ++p) *reinterpret_cast<uint4*>(pieces + p * pstride + off) = blk_pack(v[p]);
  }
}

// ws[ks][i][j] (FP32) = sum over p + q < P of L_p[i, K_ks] R_q[j, K_ks]^T.
// K steps stream through three LDS buffers by direct global -> LDS loads
// (global_load_lds_dwordx4: a wavefront moves one 1 KiB block per
// instruction, no registers involved), two steps ahead.  Direct-to-LDS loads
// of one wavefront can complete out of issue order, so only vmcnt(0) says
// that a wavefront's loads are in.  The steps therefore belong alternately to
// two groups of wavefronts (the halves of the workgroup): the group that owns
// step s loads all of its blocks, and at step s it waits with vmcnt(0) for
// them alone (its next loads, for step s + 2, are issued after that wait),
// then the barrier makes them visible to all.  Two steps stay in flight, and
// every wait is a full one.  The buffers are separate arrays and the loop is unrolled by three,
// so every access names its buffer at compile time.
template <int P, int WM, int WN, int MB, int NB>
__global__ __launch_bounds__(64 * WM * WN) void sg_gemm_kernel(int64_t kp, int64_t kc, int mtiles, int ntiles,
                                                             int nsplit, const uint16_t* __restrict__ L,
                                                             int64_t lps, const uint16_t* __restrict__ R,
                                                             int64_t rps, int64_t M, float* __restrict__ ws,
                                                             int64_t npad) {
  constexpr int TM = WM * MB * 16, TN = WN * NB * 16;
  static_assert(TM == SG_TM && (WM * WN == 4 || WM * WN == 8), "four or eight wavefronts on a 128-row tile");
  constexpr int LB = TM / 16, RB = TN / 16;  // 1 KiB blocks per piece per K step
  constexpr int NBLK = P * (LB + RB);        // blocks per K step: L pieces, then R pieces
  constexpr int G = 2, NWG = WM * WN / G;    // loading groups, wavefronts per group
  static_assert(LB % NWG == 0, "row blocks of L split evenly over a group");
  __shared__ uint4 S0[NBLK][64], S1[NBLK][64], S2[NBLK][64];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: block choices are scalar
  const int wm = wave % WM, wn = wave / WM;
  const int grp = wave / NWG, wl = wave % NWG;
  const int frag = blk_frag(lane);
  const int lr = lane & 15, lk = lane >> 4;
  const uint32_t total = (uint32_t)mtiles * ntiles * nsplit;
  const uint32_t t = xcd_tile(blockIdx.x, total);
  const int ks = (int)(t / ((uint32_t)mtiles * ntiles));
  const uint32_t rem = t % ((uint32_t)mtiles * ntiles);
  const int nt = (int)(rem / mtiles), mt = (int)(rem % mtiles);
  const int64_t kbn = kp >> 5;  // K blocks per row block
  const int64_t kb0 = ((int64_t)ks * kc) >> 5;
  const int nsteps = (int)((imin64(kp, (int64_t)(ks + 1) * kc) - (int64_t)ks * kc) >> 5);
  const uint4* Lg = reinterpret_cast<const uint4*>(L);
  const uint4* Rg = reinterpret_cast<const uint4*>(R);
  const int64_t lps4 = lps >> 3, rps4 = rps >> 3;

  // the calling wavefront (of the group owning `step`) loads rows wl, wl + NWG,
  // ... of every piece's block column of the step into buffer `dst`
  auto issue = [&](int step, uint4 (*dst)[64]) {
    const int64_t kb = kb0 + step;
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int r = 0; r < LB / NWG; ++r) {
        const int rb = wl + NWG * r;
        blk_load_lds(Lg + p * lps4 + (((int64_t)mt * LB + rb) * kbn + kb) * 64, &dst[p * LB + rb][0], lane);
      }
#pragma unroll
      for (int r = 0; r < (RB + NWG - 1) / NWG; ++r) {
        // an odd RB: the last row is loaded by both wavefronts (same bytes, same place), which keeps
        // every load unconditional
        const int rb = (RB % NWG == 0) ? wl + NWG * r : (wl + NWG * r < RB ? wl + NWG * r : RB - 1);
        blk_load_lds(Rg + p * rps4 + (((int64_t)nt * RB + rb) * kbn + kb) * 64, &dst[P * LB + p * RB + rb][0], lane);
      }
    }
  };

  sg_f4 acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = sg_f4{0.f, 0.f, 0.f, 0.f};

  // one K step: the owning group waits for its loads (the step's blocks),
  // barrier (every block is in, and the buffer read in the previous step is
  // free), the owner issues the loads of step s + 2 into that buffer, then
  // every wavefront reads its fragments and runs the MFMAs
  auto step = [&](int s, uint4 (*cur)[64], uint4 (*ahead)[64]) {
    const bool own = (s % G) == grp;
    if (own) __builtin_amdgcn_s_waitcnt(0x3F70);  // vmcnt(0)
    asm volatile("s_barrier" ::: "memory");
    sg_bf16x8 a[P][MB], b[P][NB];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        a[p][mb] = *reinterpret_cast<const sg_bf16x8*>(&cur[p * LB + wm * MB + mb][frag]);
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        b[q][nb] = *reinterpret_cast<const sg_bf16x8*>(&cur[P * LB + q * RB + wn * NB + nb][frag]);
    asm volatile("" ::: "memory");
    if (own && s + G < nsteps) issue(s + G, ahead);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int q = 0; q < P; ++q) {
          if (p + q >= P) continue;
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p][mb], b[q][nb], acc[mb][nb], 0, 0, 0);
        }
  };

  if (grp == 0 && nsteps > 0) issue(0, S0);
  if (grp == 1 && nsteps > 1) issue(1, S1);
  for (int s = 0; s < nsteps; s += 3) {
    step(s, S0, S2);
    if (s + 1 < nsteps) step(s + 1, S1, S0);
    if (s + 2 < nsteps) step(s + 2, S2, S1);
  }
  // D layout of v_mfma_f32_16x16x32_bf16: lane l, item r -> row 4 (l / 16) + r, column l % 16
  float* w = ws + (int64_t)ks * M * npad;
  const int64_t i0 = (int64_t)mt * TM, j0 = (int64_t)nt * TN;
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
// 64 x 64 tiles through LDS so both the partial reads (four columns per
// thread) and the stores of either orientation are coalesced
__global__ __launch_bounds__(256) void sg_reduce_kernel(int64_t M, int64_t N, int64_t npad, int nsplit,
                                                        const float* __restrict__ ws, double* __restrict__ C,
                                                        int64_t lci, int64_t lcj) {
  __shared__ double T[64][65];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
  const int jq = (tid & 15) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int ii = (tid >> 4) + 16 * q;
    const int64_t i = i0 + ii, j = j0 + jq;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (i < M && j < npad) {
      const float4* src = reinterpret_cast<const float4*>(ws + i * npad + j);
      const int64_t stride = M * npad / 4;
      for (int k = 0; k < nsplit; ++k) {
        const float4 v = src[k * stride];
        s0 += (double)v.x;
        s1 += (double)v.y;
        s2 += (double)v.z;
        s3 += (double)v.w;
      }
    }
    T[ii][jq] = s0;
    T[ii][jq + 1] = s1;
    T[ii][jq + 2] = s2;
    T[ii][jq + 3] = s3;
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
  int64_t kp, kc, mpad, npad;
};

SgShape sg_shape(int P, int64_t M, int64_t N, int64_t K) {
  SgShape s;
  const int tn = (P == 2) ? 288 : 144;
  s.kp = pamd_cdiv(imax64(K, 1), SG_KALIGN) * SG_KALIGN;
  s.mpad = pamd_cdiv(imax64(M, 1), SG_TM) * SG_TM;
  s.npad = pamd_cdiv(imax64(N, 1), SG_NALIGN) * SG_NALIGN;
  s.mtiles = (int)(s.mpad / SG_TM);
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
extern "C" int pamd_split_layout(int npieces, int64_t M, int64_t N, int64_t K, int64_t* kp, int64_t* mpad,
                                 int64_t* npad, int64_t* workspace_bytes) {
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 1);
  PAMD_ARG_CHECK(M >= 0, 2);
  PAMD_ARG_CHECK(N >= 0, 3);
  PAMD_ARG_CHECK(K >= 0, 4);
  PAMD_ARG_CHECK(kp != nullptr, 5);
  PAMD_ARG_CHECK(mpad != nullptr, 6);
  PAMD_ARG_CHECK(npad != nullptr, 7);
  PAMD_ARG_CHECK(workspace_bytes != nullptr, 8);
  const SgShape s = sg_shape(npieces, M, N, K);
  *kp = s.kp;
  *mpad = s.mpad;
  *npad = s.npad;
  *workspace_bytes = (int64_t)s.nsplit * imax64(M, 1) * s.npad * (int64_t)sizeof(float);
  return PAMD_OK;
}

extern "C" int pamd_split_bf16(int transpose, int npieces, int64_t rows, int64_t cols, const double* X, int64_t ldx,
                               int64_t rows_out, int64_t kp, uint16_t* pieces, int64_t pstride, void* stream) {
  PAMD_ARG_CHECK(transpose == 0 || transpose == 1, 1);
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 2);
  PAMD_ARG_CHECK(rows >= 0, 3);
  PAMD_ARG_CHECK(cols >= 0, 4);
  PAMD_ARG_CHECK(X != nullptr || rows * cols == 0, 5);
  PAMD_ARG_CHECK(ldx >= cols && ldx >= 1, 6);
  // rows_out: rows of the piece array (X's rows, or its columns when transposed)
  PAMD_ARG_CHECK(rows_out >= (transpose ? cols : rows) && rows_out % 16 == 0, 7);
  PAMD_ARG_CHECK(kp % SG_KALIGN == 0 && kp >= (transpose ? rows : cols), 8);
  PAMD_ARG_CHECK(pieces != nullptr, 9);
  PAMD_ARG_CHECK(pstride >= rows_out * kp && pstride % 8 == 0, 10);
  if (rows_out == 0 || kp == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!transpose) {
    const int64_t kbn = kp / 32, nblk = (rows_out / 16) * kbn;
    PAMD_ARG_CHECK(pamd_cdiv(nblk, 4) < (1LL << 31), 7);
    const unsigned grid = (unsigned)pamd_cdiv(nblk, 4);
    if (npieces == 2)
      sg_split_rows_kernel<2><<<grid, 256, 0, st>>>(rows, cols, X, ldx, nblk, kbn, pieces, pstride);
    else
      sg_split_rows_kernel<3><<<grid, 256, 0, st>>>(rows, cols, X, ldx, nblk, kbn, pieces, pstride);
  } else {
    PAMD_ARG_CHECK(pamd_cdiv(rows_out, 64) <= 65535, 7);
    dim3 grid((unsigned)(kp / 64), (unsigned)pamd_cdiv(rows_out, 64));
    if (npieces == 2)
      sg_split_t_kernel<2><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, kp, pieces, pstride);
    else
      sg_split_t_kernel<3><<<grid, 256, 0, st>>>(rows, cols, X, ldx, rows_out, kp, pieces, pstride);
  }
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_split_gemm_f64(int npieces, int64_t M, int64_t N, int64_t K, const uint16_t* L, int64_t lps,
                                   const uint16_t* R, int64_t rps, double* C, int64_t lci, int64_t lcj,
                                   void* workspace, void* stream) {
  PAMD_ARG_CHECK(npieces == 2 || npieces == 3, 1);
  PAMD_ARG_CHECK(M >= 0, 2);
  PAMD_ARG_CHECK(N >= 0, 3);
  PAMD_ARG_CHECK(K >= 0, 4);
  if (M == 0 || N == 0) return PAMD_OK;
  const SgShape s = sg_shape(npieces, M, N, K);
  PAMD_ARG_CHECK(L != nullptr, 5);
  PAMD_ARG_CHECK(lps >= s.mpad * s.kp && lps % 8 == 0, 6);
  PAMD_ARG_CHECK(R != nullptr, 7);
  PAMD_ARG_CHECK(rps >= s.npad * s.kp && rps % 8 == 0, 8);
  PAMD_ARG_CHECK(C != nullptr, 9);
  PAMD_ARG_CHECK((lcj == 1 && lci >= N) || (lci == 1 && lcj >= M), 10);
  PAMD_ARG_CHECK(workspace != nullptr, 12);
  hipStream_t st = (hipStream_t)stream;
  const int64_t total = (int64_t)s.mtiles * s.ntiles * s.nsplit;
  PAMD_ARG_CHECK(total < (1LL << 31), 2);
  float* ws = (float*)workspace;
  // 2 pieces: eight wavefronts of 32 x 144 (two per SIMD) beat four of
  // 64 x 144 by 5-10% (QB1 A S 75.8 -> 71.7 us); 3 pieces: four of 32 x 144
  // and eight of 16 x 144 are within 1%, four kept
  if (npieces == 2)
    sg_gemm_kernel<2, 4, 2, 2, 9><<<(unsigned)total, 512, 0, st>>>(
        s.kp, s.kc, s.mtiles, s.ntiles, s.nsplit, L, lps, R, rps, M, ws, s.npad);
  else
    sg_gemm_kernel<3, 4, 1, 2, 9><<<(unsigned)total, 256, 0, st>>>(
        s.kp, s.kc, s.mtiles, s.ntiles, s.nsplit, L, lps, R, rps, M, ws, s.npad);
  PAMD_LAUNCH_CHECK();
  dim3 rgrid((unsigned)pamd_cdiv(M, 64), (unsigned)pamd_cdiv(N, 64));
  sg_reduce_kernel<<<rgrid, 256, 0, st>>>(M, N, s.npad, s.nsplit, ws, C, lci, lcj);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
