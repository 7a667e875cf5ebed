// Mixed-precision sketch-and-precondition kernels.
//
// 1. Exact BF16 sketch (error-free splitting, after Ozaki, Ogita, Oishi and
//    Rump, "Error-free transformations of matrix multiplication by using fast
//    routines of matrix multiplication", Numer. Algorithms 59, 2012).
//    The Gaussian sketch is drawn directly on an integer grid,
//        S = 2^-5 / sqrt(min(d, m)) * Q,   Q[i][k] = rint(32 z_ik)  (|Q| <= 256),
//    so every entry of Q is an exact BF16 number.  A (and b, as one more
//    column) is cut into chunks of 256 rows; in chunk c, column j is scaled by
//    2^(8 - e_cj) so that it lies in (-256, 256) and split into P BF16 pieces
//        A[k][j] = 2^(e_cj - 8) * sum_p 2^(-9p) a_p[k][j] + (remainder < 2^(e_cj - 9P)),
//    every a_p an integer in [-256, 256].  A product Q[i][k] a_p[k][j] is an
//    integer below 2^16 in magnitude, so a 256-term sum of them is below 2^24
//    and v_mfma_f32_16x16x32_bf16 accumulates it in FP32 with no rounding at
//    all.  Each chunk's FP32 sums are scaled by 2^(e_cj - 8 - 9p) and added to
//    FP64 accumulators.  The only errors are the truncation of A after P
//    pieces (relative 2^-9P to the chunk-column maximum) and the FP64
//    additions of the chunk sums.
// 2. q = B^T (B p) for an FP32- or FP64-stored B with FP64 vectors: one read
//    of B per call (the operator of CG on the preconditioned normal equations).
// 3. Compensated residual: r = b - A x and g = A^T r with error-free products
//    and sums (Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J.
//    Sci. Comput. 26, 2005), one read of A.
// 4. Right triangular solve X R = A for large m, recursive: 32-column leaves
//    by substitution (one thread per row), the off-diagonal blocks by the
//    FP64 MFMA GEMM.
#include "blocked.h"
#include "common.h"
#include "rng.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int OZ_CHUNK = 256;       // K rows per exact FP32 accumulation
constexpr int OZ_MAX_PIECES = 5;
constexpr int OZ_TM = 128, OZ_TN = 64;
// 16-row blocks per wavefront: 4 (M) x 2 (N) wavefronts of 32 x 32 outputs.  Four
// wavefronts of 64 x 32 read less LDS per MFMA but, one per SIMD, hide no
// latency (1.6x slower at 4096 x 262144 x 1025).
constexpr int OZ_WAVE_MB = 2;
constexpr int OZ_THREADS = 64 * 2 * (OZ_TM / (16 * OZ_WAVE_MB));
constexpr int OZ_TARGET_WG = 4096;  // K groups until the grid has about this many workgroups
constexpr int OZ_BAND = 8;          // row tiles per band of the tile order

__device__ __forceinline__ uint16_t bf16_bits_of_int(double a) {
  // a is an integer with |a| <= 256: exact in BF16 (8 significant bits)
  const float f = (float)a;
  return (uint16_t)(__float_as_uint(f) >> 16);
}

// ---------------------------------------------------------------------------
// quantised Gaussian, rows x kp (zeros for k >= m); entry (i, k < m) uses the
// same Philox counter as the FP64 fill of a rows x m matrix, so the FP64 fill
// times 32, rounded to the nearest integer, is this operator
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t gauss_q_value(double z) {
  return bf16_bits_of_int(fmin(256.0, fmax(-256.0, rint(32.0 * z))) + 0.0);  // + 0.0: zero is +0
}

// The quantised values of normal_pair(seed, pair): the pair is first formed
// in FP32 from the same uniforms.  With the uniforms rounded to FP32 (2^-24
// relative), logf and sincospif within 2 ulp and sqrtf and the products
// within 1 ulp, T = -2 log u1 is off by at most 1.2e-7 + 2.4e-7 T, the radius
// r = sqrt(T) by 6e-8 / r + 2.4e-7 r, cos / sin by 6.2e-7, and so
// |z32 - z| <= 6e-8 / r + 9.8e-7 r.  When either 32 z32 lies within twice
// that bound (in units of 1/32) of a half-integer, the pair is recomputed in
// FP64, so the result is always that of gauss_q_value on the FP64 pair.
__device__ __forceinline__ void normal_pair_q(uint64_t seed, uint64_t pair, uint16_t& q0, uint16_t& q1) {
  const U4 r = philox4x32_10(U4{(uint32_t)pair, (uint32_t)(pair >> 32), 0x6A09E667u, 0u}, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  const float u1 = (float)u53_open0(r.x, r.y);
  const float u2 = (float)u53_open0(r.z, r.w);
  const float rad = sqrtf(-2.0f * logf(u1));
  float sf, cf;
  sincospif(2.0f * u2, &sf, &cf);
  const float t0 = 32.0f * (rad * cf), t1 = 32.0f * (rad * sf);
  const float d0 = fabsf(t0 - (floorf(t0) + 0.5f)), d1 = fabsf(t1 - (floorf(t1) + 0.5f));
  const float margin = 64.0f * (6e-8f / fmaxf(rad, 1e-30f) + 1e-6f * rad);  // 2 x 32 x the bound
  if (d0 < margin || d1 < margin) {
    double z0, z1;
    normal_pair(seed, pair, z0, z1);
    q0 = gauss_q_value(z0);
    q1 = gauss_q_value(z1);
  } else {
    q0 = bf16_bits_of_int((double)fminf(256.0f, fmaxf(-256.0f, rintf(t0))) + 0.0);
    q1 = bf16_bits_of_int((double)fminf(256.0f, fmaxf(-256.0f, rintf(t1))) + 0.0);
  }
}

// Blocked Q (rows_out x kp, the layout of blocked.h; zeros for i >= rows or
// k >= m): one wavefront per 16 x 32 block, lane l = row l / 4, eight
// consecutive k; the Philox pair of two consecutive entries is drawn once.
__global__ __launch_bounds__(256) void gauss_q_kernel(int64_t rows, int64_t m, int64_t nblk, int64_t kbn,
                                                      uint64_t seed, uint64_t offset, uint16_t* __restrict__ Q) {
  const int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nblk) return;
  const int lane = threadIdx.x & 63, r = lane >> 2, c = lane & 3;
  const int64_t i = (blk / kbn) * 16 + r;
  const int64_t k0 = (blk % kbn) * 32 + 8 * c;
  uint16_t v[8];
  uint64_t pair = ~0ull;
  uint16_t q0 = 0, q1 = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    v[t] = 0;
    const int64_t k = k0 + t;
    if (i < rows && k < m) {
      const uint64_t e = offset + (uint64_t)(i * m + k);
      if ((e >> 1) != pair) {
        pair = e >> 1;
        normal_pair_q(seed, pair, q0, q1);
      }
      v[t] = (e & 1) ? q1 : q0;
    }
  }
  *reinterpret_cast<uint4*>(Q + blk * 512 + r * 32 + ((c ^ (r >> 2)) << 3)) = blk_pack(v);
}

// ---------------------------------------------------------------------------
// split: block (chunk c, 32 columns); thread (column jc, row group rg of 32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double oz_load(int64_t k, int64_t j, int64_t m, int64_t n, const double* A, int64_t lda,
                                          const double* b) {
  if (k >= m) return 0.0;
  if (j < n) return A[k * lda + j];
  if (j == n && b != nullptr) return b[k];
  return 0.0;
}

__global__ __launch_bounds__(256) void oz_split_kernel(int64_t m, int64_t n, const double* __restrict__ A, int64_t lda,
                                                       const double* __restrict__ b, int P,
                                                       uint16_t* __restrict__ pieces, int64_t ldk, int64_t pstride,
                                                       int* __restrict__ exps, int64_t lde) {
  __shared__ double smax[8][32];
  __shared__ int sexp[32];
  const int jc = threadIdx.x & 31, rg = threadIdx.x >> 5;
  const int64_t c = blockIdx.x;
  const int64_t j = (int64_t)blockIdx.y * 32 + jc;
  const int64_t k0 = c * OZ_CHUNK + rg * 32;
  double x[32];  // this thread's 32 rows of column j, read once
  double mx = 0.0;
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    x[r] = oz_load(k0 + r, j, m, n, A, lda, b);
    mx = fmax(mx, fabs(x[r]));
  }
  smax[rg][jc] = mx;
  __syncthreads();
  if (rg == 0) {
    double t = 0.0;
    for (int g = 0; g < 8; ++g) t = fmax(t, smax[g][jc]);
    int e = 0;
    if (t > 0.0) frexp(t, &e);  // t < 2^e
    sexp[jc] = e;
    if (j < lde) exps[c * lde + j] = e;
  }
  __syncthreads();
  const int e = sexp[jc];
  // rows k0 + 8q .. k0 + 8q + 7: all P pieces in one pass over the values,
  // piece p = rint of y after p steps of y <- (y - rint(y)) * 512
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint16_t h[OZ_MAX_PIECES][8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double y = ldexp(x[8 * q + i], 8 - e);
#pragma unroll
      for (int p = 0; p < OZ_MAX_PIECES; ++p) {
        if (p < P) {
          const double a = rint(y);
          h[p][i] = bf16_bits_of_int(a);
          y = (y - a) * 512.0;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < OZ_MAX_PIECES; ++p) {
      if (p < P) {
        uint4 w;
        w.x = (uint32_t)h[p][0] | ((uint32_t)h[p][1] << 16);
        w.y = (uint32_t)h[p][2] | ((uint32_t)h[p][3] << 16);
        w.z = (uint32_t)h[p][4] | ((uint32_t)h[p][5] << 16);
        w.w = (uint32_t)h[p][6] | ((uint32_t)h[p][7] << 16);
        *reinterpret_cast<uint4*>(pieces + p * pstride + blk_off(j, k0 + 8 * q, ldk)) = w;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// exact sketch GEMM:
//   part[g][i][j] = sum over the chunks c of group g (in order) of
//                   sum_p 2^(e_cj - 8 - 9p) Q[i, chunk c] . a_p[chunk c, j]
// and out = scale * sum_g part[g] (groups in order).  Q (rows padded to 128)
// and the pieces (rows = columns of [A | b], padded to 64) are blocked
// (blocked.h).  One 512-thread workgroup per 128 x 64 tile and K group.
//
// The K steps of 32 stream through LDS in batches of oz_batch(P) steps by
// direct global -> LDS loads, with the 64 chunk exponents of the tile's
// columns beside each batch.  Direct-to-LDS loads of one wavefront can
// complete out of issue order, so a partial vmcnt does not say which of them
// are in; instead the two groups of four wavefronts own alternate batches.
// The owner of batch j issues its loads two batches ahead and issues nothing
// else until batch j, where it waits with vmcnt(0); a barrier then publishes
// the batch to all eight wavefronts.  Three buffers keep two batches in
// flight while one is multiplied.
// ---------------------------------------------------------------------------
constexpr int oz_batch(int P) { return P <= 4 ? 2 : 1; }  // three buffers of 8 + 4 P KiB per step fit in LDS

template <int P>
struct OzBuf {
  uint4 blk[oz_batch(P)][OZ_TM / 16 + P * (OZ_TN / 16)][64];
  int e[OZ_TN];
};

template <int P>
__global__ __launch_bounds__(OZ_THREADS) void oz_gemm_kernel(int64_t kp, int64_t group_chunks, int mtiles,
                                                             int ntiles, int ngroups,
                                                             const uint16_t* __restrict__ Q,
                                                             const uint16_t* __restrict__ pieces,
                                                             int64_t pstride, const int* __restrict__ exps,
                                                             int64_t lde, int64_t d, int64_t ncols,
                                                             int64_t ncp, double* __restrict__ part) {
  constexpr int LB = OZ_TM / 16, RB = OZ_TN / 16;  // 1 KiB blocks per K step: Q rows, piece rows
  constexpr int NBLK = LB + P * RB;
  constexpr int MB = OZ_WAVE_MB, WM = LB / MB;    // row blocks per wavefront, wavefronts along M
  constexpr int G = 2, NWG = (OZ_THREADS / 64) / G;  // loading groups, wavefronts per group
  static_assert(LB % NWG == 0 && RB % NWG == 0, "row blocks split evenly over a group");
  constexpr int B = oz_batch(P);
  constexpr int CS = OZ_CHUNK / 32;  // K steps per chunk
  static_assert(CS % B == 0, "a batch lies within one chunk");
  __shared__ OzBuf<P> S[3];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: block choices are scalar
  const int wm = wave % WM, wn = wave / WM;
  const int grp = wave / NWG, wl = wave % NWG;
  const int lr = lane & 15, lk = lane >> 4;
  const int frag = blk_frag(lane);
  const uint32_t total = (uint32_t)mtiles * ntiles * ngroups;
  const uint32_t t = xcd_tile(blockIdx.x, total);
  const int g = (int)(t / ((uint32_t)mtiles * ntiles));
  // within a K group, bands of OZ_BAND row tiles, column-major inside a band:
  // the 32 workgroups an XCD runs at once share 8 Q row tiles and 4 column
  // tiles of pieces in its L2 instead of 32 Q row tiles and one column tile
  const uint32_t rem = t % ((uint32_t)mtiles * ntiles);
  const uint32_t band = rem / ((uint32_t)OZ_BAND * ntiles);
  const uint32_t bw = imin64(OZ_BAND, (int64_t)mtiles - (int64_t)band * OZ_BAND);
  const uint32_t v = rem - band * (uint32_t)OZ_BAND * ntiles;
  const int nt = (int)(v / bw), mt = (int)(band * OZ_BAND + v % bw);
  const int64_t kbn = kp >> 5;
  const int64_t nchunks = kp / OZ_CHUNK;
  const int64_t c0 = (int64_t)g * group_chunks;
  const int64_t c1 = imin64(nchunks, c0 + group_chunks);
  const int nbatch = (int)((c1 - c0) * CS / B);
  const int64_t kb0 = c0 * CS;
  const uint4* Qg = reinterpret_cast<const uint4*>(Q);
  const uint4* Pg = reinterpret_cast<const uint4*>(pieces);
  const int64_t ps4 = pstride >> 3;
  // this lane's exponent column for the exponent load (clamped: the padded
  // columns have zero pieces, so any finite exponent leaves them zero)
  const int64_t ecol = imin64((int64_t)nt * OZ_TN + lane, ncols - 1);

  // the calling wavefront (of the group owning batch j) loads rows wl,
  // wl + NWG, ... of Q and of every piece for each step of the batch, and
  // wavefront 0 of the group the batch's chunk exponents
  auto issue = [&](int j, OzBuf<P>& dst) {
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int64_t kb = kb0 + (int64_t)j * B + u;
#pragma unroll
      for (int r = 0; r < LB / NWG; ++r) {
        const int rb = wl + NWG * r;
        blk_load_lds(Qg + (((int64_t)mt * LB + rb) * kbn + kb) * 64, &dst.blk[u][rb][0], lane);
      }
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int r = 0; r < RB / NWG; ++r) {
          const int rb = wl + NWG * r;
          blk_load_lds(Pg + p * ps4 + (((int64_t)nt * RB + rb) * kbn + kb) * 64, &dst.blk[u][LB + p * RB + rb][0],
                       lane);
        }
    }
    if (wl == 0) {
      const int64_t cidx = c0 + ((int64_t)j * B) / CS;
      __builtin_amdgcn_global_load_lds((blk_glb_void*)(exps + cidx * lde + ecol), (blk_lds_void*)&dst.e[0], 4, 0, 0);
    }
  };

  f4 acc[P][MB][2];
  double acc64[MB][2][4];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
      for (int bq = 0; bq < 2; ++bq) acc[p][a][bq] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int bq = 0; bq < 2; ++bq)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[a][bq][r] = 0.0;

  // batch j from buffer cur; ahead is the buffer read in batch j - 1, which
  // the barrier frees for the owner's loads of batch j + 2
  auto batch = [&](int j, const OzBuf<P>& cur, OzBuf<P>& ahead) {
    const bool own = (j % G) == grp;
    if (own) __builtin_amdgcn_s_waitcnt(0x3F70);  // vmcnt(0): this batch's blocks and exponents are in
    asm volatile("s_barrier" ::: "memory");
    if (own && j + G < nbatch) issue(j + G, ahead);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      bf16x8 a[MB], bv[P][2];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) a[mb] = *reinterpret_cast<const bf16x8*>(&cur.blk[u][wm * MB + mb][frag]);
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
          bv[p][nb] = *reinterpret_cast<const bf16x8*>(&cur.blk[u][LB + p * RB + wn * 2 + nb][frag]);
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int nb = 0; nb < 2; ++nb)
            acc[p][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], bv[p][nb], acc[p][mb][nb], 0, 0, 0);
    }
    if (((j + 1) * B) % CS == 0) {  // end of a chunk: exact FP32 sums -> FP64, pieces in order
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const int e = cur.e[32 * wn + 16 * nb + lr];
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const double sc = ldexp(1.0, e - 8 - 9 * p);
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[mb][nb][r] += (double)acc[p][mb][nb][r] * sc;
        }
      }
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int nb = 0; nb < 2; ++nb) acc[p][mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
    }
  };

  if (grp == 0 && nbatch > 0) issue(0, S[0]);
  if (grp == 1 && nbatch > 1) issue(1, S[1]);
  for (int j = 0; j < nbatch; j += 3) {
    batch(j, S[0], S[2]);
    if (j + 1 < nbatch) batch(j + 1, S[1], S[0]);
    if (j + 2 < nbatch) batch(j + 2, S[2], S[1]);
  }
  // D layout of v_mfma_f32_16x16x32_bf16: lane l, item r -> row 4 (l / 16) + r, col l % 16
  double* w = part + (int64_t)g * d * ncp;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = (int64_t)mt * OZ_TM + 16 * (MB * wm + mb) + 4 * lk + r;
        const int64_t j = (int64_t)nt * OZ_TN + 32 * wn + 16 * nb + lr;
        if (i < d) w[i * ncp + j] = acc64[mb][nb][r];
      }
}

// out[i][j] = scale * sum_g part[g][i][j], groups in order
__global__ __launch_bounds__(256) void oz_reduce_kernel(int64_t d, int64_t ncols, int64_t ncp, int ngroups,
                                                        const double* __restrict__ part, double scale,
                                                        double* __restrict__ out, int64_t ldo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= d * ncols) return;
  const int64_t i = idx / ncols, j = idx % ncols;
  double s = 0.0;
  for (int g = 0; g < ngroups; ++g) s += part[((int64_t)g * d + i) * ncp + j];
  out[i * ldo + j] = scale * s;
}

struct OzShape {
  int mtiles, ntiles, ngroups;
  int64_t kp, ncp, mpad, group_chunks;
};

// K groups: about OZ_TARGET_WG workgroups in all, at least two chunks per
// group (the group size is part of the summation order)
OzShape oz_shape(int64_t d, int64_t m, int64_t ncols) {
  OzShape s;
  s.kp = pamd_cdiv(imax64(m, 1), OZ_CHUNK) * OZ_CHUNK;
  s.ncp = pamd_cdiv(imax64(ncols, 1), OZ_TN) * OZ_TN;
  s.mpad = pamd_cdiv(imax64(d, 1), OZ_TM) * OZ_TM;
  s.mtiles = (int)(s.mpad / OZ_TM);
  s.ntiles = (int)(s.ncp / OZ_TN);
  const int64_t nchunks = s.kp / OZ_CHUNK;
  const int64_t want = imax64(1, pamd_cdiv(OZ_TARGET_WG, (int64_t)s.mtiles * s.ntiles));
  s.group_chunks = imax64(2, pamd_cdiv(nchunks, want));
  s.ngroups = (int)pamd_cdiv(nchunks, s.group_chunks);
  return s;
}

// ---------------------------------------------------------------------------
// q = B^T (B p): row per wavefront, partials per workgroup, ordered reduce
// ---------------------------------------------------------------------------
constexpr int NP_THREADS = 256;
constexpr int NP_MAX_GRID = 1024;

// W wavefronts per workgroup, D rows in flight per wavefront.  Rows of at
// most 256: 16 wavefronts (more rows in flight for the same number of
// partials); where the grid is at its cap (m >= 64 NP_MAX_GRID) each
// wavefront walks many rows, and 4 in flight from a quarter of the
// workgroups is faster (131072 x 200 FP32 with its reduction: 22.8 against
// 27.3 us; at 20000 rows it is slower, 14.4 against 13.5).  Wider rows: 4
// wavefronts (LDS, registers), one row ahead.
template <typename T, int NPL, int W, int D>
__global__ __launch_bounds__(64 * W) void normal_pass_kernel(int64_t m, int64_t n, const T* __restrict__ B,
                                                             int64_t ldb, const double* __restrict__ p,
                                                             double* __restrict__ part) {
  __shared__ double red[W][64 * NPL];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t gw = (int64_t)gridDim.x * W;
  double pr[NPL], acc[NPL];
#pragma unroll
  for (int q = 0; q < NPL; ++q) {
    const int64_t c = lane + 64 * q;
    pr[q] = (c < n) ? p[c] : 0.0;
    acc[q] = 0.0;
  }
  int64_t row = (int64_t)blockIdx.x * W + wave;
  // a[k]: row + k gw (zeros past the end; such a row is never reduced)
  T a[D][NPL];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int64_t r = row + k * gw;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      a[k][q] = (r < m && c < n) ? B[r * ldb + c] : T(0);
    }
  }
  for (; row < m; row += gw) {
    // row + D gw is loaded while this row is reduced
    const int64_t nrow = row + D * gw;
    T an[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      an[q] = (nrow < m && c < n) ? B[nrow * ldb + c] : T(0);
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NPL; ++q) s += (double)a[0][q] * pr[q];
    s = wave_sum(s);
#pragma unroll
    for (int q = 0; q < NPL; ++q) acc[q] += s * (double)a[0][q];
#pragma unroll
    for (int k = 0; k + 1 < D; ++k)
#pragma unroll
      for (int q = 0; q < NPL; ++q) a[k][q] = a[k + 1][q];
#pragma unroll
    for (int q = 0; q < NPL; ++q) a[D - 1][q] = an[q];
  }
#pragma unroll
  for (int q = 0; q < NPL; ++q) red[wave][lane + 64 * q] = acc[q];
  __syncthreads();
  for (int64_t c = threadIdx.x; c < n; c += 64 * W) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) t += red[w][c];
    part[(int64_t)blockIdx.x * n + c] = t;
  }
}

// out[c] = sum_g part[g][c] in a fixed order: workgroup per 64 columns, its 16
// wavefronts take the partials g = w, w + 16, ... (coalesced 64-column rows),
// and their sums are added in wavefront order.
constexpr int RED_WAVES = 16;

__global__ __launch_bounds__(64 * RED_WAVES) void reduce_partials_kernel(int64_t n, int64_t nparts,
                                                                         const double* __restrict__ part,
                                                                         double* __restrict__ out) {
  __shared__ double red[RED_WAVES][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  double t = 0.0;
  if (c < n) {
#pragma unroll 8
    for (int64_t g = wave; g < nparts; g += RED_WAVES) t += part[g * n + c];
  }
  red[wave][lane] = t;
  __syncthreads();
  if (wave == 0 && c < n) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < RED_WAVES; ++w) s += red[w][lane];
    out[c] = s;
  }
}

// ---------------------------------------------------------------------------
// compensated residual
// ---------------------------------------------------------------------------
// Error-free transformations.  FP contraction is switched off in their
// bodies: HIP contracts across statements by default and would fuse s + a * b
// (the product from two_prod, the sum in two_sum) into one FMA, after which
// the recovered error terms are no longer exact.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}

// p + e == a * b exactly
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
#pragma clang fp contract(off)
  p = a * b;
  e = fma(a, b, -p);
}

// a + b, never fused with a neighbouring product
__device__ __forceinline__ double add_nc(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

// (s, c) pair sum over the wavefront, error-free at every step
__device__ __forceinline__ void wave_sum2(double& s, double& c) {
#pragma clang fp contract(off)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double so = __shfl_xor(s, off, 64);
    const double co = __shfl_xor(c, off, 64);
    double t, e;
    two_sum(s, so, t, e);
    s = t;
    c = (c + co) + e;
  }
}

template <int NPL>
__global__ __launch_bounds__(NP_THREADS) void residual_kernel(int64_t m, int64_t n, const double* __restrict__ A,
                                                              int64_t lda, const double* __restrict__ x,
                                                              const double* __restrict__ b, double* __restrict__ r,
                                                              double* __restrict__ part) {
  // part layout per workgroup: [n sums][n compensations][2: sum r^2, sum of its errors]
  __shared__ double red_s[NP_THREADS / 64][64 * NPL];
  __shared__ double red_c[NP_THREADS / 64][64 * NPL];
  __shared__ double red_r[NP_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t gw = (int64_t)gridDim.x * (NP_THREADS / 64);
  double xr[NPL], gs[NPL], gc[NPL];
#pragma unroll
  for (int q = 0; q < NPL; ++q) {
    const int64_t c = lane + 64 * q;
    xr[q] = (c < n) ? x[c] : 0.0;
    gs[q] = 0.0;
    gc[q] = 0.0;
  }
  double rr = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (NP_THREADS / 64) + wave; row < m; row += gw) {
    double a[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      a[q] = (c < n) ? A[row * lda + c] : 0.0;
    }
    double s = 0.0, cs = 0.0;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      double pq, ep, t, e;
      two_prod(a[q], xr[q], pq, ep);
      two_sum(s, pq, t, e);
      s = t;
      cs = add_nc(cs, add_nc(ep, e));
    }
    wave_sum2(s, cs);
    // r_i = b_i - (s + cs)
    double t, e;
    two_sum(b[row], -s, t, e);
    const double ri = add_nc(t, add_nc(e, -cs));
    if (lane == 0 && r != nullptr) r[row] = ri;
    rr += ri * ri;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      double pq, ep, u, f;
      two_prod(ri, a[q], pq, ep);
      two_sum(gs[q], pq, u, f);
      gs[q] = u;
      gc[q] = add_nc(gc[q], add_nc(ep, f));
    }
  }
#pragma unroll
  for (int q = 0; q < NPL; ++q) {
    red_s[wave][lane + 64 * q] = gs[q];
    red_c[wave][lane + 64 * q] = gc[q];
  }
  if (lane == 0) red_r[wave] = rr;
  __syncthreads();
  double* pw = part + (int64_t)blockIdx.x * (2 * n + 2);
  for (int64_t c = threadIdx.x; c < n; c += NP_THREADS) {
    double s = 0.0, cc = 0.0;
#pragma unroll
    for (int w = 0; w < NP_THREADS / 64; ++w) {
      double t, e;
      two_sum(s, red_s[w][c], t, e);
      s = t;
      cc = add_nc(cc, add_nc(red_c[w][c], e));
    }
    pw[c] = s;
    pw[n + c] = cc;
  }
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NP_THREADS / 64; ++w) t += red_r[w];
    pw[2 * n] = t;
    pw[2 * n + 1] = 0.0;
  }
}

// g[c] = sum_q (s, comp) pairs of part in a fixed order (as reduce_partials_kernel,
// pairs combined error-free); block n / 64 (one extra) also sums ||r||^2.
__global__ __launch_bounds__(64 * RED_WAVES) void residual_reduce_kernel(int64_t n, int64_t nparts,
                                                                         const double* __restrict__ part,
                                                                         double* __restrict__ g,
                                                                         double* __restrict__ rss) {
  __shared__ double rs[RED_WAVES][64];
  __shared__ double rc[RED_WAVES][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  const int64_t ps = 2 * n + 2;
  double s = 0.0, cc = 0.0;
  if (c < n) {
    for (int64_t q = wave; q < nparts; q += RED_WAVES) {
      double t, e;
      two_sum(s, part[q * ps + c], t, e);
      s = t;
      cc = add_nc(cc, add_nc(part[q * ps + n + c], e));
    }
  } else if (c == n && rss != nullptr) {
    for (int64_t q = wave; q < nparts; q += RED_WAVES) s += part[q * ps + 2 * n];
  }
  rs[wave][lane] = s;
  rc[wave][lane] = cc;
  __syncthreads();
  if (wave == 0) {
    double S = 0.0, C = 0.0;
    for (int w = 0; w < RED_WAVES; ++w) {
      double t, e;
      two_sum(S, rs[w][lane], t, e);
      S = t;
      C = add_nc(C, add_nc(rc[w][lane], e));
    }
    if (c < n) g[c] = add_nc(S, C);
    else if (c == n && rss != nullptr) rss[0] = S;
  }
}

// ---------------------------------------------------------------------------
// TRSM leaf: X (m x nb, nb <= LEAF) = Xin R^{-1}, R upper nb x nb; thread per
// row; with X32, the FP32 rounding of X is written too.  Xin may be X: each
// thread reads its whole row before writing it and touches no other row, so
// the pointers are restrict-qualified either way.
// ---------------------------------------------------------------------------
constexpr int LEAF = 32;

// FULL: nb == LEAF, rows moved straight between memory and registers; else the
// partial last leaf goes through an LDS copy (a register array may only be
// indexed with compile-time indices, and nb is a run-time value)
template <bool FULL, bool W32>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int64_t m, int nb, const double* __restrict__ R, int64_t ldr,
                                                        const double* __restrict__ Xin, int64_t ldxin,
                                                        double* __restrict__ X, int64_t ldx, float* __restrict__ X32,
                                                        int64_t ld32) {
  __shared__ double Rl[LEAF][LEAF + 1];
  __shared__ double rinv[LEAF];
  __shared__ double xs[FULL ? 1 : 256][LEAF + 1];
  for (int e = threadIdx.x; e < LEAF * LEAF; e += blockDim.x) {
    const int i = e / LEAF, j = e % LEAF;
    const double v = (i < nb && j < nb) ? R[(int64_t)i * ldr + j] : ((i == j) ? 1.0 : 0.0);
    Rl[i][j] = v;
    if (i == j) rinv[i] = 1.0 / v;
  }
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m) return;
  const double* xi = Xin + row * ldxin;
  double* xr = X + row * ldx;
  double x[LEAF];
  if constexpr (FULL) {
#pragma unroll
    for (int j = 0; j < LEAF; ++j) x[j] = xi[j];
  } else {
    for (int j = 0; j < LEAF; ++j) xs[threadIdx.x][j] = (j < nb) ? xi[j] : 0.0;
#pragma unroll
    for (int j = 0; j < LEAF; ++j) x[j] = xs[threadIdx.x][j];
  }
#pragma unroll
  for (int j = 0; j < LEAF; ++j) {
    double rowj[LEAF];
#pragma unroll
    for (int i = 0; i < LEAF; ++i) rowj[i] = (i > j) ? Rl[j][i] : 0.0;
    const double rj = rinv[j];
    __builtin_amdgcn_sched_barrier(0);  // the row's LDS reads in one batch
    const double y = x[j] * rj;
    x[j] = y;
#pragma unroll
    for (int i = 0; i < LEAF; ++i)
      if (i > j) x[i] -= rowj[i] * y;
  }
  // W32 a template argument: a run-time test here makes the compiler hoist the
  // whole block's R reads into registers and spill
  if constexpr (FULL) {
#pragma unroll
    for (int j = 0; j < LEAF; ++j) xr[j] = x[j];
    if constexpr (W32) {
#pragma unroll
      for (int j = 0; j < LEAF; ++j) X32[row * ld32 + j] = (float)x[j];
    }
  } else {
#pragma unroll
    for (int j = 0; j < LEAF; ++j) xs[threadIdx.x][j] = x[j];
    for (int j = 0; j < nb; ++j) xr[j] = xs[threadIdx.x][j];
    if constexpr (W32)
      for (int j = 0; j < nb; ++j) X32[row * ld32 + j] = (float)xs[threadIdx.x][j];
  }
}

// X = Xin R^{-1} (Xin may be X), recursively over column halves: X1 = Xin1
// R11^{-1}, X2 = (Xin2 - X1 R12) R22^{-1}.  Only the leftmost leaf of each
// level and the updates read Xin, so an out-of-place solve costs no copy of
// Xin.  With X32, every leaf also writes its columns' FP32 rounding.
int trsm_rec(int64_t m, int64_t n, const double* R, int64_t ldr, const double* Xin, int64_t ldxin, double* X,
             int64_t ldx, float* X32, int64_t ld32, void* ws, int64_t wsb, hipStream_t st) {
  if (n <= LEAF) {
    const unsigned g = (unsigned)pamd_cdiv(m, 256);
    const int nb = (int)n;
    if (n == LEAF && X32)
      trsm_leaf_kernel<true, true><<<g, 256, 0, st>>>(m, nb, R, ldr, Xin, ldxin, X, ldx, X32, ld32);
    else if (n == LEAF)
      trsm_leaf_kernel<true, false><<<g, 256, 0, st>>>(m, nb, R, ldr, Xin, ldxin, X, ldx, X32, ld32);
    else if (X32)
      trsm_leaf_kernel<false, true><<<g, 256, 0, st>>>(m, nb, R, ldr, Xin, ldxin, X, ldx, X32, ld32);
    else
      trsm_leaf_kernel<false, false><<<g, 256, 0, st>>>(m, nb, R, ldr, Xin, ldxin, X, ldx, X32, ld32);
    PAMD_LAUNCH_CHECK();
    return PAMD_OK;
  }
  const int64_t n1 = ((n / 2 + LEAF - 1) / LEAF) * LEAF;
  int rc = trsm_rec(m, n1, R, ldr, Xin, ldxin, X, ldx, X32, ld32, ws, wsb, st);
  if (rc) return rc;
  // X2 = Xin2 - X1 R12
  rc = pamd_gemm_cin_f64(0, 0, m, n - n1, n1, -1.0, X, ldx, R + n1, ldr, 1.0, Xin + n1, ldxin, X + n1, ldx, ws, wsb,
                         st);
  if (rc) return rc;
  return trsm_rec(m, n - n1, R + n1 * ldr + n1, ldr, X + n1, ldx, X + n1, ldx, X32 ? X32 + n1 : nullptr, ld32, ws,
                  wsb, st);
}

inline int64_t np_grid(int64_t m) { return imax64(1, imin64(NP_MAX_GRID, pamd_cdiv(m, 64))); }

// the normal pass's launch shape for rows of 64 NPL values (see normal_pass_kernel)
template <typename T, int NPL>
void np_launch(unsigned grid, bool deep, hipStream_t st, int64_t m, int64_t n, const T* B, int64_t ldb,
               const double* p, double* part) {
  if constexpr (NPL <= 4) {
    if (deep)
      normal_pass_kernel<T, NPL, 16, 4><<<grid, 1024, 0, st>>>(m, n, B, ldb, p, part);
    else
      normal_pass_kernel<T, NPL, 16, 1><<<grid, 1024, 0, st>>>(m, n, B, ldb, p, part);
  } else {
    normal_pass_kernel<T, NPL, 4, 1><<<grid, 256, 0, st>>>(m, n, B, ldb, p, part);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int pamd_ozaki_layout(int64_t d, int64_t m, int64_t ncols, int64_t* kp, int64_t* dpad, int64_t* ncp,
                                 int64_t* group_chunks, int64_t* workspace_bytes) {
  PAMD_ARG_CHECK(d >= 0, 1);
  PAMD_ARG_CHECK(m >= 0, 2);
  PAMD_ARG_CHECK(ncols >= 0, 3);
  PAMD_ARG_CHECK(kp && dpad && ncp && group_chunks && workspace_bytes, 4);
  const OzShape s = oz_shape(d, m, ncols);
  *kp = s.kp;
  *dpad = s.mpad;
  *ncp = s.ncp;
  *group_chunks = s.group_chunks;
  *workspace_bytes = (int64_t)s.ngroups * imax64(d, 1) * s.ncp * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_gauss_q_bf16(int64_t rows, int64_t rows_out, int64_t m, uint64_t seed, uint64_t offset,
                                 uint16_t* Q, void* stream) {
  PAMD_ARG_CHECK(rows >= 0, 1);
  PAMD_ARG_CHECK(rows_out >= rows && rows_out % 16 == 0, 2);
  PAMD_ARG_CHECK(m >= 0, 3);
  const int64_t kp = pamd_cdiv(imax64(m, 1), OZ_CHUNK) * OZ_CHUNK;
  PAMD_ARG_CHECK(Q != nullptr || rows_out == 0, 6);
  if (rows_out == 0) return PAMD_OK;
  const int64_t kbn = kp / 32, nblk = (rows_out / 16) * kbn;
  PAMD_ARG_CHECK(pamd_cdiv(nblk, 4) < (1LL << 31), 2);
  gauss_q_kernel<<<(unsigned)pamd_cdiv(nblk, 4), 256, 0, (hipStream_t)stream>>>(rows, m, nblk, kbn, seed, offset, Q);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_ozaki_split_bf16(int64_t m, int64_t n, const double* A, int64_t lda, const double* b,
                                     int npieces, uint16_t* pieces, int64_t pstride, int* exps, int64_t lde,
                                     void* stream) {
  const int64_t ncols = n + (b != nullptr ? 1 : 0);
  const int64_t kp = pamd_cdiv(imax64(m, 1), OZ_CHUNK) * OZ_CHUNK;
  const int64_t ncp = pamd_cdiv(imax64(ncols, 1), OZ_TN) * OZ_TN;
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0, 2);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 3);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 4);
  PAMD_ARG_CHECK(npieces >= 1 && npieces <= OZ_MAX_PIECES, 6);
  PAMD_ARG_CHECK(pieces != nullptr, 7);
  PAMD_ARG_CHECK(pstride >= ncp * kp && pstride % 8 == 0, 8);
  PAMD_ARG_CHECK(exps != nullptr, 9);
  PAMD_ARG_CHECK(lde >= ncols, 10);
  dim3 grid((unsigned)(kp / OZ_CHUNK), (unsigned)(ncp / 32));
  oz_split_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(m, n, A, lda, b, npieces, pieces, kp, pstride, exps, lde);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_ozaki_sketch_f64(int64_t d, int64_t m, int64_t ncols, int npieces, const uint16_t* Q,
                                     const uint16_t* pieces, int64_t pstride, const int* exps, int64_t lde,
                                     double scale, double* out, int64_t ldo, void* workspace, void* stream) {
  PAMD_ARG_CHECK(d >= 0, 1);
  PAMD_ARG_CHECK(m >= 1, 2);
  PAMD_ARG_CHECK(ncols >= 0, 3);
  PAMD_ARG_CHECK(npieces >= 1 && npieces <= OZ_MAX_PIECES, 4);
  if (d == 0 || ncols == 0) return PAMD_OK;
  const OzShape s = oz_shape(d, m, ncols);
  PAMD_ARG_CHECK(Q != nullptr, 5);
  PAMD_ARG_CHECK(pieces != nullptr, 6);
  PAMD_ARG_CHECK(pstride >= s.ncp * s.kp && pstride % 8 == 0, 7);
  PAMD_ARG_CHECK(exps != nullptr, 8);
  PAMD_ARG_CHECK(lde >= ncols, 9);
  PAMD_ARG_CHECK(out != nullptr, 11);
  PAMD_ARG_CHECK(ldo >= ncols && ldo >= 1, 12);
  PAMD_ARG_CHECK(workspace != nullptr, 13);
  const int64_t total = (int64_t)s.mtiles * s.ntiles * s.ngroups;
  PAMD_ARG_CHECK(total < (1LL << 31), 1);
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)workspace;
#define PAMD_OZ(P)                                                                                       \
  oz_gemm_kernel<P><<<(unsigned)total, OZ_THREADS, 0, st>>>(s.kp, s.group_chunks, s.mtiles, s.ntiles,    \
                                                            s.ngroups, Q, pieces, pstride, exps, lde, d,  \
                                                            ncols, s.ncp, part)
  switch (npieces) {
    case 1: PAMD_OZ(1); break;
    case 2: PAMD_OZ(2); break;
    case 3: PAMD_OZ(3); break;
    case 4: PAMD_OZ(4); break;
    default: PAMD_OZ(5); break;
  }
#undef PAMD_OZ
  PAMD_LAUNCH_CHECK();
  oz_reduce_kernel<<<(unsigned)pamd_cdiv(d * ncols, 256), 256, 0, st>>>(d, ncols, s.ncp, s.ngroups, part, scale,
                                                                       out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_normal_pass_workspace(int64_t m, int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= 1024, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 3);
  *bytes = np_grid(m) * n * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_normal_pass(int dtype, int64_t m, int64_t n, const void* B, int64_t ldb, const double* p,
                                double* q, void* workspace, void* stream) {
  PAMD_ARG_CHECK(dtype == 0 || dtype == 1, 1);
  PAMD_ARG_CHECK(m >= 0, 2);
  PAMD_ARG_CHECK(n >= 1 && n <= 1024, 3);
  PAMD_ARG_CHECK(B != nullptr || m == 0, 4);
  PAMD_ARG_CHECK(ldb >= n, 5);
  PAMD_ARG_CHECK(p != nullptr, 6);
  PAMD_ARG_CHECK(q != nullptr, 7);
  PAMD_ARG_CHECK(workspace != nullptr, 8);
  hipStream_t st = (hipStream_t)stream;
  const int npl = (int)pamd_cdiv(n, 64);
  // rows of at most 256 with the grid at its cap: 4 rows in flight per wavefront, a quarter of the workgroups
  const bool deep = npl <= 4 && np_grid(m) == NP_MAX_GRID;
  const int64_t grid = deep ? np_grid(m) / 4 : np_grid(m);
  double* part = (double*)workspace;
#define PAMD_NP(T, NPL) np_launch<T, NPL>((unsigned)grid, deep, st, m, n, (const T*)B, ldb, p, part)
#define PAMD_NP_T(T)                 \
  if (npl <= 1) PAMD_NP(T, 1);       \
  else if (npl <= 2) PAMD_NP(T, 2);  \
  else if (npl <= 4) PAMD_NP(T, 4);  \
  else if (npl <= 8) PAMD_NP(T, 8);  \
  else PAMD_NP(T, 16);
  if (dtype == 0) {
    PAMD_NP_T(double)
  } else {
    PAMD_NP_T(float)
  }
#undef PAMD_NP_T
#undef PAMD_NP
  PAMD_LAUNCH_CHECK();
  reduce_partials_kernel<<<(unsigned)pamd_cdiv(n, 64), 64 * RED_WAVES, 0, st>>>(n, grid, part, q);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_residual_workspace(int64_t m, int64_t n, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= 1024, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 3);
  *bytes = np_grid(m) * (2 * n + 2) * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_residual_f64(int64_t m, int64_t n, const double* A, int64_t lda, const double* x,
                                 const double* b, double* r, double* g, double* rss, void* workspace,
                                 void* stream) {
  PAMD_ARG_CHECK(m >= 1, 1);
  PAMD_ARG_CHECK(n >= 1 && n <= 1024, 2);
  PAMD_ARG_CHECK(A != nullptr, 3);
  PAMD_ARG_CHECK(lda >= n, 4);
  PAMD_ARG_CHECK(x != nullptr, 5);
  PAMD_ARG_CHECK(b != nullptr, 6);
  PAMD_ARG_CHECK(g != nullptr, 8);
  PAMD_ARG_CHECK(workspace != nullptr, 10);
  hipStream_t st = (hipStream_t)stream;
  const int64_t grid = np_grid(m);
  double* part = (double*)workspace;
  const int npl = (int)pamd_cdiv(n, 64);
#define PAMD_RK(NPL) residual_kernel<NPL><<<(unsigned)grid, NP_THREADS, 0, st>>>(m, n, A, lda, x, b, r, part)
  if (npl <= 1) PAMD_RK(1);
  else if (npl <= 2) PAMD_RK(2);
  else if (npl <= 4) PAMD_RK(4);
  else if (npl <= 8) PAMD_RK(8);
  else PAMD_RK(16);
#undef PAMD_RK
  PAMD_LAUNCH_CHECK();
  residual_reduce_kernel<<<(unsigned)pamd_cdiv(n + 1, 64), 64 * RED_WAVES, 0, st>>>(n, grid, part, g, rss);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_trsm_right_upper_rec_f64(int64_t m, int64_t n, const double* R, int64_t ldr, double* X,
                                             int64_t ldx, void* workspace, int64_t workspace_bytes, void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0, 2);
  PAMD_ARG_CHECK(R != nullptr || n == 0, 3);
  PAMD_ARG_CHECK(ldr >= n && ldr >= 1, 4);
  PAMD_ARG_CHECK(X != nullptr || m * n == 0, 5);
  PAMD_ARG_CHECK(ldx >= n && ldx >= 1, 6);
  if (m == 0 || n == 0) return PAMD_OK;
  return trsm_rec(m, n, R, ldr, X, ldx, X, ldx, nullptr, 0, workspace, workspace_bytes, (hipStream_t)stream);
}

extern "C" int pamd_trsm_right_upper_rec_f32out(int64_t m, int64_t n, const double* R, int64_t ldr, const double* A,
                                                int64_t lda, double* X, int64_t ldx, float* X32, int64_t ld32,
                                                void* workspace, int64_t workspace_bytes, void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0, 2);
  PAMD_ARG_CHECK(R != nullptr || n == 0, 3);
  PAMD_ARG_CHECK(ldr >= n && ldr >= 1, 4);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 5);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 6);
  PAMD_ARG_CHECK(X != nullptr || m * n == 0, 7);
  PAMD_ARG_CHECK(ldx >= n && ldx >= 1, 8);
  PAMD_ARG_CHECK(X32 != nullptr || m * n == 0, 9);
  PAMD_ARG_CHECK(ld32 >= n && ld32 >= 1, 10);
  if (m == 0 || n == 0) return PAMD_OK;
  return trsm_rec(m, n, R, ldr, A, lda, X, ldx, X32, ld32, workspace, workspace_bytes, (hipStream_t)stream);
}
