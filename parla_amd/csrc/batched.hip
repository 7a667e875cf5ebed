// Batched least squares over many independent problems of uneven size.
//
// Every problem i is described by a pamd_ls_desc (m_i, n_i <= 512 and
// offsets into shared pools): its FP32 matrix A_i (row-major, lda = n_i), its
// m-vectors (b_i, FP64) and n-vectors (FP64), and its upper-triangular
// preconditioner R_i (n_i x n_i FP64).  The kernels below each cover ALL
// problems in one launch:
//   * row-block passes (normal operator A^T A p, compensated residual): the
//     unit of work is a fixed 64-row block of one problem, handled by one
//     wavefront; each block writes its own partial vector and a second kernel
//     sums a problem's blocks in a fixed order.  The summation order therefore
//     does not depend on how blocks are assigned to workgroups, which is the
//     only thing the autotuned launch shape changes: results are bit-equal
//     for every tile choice.  Blocks are grouped by NPL = ceil(n_i / 64) so a
//     launch is specialised to its row length.
//   * per-problem kernels (triangular solves with R_i, the CG recurrences):
//     one workgroup per problem.
// Nothing here allocates or synchronises with the host, so a whole
// refinement step can be captured in one HIP graph and replayed.
#include "common.h"
#include "sjlt_dev.h"
#include "trsm_dev.h"
#include "trsv_dev.h"

namespace {

constexpr int BT_MAXN = 512;
constexpr int BT_BLOCK_ROWS = 64;

struct Desc {
  int64_t m, n;
  int64_t a_off;   // A pool element offset (FP32, lda = n)
  int64_t mv_off;  // m-vector pool offset
  int64_t nv_off;  // n-vector pool offset
  int64_t r_off;   // R pool offset (ld = n)
  int64_t blk0;    // index of the problem's first 64-row block
  int64_t part_off;  // partial pool offset: block b of this problem at part_off + (b - blk0) * pstride_i
};

// work item: one 64-row block
struct Blk {
  int32_t prob;
  int32_t row0;
};

// ---------------------------------------------------------------------------
// q_i = A_i^T (A_i p_i), partial per block
// ---------------------------------------------------------------------------
template <int NPL>
__global__ __launch_bounds__(256) void bt_normal_kernel(const Desc* __restrict__ D, const Blk* __restrict__ blks,
                                                        int64_t nblk, int blocks_per_wave,
                                                        const float* __restrict__ A, const double* __restrict__ pv,
                                                        const double* __restrict__ state, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int k = 0; k < blocks_per_wave; ++k) {
    const int64_t bi = wid * blocks_per_wave + k;
    if (bi >= nblk) return;
    const Blk bk = blks[bi];
    if (state != nullptr && state[4 * bk.prob + 1] == 0.0) continue;  // frozen problem: partial stays
    const Desc d = D[bk.prob];
    const int64_t n = d.n;
    const float* Ai = A + d.a_off;
    double pr[NPL], acc[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      pr[q] = (c < n) ? pv[d.nv_off + c] : 0.0;
      acc[q] = 0.0;
    }
    const int64_t r1 = imin64(d.m, (int64_t)bk.row0 + BT_BLOCK_ROWS);
    float a[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      a[q] = (c < n) ? Ai[(int64_t)bk.row0 * n + c] : 0.f;
    }
    for (int64_t row = bk.row0; row < r1; ++row) {
      // the next row's loads are in flight while this row is reduced
      float an[NPL];
      if (row + 1 < r1) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
          const int64_t c = lane + 64 * q;
          an[q] = (c < n) ? Ai[(row + 1) * n + c] : 0.f;
        }
      }
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NPL; ++q) s += (double)a[q] * pr[q];
      s = wave_sum(s);
#pragma unroll
      for (int q = 0; q < NPL; ++q) acc[q] += s * (double)a[q];
      if (row + 1 < r1) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) a[q] = an[q];
      }
    }
    double* pw = part + d.part_off + (int64_t)(bk.row0 / BT_BLOCK_ROWS) * (2 * n + 2);
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      if (c < n) pw[c] = acc[q];
    }
  }
}

// ---------------------------------------------------------------------------
// compensated residual r_i = b_i - A_i x_i, g_i = A_i^T r_i, ||r_i||^2
// (error-free transformations; FP contraction off where the terms are formed)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bt_two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}

__device__ __forceinline__ void bt_two_prod(double a, double b, double& p, double& e) {
#pragma clang fp contract(off)
  p = a * b;
  e = fma(a, b, -p);
}

__device__ __forceinline__ double bt_add(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

template <int NPL>
__global__ __launch_bounds__(256) void bt_residual_kernel(const Desc* __restrict__ D, const Blk* __restrict__ blks,
                                                          int64_t nblk, int blocks_per_wave,
                                                          const float* __restrict__ A, const double* __restrict__ b,
                                                          const double* __restrict__ xv,
                                                          const double* __restrict__ state, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int k = 0; k < blocks_per_wave; ++k) {
    const int64_t bi = wid * blocks_per_wave + k;
    if (bi >= nblk) return;
    const Blk bk = blks[bi];
    if (state != nullptr && state[4 * bk.prob + 1] == 0.0) continue;  // frozen problem: partial stays
    const Desc d = D[bk.prob];
    const int64_t n = d.n;
    const float* Ai = A + d.a_off;
    double xr[NPL], gs[NPL], gc[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      xr[q] = (c < n) ? xv[d.nv_off + c] : 0.0;
      gs[q] = 0.0;
      gc[q] = 0.0;
    }
    double rr = 0.0;
    const int64_t r1 = imin64(d.m, (int64_t)bk.row0 + BT_BLOCK_ROWS);
    float af[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      af[q] = (c < n) ? Ai[(int64_t)bk.row0 * n + c] : 0.f;
    }
    for (int64_t row = bk.row0; row < r1; ++row) {
      double a[NPL];
#pragma unroll
      for (int q = 0; q < NPL; ++q) a[q] = (double)af[q];
      if (row + 1 < r1) {  // prefetch the next row
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
          const int64_t c = lane + 64 * q;
          af[q] = (c < n) ? Ai[(row + 1) * n + c] : 0.f;
        }
      }
      double s = 0.0, cs = 0.0;
#pragma unroll
      for (int q = 0; q < NPL; ++q) {
        double p, ep, t, e;
        bt_two_prod(a[q], xr[q], p, ep);
        bt_two_sum(s, p, t, e);
        s = t;
        cs = bt_add(cs, bt_add(ep, e));
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double so = __shfl_xor(s, off, 64);
        const double co = __shfl_xor(cs, off, 64);
        double t, e;
        bt_two_sum(s, so, t, e);
        s = t;
        cs = bt_add(bt_add(cs, co), e);
      }
      double t, e;
      bt_two_sum(b[d.mv_off + row], -s, t, e);
      const double ri = bt_add(t, bt_add(e, -cs));
      rr += ri * ri;
#pragma unroll
      for (int q = 0; q < NPL; ++q) {
        double p, ep, u, f;
        bt_two_prod(ri, a[q], p, ep);
        bt_two_sum(gs[q], p, u, f);
        gs[q] = u;
        gc[q] = bt_add(gc[q], bt_add(ep, f));
      }
    }
    double* pw = part + d.part_off + (int64_t)(bk.row0 / BT_BLOCK_ROWS) * (2 * n + 2);
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      if (c < n) {
        pw[c] = gs[q];
        pw[n + c] = gc[q];
      }
    }
    if (lane == 0) pw[2 * n] = rr;
  }
}

// sum a problem's block partials in a fixed order: one workgroup per problem,
// 64-column chunks, its 16 wavefronts take the blocks bb = w, w + 16, ... and
// their sums are combined in wavefront order; out (n-vector pool) and, for
// the residual, the compensation terms folded in and ||r_i||^2 in scal[prob]
constexpr int BT_RED_WAVES = 16;

__global__ __launch_bounds__(64 * BT_RED_WAVES) void bt_reduce_kernel(const Desc* __restrict__ D,
                                                                      const double* __restrict__ part,
                                                                      int compensated, double* __restrict__ out,
                                                                      double* __restrict__ scal) {
  __shared__ double rs[BT_RED_WAVES][64];
  __shared__ double rc[BT_RED_WAVES][64];
  const Desc d = D[blockIdx.x];
  const int64_t n = d.n;
  const int64_t nb = pamd_cdiv(d.m, BT_BLOCK_ROWS);
  const double* p0 = part + d.part_off;
  const int64_t ps = 2 * n + 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t cb = 0; cb <= n; cb += 64) {  // the chunk holding column n carries ||r||^2
    const int64_t c = cb + lane;
    double s = 0.0, cc = 0.0;
    if (c < n) {
      for (int64_t bb = wave; bb < nb; bb += BT_RED_WAVES) {
        if (compensated) {
          double t, e;
          bt_two_sum(s, p0[bb * ps + c], t, e);
          s = t;
          cc = bt_add(cc, bt_add(p0[bb * ps + n + c], e));
        } else {
          s += p0[bb * ps + c];
        }
      }
    } else if (c == n && compensated) {
      for (int64_t bb = wave; bb < nb; bb += BT_RED_WAVES) s += p0[bb * ps + 2 * n];
    }
    rs[wave][lane] = s;
    rc[wave][lane] = cc;
    __syncthreads();
    if (wave == 0) {
      double S = 0.0, C = 0.0;
      for (int w = 0; w < BT_RED_WAVES; ++w) {
        if (compensated && c < n) {
          double t, e;
          bt_two_sum(S, rs[w][lane], t, e);
          S = t;
          C = bt_add(C, bt_add(rc[w][lane], e));
        } else {
          S += rs[w][lane];
        }
      }
      if (c < n) out[d.nv_off + c] = compensated ? bt_add(S, C) : S;
      else if (c == n && compensated) scal[blockIdx.x] = S;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// per-problem kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bt_trsv_kernel(const Desc* __restrict__ D, int trans,
                                                      const double* __restrict__ Rp, double* __restrict__ v) {
  extern __shared__ double smem[];
  const Desc d = D[blockIdx.x];
  const int n = (int)d.n;
  double* xs = smem;
  for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = v[d.nv_off + i];
  __syncthreads();
  trsv_upper_block<8>(trans, n, Rp + d.r_off, n, xs, smem + BT_MAXN);
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[d.nv_off + i] = xs[i];
}

// v_i <- M_i^T v_i, M_i = R_i^{-1} (upper triangular) or its transpose
// (lower), row-major with ld n_i, so that both R^{-1} v and R^{-T} v are a
// product with a transposed matrix: one workgroup of four wavefronts per
// problem, each lane owning eight columns, each wavefront every fourth row
// (two rows' coalesced segments loaded at a time), the wavefronts' column
// sums added in wavefront order.
__global__ __launch_bounds__(256) void bt_trmv_kernel(const Desc* __restrict__ D, int lower,
                                                      const double* __restrict__ Mp, double* __restrict__ v) {
  __shared__ double xs[BT_MAXN];
  __shared__ double red[4][BT_MAXN];
  const Desc d = D[blockIdx.x];
  const int n = (int)d.n;
  const double* M = Mp + d.r_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < n; i += blockDim.x) xs[i] = v[d.nv_off + i];
  __syncthreads();
  double acc[BT_MAXN / 64];
#pragma unroll
  for (int k = 0; k < BT_MAXN / 64; ++k) acc[k] = 0.0;
  for (int i0 = wave; i0 < n; i0 += 8) {
    double mv[2][BT_MAXN / 64];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = i0 + 4 * h;
      const double* row = M + (int64_t)(i < n ? i : 0) * n;
#pragma unroll
      for (int k = 0; k < BT_MAXN / 64; ++k) {
        const int j = lane + 64 * k;
        const bool in = i < n && j < n && (lower ? (j <= i) : (j >= i));
        mv[h][k] = in ? row[j] : 0.0;
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = i0 + 4 * h;
      const double xi = (i < n) ? xs[i] : 0.0;
#pragma unroll
      for (int k = 0; k < BT_MAXN / 64; ++k) acc[k] += mv[h][k] * xi;
    }
  }
#pragma unroll
  for (int k = 0; k < BT_MAXN / 64; ++k) red[wave][lane + 64 * k] = acc[k];
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    v[d.nv_off + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
}

__device__ inline double bt_block_dot(const double* a, const double* b, int64_t n, double* red) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += a[i] * b[i];
  return block_sum(s, red);
}

// CG state per problem: st[4 * i + 0] = rr, +1 = active flag (1.0 / 0.0)
// mode 0 (start): r = rhs, p = r, z = 0, rr = r.r
// mode 1 (step, q = M p given): alpha = rr / p.q; z += alpha p; r -= alpha q;
//                               rr' = r.r; p = r + (rr' / rr) p; rr = rr'
__global__ __launch_bounds__(256) void bt_cg_kernel(const Desc* __restrict__ D, int mode, double* __restrict__ st,
                                                    const double* __restrict__ rhs, double* __restrict__ r,
                                                    double* __restrict__ p, const double* __restrict__ q,
                                                    double* __restrict__ z) {
  __shared__ double red[16];
  __shared__ double sh[2];
  const Desc d = D[blockIdx.x];
  const int64_t n = d.n, o = d.nv_off;
  double* S = st + 4 * blockIdx.x;
  if (S[1] == 0.0) return;  // problem converged: its vectors stay as they are
  if (mode == 0) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      r[o + i] = rhs[o + i];
      p[o + i] = rhs[o + i];
      z[o + i] = 0.0;
    }
    const double rr = bt_block_dot(rhs + o, rhs + o, n, red);
    if (threadIdx.x == 0) S[0] = rr;
    return;
  }
  const double pq = bt_block_dot(p + o, q + o, n, red);
  if (threadIdx.x == 0) sh[0] = (pq > 0.0) ? S[0] / pq : 0.0;
  __syncthreads();
  const double alpha = sh[0];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    z[o + i] += alpha * p[o + i];
    r[o + i] -= alpha * q[o + i];
  }
  __syncthreads();
  const double rr_new = bt_block_dot(r + o, r + o, n, red);
  if (threadIdx.x == 0) sh[1] = (S[0] > 0.0) ? rr_new / S[0] : 0.0;
  __syncthreads();
  const double beta = sh[1];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[o + i] = r[o + i] + beta * p[o + i];
  if (threadIdx.x == 0) S[0] = rr_new;
}

// x += z (active problems only)
__global__ __launch_bounds__(256) void bt_axpy_kernel(const Desc* __restrict__ D, const double* __restrict__ st,
                                                      double* __restrict__ x, const double* __restrict__ z) {
  const Desc d = D[blockIdx.x];
  if (st[4 * blockIdx.x + 1] == 0.0) return;
  for (int64_t i = threadIdx.x; i < d.n; i += blockDim.x) x[d.nv_off + i] += z[d.nv_off + i];
}

// v = a (n-vector copy)
__global__ __launch_bounds__(256) void bt_copy_kernel(const Desc* __restrict__ D, const double* __restrict__ a,
                                                      double* __restrict__ v) {
  const Desc d = D[blockIdx.x];
  for (int64_t i = threadIdx.x; i < d.n; i += blockDim.x) v[d.nv_off + i] = a[d.nv_off + i];
}


// ---------------------------------------------------------------------------
// Batched sketch and factorisation.  Y_i = S_i A_i with a per-problem SJLT
// (the same column draws as the single-problem sketch), then the shifted
// CholeskyQR3 of Y_i:  G = Y^T Y (+ shift I in the first round),
// G = R_k^T R_k, Y <- Y R_k^{-1}, and R = R_2 R_1 R_0.  Every kernel covers
// all problems; its work items (column chunks, sketch rows, 64 x 64 upper
// tiles of n_i x n_i, 16-row strips of Y_i, 16 x 16 diagonal blocks of R_i)
// are numbered problem after problem, and a work item finds its problem by a
// binary search over the problems' first indices in the sketch descriptors.
// ---------------------------------------------------------------------------
struct SkDesc {
  int64_t d;        // sketch rows
  int64_t y_off;    // FP64 pool offset of Y_i (d x n, ld n)
  int64_t rp_off;   // rowptr pool offset (d + 1 entries)
  int64_t cnt_off;  // chunk histogram pool offset (nchunk x d int32)
  int64_t chunk0, row0, tile0, strip0, dblk0;  // first work item of each kind
  uint64_t seed;
};

constexpr int SK_CHUNK = 2048;  // columns of S_i (rows of A_i) per chunk
constexpr int SK_MAXD = 4096;
constexpr int SK_KMAX = 16;
constexpr int GT = 64;          // Gram / product tile

// the problem p with first(p) <= idx < first(p + 1)
template <typename F>
__device__ __forceinline__ int64_t bt_find(int64_t nprob, int64_t idx, F first) {
  int64_t lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (first(mid) <= idx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// upper tile t (row-major over ti <= tj) of a T x T tile grid
__device__ __forceinline__ void bt_upper_tile(int64_t t, int T, int& ti, int& tj) {
  ti = 0;
  while (t >= T - ti) {
    t -= T - ti;
    ++ti;
  }
  tj = ti + (int)t;
}

// row histograms of one column chunk
__global__ __launch_bounds__(256) void bt_sk_count_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                          int64_t nprob, int k, int32_t* __restrict__ cnt) {
  __shared__ int32_t hist[SK_MAXD];
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].chunk0; });
  const int64_t m = D[p].m;
  const SkDesc s = S[p];
  const int64_t c = blockIdx.x - s.chunk0;
  for (int r = threadIdx.x; r < s.d; r += blockDim.x) hist[r] = 0;
  __syncthreads();
  const int64_t j1 = imin64(m, (c + 1) * SK_CHUNK);
  for (int64_t j = c * SK_CHUNK + threadIdx.x; j < j1; j += blockDim.x) {
    int32_t rows[SK_KMAX];
    uint64_t signs;
    sjlt_column<SK_KMAX>(s.d, k, s.seed, j, rows, signs);
#pragma unroll
    for (int l = 0; l < SK_KMAX; ++l)
      if (l < k) atomicAdd(&hist[rows[l]], 1);
  }
  __syncthreads();
  int32_t* out = cnt + s.cnt_off + c * s.d;
  for (int r = threadIdx.x; r < s.d; r += blockDim.x) out[r] = hist[r];
}

// rowptr_i and the chunks' first slots, one workgroup per problem
__global__ __launch_bounds__(1024) void bt_sk_scan_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                          int32_t* __restrict__ cnt, int64_t* __restrict__ rowptr) {
  const SkDesc s = S[blockIdx.x];
  const int nchunk = (int)pamd_cdiv(D[blockIdx.x].m, SK_CHUNK);
  csr_scan_body(s.d, nchunk, cnt + s.cnt_off, rowptr + s.rp_off);
}

// CSR entries of one chunk in column order: the wavefront draws 64 columns
// at once (a lane each) into LDS, then places them one column at a time
// (the k rows of a column are distinct, so its lanes never share a row).
__global__ __launch_bounds__(64) void bt_sk_fill_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                        int64_t nprob, int k, const int32_t* __restrict__ cnt,
                                                        int32_t* __restrict__ cols, double* __restrict__ vals) {
  __shared__ int32_t cur[SK_MAXD];
  __shared__ int32_t rl[64][SK_KMAX + 1];
  __shared__ uint32_t sl[64];
  const int lane = threadIdx.x;
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].chunk0; });
  const Desc d = D[p];
  const SkDesc s = S[p];
  const int64_t c = blockIdx.x - s.chunk0;
  const int32_t* off = cnt + s.cnt_off + c * s.d;
  for (int r = lane; r < s.d; r += 64) cur[r] = off[r];
  __syncthreads();
  const int64_t base = (int64_t)k * d.mv_off;
  const double mag = 1.0 / sqrt((double)k);
  const bool distinct = s.d >= k;
  const int64_t j0 = c * SK_CHUNK, j1 = imin64(d.m, j0 + SK_CHUNK);
  for (int64_t jb = j0; jb < j1; jb += 64) {
    if (jb + lane < j1) {
      int32_t rows[SK_KMAX];
      uint64_t signs;
      sjlt_column<SK_KMAX>(s.d, k, s.seed, jb + lane, rows, signs);
#pragma unroll
      for (int l = 0; l < SK_KMAX; ++l)
        if (l < k) rl[lane][l] = rows[l];
      sl[lane] = (uint32_t)signs;
    }
    PAMD_WAVE_SYNC();
    const int nq = (int)imin64(64, j1 - jb);
    for (int q = 0; q < nq; ++q) {
      if (distinct) {
        if (lane < k) {
          const int32_t r = rl[q][lane];
          const int32_t pos = cur[r];
          cur[r] = pos + 1;
          cols[base + pos] = (int32_t)(jb + q);
          vals[base + pos] = ((sl[q] >> lane) & 1u) ? -mag : mag;
        }
      } else if (lane == 0) {
        for (int l = 0; l < k; ++l) {
          const int32_t r = rl[q][l];
          const int32_t pos = cur[r]++;
          cols[base + pos] = (int32_t)(jb + q);
          vals[base + pos] = ((sl[q] >> l) & 1u) ? -mag : mag;
        }
      }
      PAMD_WAVE_SYNC();
    }
  }
}

// Y_i[r, :] = sum over row r of S_i of v * A_i[j, :] (FP32 A, FP64 sums in
// column order, as the single-problem left product)
__global__ __launch_bounds__(256) void bt_sk_left_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                         int64_t nprob, int k, const int64_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ cols,
                                                         const double* __restrict__ vals, const float* __restrict__ A,
                                                         double* __restrict__ Y) {
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].row0; });
  const Desc d = D[p];
  const SkDesc s = S[p];
  const int64_t r = blockIdx.x - s.row0;
  const int64_t base = (int64_t)k * d.mv_off;
  const int64_t p0 = base + rowptr[s.rp_off + r], p1 = base + rowptr[s.rp_off + r + 1];
  const int64_t n = d.n;
  const float* Ai = A + d.a_off;
  double* y = Y + s.y_off + r * n;
  for (int64_t c = threadIdx.x; c < n; c += blockDim.x) {
    double acc = 0.0;
    int64_t q = p0;
    for (; q + 4 <= p1; q += 4) {
      const double a0 = (double)Ai[(int64_t)cols[q] * n + c];
      const double a1 = (double)Ai[(int64_t)cols[q + 1] * n + c];
      const double a2 = (double)Ai[(int64_t)cols[q + 2] * n + c];
      const double a3 = (double)Ai[(int64_t)cols[q + 3] * n + c];
      acc += vals[q] * a0;
      acc += vals[q + 1] * a1;
      acc += vals[q + 2] * a2;
      acc += vals[q + 3] * a3;
    }
    for (; q < p1; ++q) acc += vals[q] * (double)Ai[(int64_t)cols[q] * n + c];
    y[c] = acc;
  }
}

// G_i = Y_i^T Y_i on the upper 64 x 64 tiles (diagonal tiles whole), FP64
// MFMA: four wavefronts of 32 x 32, K staged through LDS 16 rows at a time
__global__ __launch_bounds__(256) void bt_gram_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                      int64_t nprob, const double* __restrict__ Y,
                                                      double* __restrict__ G) {
  __shared__ double Ya[16][GT + 1], Yb[16][GT + 1];
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].tile0; });
  const int64_t n = D[p].n;
  const SkDesc s = S[p];
  int ti, tj;
  bt_upper_tile(blockIdx.x - s.tile0, (int)pamd_cdiv(n, GT), ti, tj);
  const double* y = Y + s.y_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1, lr = lane & 15, lk = lane >> 4;
  const int64_t ci = (int64_t)ti * GT, cj = (int64_t)tj * GT;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = 0; k0 < s.d; k0 += 16) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int e = tid + 256 * h, kr = e >> 6, c = e & 63;
      const int64_t row = k0 + kr;
      Ya[kr][c] = (row < s.d && ci + c < n) ? y[row * n + ci + c] : 0.0;
      Yb[kr][c] = (row < s.d && cj + c < n) ? y[row * n + cj + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = Ya[kk + lk][32 * wm + 16 * t + lr];
        b[t] = Yb[kk + lk][32 * wn + 16 * t + lr];
      }
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
    __syncthreads();
  }
  double* g = G + D[p].r_off;
  // D layout of v_mfma_f64_16x16x4f64: lane l, item r -> row l / 16 + 4 r, column l % 16
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = ci + 32 * wm + 16 * mt + lk + 4 * r;
        const int64_t j = cj + 32 * wn + 16 * nt + lr;
        if (i < n && j < n) g[i * n + j] = acc[mt][nt][r];
      }
}

// G_i += 11 (d n + n (n + 1)) u trace(G_i) I: the shift of shifted CholeskyQR
// (trace(G) = ||Y||_F^2 bounds ||Y||_2^2)
__global__ __launch_bounds__(256) void bt_shift_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                       double* __restrict__ G) {
  __shared__ double red[4];
  const int64_t n = D[blockIdx.x].n, dd = S[blockIdx.x].d;
  double* g = G + D[blockIdx.x].r_off;
  double t = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) t += g[i * n + i];
  t = block_sum(t, red);
  const double shift = 11.0 * ((double)dd * (double)n + (double)n * (double)(n + 1)) * 0x1p-53 * t;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) g[i * n + i] += shift;
}

__global__ __launch_bounds__(64) void bt_dinv_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                     int64_t nprob, const double* __restrict__ R,
                                                     double* __restrict__ Dinv) {
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].dblk0; });
  const int64_t n = D[p].n;
  trsm_dinv_body((int)(blockIdx.x - S[p].dblk0), (int)n, R + D[p].r_off, n, Dinv + S[p].dblk0 * TS_B * TS_B);
}

// Y_i <- Y_i R_i^{-1}, one 16-row strip per workgroup
__global__ __launch_bounds__(256) void bt_trsm_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                      int64_t nprob, const double* __restrict__ R,
                                                      const double* __restrict__ Dinv, double* __restrict__ Y) {
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].strip0; });
  const int64_t n = D[p].n;
  const SkDesc s = S[p];
  double* y = Y + s.y_off;
  trsm_strip_body(blockIdx.x - s.strip0, s.d, (int)n, y, n, R + D[p].r_off, n, Dinv + s.dblk0 * TS_B * TS_B, y, n);
}

// C_i = A_i B_i for upper-triangular A_i, B_i (entries below their diagonals
// are ignored); C's upper tiles are written with zeros below the diagonal.
// Thread (ty, tx) owns a 4 x 4 block of the 64 x 64 tile; FP64 FMAs in
// ascending k.
__global__ __launch_bounds__(256) void bt_trmm_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                      int64_t nprob, const double* __restrict__ Ap,
                                                      const double* __restrict__ Bp, double* __restrict__ Cp) {
  __shared__ double As[GT][17], Bs[16][GT + 1];
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].tile0; });
  const int64_t n = D[p].n;
  int ti, tj;
  bt_upper_tile(blockIdx.x - S[p].tile0, (int)pamd_cdiv(n, GT), ti, tj);
  const int64_t ro = D[p].r_off;
  const double* a = Ap + ro;
  const double* b = Bp + ro;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int64_t i0 = (int64_t)ti * GT, j0 = (int64_t)tj * GT;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  const int64_t kend = imin64(n, j0 + GT);
  for (int64_t k0 = i0; k0 < kend; k0 += 16) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int e = tid + 256 * h;
      {  // A tile 64 x 16: e -> (row e / 16, k e % 16)
        const int64_t i = i0 + (e >> 4), kk = k0 + (e & 15);
        As[e >> 4][e & 15] = (i < n && kk < n && kk >= i) ? a[i * n + kk] : 0.0;
      }
      {  // B tile 16 x 64: e -> (k e / 64, column e % 64)
        const int64_t kk = k0 + (e >> 6), j = j0 + (e & 63);
        Bs[e >> 6][e & 63] = (kk < n && j < n && kk <= j) ? b[kk * n + j] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) av[u] = As[4 * ty + u][kk];
#pragma unroll
      for (int v = 0; v < 4; ++v) bv[v] = Bs[kk][4 * tx + v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
    }
    __syncthreads();
  }
  double* c = Cp + ro;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t i = i0 + 4 * ty + u, j = j0 + 4 * tx + v;
      if (i < n && j < n) c[i * n + j] = (i <= j) ? acc[u][v] : 0.0;
    }
}

// ||A_i||_F^2: a partial per column chunk (rows of A_i), summed in chunk order
__global__ __launch_bounds__(256) void bt_fro2_partial_kernel(const Desc* __restrict__ D,
                                                              const SkDesc* __restrict__ S, int64_t nprob,
                                                              const float* __restrict__ A,
                                                              double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t p = bt_find(nprob, blockIdx.x, [&](int64_t q) { return S[q].chunk0; });
  const Desc d = D[p];
  const int64_t c = blockIdx.x - S[p].chunk0;
  const int64_t e0 = c * SK_CHUNK * d.n, e1 = imin64(d.m, (c + 1) * SK_CHUNK) * d.n;
  const float* Ai = A + d.a_off;
  double t = 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const double v = (double)Ai[e];
    t += v * v;
  }
  t = block_sum(t, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(64) void bt_fro2_final_kernel(const Desc* __restrict__ D, const SkDesc* __restrict__ S,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const int64_t nc = pamd_cdiv(D[blockIdx.x].m, SK_CHUNK), c0 = S[blockIdx.x].chunk0;
  double t = 0.0;
  for (int64_t c = 0; c < nc; ++c) t += part[c0 + c];
  out[blockIdx.x] = t;
}

// out_i = a_i . b_i over the n-vector pool
__global__ __launch_bounds__(256) void bt_dot_kernel(const Desc* __restrict__ D, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ out) {
  __shared__ double red[4];
  const Desc d = D[blockIdx.x];
  const double t = bt_block_dot(a + d.nv_off, b + d.nv_off, d.n, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static_assert(sizeof(Desc) == sizeof(pamd_ls_desc), "descriptor layout");
static_assert(sizeof(Blk) == sizeof(pamd_ls_block), "block layout");

#define PAMD_BT_NPL(KERNEL, NPL, ...)                                          \
  switch (NPL) {                                                               \
    case 1: KERNEL<1><<<grid, threads, 0, st>>>(__VA_ARGS__); break;           \
    case 2: KERNEL<2><<<grid, threads, 0, st>>>(__VA_ARGS__); break;           \
    case 4: KERNEL<4><<<grid, threads, 0, st>>>(__VA_ARGS__); break;           \
    default: KERNEL<8><<<grid, threads, 0, st>>>(__VA_ARGS__); break;          \
  }

static int bt_shape(int npl, int64_t nblk, int waves_per_wg, int blocks_per_wave, unsigned* grid, unsigned* threads) {
  if (!(npl == 1 || npl == 2 || npl == 4 || npl == 8)) return -1;
  if (!(waves_per_wg == 1 || waves_per_wg == 2 || waves_per_wg == 4)) return -2;
  if (blocks_per_wave < 1 || blocks_per_wave > 64) return -3;
  *threads = 64u * (unsigned)waves_per_wg;
  *grid = (unsigned)imax64(1, pamd_cdiv(nblk, (int64_t)waves_per_wg * blocks_per_wave));
  return 0;
}

extern "C" int pamd_bt_normal(int npl, const pamd_ls_desc* descs, const pamd_ls_block* blocks, int64_t nblk,
                              int waves_per_wg, int blocks_per_wave, const float* A, const double* p,
                              const double* state, double* part, void* stream) {
  unsigned grid, threads;
  const int rc = bt_shape(npl, nblk, waves_per_wg, blocks_per_wave, &grid, &threads);
  PAMD_ARG_CHECK(rc != -1, 1);
  PAMD_ARG_CHECK(descs != nullptr, 2);
  PAMD_ARG_CHECK(blocks != nullptr || nblk == 0, 3);
  PAMD_ARG_CHECK(nblk >= 0, 4);
  PAMD_ARG_CHECK(rc != -2, 5);
  PAMD_ARG_CHECK(rc != -3, 6);
  if (nblk == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  PAMD_BT_NPL(bt_normal_kernel, npl, (const Desc*)descs, (const Blk*)blocks, nblk, blocks_per_wave, A, p, state,
              part)
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_residual(int npl, const pamd_ls_desc* descs, const pamd_ls_block* blocks, int64_t nblk,
                                int waves_per_wg, int blocks_per_wave, const float* A, const double* b,
                                const double* x, const double* state, double* part, void* stream) {
  unsigned grid, threads;
  const int rc = bt_shape(npl, nblk, waves_per_wg, blocks_per_wave, &grid, &threads);
  PAMD_ARG_CHECK(rc != -1, 1);
  PAMD_ARG_CHECK(descs != nullptr, 2);
  PAMD_ARG_CHECK(blocks != nullptr || nblk == 0, 3);
  PAMD_ARG_CHECK(nblk >= 0, 4);
  PAMD_ARG_CHECK(rc != -2, 5);
  PAMD_ARG_CHECK(rc != -3, 6);
  if (nblk == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  PAMD_BT_NPL(bt_residual_kernel, npl, (const Desc*)descs, (const Blk*)blocks, nblk, blocks_per_wave, A, b, x,
              state, part)
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_reduce(int64_t nprob, const pamd_ls_desc* descs, const double* part, int compensated,
                              double* out, double* scal, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(part != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(compensated == 0 || compensated == 1, 4);
  PAMD_ARG_CHECK(out != nullptr || nprob == 0, 5);
  PAMD_ARG_CHECK(scal != nullptr || !compensated || nprob == 0, 6);
  if (nprob == 0) return PAMD_OK;
  bt_reduce_kernel<<<(unsigned)nprob, 64 * BT_RED_WAVES, 0, (hipStream_t)stream>>>((const Desc*)descs, part,
                                                                                  compensated, out, scal);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_trsv(int64_t nprob, const pamd_ls_desc* descs, int trans, const double* R, double* v,
                            void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(trans == 0 || trans == 1, 3);
  PAMD_ARG_CHECK(R != nullptr || nprob == 0, 4);
  PAMD_ARG_CHECK(v != nullptr || nprob == 0, 5);
  if (nprob == 0) return PAMD_OK;
  const size_t shmem = (size_t)(BT_MAXN + TRSV_SCRATCH_DOUBLES) * sizeof(double);
  bt_trsv_kernel<<<(unsigned)nprob, 256, shmem, (hipStream_t)stream>>>((const Desc*)descs, trans, R, v);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_trmv(int64_t nprob, const pamd_ls_desc* descs, int trans, const double* Rinv,
                            const double* RinvT, double* v, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(trans == 0 || trans == 1, 3);
  PAMD_ARG_CHECK(Rinv != nullptr || nprob == 0, 4);
  PAMD_ARG_CHECK(RinvT != nullptr || nprob == 0, 5);
  PAMD_ARG_CHECK(v != nullptr || nprob == 0, 6);
  if (nprob == 0) return PAMD_OK;
  // R^{-T} v = Rinv^T v; R^{-1} v = (Rinv^T)^T v
  bt_trmv_kernel<<<(unsigned)nprob, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, trans ? 0 : 1,
                                                                   trans ? Rinv : RinvT, v);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_cg(int64_t nprob, const pamd_ls_desc* descs, int mode, double* state, const double* rhs,
                          double* r, double* p, const double* q, double* z, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(mode == 0 || mode == 1, 3);
  PAMD_ARG_CHECK(state != nullptr || nprob == 0, 4);
  if (nprob == 0) return PAMD_OK;
  bt_cg_kernel<<<(unsigned)nprob, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, mode, state, rhs, r, p, q, z);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_axpy(int64_t nprob, const pamd_ls_desc* descs, const double* state, double* x,
                            const double* z, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  if (nprob == 0) return PAMD_OK;
  bt_axpy_kernel<<<(unsigned)nprob, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, state, x, z);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_copy(int64_t nprob, const pamd_ls_desc* descs, const double* a, double* v, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  if (nprob == 0) return PAMD_OK;
  bt_copy_kernel<<<(unsigned)nprob, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, a, v);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

// ---------------------------------------------------------------------------
// C ABI: batched sketch and factorisation
// ---------------------------------------------------------------------------
static_assert(sizeof(SkDesc) == sizeof(pamd_sk_desc), "sketch descriptor layout");

extern "C" int pamd_bt_layout(int* chunk_rows, int* tile, int* strip_rows, int* max_d, int* max_vec_nnz) {
  PAMD_ARG_CHECK(chunk_rows != nullptr, 1);
  PAMD_ARG_CHECK(tile != nullptr, 2);
  PAMD_ARG_CHECK(strip_rows != nullptr, 3);
  PAMD_ARG_CHECK(max_d != nullptr, 4);
  PAMD_ARG_CHECK(max_vec_nnz != nullptr, 5);
  *chunk_rows = SK_CHUNK;
  *tile = GT;
  *strip_rows = TS_B;
  *max_d = SK_MAXD;
  *max_vec_nnz = SK_KMAX;
  return PAMD_OK;
}

extern "C" int pamd_bt_sjlt_csr(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t nchunk,
                                int k, int32_t* cnt, int64_t* rowptr, int32_t* cols, double* vals, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(nchunk >= nprob && nchunk < (1LL << 31), 4);
  PAMD_ARG_CHECK(k >= 1 && k <= SK_KMAX, 5);
  PAMD_ARG_CHECK(cnt != nullptr || nprob == 0, 6);
  PAMD_ARG_CHECK(rowptr != nullptr || nprob == 0, 7);
  PAMD_ARG_CHECK(cols != nullptr || nprob == 0, 8);
  PAMD_ARG_CHECK(vals != nullptr || nprob == 0, 9);
  if (nprob == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const Desc* D = (const Desc*)descs;
  const SkDesc* S = (const SkDesc*)sk;
  bt_sk_count_kernel<<<(unsigned)nchunk, 256, 0, st>>>(D, S, nprob, k, cnt);
  PAMD_LAUNCH_CHECK();
  bt_sk_scan_kernel<<<(unsigned)nprob, 1024, 0, st>>>(D, S, cnt, rowptr);
  PAMD_LAUNCH_CHECK();
  bt_sk_fill_kernel<<<(unsigned)nchunk, 64, 0, st>>>(D, S, nprob, k, cnt, cols, vals);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_sjlt_left(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t nrows,
                                 int k, const int64_t* rowptr, const int32_t* cols, const double* vals,
                                 const float* A, double* Y, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(nrows >= nprob && nrows < (1LL << 31), 4);
  PAMD_ARG_CHECK(k >= 1 && k <= SK_KMAX, 5);
  PAMD_ARG_CHECK(rowptr != nullptr || nprob == 0, 6);
  PAMD_ARG_CHECK(cols != nullptr || nprob == 0, 7);
  PAMD_ARG_CHECK(vals != nullptr || nprob == 0, 8);
  PAMD_ARG_CHECK(A != nullptr || nprob == 0, 9);
  PAMD_ARG_CHECK(Y != nullptr || nprob == 0, 10);
  if (nprob == 0) return PAMD_OK;
  bt_sk_left_kernel<<<(unsigned)nrows, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, (const SkDesc*)sk, nprob,
                                                                      k, rowptr, cols, vals, A, Y);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_gram(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t ntiles,
                            int shifted, const double* Y, double* G, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(ntiles >= nprob && ntiles < (1LL << 31), 4);
  PAMD_ARG_CHECK(shifted == 0 || shifted == 1, 5);
  PAMD_ARG_CHECK(Y != nullptr || nprob == 0, 6);
  PAMD_ARG_CHECK(G != nullptr || nprob == 0, 7);
  if (nprob == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  bt_gram_kernel<<<(unsigned)ntiles, 256, 0, st>>>((const Desc*)descs, (const SkDesc*)sk, nprob, Y, G);
  PAMD_LAUNCH_CHECK();
  if (shifted) {
    bt_shift_kernel<<<(unsigned)nprob, 256, 0, st>>>((const Desc*)descs, (const SkDesc*)sk, G);
    PAMD_LAUNCH_CHECK();
  }
  return PAMD_OK;
}

extern "C" int pamd_bt_trsm(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t nstrips,
                            int64_t ndblk, int nmax, const double* R, double* dinv, double* Y, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(nstrips >= nprob && nstrips < (1LL << 31), 4);
  PAMD_ARG_CHECK(ndblk >= nprob && ndblk < (1LL << 31), 5);
  PAMD_ARG_CHECK(nmax >= 1 && nmax <= BT_MAXN, 6);
  PAMD_ARG_CHECK(R != nullptr || nprob == 0, 7);
  PAMD_ARG_CHECK(dinv != nullptr || nprob == 0, 8);
  PAMD_ARG_CHECK(Y != nullptr || nprob == 0, 9);
  if (nprob == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const Desc* D = (const Desc*)descs;
  const SkDesc* S = (const SkDesc*)sk;
  bt_dinv_kernel<<<(unsigned)ndblk, 64, 0, st>>>(D, S, nprob, R, dinv);
  PAMD_LAUNCH_CHECK();
  bt_trsm_kernel<<<(unsigned)nstrips, 256, trsm_strip_lds(nmax), st>>>(D, S, nprob, R, dinv, Y);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_trmm_upper(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t ntiles,
                                  const double* A, const double* B, double* C, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(ntiles >= nprob && ntiles < (1LL << 31), 4);
  PAMD_ARG_CHECK(A != nullptr || nprob == 0, 5);
  PAMD_ARG_CHECK(B != nullptr || nprob == 0, 6);
  PAMD_ARG_CHECK(C != nullptr || nprob == 0, 7);
  PAMD_ARG_CHECK(C != A && C != B, 7);
  if (nprob == 0) return PAMD_OK;
  bt_trmm_kernel<<<(unsigned)ntiles, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, (const SkDesc*)sk, nprob,
                                                                    A, B, C);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_fro2(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk, int64_t nchunk,
                            const float* A, double* part, double* out, void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(sk != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(nchunk >= nprob && nchunk < (1LL << 31), 4);
  PAMD_ARG_CHECK(A != nullptr || nprob == 0, 5);
  PAMD_ARG_CHECK(part != nullptr || nprob == 0, 6);
  PAMD_ARG_CHECK(out != nullptr || nprob == 0, 7);
  if (nprob == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  bt_fro2_partial_kernel<<<(unsigned)nchunk, 256, 0, st>>>((const Desc*)descs, (const SkDesc*)sk, nprob, A, part);
  PAMD_LAUNCH_CHECK();
  bt_fro2_final_kernel<<<(unsigned)nprob, 64, 0, st>>>((const Desc*)descs, (const SkDesc*)sk, part, out);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_dot(int64_t nprob, const pamd_ls_desc* descs, const double* a, const double* b, double* out,
                           void* stream) {
  PAMD_ARG_CHECK(nprob >= 0, 1);
  PAMD_ARG_CHECK(descs != nullptr || nprob == 0, 2);
  PAMD_ARG_CHECK(a != nullptr || nprob == 0, 3);
  PAMD_ARG_CHECK(b != nullptr || nprob == 0, 4);
  PAMD_ARG_CHECK(out != nullptr || nprob == 0, 5);
  if (nprob == 0) return PAMD_OK;
  bt_dot_kernel<<<(unsigned)nprob, 256, 0, (hipStream_t)stream>>>((const Desc*)descs, a, b, out);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
