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
                                                        double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int k = 0; k < blocks_per_wave; ++k) {
    const int64_t bi = wid * blocks_per_wave + k;
    if (bi >= nblk) return;
    const Blk bk = blks[bi];
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
                                                          const double* __restrict__ xv, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int k = 0; k < blocks_per_wave; ++k) {
    const int64_t bi = wid * blocks_per_wave + k;
    if (bi >= nblk) return;
    const Blk bk = blks[bi];
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
  trsv_upper_block(trans, n, Rp + d.r_off, n, xs, smem + BT_MAXN);
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[d.nv_off + i] = xs[i];
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
                              int waves_per_wg, int blocks_per_wave, const float* A, const double* p, double* part,
                              void* stream) {
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
  PAMD_BT_NPL(bt_normal_kernel, npl, (const Desc*)descs, (const Blk*)blocks, nblk, blocks_per_wave, A, p, part)
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_bt_residual(int npl, const pamd_ls_desc* descs, const pamd_ls_block* blocks, int64_t nblk,
                                int waves_per_wg, int blocks_per_wave, const float* A, const double* b,
                                const double* x, double* part, void* stream) {
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
  PAMD_BT_NPL(bt_residual_kernel, npl, (const Desc*)descs, (const Blk*)blocks, nblk, blocks_per_wave, A, b, x, part)
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
