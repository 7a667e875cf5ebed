// Sketching operators: counter-based Gaussian fill, SJLT draw, deterministic
// CSC -> CSR transpose, and the sparse sketch applications S @ A and A @ S.
//
// The SJLT products are gathers over the CSR of the wide operator W, so every
// output element is written once, with a fixed summation order (column index
// ascending, the order in which scipy's CSC product accumulates), and without
// atomics:
//   left   out[i, :] = sum_{(j, v) in row i of W} v * A[j, :]     (W @ A)
//   right  out[r, c] = sum_{(j, v) in row c of W} v * A[r, j]     (A @ W^T)
// The right product (k <= 288) stages 32-row tiles of A in LDS, transposed,
// with one lane per row and each output column's CSR list walked by a
// half-wavefront, so W's index arrays are read once per 32 rows of A; wider
// outputs use one thread per output column over 16-row tiles.
#include "common.h"
#include "sjlt_dev.h"

namespace {

__global__ void gaussian_fill_kernel(int64_t rows, int64_t cols, double scale, uint64_t seed, uint64_t offset,
                                     double* __restrict__ X, int64_t ld) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int64_t i = idx / cols, j = idx % cols;
  const uint64_t e = offset + (uint64_t)idx;
  double z0, z1;
  normal_pair(seed, e >> 1, z0, z1);
  X[i * ld + j] = scale * ((e & 1) ? z1 : z0);
}

// Column j: k distinct rows in [0, n_short) (rejection on repeats), signs.
template <int KMAX>
__global__ void sjlt_draw_kernel(int64_t n_short, int64_t n_long, int k, uint64_t seed, int32_t* __restrict__ rows,
                                 double* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_long) return;
  const double mag = 1.0 / sqrt((double)k);
  int32_t picked[KMAX];
  uint64_t signs;
  sjlt_column<KMAX>(n_short, k, seed, j, picked, signs);
#pragma unroll
  for (int l = 0; l < KMAX; ++l) {
    if (l >= k) continue;
    vals[j * k + l] = ((signs >> l) & 1) ? -mag : mag;
    rows[j * k + l] = picked[l];
  }
}

// ---------------------------------------------------------------------------
// CSC -> CSR (stable by column) in three passes over blocks of columns.
// ---------------------------------------------------------------------------
constexpr int CSR_CB = 64;  // columns per block

__global__ __launch_bounds__(256) void csr_count_kernel(int64_t n_short, int64_t n_long, int k,
                                                        const int32_t* __restrict__ rows, int32_t* __restrict__ cnt) {
  extern __shared__ int32_t hist[];
  for (int64_t r = threadIdx.x; r < n_short; r += blockDim.x) hist[r] = 0;
  __syncthreads();
  const int64_t j0 = (int64_t)blockIdx.x * CSR_CB;
  const int64_t j1 = min(n_long, j0 + CSR_CB);
  for (int64_t e = j0 * k + threadIdx.x; e < j1 * k; e += blockDim.x) atomicAdd(&hist[rows[e]], 1);
  __syncthreads();
  for (int64_t r = threadIdx.x; r < n_short; r += blockDim.x) cnt[(int64_t)blockIdx.x * n_short + r] = hist[r];
}

// rowptr = exclusive scan of row counts; cnt[b][r] <- start offset of block b in row r
__global__ __launch_bounds__(1024) void csr_scan_kernel(int64_t n_short, int nblk, int32_t* __restrict__ cnt,
                                                        int64_t* __restrict__ rowptr) {
  csr_scan_body(n_short, nblk, cnt, rowptr);
}

// With many column blocks the scan goes through groups of CSR_GB blocks, so
// that no thread walks all of them: gsum[g][r] = the count of row r in group
// g; the one-workgroup scan turns gsum into the groups' start offsets; then
// every (group, row) hands out the offsets of its blocks.  Integer sums: the
// offsets equal the direct scan's.
constexpr int CSR_GB = 32;

__global__ __launch_bounds__(256) void csr_group_sum_kernel(int64_t n_short, int nblk, const int32_t* __restrict__ cnt,
                                                            int32_t* __restrict__ gsum) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n_short) return;
  const int b0 = blockIdx.y * CSR_GB, b1 = min(nblk, b0 + CSR_GB);
  int32_t t = 0;
  for (int b = b0; b < b1; ++b) t += cnt[(int64_t)b * n_short + r];
  gsum[(int64_t)blockIdx.y * n_short + r] = t;
}

__global__ __launch_bounds__(256) void csr_group_offsets_kernel(int64_t n_short, int nblk,
                                                                const int32_t* __restrict__ gsum,
                                                                int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n_short) return;
  const int b0 = blockIdx.y * CSR_GB, b1 = min(nblk, b0 + CSR_GB);
  int32_t run = gsum[(int64_t)blockIdx.y * n_short + r];
  for (int b = b0; b < b1; ++b) {
    const int32_t t = cnt[(int64_t)b * n_short + r];
    cnt[(int64_t)b * n_short + r] = run;
    run += t;
  }
}

// Thread t owns rows r with r % 256 == t and places them in column order.
// One wavefront per block of CSR_CB columns walks its columns in order; the k
// entries of a column go to their rows' next free slots together (the rows of
// one column are distinct), so every row's entries land in column order, as
// the sequential transpose would place them.  Columns drawn with replacement
// (n_short < k) may repeat a row: their entries are placed one at a time.
__global__ __launch_bounds__(64) void csr_fill_kernel(int64_t n_short, int64_t n_long, int k,
                                                      const int32_t* __restrict__ rows,
                                                      const double* __restrict__ vals,
                                                      const int32_t* __restrict__ cnt, int32_t* __restrict__ cols,
                                                      double* __restrict__ cvals) {
  extern __shared__ int32_t cur[];  // n_short next-slot cursors
  const int lane = threadIdx.x;
  const int32_t* off = cnt + (int64_t)blockIdx.x * n_short;
  for (int64_t r = lane; r < n_short; r += 64) cur[r] = off[r];
  __syncthreads();
  const int64_t j0 = (int64_t)blockIdx.x * CSR_CB;
  const int64_t j1 = min(n_long, j0 + CSR_CB);
  const bool distinct = n_short >= k;
  for (int64_t j = j0; j < j1; ++j) {
    if (distinct) {
      if (lane < k) {
        const int64_t e = j * k + lane;
        const int32_t r = rows[e];
        const int32_t p = cur[r];
        cur[r] = p + 1;
        cols[p] = (int32_t)j;
        cvals[p] = vals[e];
      }
    } else if (lane == 0) {
      for (int l = 0; l < k; ++l) {
        const int64_t e = j * k + l;
        const int32_t r = rows[e];
        const int32_t p = cur[r]++;
        cols[p] = (int32_t)j;
        cvals[p] = vals[e];
      }
    }
    PAMD_WAVE_SYNC();  // this column's cursor updates before the next column reads them
  }
}

// ---------------------------------------------------------------------------
// SJLT applications
// ---------------------------------------------------------------------------

// out[i, :] = sum_p vals[p] * A[cols[p], :], one 256-thread block per output row
__global__ __launch_bounds__(256) void sjlt_left_kernel(int64_t d, int64_t n, const int64_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ cols,
                                                        const double* __restrict__ vals,
                                                        const double* __restrict__ A, int64_t lda,
                                                        double* __restrict__ out, int64_t ldo) {
  const int64_t i = blockIdx.x;
  const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
  for (int64_t c0 = (int64_t)blockIdx.y * 1024; c0 < min(n, (int64_t)(blockIdx.y + 1) * 1024); c0 += 256) {
    const int64_t c = c0 + threadIdx.x;
    if (c >= n) break;
    double s = 0.0;
    int64_t p = p0;
    for (; p + 4 <= p1; p += 4) {
      const double a0 = A[(int64_t)cols[p] * lda + c];
      const double a1 = A[(int64_t)cols[p + 1] * lda + c];
      const double a2 = A[(int64_t)cols[p + 2] * lda + c];
      const double a3 = A[(int64_t)cols[p + 3] * lda + c];
      s += vals[p] * a0;
      s += vals[p + 1] * a1;
      s += vals[p + 2] * a2;
      s += vals[p + 3] * a3;
    }
    for (; p < p1; ++p) s += vals[p] * A[(int64_t)cols[p] * lda + c];
    out[i * ldo + c] = s;
  }
}

constexpr int RT_ROWS = 16;   // rows of A per block
constexpr int RT_COLS = 512;  // columns of A per LDS tile

// out[r, c] = sum_{p in row c of W} vals[p] * A[r, cols[p]]: one thread per
// output column c (blockDim = k rounded up to 64), RT_ROWS rows of A per
// workgroup staged through LDS RT_COLS columns at a time.  A column's CSR
// entries are sorted by j, so the entries falling in the current tile are a
// prefix of what remains; they are fetched eight at a time, so a thread waits
// for the index arrays once per eight entries rather than once per entry.
__global__ __launch_bounds__(1024) void sjlt_right_kernel(int64_t m, int64_t n, int64_t k,
                                                          const int64_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ cols,
                                                          const double* __restrict__ vals,
                                                          const double* __restrict__ A, int64_t lda,
                                                          double* __restrict__ out, int64_t ldo) {
  __shared__ double As[RT_ROWS][RT_COLS];
  const int64_t r0 = (int64_t)blockIdx.x * RT_ROWS;
  const int nr = (int)imin64(RT_ROWS, m - r0);
  const int c = threadIdx.x;
  const bool active = c < k;
  int64_t p = active ? rowptr[c] : 0;
  const int64_t pend = active ? rowptr[c + 1] : 0;
  double acc[RT_ROWS];
#pragma unroll
  for (int rr = 0; rr < RT_ROWS; ++rr) acc[rr] = 0.0;
  for (int64_t j0 = 0; j0 < n; j0 += RT_COLS) {
    const int64_t jn = imin64(RT_COLS, n - j0);
    __syncthreads();
#pragma unroll 8
    for (int e = threadIdx.x; e < RT_ROWS * RT_COLS; e += blockDim.x) {
      const int rr = e / RT_COLS, jj = e % RT_COLS;
      As[rr][jj] = (rr < nr && jj < jn) ? A[(r0 + rr) * lda + j0 + jj] : 0.0;
    }
    __syncthreads();
    const int64_t j1 = j0 + jn;
    while (p < pend) {
      int32_t jv[8];
      double vv[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const bool in = p + e < pend;
        jv[e] = in ? cols[p + e] : INT32_MAX;
        vv[e] = in ? vals[p + e] : 0.0;
      }
      int used = 0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (jv[e] < j1) {  // a prefix of the eight (sorted entries)
          const int jl = jv[e] - (int)j0;
#pragma unroll
          for (int rr = 0; rr < RT_ROWS; ++rr) acc[rr] = fma(vv[e], As[rr][jl], acc[rr]);
          used = e + 1;
        }
      }
      p += used;
      if (used < 8) break;  // the rest of this column lies in later tiles
    }
  }
  if (active)
#pragma unroll
    for (int rr = 0; rr < RT_ROWS; ++rr)
      if (rr < nr) out[(r0 + rr) * ldo + c] = acc[rr];
}

// A * W^T for k <= 288 output columns, one lane per row of A.  A workgroup of
// 1024 threads owns RS_ROWS = 32 rows of A and every output column: each
// half-wavefront (32 lanes, lane = row) owns the output columns c = h, h + 32,
// h + 64, ... and keeps their accumulators and CSR cursors in registers.  A
// streams through LDS in tiles of 32 rows x RS_TJ columns, stored transposed
// (column j of the tile is 32 consecutive doubles, one per row), so for an
// entry (j, v) of column c the half-wave reads the 32 rows of column j in one
// conflict-free 256-byte access and each lane does one FMA.  The next tile is
// loaded into registers while the current one is used.  Every output sums its
// entries in ascending j with one accumulator, the order of sjlt_right_kernel,
// so both give the same bits.  A is read from HBM once; the CSR (index and
// value arrays, L2-resident) once per workgroup.
constexpr int RS_ROWS = 32;
constexpr int RS_TJ = 256;
constexpr int RS_THREADS = 1024;
constexpr int RS_HALVES = RS_THREADS / 32;                   // 32 half-wavefronts
constexpr int RS_PRE = RS_ROWS * RS_TJ / RS_THREADS;          // tile elements per thread

template <int CPH>  // output columns per half-wavefront: k <= 32 * CPH
__global__ __launch_bounds__(RS_THREADS) void sjlt_right_rows_kernel(int64_t m, int64_t n, int64_t k,
                                                                     const int64_t* __restrict__ rowptr,
                                                                     const int32_t* __restrict__ cols,
                                                                     const double* __restrict__ vals,
                                                                     const double* __restrict__ A, int64_t lda,
                                                                     double* __restrict__ out, int64_t ldo) {
  __shared__ double As[RS_TJ][RS_ROWS + 1];
  const int tid = threadIdx.x;
  const int r = tid & 31;           // row within the block (lane within the half-wave)
  const int h = tid >> 5;           // half-wavefront
  const int64_t r0 = (int64_t)blockIdx.x * RS_ROWS;
  int p[CPH], pe[CPH];
  double acc[CPH];
#pragma unroll
  for (int q = 0; q < CPH; ++q) {
    const int64_t c = h + (int64_t)RS_HALVES * q;
    p[q] = (c < k) ? (int)rowptr[c] : 0;
    pe[q] = (c < k) ? (int)rowptr[c + 1] : 0;
    acc[q] = 0.0;
  }
  // tile element e of this thread: row e / RS_TJ, column e % RS_TJ (coalesced along j)
  double pre[RS_PRE];
  auto load_tile = [&](int64_t j0) {
#pragma unroll
    for (int t = 0; t < RS_PRE; ++t) {
      const int e = tid + RS_THREADS * t;
      const int64_t row = r0 + e / RS_TJ, col = j0 + e % RS_TJ;
      pre[t] = (row < m && col < n) ? A[row * lda + col] : 0.0;
    }
  };
  load_tile(0);
  for (int64_t j0 = 0; j0 < n; j0 += RS_TJ) {
    __syncthreads();  // every half-wave is done with the previous tile
#pragma unroll
    for (int t = 0; t < RS_PRE; ++t) {
      const int e = tid + RS_THREADS * t;
      As[e % RS_TJ][e / RS_TJ] = pre[t];
    }
    __syncthreads();
    const int j1 = (int)imin64(n, j0 + RS_TJ);
    // the entries of column q in [j0, j1) are a prefix of what remains
    // (sorted by j): the 32 lanes of the half-wave fetch the next 32 entries
    // of every column at once (one memory round trip per tile), the ballot
    // counts those inside this tile, and the lanes read them in order from
    // the fetching lanes.  They are issued before the next tile's loads:
    // loads complete in order, so waiting for them then does not wait for
    // the next tile.
    int jv[CPH];
    double vv[CPH];
#pragma unroll
    for (int q = 0; q < CPH; ++q) {
      const bool in = p[q] + r < pe[q];
      jv[q] = in ? cols[p[q] + r] : INT32_MAX;
      vv[q] = in ? vals[p[q] + r] : 0.0;
    }
    if (j0 + RS_TJ < n) load_tile(j0 + RS_TJ);  // in flight while this tile is used
    const int hb = (tid & 63) & 32;  // first lane of this half-wave
#pragma unroll
    for (int q = 0; q < CPH; ++q) {
      for (;;) {
        const uint64_t mask = __ballot(jv[q] < j1);
        const int cnt = __popcll((mask >> hb) & 0xffffffffull);
        for (int e = 0; e < cnt; ++e) {
          const int je = __shfl(jv[q], hb + e, 64);
          const double ve = __shfl(vv[q], hb + e, 64);
          acc[q] = fma(ve, As[je - (int)j0][r], acc[q]);
        }
        p[q] += cnt;
        if (cnt < 32) break;  // the rest of this column lies in later tiles
        const bool in = p[q] + r < pe[q];
        jv[q] = in ? cols[p[q] + r] : INT32_MAX;
        vv[q] = in ? vals[p[q] + r] : 0.0;
      }
    }
  }
  const int64_t row = r0 + r;
  if (row < m) {
#pragma unroll
    for (int q = 0; q < CPH; ++q) {
      const int64_t c = h + (int64_t)RS_HALVES * q;
      if (c < k) out[row * ldo + c] = acc[q];
    }
  }
}

__global__ void transpose_kernel(int64_t rows, int64_t cols, const double* __restrict__ A, int64_t lda,
                                 double* __restrict__ B, int64_t ldb) {
  __shared__ double tile[32][33];
  const int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = by + r, j = bx + tx;
    if (i < rows && j < cols) tile[r][tx] = A[i * lda + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t j = bx + r, i = by + tx;  // B[j][i] = A[i][j]
    if (i < rows && j < cols) B[j * ldb + i] = tile[tx][r];
  }
}

}  // namespace

extern "C" int pamd_gaussian_fill_f64(int64_t rows, int64_t cols, double scale, uint64_t seed, uint64_t offset,
                                      double* X, int64_t ld, void* stream) {
  PAMD_ARG_CHECK(rows >= 0, 1);
  PAMD_ARG_CHECK(cols >= 0, 2);
  PAMD_ARG_CHECK(X != nullptr || rows * cols == 0, 6);
  PAMD_ARG_CHECK(ld >= cols && ld >= 1, 7);
  const int64_t tot = rows * cols;
  if (tot == 0) return PAMD_OK;
  gaussian_fill_kernel<<<(unsigned)pamd_cdiv(tot, 256), 256, 0, (hipStream_t)stream>>>(rows, cols, scale, seed,
                                                                                       offset, X, ld);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_sjlt_draw(int64_t n_short, int64_t n_long, int k, uint64_t seed, int32_t* rows, double* vals,
                              void* stream) {
  PAMD_ARG_CHECK(n_short >= 1 && n_short < (1LL << 31), 1);
  PAMD_ARG_CHECK(n_long >= 0, 2);
  PAMD_ARG_CHECK(k >= 1 && k <= 64, 3);
  PAMD_ARG_CHECK(rows != nullptr || n_long == 0, 5);
  PAMD_ARG_CHECK(vals != nullptr || n_long == 0, 6);
  if (n_long == 0) return PAMD_OK;
  const unsigned grid = (unsigned)pamd_cdiv(n_long, 256);
  if (k <= 16)
    sjlt_draw_kernel<16><<<grid, 256, 0, (hipStream_t)stream>>>(n_short, n_long, k, seed, rows, vals);
  else
    sjlt_draw_kernel<64><<<grid, 256, 0, (hipStream_t)stream>>>(n_short, n_long, k, seed, rows, vals);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

#define PAMD_CSR_MAX_SHORT 16384

extern "C" int pamd_sjlt_csr_workspace(int64_t n_short, int64_t n_long, int64_t* bytes) {
  PAMD_ARG_CHECK(n_short >= 1 && n_short <= PAMD_CSR_MAX_SHORT, 1);
  PAMD_ARG_CHECK(n_long >= 0, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 3);
  const int64_t nblk = imax64(1, pamd_cdiv(n_long, CSR_CB));
  const int64_t ngrp = nblk > CSR_GB ? pamd_cdiv(nblk, CSR_GB) : 0;  // group sums of the two-level scan
  *bytes = (nblk + ngrp) * n_short * (int64_t)sizeof(int32_t);
  return PAMD_OK;
}

extern "C" int pamd_sjlt_csr(int64_t n_short, int64_t n_long, int k, const int32_t* rows, const double* vals,
                             int64_t* rowptr, int32_t* cols, double* cvals, void* workspace, void* stream) {
  PAMD_ARG_CHECK(n_short >= 1 && n_short <= PAMD_CSR_MAX_SHORT, 1);
  PAMD_ARG_CHECK(n_long >= 0 && n_long * k < (1LL << 31), 2);
  PAMD_ARG_CHECK(k >= 1 && k <= 64, 3);
  PAMD_ARG_CHECK(rows != nullptr || n_long == 0, 4);
  PAMD_ARG_CHECK(vals != nullptr || n_long == 0, 5);
  PAMD_ARG_CHECK(rowptr != nullptr, 6);
  PAMD_ARG_CHECK(cols != nullptr || n_long == 0, 7);
  PAMD_ARG_CHECK(cvals != nullptr || n_long == 0, 8);
  PAMD_ARG_CHECK(workspace != nullptr, 9);
  hipStream_t st = (hipStream_t)stream;
  const int nblk = (int)imax64(1, pamd_cdiv(n_long, CSR_CB));
  int32_t* cnt = (int32_t*)workspace;
  csr_count_kernel<<<nblk, 256, n_short * sizeof(int32_t), st>>>(n_short, n_long, k, rows, cnt);
  PAMD_LAUNCH_CHECK();
  if (nblk > CSR_GB) {
    const int ngrp = (int)pamd_cdiv(nblk, CSR_GB);
    int32_t* gsum = cnt + (int64_t)nblk * n_short;
    const dim3 grid((unsigned)pamd_cdiv(n_short, 256), (unsigned)ngrp);
    csr_group_sum_kernel<<<grid, 256, 0, st>>>(n_short, nblk, cnt, gsum);
    PAMD_LAUNCH_CHECK();
    csr_scan_kernel<<<1, 1024, 0, st>>>(n_short, ngrp, gsum, rowptr);
    PAMD_LAUNCH_CHECK();
    csr_group_offsets_kernel<<<grid, 256, 0, st>>>(n_short, nblk, gsum, cnt);
  } else {
    csr_scan_kernel<<<1, 1024, 0, st>>>(n_short, nblk, cnt, rowptr);
  }
  PAMD_LAUNCH_CHECK();
  if (n_long > 0) {
    csr_fill_kernel<<<nblk, 64, n_short * sizeof(int32_t), st>>>(n_short, n_long, k, rows, vals, cnt, cols, cvals);
    PAMD_LAUNCH_CHECK();
  }
  return PAMD_OK;
}

namespace {

// launch sjlt_right_rows_kernel<cph> (cph = ceil(k / 32) <= 9: beyond, its registers spill); false otherwise
bool sjlt_right_rows_launch(int cph, int64_t m, int64_t n, int64_t k, const int64_t* rowptr, const int32_t* cols,
                            const double* vals, const double* A, int64_t lda, double* out, int64_t ldo,
                            hipStream_t st) {
  const unsigned nb = (unsigned)pamd_cdiv(m, RS_ROWS);
#define PAMD_RS(C)                                                                                         \
  case C:                                                                                                  \
    sjlt_right_rows_kernel<C><<<nb, RS_THREADS, 0, st>>>(m, n, k, rowptr, cols, vals, A, lda, out, ldo); \
    return true;
  switch (cph) {
    PAMD_RS(1) PAMD_RS(2) PAMD_RS(3) PAMD_RS(4) PAMD_RS(5) PAMD_RS(6) PAMD_RS(7) PAMD_RS(8) PAMD_RS(9)
    default:
      return false;
  }
#undef PAMD_RS
}

}  // namespace

extern "C" int pamd_sjlt_left_f64(int64_t d, int64_t m, int64_t n, const int64_t* rowptr, const int32_t* cols,
                                  const double* vals, const double* A, int64_t lda, double* out, int64_t ldo,
                                  void* stream) {
  PAMD_ARG_CHECK(d >= 0 && d < (1LL << 31), 1);
  PAMD_ARG_CHECK(m >= 0, 2);
  PAMD_ARG_CHECK(n >= 0, 3);
  PAMD_ARG_CHECK(rowptr != nullptr, 4);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 7);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 8);
  PAMD_ARG_CHECK(out != nullptr || d * n == 0, 9);
  PAMD_ARG_CHECK(ldo >= n && ldo >= 1, 10);
  if (d == 0 || n == 0) return PAMD_OK;
  dim3 grid((unsigned)d, (unsigned)pamd_cdiv(n, 1024));
  sjlt_left_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(d, n, rowptr, cols, vals, A, lda, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_sjlt_right_f64(int64_t m, int64_t n, int64_t k, const int64_t* rowptr, const int32_t* cols,
                                   const double* vals, const double* A, int64_t lda, double* out, int64_t ldo,
                                   void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n < (1LL << 31), 2);
  PAMD_ARG_CHECK(k >= 0 && k <= 1024, 3);
  PAMD_ARG_CHECK(rowptr != nullptr, 4);
  PAMD_ARG_CHECK(A != nullptr || m * n == 0, 7);
  PAMD_ARG_CHECK(lda >= n && lda >= 1, 8);
  PAMD_ARG_CHECK(out != nullptr || m * k == 0, 9);
  PAMD_ARG_CHECK(ldo >= k && ldo >= 1, 10);
  if (m == 0 || k == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const int cph = (int)pamd_cdiv(k, RS_HALVES);
  if (cph <= 9 && k * n < (1LL << 31) && sjlt_right_rows_launch(cph, m, n, k, rowptr, cols, vals, A, lda, out, ldo,
                                                                st)) {
    PAMD_LAUNCH_CHECK();
    return PAMD_OK;
  }
  const unsigned nb = (unsigned)pamd_cdiv(m, RT_ROWS);
  const unsigned threads = (unsigned)(pamd_cdiv(k, 64) * 64);
  sjlt_right_kernel<<<nb, threads, 0, st>>>(m, n, k, rowptr, cols, vals, A, lda, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_transpose_f64(int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb,
                                  void* stream) {
  PAMD_ARG_CHECK(rows >= 0, 1);
  PAMD_ARG_CHECK(cols >= 0, 2);
  PAMD_ARG_CHECK(A != nullptr || rows * cols == 0, 3);
  PAMD_ARG_CHECK(lda >= cols && lda >= 1, 4);
  PAMD_ARG_CHECK(B != nullptr || rows * cols == 0, 5);
  PAMD_ARG_CHECK(ldb >= rows && ldb >= 1, 6);
  if (rows == 0 || cols == 0) return PAMD_OK;
  dim3 grid((unsigned)pamd_cdiv(cols, 32), (unsigned)pamd_cdiv(rows, 32));
  transpose_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(rows, cols, A, lda, B, ldb);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

namespace {

// out[i][j] = A[i][j] * s[j] (invert = 0) or A[i][j] / s[j] (invert = 1)
__global__ void scale_cols_kernel(int64_t rows, int64_t cols, const double* __restrict__ A, int64_t lda,
                                  const double* __restrict__ s, int invert, double* __restrict__ out, int64_t ldo) {
  const int64_t i = blockIdx.y;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cols; j += (int64_t)gridDim.x * blockDim.x) {
    const double a = A[i * lda + j];
    out[i * ldo + j] = invert ? a / s[j] : a * s[j];
  }
}

}  // namespace

extern "C" int pamd_scale_cols_f64(int64_t rows, int64_t cols, const double* A, int64_t lda, const double* s,
                                   int invert, double* out, int64_t ldo, void* stream) {
  PAMD_ARG_CHECK(rows >= 0 && rows < (int64_t)1 << 31, 1);
  PAMD_ARG_CHECK(cols >= 0, 2);
  PAMD_ARG_CHECK(A != nullptr || rows * cols == 0, 3);
  PAMD_ARG_CHECK(lda >= cols && lda >= 1, 4);
  PAMD_ARG_CHECK(s != nullptr || cols == 0, 5);
  PAMD_ARG_CHECK(invert == 0 || invert == 1, 6);
  PAMD_ARG_CHECK(out != nullptr || rows * cols == 0, 7);
  PAMD_ARG_CHECK(ldo >= cols && ldo >= 1, 8);
  if (rows == 0 || cols == 0) return PAMD_OK;
  dim3 grid((unsigned)imin64(pamd_cdiv(cols, 256), 64), (unsigned)rows);
  scale_cols_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(rows, cols, A, lda, s, invert, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

namespace {

// Rows r_i of the orthonormal DCT-II of length N applied after the sign flip
// and permutation of the SRCT: out[i][j] = s(r_i) cos(pi r_i (2 q + 1) / (2N)) e[q]
// with q = pinv[j], s(0) = sqrt(1/N), s(k) = sqrt(2/N).  The argument is
// reduced exactly in integers (period 4N) before the cosine.
__global__ void srct_fill_kernel(int64_t N, const int64_t* __restrict__ r, const double* __restrict__ e,
                                 const int64_t* __restrict__ pinv, double* __restrict__ out, int64_t ld) {
  const int64_t i = blockIdx.y;
  const int64_t ri = r[i];
  const double si = (ri == 0) ? sqrt(1.0 / (double)N) : sqrt(2.0 / (double)N);
  const int64_t period = 4 * N;
  const double inv2n = 1.0 / (double)(2 * N);
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = pinv[j];
    const int64_t t = (ri * (2 * q + 1)) % period;
    out[i * ld + j] = si * cospi((double)t * inv2n) * e[q];
  }
}

}  // namespace

extern "C" int pamd_srct_fill_f64(int64_t d, int64_t N, const int64_t* r, const double* e, const int64_t* pinv,
                                  double* out, int64_t ld, void* stream) {
  PAMD_ARG_CHECK(d >= 0 && d <= N && d < (int64_t)1 << 31, 1);
  PAMD_ARG_CHECK(N >= 0 && N < (int64_t)1 << 30, 2);
  PAMD_ARG_CHECK(r != nullptr || d == 0, 3);
  PAMD_ARG_CHECK(e != nullptr || N == 0, 4);
  PAMD_ARG_CHECK(pinv != nullptr || N == 0, 5);
  PAMD_ARG_CHECK(out != nullptr || d * N == 0, 6);
  PAMD_ARG_CHECK(ld >= N && ld >= 1, 7);
  if (d == 0 || N == 0) return PAMD_OK;
  dim3 grid((unsigned)imin64(pamd_cdiv(N, 256), 64), (unsigned)d);
  srct_fill_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(N, r, e, pinv, out, ld);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
