// R factor of a tall matrix by communication-avoiding Householder QR (CAQR),
// FP64.  This is the sketch factorisation of sketch-and-precondition, whose
// callers use R and Q^T applied to appended columns (z = Q^T S b), never Q.
//
// Panels of CQ_NB = 32 columns.  For a panel with mr rows left, the rows are
// cut into P = ceil(mr / 512) blocks of at most 512 rows:
//   1. cq_qr_kernel, one workgroup per block, factors its block in registers
//      (a thread per two columns and 32 rows): 32 column steps, each a DPP
//      broadcast of the pivot column and one fused sum per thread (the
//      column's squared norm below the diagonal, its products with the later
//      columns, and with the earlier reflectors for T), one barrier.  It
//      writes the block's R_b into the block's first rows and its reflectors
//      V_b (explicit unit diagonal) and compact-WY factor T_b (LAPACK dlarft,
//      forward / columnwise) to the workspace.
//   2. for P > 1, the same kernel factors the stack [R_0; ...; R_{P-1}] (the
//      blocks' first nb rows, read in place): the panel's R lands in the first
//      block's rows.
//   3. cq_apply_kernel applies Q_b^T = I - V_b T_b^T V_b^T to each block of the
//      trailing columns (one workgroup per block and 32-column tile, on FP64
//      MFMA), then, for P > 1, the stack's Q^T to the trailing columns' stacked
//      rows.
// The next panel works on the remaining rows where they lie: R does not
// depend on the order of the rows below a panel (for a row permutation Pi,
// Pi A = (Pi Q) R), so stacked rows are never moved.  At the end, rows of R
// (and of the appended columns) are negated where R's diagonal is negative:
// R is the thin QR's R with a non-negative diagonal, unique for full rank.
// Every reduction has a fixed order, so results are run-to-run identical.
#include "common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int CQ_NB = 32;                 // panel width
constexpr int CQ_QT = 256;                // threads of the QR kernel: 16 groups of 32 rows x 32 columns
constexpr int CQ_NW = CQ_QT / 64;
constexpr int CQ_AT = 256;                // threads of the apply kernel
constexpr int CQ_CT = 32;                 // trailing columns per apply workgroup
constexpr int CQ_MAXROWS = 512;           // rows of one QR (block or stack)
constexpr int64_t CQ_MAXM = 8192;         // rows supported: at most 16 blocks, a stack of 512 rows

// number of row blocks for mr remaining rows (mr <= CQ_MAXM)
inline int cq_blocks(int64_t mr) { return (int)pamd_cdiv(mr, CQ_MAXROWS); }

// first row of block b of P over mr rows (balanced: every block has
// floor(mr / P) or ceil(mr / P) rows)
__host__ __device__ inline int cq_start(int mr, int P, int b) { return (int)(((int64_t)mr * b) / P); }

// the rows one QR or apply works on, relative to the panel's first row:
// a contiguous block, or the stack of every block's first nb rows
struct CqRows {
  int mr, P, nb, tree, base, nrows;
  __device__ CqRows(int mr_, int P_, int nb_, int tree_, int b) : mr(mr_), P(P_), nb(nb_), tree(tree_) {
    base = tree ? 0 : cq_start(mr, P, b);
    nrows = tree ? P * nb : cq_start(mr, P, b + 1) - base;
  }
  __device__ int row(int k) const { return tree ? cq_start(mr, P, k / nb) + k % nb : base + k; }
};

__global__ void cq_zero_flag_kernel(int* flag) { flag[0] = 0; }

// Thread t of cq_qr_kernel holds columns c and c + 16, c = t % 16, of the
// rows 32 g .. 32 g + 31, g = t / 16, of its block (or stack): a row group
// is one 16-lane row of a wavefront, four per wavefront.  What a column step
// shares through LDS, double-buffered by the column's parity:
struct CqShared {
  double (*red)[CQ_NW][CQ_NB];  // [2][CQ_NW][CQ_NB] per-wavefront sums
  double (*prow)[CQ_NB];        // [2][CQ_NB] the pivot row
  double* taus;                 // [CQ_NB]
  double (*Gs)[CQ_NB + 1];      // Gs[l][j] = v_l . v_j, l < j
  int g, c, lane, wave, tid;
};

// v in lane N of this lane's 16-lane row, for every lane (DPP row_newbcast)
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
  return dpp_f64<0x150 + N>(v);
}

// 1 / x and 1 / sqrt(x) from the hardware estimates and two Newton steps
// each (the estimates' relative errors are squared twice: a few ulps)
__device__ __forceinline__ double cq_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * fma(-x, y, 2.0);
  return y * fma(-x, y, 2.0);
}

__device__ __forceinline__ double cq_rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// Column J (a template argument, so every register index is static).  Rows
// r > J are the column's part below the diagonal; row J, the pivot row, is
// row J of group 0.  The pivot column x reaches the lanes of a row group
// from the lane that holds it by a DPP broadcast inside the 16-lane row, so
// the step has one barrier: after the wavefronts' sums.  The reflector
// v_J = [1; scale x] is kept unscaled (x in the registers of column J, scale
// in that column's `vs`): every update is one multiply-add per element, the
// same for all lanes.
template <int J>
__device__ __forceinline__ void cq_column(double (&a)[2][CQ_NB], double (&vs)[2], const CqShared& s) {
  constexpr int buf = J & 1, JS = J / 16, JL = J % 16;
  if (s.g == 0) {
    s.prow[buf][s.c] = a[0][J];
    s.prow[buf][s.c + 16] = a[1][J];
  }
  // x: column J of this lane's rows, zero at and above the diagonal (group 0)
  double x[CQ_NB];
#pragma unroll
  for (int i = 0; i < CQ_NB; ++i) {
    x[i] = row_bcast<JL>(a[JS][i]);
    if (i <= J) x[i] = (s.g == 0) ? 0.0 : x[i];
  }
  // x . columns c, c + 16 over this thread's rows (column J: the squared norm,
  // later columns, earlier (unscaled) reflectors for T), four partial sums each
  double part[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < CQ_NB; ++i) p[i & 3] += x[i] * a[h][i];
    part[h] = (p[0] + p[1]) + (p[2] + p[3]);
    // the wavefront's four row groups: rows 0 + 1 and 2 + 3, then the two pairs
    part[h] += __shfl_xor(part[h], 16, 64);
    part[h] += __shfl_xor(part[h], 32, 64);
  }
  if (s.lane < 16) {
    s.red[buf][s.wave][s.c] = part[0];
    s.red[buf][s.wave][s.c + 16] = part[1];
  }
  __syncthreads();  // the one barrier of the column
  double tot[2], sq;
  {
    double t0[CQ_NW], t1[CQ_NW], sv[CQ_NW];
#pragma unroll
    for (int w = 0; w < CQ_NW; ++w) {
      t0[w] = s.red[buf][w][s.c];
      t1[w] = s.red[buf][w][s.c + 16];
      sv[w] = s.red[buf][w][J];
    }
    tot[0] = (t0[0] + t0[1]) + (t0[2] + t0[3]);
    tot[1] = (t1[0] + t1[1]) + (t1[2] + t1[3]);
    sq = (sv[0] + sv[1]) + (sv[2] + sv[3]);
  }
  const double alpha = s.prow[buf][J];
  const double pc[2] = {s.prow[buf][s.c], s.prow[buf][s.c + 16]};
  // beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha) / beta = 1 + |alpha| / ||.||,
  // scale = 1 / (alpha - beta) = sign(alpha) / (|alpha| + ||.||)
  double beta, tau, scale;
  if (sq == 0.0) {
    beta = alpha;
    tau = 0.0;
    scale = 0.0;
  } else {
    const double s2 = fma(alpha, alpha, sq);
    const double rn = cq_rsq(s2);
    const double nrm = s2 * rn;
    beta = -copysign(nrm, alpha);
    tau = fma(fabs(alpha), rn, 1.0);
    scale = copysign(cq_rcp(fabs(alpha) + nrm), alpha);
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int col = s.c + 16 * h;
    // col > J: a_col -= tau (v . a_col) v = w [1 at row J; scale x below]; other columns: w = 0
    const double w = (col > J) ? tau * (pc[h] + scale * tot[h]) : 0.0;
    const double ws = w * scale;
#pragma unroll
    for (int i = 0; i < CQ_NB; ++i) a[h][i] -= ws * x[i];
    if (s.g == 0) {
      if (col > J) a[h][J] -= w;
      if (col == J) a[h][J] = beta;
      // G[col][J] = v_col . v_J = vs_col (v_col[J] + scale sum_{r > J} a_col[r] x_r), col < J
      if (col < J) s.Gs[col][J] = vs[h] * (pc[h] + scale * tot[h]);
    }
    if (col == J) vs[h] = scale;
  }
  if (s.tid == 0) s.taus[J] = tau;
}

// columns J .. nb - 1 (nb is uniform, so every barrier is reached by all threads)
template <int J>
__device__ __forceinline__ void cq_columns(double (&a)[2][CQ_NB], double (&vs)[2], const CqShared& s, int nb) {
  if constexpr (J < CQ_NB) {
    if (J >= nb) return;
    cq_column<J>(a, vs, s);
    cq_columns<J + 1>(a, vs, s, nb);
  }
}

// Householder QR of one block (tree = 0, block blockIdx.x) or of the stack
// (tree = 1) of the panel whose top-left element is A[0].  V (nrows x 32, the
// block's rows from V + base * 32) and T (32 x 32 at T + b * 1024) go to the
// workspace; R_b's rows are written over the first nb rows.
__global__ __launch_bounds__(CQ_QT) void cq_qr_kernel(int mr, int P, int nb, int tree, double* __restrict__ A,
                                                      int64_t lda, double* __restrict__ Vw,
                                                      double* __restrict__ Tw) {
  __shared__ double red[2][CQ_NW][CQ_NB];
  __shared__ double prow[2][CQ_NB];
  __shared__ double taus[CQ_NB];
  __shared__ double Gs[CQ_NB][CQ_NB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = tid / 16, c = tid % 16;
  const int b = tree ? 0 : (int)blockIdx.x;
  const CqRows rows(mr, P, nb, tree, b);
  const int nrows = rows.nrows;
  double* V = Vw + (int64_t)rows.base * CQ_NB;
  double* T = Tw + (int64_t)b * CQ_NB * CQ_NB;
  for (int e = tid; e < CQ_NB * (CQ_NB + 1); e += CQ_QT) Gs[e / (CQ_NB + 1)][e % (CQ_NB + 1)] = 0.0;
  if (tid < CQ_NB) taus[tid] = 0.0;

  // a wavefront's loads cover four rows of 16 contiguous values
  double a[2][CQ_NB];
#pragma unroll
  for (int i = 0; i < CQ_NB; ++i) {
    const int k = CQ_NB * g + i;
    const int64_t ro = (k < nrows) ? (int64_t)rows.row(k) * lda : 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) a[h][i] = (k < nrows && c + 16 * h < nb) ? A[ro + c + 16 * h] : 0.0;
  }
  __syncthreads();

  const CqShared sh{red, prow, taus, Gs, g, c, lane, wave, tid};
  double vs[2] = {0.0, 0.0};  // the scales of this thread's columns' reflectors
  cq_columns<0>(a, vs, sh, nb);

  // R_b: the first nb rows (group 0), zeros left of the diagonal
  if (g == 0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = c + 16 * h;
      if (col < nb) {
#pragma unroll
        for (int i = 0; i < CQ_NB; ++i)
          if (i < nb) A[(int64_t)rows.row(i) * lda + col] = (col >= i) ? a[h][i] : 0.0;
      }
    }
  }
  // V with its explicit unit diagonal, rows of 32
#pragma unroll
  for (int i = 0; i < CQ_NB; ++i) {
    const int k = CQ_NB * g + i;
    if (k < nrows) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int col = c + 16 * h;
        V[(int64_t)k * CQ_NB + col] = (col >= nb || col > k) ? 0.0 : ((col == k) ? 1.0 : vs[h] * a[h][i]);
      }
    }
  }
  // T: thread i forms row i (T[i][j] = -tau_j sum_{i <= l < j} T[i][l] G[l][j]).
  // Gs and taus are written by wavefront 0 alone (group 0 and thread 0), so
  // T needs no block barrier: the other wavefronts' stores drain meanwhile.
  if (wave == 0) PAMD_WAVE_SYNC();
  if (tid < CQ_NB) {
    const int i = tid;
    double Trow[CQ_NB];
#pragma unroll
    for (int jj = 0; jj < CQ_NB; ++jj) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < jj; ++l) s += Trow[l] * Gs[l][jj];
      Trow[jj] = (jj < i) ? 0.0 : ((jj == i) ? taus[jj] : -taus[jj] * s);
    }
#pragma unroll
    for (int jj = 0; jj < CQ_NB; ++jj) T[i * CQ_NB + jj] = Trow[jj];
  }
}

// C <- (I - V T^T V^T) C on the rows of one block (tree = 0) or of the
// stack (tree = 1), for 32 trailing columns (nc columns in all; C is the
// trailing block's top-left element); a workgroup per block and column tile.  Wavefront w owns a quarter of the rows (at most 128):
// it forms its part of W = V^T C (all four 16 x 16 blocks) on
// v_mfma_f64_16x16x4f64, keeping its C values in registers, and after
// W2 = T^T W writes its rows of C - V W2: the B-operand layout of the first
// product and the result layout of the second put the same rows of C in the
// same lane, so C is read once.  V's next batch is loaded before the current
// batch's products.  The quarters' parts of W are added in wavefront order.
__global__ __launch_bounds__(CQ_AT) void cq_apply_kernel(int mr, int P, int nb, int tree, int64_t nc,
                                                         const double* __restrict__ Vw,
                                                         const double* __restrict__ Tw, double* __restrict__ C,
                                                         int64_t ldc) {
  constexpr int NWA = CQ_AT / 64;
  constexpr int QROWS = CQ_MAXROWS / NWA;  // rows of a wavefront's quarter, at most
  constexpr int NS = QROWS / 4;            // its K steps of 4 rows
  constexpr int U = 8;                     // K steps per batch
  __shared__ double Wp[NWA][CQ_NB][CQ_CT + 1];
  __shared__ double W2s[CQ_NB][CQ_CT + 1];
  __shared__ double Ts[CQ_NB][CQ_NB + 1];
  __shared__ int64_t roff[CQ_MAXROWS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  // tiles block-major, contiguous per XCD: the column tiles of a block share
  // its V in one L2
  const int ct = (int)pamd_cdiv(nc, CQ_CT);
  const int tile = (int)xcd_tile(blockIdx.x, gridDim.x);
  const int b = tree ? 0 : tile / ct;
  const CqRows rows(mr, P, nb, tree, b);
  const int K = rows.nrows;
  const double* V = Vw + (int64_t)rows.base * CQ_NB;
  const double* T = Tw + (int64_t)b * CQ_NB * CQ_NB;
  const int64_t col0 = (int64_t)(tile % ct) * CQ_CT;
  for (int k = tid; k < K; k += CQ_AT) roff[k] = (int64_t)rows.row(k) * ldc;
  for (int e = tid; e < CQ_NB * CQ_NB; e += CQ_AT) Ts[e / CQ_NB][e % CQ_NB] = T[e];
  __syncthreads();
  const int quarter = (int)pamd_cdiv(pamd_cdiv(K, NWA), 16) * 16;
  const int kb = wave * quarter, ke = min(K, kb + quarter);
  const int64_t colh[2] = {col0 + lr, col0 + 16 + lr};
  const bool cok[2] = {colh[0] < nc, colh[1] < nc};

  // cv[s][h] = C[kb + 4 s + lk][col0 + 16 h + lr]: the B operand of K step s,
  // and item s % 4 of the result block of rows kb + 16 (s / 4) ..
  double cv[NS][2];
  {
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    double va[2][U][2];  // V (A operand, V^T) of two batches
    auto loadv = [&](int bt, double (&v)[U][2]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = kb + 4 * (bt * U + u) + lk;
#pragma unroll
        for (int h = 0; h < 2; ++h) v[u][h] = (k < ke) ? V[(int64_t)k * CQ_NB + 16 * h + lr] : 0.0;
      }
    };
    auto loadc = [&](int bt) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = kb + 4 * (bt * U + u) + lk;
#pragma unroll
        for (int h = 0; h < 2; ++h) cv[bt * U + u][h] = (k < ke && cok[h]) ? C[roff[k] + colh[h]] : 0.0;
      }
    };
    loadv(0, va[0]);
    loadc(0);
#pragma unroll
    for (int bt = 0; bt < NS / U; ++bt) {
      if (kb + 4 * U * bt >= ke) break;  // uniform per wavefront
      if (bt + 1 < NS / U && kb + 4 * U * (bt + 1) < ke) {
        loadv(bt + 1, va[(bt + 1) & 1]);
        loadc(bt + 1);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[bt & 1][u][i], cv[bt * U + u][j], acc[i][j], 0, 0, 0);
    }
    // D layout of v_mfma_f64_16x16x4f64: lane l, item r -> row l / 16 + 4 r, column l % 16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) Wp[wave][16 * i + lk + 4 * r][16 * j + lr] = acc[i][j][r];
  }
  __syncthreads();
  // W2 = T^T W, W the quarters' sum in wavefront order (T upper triangular)
  {
    const int c = tid & (CQ_CT - 1), ag = tid >> 5;
    constexpr int APT = CQ_NB / (CQ_AT / CQ_CT);
    double wcol[CQ_NB];
#pragma unroll
    for (int a = 0; a < CQ_NB; ++a) {
      double s = Wp[0][a][c];
#pragma unroll
      for (int w = 1; w < NWA; ++w) s += Wp[w][a][c];
      wcol[a] = s;
    }
#pragma unroll
    for (int u = 0; u < APT; ++u) {
      const int a = ag * APT + u;
      double s = 0.0;
#pragma unroll
      for (int ap = 0; ap < CQ_NB; ++ap)
        if (ap <= a) s += Ts[ap][a] * wcol[ap];
      W2s[a][c] = s;
    }
  }
  __syncthreads();
  // C - V W2 on this wavefront's rows, 16-row blocks, the next block's V in flight
  {
    double wb[2][CQ_NB / 4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < CQ_NB / 4; ++q) wb[h][q] = W2s[4 * q + lk][16 * h + lr];
    double va[2][CQ_NB / 4];
    auto loadv = [&](int t, double (&v)[CQ_NB / 4]) {
      const int kr = kb + 16 * t + lr;
#pragma unroll
      for (int q = 0; q < CQ_NB / 4; ++q) v[q] = (kr < ke) ? V[(int64_t)kr * CQ_NB + 4 * q + lk] : 0.0;
    };
    loadv(0, va[0]);
#pragma unroll
    for (int t = 0; t < NS / 4; ++t) {
      if (kb + 16 * t >= ke) break;  // uniform per wavefront
      if (t + 1 < NS / 4 && kb + 16 * (t + 1) < ke) loadv(t + 1, va[(t + 1) & 1]);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < CQ_NB / 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(va[t & 1][q], wb[h][q], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = kb + 16 * t + lk + 4 * r;
          if (k < ke && cok[h]) C[roff[k] + colh[h]] = cv[4 * t + r][h] - acc[r];
        }
      }
    }
  }
}

// Negate row j of A[:, j:ncols] where R[j][j] < 0; flag an exact zero diagonal.
__global__ __launch_bounds__(256) void cq_sign_kernel(int64_t ncols, double* __restrict__ A, int64_t lda,
                                                      int* __restrict__ flag) {
  const int64_t j = blockIdx.x;
  double* row = A + j * lda;
  const double d = row[j];
  if (d == 0.0 && threadIdx.x == 0) flag[0] = 1;
  if (d < 0.0)
    for (int64_t c = j + threadIdx.x; c < ncols; c += blockDim.x) row[c] = -row[c];
}

struct CqLayout {
  int64_t vloc, tloc, vtree, ttree, flag, total;  // offsets in doubles
};

CqLayout cq_layout(int64_t m) {
  CqLayout L;
  L.vloc = 0;
  L.tloc = L.vloc + imax64(m, 1) * CQ_NB;
  L.vtree = L.tloc + (CQ_MAXM / CQ_MAXROWS) * CQ_NB * CQ_NB;
  L.ttree = L.vtree + (int64_t)CQ_MAXROWS * CQ_NB;
  L.flag = L.ttree + CQ_NB * CQ_NB;
  L.total = L.flag + 4;
  return L;
}

}  // namespace

extern "C" int pamd_qr_r_workspace(int64_t m, int64_t n, int64_t ncols, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= m, 2);
  PAMD_ARG_CHECK(ncols >= n, 3);
  PAMD_ARG_CHECK(bytes != nullptr, 4);
  *bytes = cq_layout(m).total * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_qr_r_f64(int64_t m, int64_t n, int64_t ncols, double* A, int64_t lda, void* workspace,
                             void* stream) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 0 && n <= m, 2);
  PAMD_ARG_CHECK(ncols >= n, 3);
  PAMD_ARG_CHECK(A != nullptr || m * ncols == 0, 4);
  PAMD_ARG_CHECK(lda >= ncols && lda >= 1, 5);
  PAMD_ARG_CHECK(workspace != nullptr, 6);
  if (m > CQ_MAXM) return PAMD_ERR_UNSUPPORTED;
  if (n == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const CqLayout L = cq_layout(m);
  double* ws = (double*)workspace;
  int* flag = (int*)(ws + L.flag);
  cq_zero_flag_kernel<<<1, 1, 0, st>>>(flag);
  PAMD_LAUNCH_CHECK();
  for (int64_t c0 = 0; c0 < n; c0 += CQ_NB) {
    const int nb = (int)imin64(CQ_NB, n - c0);
    const int mr = (int)(m - c0);
    const int P = cq_blocks(mr);
    const int64_t nc = ncols - c0 - nb;
    double* Ap = A + c0 * lda + c0;
    cq_qr_kernel<<<P, CQ_QT, 0, st>>>(mr, P, nb, 0, Ap, lda, ws + L.vloc, ws + L.tloc);
    PAMD_LAUNCH_CHECK();
    if (P > 1) {
      cq_qr_kernel<<<1, CQ_QT, 0, st>>>(mr, P, nb, 1, Ap, lda, ws + L.vtree, ws + L.ttree);
      PAMD_LAUNCH_CHECK();
    }
    if (nc > 0) {
      const unsigned ct = (unsigned)pamd_cdiv(nc, CQ_CT);
      cq_apply_kernel<<<ct * (unsigned)P, CQ_AT, 0, st>>>(mr, P, nb, 0, nc, ws + L.vloc, ws + L.tloc, Ap + nb, lda);
      PAMD_LAUNCH_CHECK();
      if (P > 1) {
        cq_apply_kernel<<<ct, CQ_AT, 0, st>>>(mr, P, nb, 1, nc, ws + L.vtree, ws + L.ttree, Ap + nb, lda);
        PAMD_LAUNCH_CHECK();
      }
    }
  }
  cq_sign_kernel<<<(unsigned)n, 256, 0, st>>>(ncols, A, lda, flag);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_qr_r_singular_flag(const void* workspace, int64_t m, int** flag_ptr) {
  PAMD_ARG_CHECK(workspace != nullptr, 1);
  PAMD_ARG_CHECK(m >= 0, 2);
  PAMD_ARG_CHECK(flag_ptr != nullptr, 3);
  *flag_ptr = (int*)((double*)workspace + cq_layout(m).flag);
  return PAMD_OK;
}
