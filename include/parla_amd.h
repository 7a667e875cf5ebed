/*
 * parla_amd.h -- C ABI of the MI355X (gfx950) randomized-NLA library.
 *
 * The library implements the hot path of PARLA 0.1.4 (BallisticLA/parla): the
 * sketching operators (dense Gaussian S*A, sparse SJLT), sketch-and-precondition
 * least squares (sketch A, Householder QR of SA, LSQR preconditioned by R) and
 * the power-iteration QB rangefinder that feeds randomized SVD / EVD.  The
 * reference is pure Python on numpy/scipy, so each entry point below replaces
 * the numpy/scipy call site named in its comment (file:line in the PARLA tree).
 *
 * Conventions
 *  - Matrices are row-major (numpy C order): element (i, j) of X is
 *    X[i * ld + j].  Leading dimensions are in elements.
 *  - Every pointer argument is a device pointer unless its comment says host.
 *  - Memory is caller-owned; query functions (*_workspace) give the scratch
 *    size in bytes.  No entry point allocates device memory.
 *  - `stream` is a hipStream_t (NULL = the default stream).  Work is enqueued
 *    asynchronously unless the comment says the call blocks.
 *  - Return value: 0 = ok; < 0 = argument -value is invalid (LAPACK info
 *    style); PAMD_ERR_* = runtime failure; > 0 = numerical event.
 *  - Entry points never abort and never throw across the ABI; they are
 *    re-entrant (no hidden global state besides the HIP runtime), except
 *    pamd_svd_jacobi_noalloc_f64, whose per-GPU static work array makes calls
 *    that overlap on one GPU race.
 */
#ifndef PARLA_AMD_H
#define PARLA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAMD_OK 0
#define PAMD_ERR_HIP (-1000)         /* a HIP runtime call failed */
#define PAMD_ERR_UNSUPPORTED (-1001) /* shape or option outside what is built */
#define PAMD_EVENT_SINGULAR 1        /* exact zero on the diagonal of R */

/* ABI version of this header (bumped on any signature change). */
int pamd_abi_version(void);

/* Static description of a status code (host string, never NULL). */
const char* pamd_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Dense kernels                                                             */
/* ------------------------------------------------------------------------ */

/* C = alpha * op(A) * op(B) + beta * C, FP64 on MFMA.  op(A) is M x K
 * (transA = 0: A is M x K; transA = 1: A is K x M), op(B) is K x N.
 * Replaces numpy matmul -> dgemm at: S @ A (drivers/least_squares.py:303),
 * A @ S and A.T @ S (comps/sketchers/aware.py:166,175,179), Y = A @ S
 * (comps/rangefinders.py:186), B = Q.T @ A (comps/qb.py:351), Q @ U
 * (drivers/svd.py:175, drivers/evd.py:288), B @ Q (drivers/evd.py:278). */
int pamd_gemm_f64(int transA, int transB, int64_t M, int64_t N, int64_t K,
                  double alpha, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double beta,
                  double* C, int64_t ldc, void* workspace,
                  int64_t workspace_bytes, void* stream);

/* Workspace (bytes) that lets pamd_gemm_f64 split K for this shape (0 when a
 * split would not help).  Passing less workspace only disables the split. */
int pamd_gemm_workspace(int64_t M, int64_t N, int64_t K, int64_t* bytes);

/* y = alpha * op(A) * x + beta * y, A is m x n (trans = 1: op(A) = A^T).
 * Replaces A @ x / A.T @ y in drivers/least_squares.py:344,361 and
 * comps/determiter/saddle.py:199. */
int pamd_gemv_f64(int trans, int64_t m, int64_t n, double alpha,
                  const double* A, int64_t lda, const double* x, double beta,
                  double* y, void* workspace, int64_t workspace_bytes,
                  void* stream);

/* Workspace (bytes) for the row-split reduction of pamd_gemv_f64 with
 * trans = 1 (0 for trans = 0). */
int pamd_gemv_workspace(int trans, int64_t m, int64_t n, int64_t* bytes);

/* result[0] = ||x||_2 (device scalar).  Replaces la.norm at
 * drivers/least_squares.py:347,366. */
int pamd_nrm2_f64(int64_t n, const double* x, double* result, void* stream);

/* result[0] = ||x||_2 for long x (e.g. ||A||_F of a contiguous matrix): one
 * partial sum of squares per workgroup (up to 1024, in `workspace`), then an
 * ordered final sum.  Replaces np.linalg.norm of the whole matrix. */
int pamd_nrm2_workspace(int64_t n, int64_t* bytes);
int pamd_nrm2_large_f64(int64_t n, const double* x, double* result,
                        void* workspace, void* stream);

/* In-place triangular solve with the upper-triangular n x n R (any n < 2^30;
 * up to 2048 the solution vector is staged in LDS, beyond that it is solved
 * in place in x): trans = 0: x <- R^{-1} x;  trans = 1: x <- R^{-T} x.
 * Replaces la.solve_triangular at drivers/least_squares.py:316,363 and
 * comps/preconditioning.py:28,37,40-41. */
int pamd_trsv_upper_f64(int trans, int64_t n, const double* R, int64_t ldr,
                        double* x, void* stream);

/* ------------------------------------------------------------------------ */
/* Sketching operators                                                      */
/* ------------------------------------------------------------------------ */

/* Fill X (rows x cols, leading dimension ld) with independent N(0, scale^2)
 * values from a counter-based generator (Philox-4x32-10, Salmon et al. 2011,
 * with the Box-Muller transform).  Element (i, j) depends only on
 * (seed, offset + i * cols + j), so any tiling reproduces the same matrix.
 * On-device replacement of rng.normal in utils/sketching.py:24. */
int pamd_gaussian_fill_f64(int64_t rows, int64_t cols, double scale,
                           uint64_t seed, uint64_t offset, double* X,
                           int64_t ld, void* stream);

/* The subsampled randomized cosine transform as a dense d x N operator:
 * out[i][j] = s(r_i) cos(pi r_i (2 q + 1) / (2N)) e[q], q = pinv[j],
 * s(0) = 1/sqrt(N), s(k) = sqrt(2/N) -- row r_i of the orthonormal DCT-II
 * (scipy.fft.dct(norm='ortho')) composed with the sign vector e and the
 * permutation perm (pinv = argsort(perm)), so out @ X equals apply_srct of
 * utils/sketching.py:119-176 (mat[perm] * e, dct along axis 0, rows r), and
 * out^T its adjoint.  The DCT argument is reduced exactly in integers
 * before the cosine.  N < 2^30.  Draws come from generate_srct
 * (utils/sketching.py:106-116), made on the host in the reference's order. */
int pamd_srct_fill_f64(int64_t d, int64_t N, const int64_t* r, const double* e,
                       const int64_t* pinv, double* out, int64_t ld, void* stream);

/* Draw a wide SJLT W (n_short x n_long) on the device: column j gets k
 * distinct rows (k <= n_short; with replacement when n_short < k) and signs
 * +-1/sqrt(k).  Writes rows[j*k + l] and vals[j*k + l] (the CSC layout of
 * utils/sketching.py:63-74).  Counter-based, so not numpy's stream: the
 * drop-in path draws on the host with numpy and uploads instead. */
int pamd_sjlt_draw(int64_t n_short, int64_t n_long, int k, uint64_t seed,
                   int32_t* rows, double* vals, void* stream);

/* Workspace (bytes) for pamd_sjlt_csr. */
int pamd_sjlt_csr_workspace(int64_t n_short, int64_t n_long, int64_t* bytes);

/* Deterministic CSC -> CSR transpose of a wide SJLT with k entries per
 * column: rowptr (n_short + 1, int64), cols / cvals (n_long * k) sorted by
 * column inside each row.  Mirrors the COO -> CSC/CSR conversions of
 * utils/sketching.py:73-79. */
int pamd_sjlt_csr(int64_t n_short, int64_t n_long, int k, const int32_t* rows,
                  const double* vals, int64_t* rowptr, int32_t* cols,
                  double* cvals, void* workspace, void* stream);

/* out (d x n) = W (d x m) * A (m x n), W given by its CSR (rowptr d + 1).
 * Replaces the scipy CSC @ dense product S @ A at
 * drivers/least_squares.py:303,314 (S = SkOpSJ) and A.T @ S at
 * comps/sketchers/aware.py:166 (computed as (W @ A)^T). */
int pamd_sjlt_left_f64(int64_t d, int64_t m, int64_t n, const int64_t* rowptr,
                       const int32_t* cols, const double* vals,
                       const double* A, int64_t lda, double* out, int64_t ldo,
                       void* stream);

/* out (m x k) = A (m x n) * W^T, W (k x n) given by its CSR (rowptr k + 1).
 * Replaces the dense @ CSR product A @ S at comps/sketchers/aware.py:175 and
 * comps/rangefinders.py:186 for a tall SJLT S = W^T. */
int pamd_sjlt_right_f64(int64_t m, int64_t n, int64_t k, const int64_t* rowptr,
                        const int32_t* cols, const double* vals,
                        const double* A, int64_t lda, double* out, int64_t ldo,
                        void* stream);

/* ------------------------------------------------------------------------ */
/* Householder QR                                                            */
/* ------------------------------------------------------------------------ */

/* Workspace (bytes) for pamd_geqrf_f64 / pamd_orgqr_f64 on an m x ncols
 * matrix whose first n columns are factored. */
int pamd_geqrf_workspace(int64_t m, int64_t n, int64_t ncols, int64_t* bytes);

/* Blocked Householder QR (LAPACK geqrf semantics, R may have negative
 * diagonal) of the first n columns of A (m x ncols, m >= n).  On exit R is in
 * the upper triangle of A[:n, :n], and columns n..ncols-1 hold Q^T times the
 * input columns (so appending S*b gives z = Q^T S b in A[:n, n]).  The
 * reflectors stay in `workspace` for pamd_orgqr_f64.  Replaces la.qr at
 * drivers/least_squares.py:311 (with the Q[:d].T @ b_ske of :315 fused) and
 * the QR inside utils/linalg_wrappers.py:7, comps/rangefinders.py:187.
 * Returns PAMD_EVENT_SINGULAR if R has an exact zero diagonal entry. */
int pamd_geqrf_f64(int64_t m, int64_t n, int64_t ncols, double* A,
                   int64_t lda, void* workspace, void* stream);

/* Form the economic Q (m x n) from the reflectors a preceding
 * pamd_geqrf_f64 left in `workspace` (LAPACK orgqr).  Replaces the Q factor
 * of la.qr(..., mode='economic')[0] at utils/linalg_wrappers.py:7 and
 * comps/rangefinders.py:187. */
int pamd_orgqr_f64(int64_t m, int64_t n, int64_t ncols, const void* workspace,
                   double* Q, int64_t ldq, void* stream);

/* Workspace (bytes) for pamd_lu_pl_f64. */
int pamd_lu_workspace(int64_t m, int64_t n, int64_t* bytes);

/* PL (m x min(m, n)) = P @ L of the partially pivoted LU of A (m x n),
 * scipy.linalg.lu(A, permute_l=True)[0]; pivots are the first largest |entry|
 * (LAPACK idamax).  Replaces the stabilizer utils/linalg_wrappers.py:10-12. */
int pamd_lu_pl_f64(int64_t m, int64_t n, const double* A, int64_t lda,
                   double* PL, int64_t ldpl, void* workspace, void* stream);

/* Pointer (device, int32) to the flag pamd_geqrf_f64 sets when R has an exact
 * zero on its diagonal (lives inside `workspace`; host call, no GPU work). */
int pamd_qr_singular_flag(const void* workspace, int64_t m, int64_t n,
                          int64_t ncols, int** flag_ptr);

/* Workspace (bytes) for pamd_qr_r_f64 on an m x ncols matrix whose first n
 * columns are factored. */
int pamd_qr_r_workspace(int64_t m, int64_t n, int64_t ncols, int64_t* bytes);

/* R factor of the first n columns of A (m x ncols, n <= m <= 8192) by
 * communication-avoiding Householder QR: on exit R, with a non-negative
 * diagonal, is in the upper triangle of A[:n, :n] and A[:n, n:ncols] holds
 * the first n rows of Q^T times the input columns (so appending S*b gives
 * z = Q^T S b in A[:n, n]).  The rest of A is overwritten; Q is not kept.
 * Replaces la.qr at drivers/least_squares.py:311 with the Q[:d].T @ b_ske of
 * :315 fused, whose callers use R and z only; R differs from LAPACK's by the
 * signs of its rows.  Returns PAMD_ERR_UNSUPPORTED for m > 8192 (use
 * pamd_geqrf_f64 there). */
int pamd_qr_r_f64(int64_t m, int64_t n, int64_t ncols, double* A, int64_t lda,
                  void* workspace, void* stream);

/* Pointer (device, int32) to the flag pamd_qr_r_f64 sets when R has an exact
 * zero on its diagonal (inside `workspace`; host call, no GPU work). */
int pamd_qr_r_singular_flag(const void* workspace, int64_t m, int** flag_ptr);

/* out = A diag(s) (invert = 0) or A diag(s)^{-1} (invert = 1), A rows x cols
 * (out may alias A).  Forms the SVD right preconditioner M = V / sigma of
 * comps/preconditioning.py:77 and its column-scaled variants
 * (comps/determiter/saddle.py:138-143). */
int pamd_scale_cols_f64(int64_t rows, int64_t cols, const double* A, int64_t lda,
                        const double* s, int invert, double* out, int64_t ldo,
                        void* stream);

/* Workspace (bytes) for pamd_svd_jacobi_f64 (1 <= n <= m <= 1024). */
int pamd_svd_jacobi_workspace(int64_t m, int64_t n, int64_t* bytes);

/* Thin SVD A = U diag(s) Vt of A (m x n, n <= m <= 1024) by one-sided block
 * Jacobi: U (m x n, may be NULL), s (n, descending), Vt (n x n, rows are the
 * right singular vectors, may be NULL).  Columns of U for zero singular values
 * are zero.  Up to max_sweeps (<= 64) sweeps are enqueued; sweeps after the
 * first that rotates nothing return at once, so nothing is read back.
 * info[0] (device int) = 0, or 1 when the last sweep still rotated.
 * Replaces la.svd at drivers/svd.py:164, comps/preconditioning.py:74 and the
 * SVD-based la.lstsq of drivers/least_squares.py:186 (on the triangular
 * factor of a Householder QR when the matrix is larger). */
int pamd_svd_jacobi_f64(int64_t m, int64_t n, const double* A, int64_t lda,
                        double* U, int64_t ldu, double* s, double* Vt, int64_t ldvt,
                        int max_sweeps, void* workspace, int* info, void* stream);

/* pamd_svd_jacobi_f64 without a workspace argument and without allocating:
 * the kernels work in a file-scope device array of the library (16.8 MB per
 * GPU, held until the process ends), so m <= 1024 is the limit (larger m
 * return -1).  No counterpart in the reference: this entry point is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads) race; the caller serialises them.  Same results, bit for
 * bit, as pamd_svd_jacobi_f64. */
int pamd_svd_jacobi_noalloc_f64(int64_t m, int64_t n, const double* A, int64_t lda,
                                double* U, int64_t ldu, double* s, double* Vt,
                                int64_t ldvt, int max_sweeps, int* info, void* stream);

/* out (n x n) = C + shift * I for the symmetric C whose lower triangle A
 * holds (the upper triangle of A is not read, as scipy.linalg.eigh).  First
 * step of the eigendecomposition that replaces la.eigh at drivers/evd.py:279. */
int pamd_sym_shift_f64(int64_t n, const double* A, int64_t lda, double shift,
                       double* out, int64_t ldo, void* stream);

/* Eigenvalues of a symmetric C from the SVD C + shift * I = U diag(s) Vt:
 * lamb_j = s_j sign(u_j . v_j) - shift (u_j column j of U, v_j row j of Vt;
 * the eigenvectors are the v_j).  info[0] (device int, may be NULL) is raised
 * to 2 when, for some s_j > 1e-10 s_0, |u_j . v_j| differs from 1 by more than
 * 1e-6: an eigenvalue pair lambda, -lambda of nearly equal magnitude mixed the
 * singular vectors, and the SVD must be redone with shift >= max |lambda|.
 * lamb must not alias s.  Second step of the la.eigh replacement
 * (drivers/evd.py:279). */
int pamd_eig_from_svd_f64(int64_t n, const double* U, int64_t ldu, const double* Vt,
                          int64_t ldvt, const double* s, double shift, double* lamb,
                          int* info, void* stream);

/* out (cols x rows) = A^T for A (rows x cols); plumbing for A.T @ S products. */
int pamd_transpose_f64(int64_t rows, int64_t cols, const double* A, int64_t lda,
                       double* B, int64_t ldb, void* stream);

/* In-place Cholesky A = R^T R of the symmetric positive definite n x n A
 * (upper triangle read and overwritten with R).  info[0] (device int) = 0, or
 * j + 1 when the j-th pivot is not positive (factorisation stopped).  Used by
 * CholeskyQR2 inside the QB stabilizer (utils/linalg_wrappers.py:6-7);
 * n <= 512. */
int pamd_potrf_upper_f64(int64_t n, double* A, int64_t lda, int* info,
                         void* stream);

/* out = A + alpha B + sigma I (upper_half 0) or H(A + alpha B + sigma I)
 * (upper_half 1: the upper triangle with the diagonal halved, zeros below),
 * k x k with k <= 512, B optional; nrm (optional, one device double) gets
 * ||A + alpha B + sigma I||_F.  One launch for the small steps of the Newton
 * second round of CholeskyQR2 inside the QB stabilizer
 * (utils/linalg_wrappers.py:6-7). */
int pamd_small_combine_f64(int64_t k, const double* A, int64_t lda, const double* B,
                           int64_t ldb, double alpha, double sigma, int upper_half,
                           double* out, int64_t ldo, double* nrm, void* stream);

/* out[0..3] (device doubles) = min |G_ii|, max |G_ii|, enorm[0] (0 when
 * enorm is NULL), info[0]: the CholeskyQR2 acceptance figures (pivot,
 * conditioning estimate, size of the second-round correction) for a single
 * read-back; 1 <= k <= 512. */
int pamd_cholqr_stats_f64(int64_t k, const double* G, int64_t ldg, const int* info,
                          const double* enorm, double* out, void* stream);

/* Rinv (n x n, zeros below the diagonal) = R^{-1} for upper-triangular R
 * (n <= 2048); one triangular solve per column, all columns in parallel. */
int pamd_trtri_upper_f64(int64_t n, const double* R, int64_t ldr, double* Rinv,
                         int64_t ldi, void* stream);

/* In place B (m x n) <- B R^{-1} with R upper triangular n x n (n <= 2048):
 * the Q = Y R^{-1} step of CholeskyQR. */
/* X = Y R^{-1} (R upper k x k, k <= 512; X may alias Y) on FP64 MFMA: one
 * workgroup per 16-row strip held in LDS, 16-column blocks solved left to
 * right with the inverses of R's 16 x 16 diagonal blocks (workspace from
 * pamd_trsm_mfma_workspace).  The Y R^{-1} step of CholeskyQR (orth,
 * utils/linalg_wrappers.py:6-7) in one launch, without forming R^{-1}. */
int pamd_trsm_mfma_workspace(int64_t k, int64_t* bytes);
int pamd_trsm_right_upper_mfma_f64(int64_t m, int64_t k, const double* R,
                                   int64_t ldr, const double* Y, int64_t ldy,
                                   double* X, int64_t ldx, void* workspace,
                                   void* stream);

int pamd_trsm_right_upper_f64(int64_t m, int64_t n, const double* R,
                              int64_t ldr, double* B, int64_t ldb, void* stream);

/* ------------------------------------------------------------------------ */
/* Preconditioned LSQR (device-resident recurrences)                         */
/* ------------------------------------------------------------------------ */

typedef struct pamd_lsqr_problem {
  int64_t m, n;          /* A is m x n */
  const double* A;       /* device, row-major */
  int64_t lda;
  double sqrt_delta;     /* > 0: operator is [A; sqrt_delta I] (a_lift) */
  const double* R;       /* device upper-triangular n x n, or NULL (no preconditioner) */
  int64_t ldr;
  const double* b;       /* device, m entries (lifted zeros are implicit) */
  const double* x0;      /* device, n entries, or NULL */
  double damp, atol, btol, conlim;
  int64_t iter_lim;
  int calc_var;
} pamd_lsqr_problem;

typedef struct pamd_lsqr_result {  /* host struct */
  int64_t istop, itn;
  double r1norm, r2norm, anorm, acond, arnorm, xnorm;
  int64_t n_log;         /* number of valid entries in arnorms */
} pamd_lsqr_result;

/* Workspace (bytes) for pamd_lsqr_run_f64. */
int pamd_lsqr_workspace(int64_t m, int64_t n, int64_t iter_lim, int64_t* bytes);

/* Run LSQR (comps/determiter/lsqr.py:98-574, the PARLA-modified Paige-Saunders
 * recurrences, same stopping tests and istop codes) on the operator
 * A_pc = [A; sqrt_delta I] M, M = R^{-1} (comps/preconditioning.py:16-67),
 * with right-hand side [b; 0], starting from x0 (preconditioned coordinates,
 * the z0 of comps/determiter/saddle.py:197).  M is applied by triangular
 * substitution with R (one R x = y and one R^T x = y solve per iteration, as
 * la.solve_triangular does at preconditioning.py:28,37), never as an explicit
 * inverse.  Writes x (n, preconditioned coordinates), x_orig (n, optional:
 * R^{-1} x, i.e. the M_fwd(result[0]) of saddle.py:198), var (n,
 * if calc_var), arnorms (iter_lim; the per-iteration ||A_pc^T r|| log of
 * lsqr.py:413) and fills *result (host).  All recurrences run on the device;
 * this call blocks, reading the device state back once per batch of
 * iterations.  Any n < 2^30: up to n = 1900 the triangular solves keep their
 * vector in LDS, beyond that in the workspace (slower, same arithmetic). */
int pamd_lsqr_run_f64(const pamd_lsqr_problem* prob, double* x, double* x_orig,
                      double* var, double* arnorms, pamd_lsqr_result* result,
                      void* workspace, void* stream);

/* The same LSQR (comps/determiter/lsqr.py:98-574 on the operator of
 * pamd_lsqr_run_f64) as launches only: no host synchronisation and no
 * read-back, so a caller can capture it in a HIP graph.  All iter_lim
 * iterations are launched; those after the stopping test fired return at
 * once.  *result is DEVICE memory (sizeof(pamd_lsqr_result) bytes), written
 * by the last launch.  Results equal pamd_lsqr_run_f64's bit for bit. */
int pamd_lsqr_launch_f64(const pamd_lsqr_problem* prob, double* x, double* x_orig,
                         double* var, double* arnorms, pamd_lsqr_result* result,
                         void* workspace, void* stream);

/* ------------------------------------------------------------------------ */
/* Mixed precision: exact BF16 sketch, compensated residual, CG operator      */
/* ------------------------------------------------------------------------ */

/* Shapes of the exact BF16 sketch of d rows over m rows of [A | b] (ncols
 * columns): kp = m rounded up to 256 (K padding), dpad = d rounded up to 128
 * (rows of Q's array), ncp = ncols rounded up to 64 (rows of each piece
 * array), group_chunks = 256-row chunks per K group (the FP64 summation
 * order), and the workspace of pamd_ozaki_sketch_f64. */
int pamd_ozaki_layout(int64_t d, int64_t m, int64_t ncols, int64_t* kp, int64_t* dpad,
                      int64_t* ncp, int64_t* group_chunks, int64_t* workspace_bytes);

/* Quantised Gaussian sketch Q: Q[i][k] = rint(32 z_ik) clipped to [-256, 256]
 * with z the same Philox stream as pamd_gaussian_fill_f64 (counter
 * offset + i*m + k), stored as BF16 bit patterns in the blocked layout of
 * pamd_split_bf16 (rows_out x kp, rows_out a multiple of 16, zeros for
 * i >= rows or k >= m).  The sketching operator is S = Q * 2^-5 /
 * sqrt(min(rows, m)), the normalisation of utils/sketching.py:23 on an
 * integer grid, so that the BF16 sketch product below is exact.  Replaces
 * rng.normal of utils/sketching.py:24 for the mixed-precision sketch. */
int pamd_gauss_q_bf16(int64_t rows, int64_t rows_out, int64_t m, uint64_t seed,
                      uint64_t offset, uint16_t* Q, void* stream);

/* Error-free split of [A | b] (m x n, plus b as column n when b != NULL) into
 * npieces (1..5) BF16 integer pieces per 256-row chunk and column:
 * A[k][j] ~= 2^(e_cj - 8) sum_p 2^(-9p) a_p[k][j], |a_p| <= 256, with the
 * exponent e_cj (max |A[chunk c, j]| < 2^e_cj) in exps[c * lde + j].
 * pieces: npieces blocked arrays (pstride >= ncp * kp elements apart) of
 * ncp x kp, row j holding column j's K run.  First half of the exact sketch
 * of A (drivers/least_squares.py:303,314 in mixed precision). */
int pamd_ozaki_split_bf16(int64_t m, int64_t n, const double* A, int64_t lda,
                          const double* b, int npieces, uint16_t* pieces,
                          int64_t pstride, int* exps, int64_t lde, void* stream);

/* out (d x ncols, FP64) = scale * Q @ [A | b] from the split above, on
 * v_mfma_f32_16x16x32_bf16: every 256-long partial sum of integer products is
 * exact in FP32 and is added in FP64 with its chunk scale, chunk after chunk
 * within a K group of group_chunks chunks; the group sums are then added in
 * group order and scaled.  The result equals scale * Q @ (A truncated to
 * 9*npieces bits below each chunk-column maximum) up to those FP64
 * additions.  Q has dpad rows (pamd_ozaki_layout).  Replaces S @ A and S @ b
 * of drivers/least_squares.py:303,314. */
int pamd_ozaki_sketch_f64(int64_t d, int64_t m, int64_t ncols, int npieces,
                          const uint16_t* Q, const uint16_t* pieces, int64_t pstride,
                          const int* exps, int64_t lde, double scale, double* out,
                          int64_t ldo, void* workspace, void* stream);

/* q = B^T (B p) in one read of B (m x n, n <= 1024, row-major; dtype 0: FP64,
 * 1: FP32 storage), FP64 vectors and accumulation, deterministic order.  The
 * operator of the refinement's inner CG on the preconditioned normal
 * equations; no counterpart in the reference, whose solver is LSQR
 * (comps/determiter/lsqr.py:98). */
int pamd_normal_pass_workspace(int64_t m, int64_t n, int64_t* bytes);
int pamd_normal_pass(int dtype, int64_t m, int64_t n, const void* B, int64_t ldb,
                     const double* p, double* q, void* workspace, void* stream);

/* Compensated residual, one read of A (m x n, n <= 1024): r = b - A x with
 * error-free products and sums (Dot2), g = A^T r likewise, rss[0] = ||r||^2.
 * r and rss may be NULL.  Replaces the FP64 residuals of the SPO log and
 * presolve (drivers/least_squares.py:344-347, 361-366) where the backward
 * error must be evaluated below the FP64 rounding floor of A x. */
int pamd_residual_workspace(int64_t m, int64_t n, int64_t* bytes);
int pamd_residual_f64(int64_t m, int64_t n, const double* A, int64_t lda,
                      const double* x, const double* b, double* r, double* g,
                      double* rss, void* workspace, void* stream);

/* X (m x n) <- X R^{-1}, R upper n x n, for large m: recursive, 32-column
 * substitution leaves and FP64 MFMA GEMM updates (backward stable like
 * substitution, unlike multiplying by an explicit inverse).  Forms the
 * explicitly preconditioned A R^{-1} that comps/preconditioning.py:16-41
 * applies implicitly; workspace as for pamd_gemm_f64 with M = m. */
int pamd_trsm_right_upper_rec_f64(int64_t m, int64_t n, const double* R,
                                  int64_t ldr, double* X, int64_t ldx,
                                  void* workspace, int64_t workspace_bytes,
                                  void* stream);

/* X = A R^{-1} out of place (A is not modified) with the FP32 rounding of X
 * in X32 (m x n, leading dimension ld32): the same recursion and results as
 * pamd_trsm_right_upper_rec_f64 on a copy of A, without the copy of A and
 * the separate conversion pass.  Forms the FP32-stored preconditioned
 * operator of the mixed-precision solver (comps/preconditioning.py:16-41
 * applies A R^{-1} implicitly); workspace as for pamd_gemm_f64 with M = m. */
int pamd_trsm_right_upper_rec_f32out(int64_t m, int64_t n, const double* R,
                                     int64_t ldr, const double* A, int64_t lda,
                                     double* X, int64_t ldx, float* X32,
                                     int64_t ld32, void* workspace,
                                     int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Learned leverage-score sketch                                             */
/* ------------------------------------------------------------------------ */

/* The MLP surrogate 16 -> 64 -> 64 -> 1 (ReLU) maps a row's 16 pre-sketch
 * features to the log of its normalised leverage score.  Weights wb are BF16
 * bit patterns W1[16][64] | W2[64][64] | w3[64]; biases FP32 b1[64] | b2[64] |
 * b3; products and sums FP32.  The flat FP32 parameter / gradient vector is
 * W1 | b1 | W2 | b2 | w3 | b3 (pamd_mlp_nparams() entries).  No counterpart
 * in the reference; it feeds the row-sampling operator that SkOpIN /
 * sampling_operator (comps/sketchers/oblivious.py:75-82,
 * utils/sketching.py:204-232) expose as an API slot. */
int pamd_mlp_nparams(void);

/* y[r] = MLP(X[r, 0:16]) for r < rows (X FP32 row-major, ldx >= 16). */
int pamd_mlp_fwd(int64_t rows, const float* X, int64_t ldx, const uint16_t* wb,
                 const float* bias, float* y, void* stream);

/* grad (nparams + 1 FP32) = d/dparams of (1/rows_total) sum_r (MLP(X_r) - t_r)^2
 * over these rows; grad[nparams] = the loss contribution.  Deterministic. */
int pamd_mlp_bwd_workspace(int64_t rows, int64_t* bytes);
int pamd_mlp_bwd(int64_t rows, int64_t rows_total, const float* X, int64_t ldx,
                 const float* t, const uint16_t* wb, const float* bias,
                 float* grad, void* workspace, void* stream);

/* One Adam step (bias-corrected, step >= 1) on the FP32 masters w with
 * moments m, v; then wb / bias are refreshed (BF16 round-to-nearest-even). */
int pamd_mlp_adam(float* w, float* m, float* v, const float* grad, float lr,
                  float beta1, float beta2, float eps, int64_t step,
                  uint16_t* wb, float* bias, void* stream);

/* out[r, :] = scale[r] * A[idx[r], :] (scale may be NULL: 1): the row-sampling
 * sketch S @ A of utils/sketching.py:204-232 with the sampling weights of a
 * leverage-score distribution. */
int pamd_row_gather_f64(int64_t d, int64_t m, int64_t n, const int64_t* idx,
                        const double* scale, const double* A, int64_t lda,
                        double* out, int64_t ldo, void* stream);

/* out[idx[r], :] = A[r, :] for r < d (out has m rows, the other rows are left
 * as they are; the indices must be distinct, 0 <= idx[r] < m): the adjoint of
 * the row selection, S^T @ Y for the wide SkOpIN operator and S @ X for the
 * tall one (comps/sketchers/oblivious.py:75-82, utils/sketching.py:204-232,
 * whose rmatvec this is). */
int pamd_row_scatter_f64(int64_t d, int64_t m, int64_t n, const int64_t* idx,
                         const double* A, int64_t lda, double* out, int64_t ldo,
                         void* stream);

/* Workspace (bytes) for pamd_sample_rows_f64 (the CDF: m doubles). */
int pamd_sample_rows_workspace(int64_t m, int64_t* bytes);

/* d row indices drawn with replacement from the weights p (m >= 1, p_i >= 0,
 * sum > 0; need not be normalised) by inverse CDF on Philox-4x32-10 uniforms
 * (counter = draw index, key = seed), and the scales 1 / sqrt(d p_idx / sum p)
 * of the sampling sketch.  The CDF is summed in a fixed order, so the draws
 * depend only on (p, d, seed).  The weighted row-sampling counterpart of
 * sampling_operator (utils/sketching.py:204-232) for the learned sketch; no
 * reference call site draws weighted rows. */
int pamd_sample_rows_f64(int64_t m, const double* p, int64_t d, uint64_t seed,
                         int64_t* idx, double* scale, void* workspace, void* stream);

/* Batched potrf / trtri: problem i is the n[i] x n[i] matrix (ld n[i]) at
 * element offset off[i] of the pool (n and off are device arrays; n[i] <= 512
 * for potrf, nmax = max n[i]).  One workgroup per problem (per column for
 * trtri), one launch for the whole batch; info[i] as pamd_potrf_upper_f64.
 * The factorisations behind the per-problem CholeskyQR of a batch of
 * sketches (drivers/least_squares.py:311, one problem per call there). */
int pamd_potrf_batched_f64(int64_t nprob, const int64_t* n, const int64_t* off,
                           double* pool, int* info, void* stream);
int pamd_trtri_batched_f64(int64_t nprob, int64_t nmax, const int64_t* n,
                           const int64_t* off, const double* R, double* Rinv,
                           double* RinvT, void* stream);  /* RinvT: the transposes, or NULL */

/* ------------------------------------------------------------------------ */
/* Batched least squares over problems of uneven size                        */
/* ------------------------------------------------------------------------ */

/* One problem of a batch (no counterpart in the reference, whose drivers
 * solve one problem per call, drivers/least_squares.py:206).  A_i is FP32
 * row-major with lda = n (n <= 512) at element offset a_off of the A pool;
 * its m-vectors (b) start at mv_off of the m-vector pool, its n-vectors at
 * nv_off of every n-vector pool, R_i (n x n FP64, ld n) at r_off; its
 * 64-row blocks write partial vectors of 2n + 2 doubles from part_off of
 * the partial pool; blk0 is the index of its first block. */
typedef struct pamd_ls_desc {
  int64_t m, n, a_off, mv_off, nv_off, r_off, blk0, part_off;
} pamd_ls_desc;

/* A 64-row block of problem `prob` starting at row `row0`. */
typedef struct pamd_ls_block {
  int32_t prob, row0;
} pamd_ls_block;

/* Row-block passes over all problems whose n has ceil(n/64) == npl
 * (npl in 1, 2, 4, 8; blocks lists only those problems' blocks).  The launch
 * shape (waves_per_wg in 1, 2, 4; blocks_per_wave 1..64) is the autotuned
 * choice and does not change any result bit.
 *   normal:   partial of A_i^T (A_i p_i) per block;
 *   residual: partial of the compensated g_i = A_i^T (b_i - A_i x_i) and
 *             ||r_i||^2 per block.  Replace LSQR's matvecs
 *             (comps/determiter/lsqr.py:345-380) for a whole batch.
 * state (the CG state of pamd_bt_cg, or NULL): blocks of problems whose
 * active flag is 0 are skipped and keep their previous partials. */
int pamd_bt_normal(int npl, const pamd_ls_desc* descs, const pamd_ls_block* blocks,
                   int64_t nblk, int waves_per_wg, int blocks_per_wave,
                   const float* A, const double* p, const double* state, double* part,
                   void* stream);
int pamd_bt_residual(int npl, const pamd_ls_desc* descs, const pamd_ls_block* blocks,
                     int64_t nblk, int waves_per_wg, int blocks_per_wave,
                     const float* A, const double* b, const double* x,
                     const double* state, double* part, void* stream);

/* Per-problem sum of the block partials in block order into out (n-vector
 * pool); compensated = 1 also folds the compensation terms and writes
 * ||r_i||^2 to scal[i]. */
int pamd_bt_reduce(int64_t nprob, const pamd_ls_desc* descs, const double* part,
                   int compensated, double* out, double* scal, void* stream);

/* v_i <- R_i^{-1} v_i (trans 0) or R_i^{-T} v_i (trans 1) for every problem:
 * the preconditioner solves of comps/preconditioning.py:28,37, batched. */
int pamd_bt_trsv(int64_t nprob, const pamd_ls_desc* descs, int trans,
                 const double* R, double* v, void* stream);

/* v_i <- Rinv_i v_i (trans 0) or Rinv_i^T v_i (trans 1) for every problem,
 * Rinv_i = R_i^{-1} (from pamd_trtri_batched_f64, laid out like R) and
 * RinvT_i its transpose (both read as coalesced row segments): the same
 * preconditioner (comps/preconditioning.py:28,37) applied as a product, for
 * the refinement's inner CG, which applies it twice per step.  The iteration
 * then runs on the fixed operator Rinv_i^T A_i^T A_i Rinv_i; the rounding of
 * Rinv_i (of order u kappa(R_i)) changes the preconditioner, not the normal
 * equations the refinement solves. */
int pamd_bt_trmv(int64_t nprob, const pamd_ls_desc* descs, int trans,
                 const double* Rinv, const double* RinvT, double* v, void* stream);

/* Conjugate gradients on every problem's preconditioned normal equations,
 * recurrences on the device; state[4 i] = r.r, state[4 i + 1] = 1 while
 * problem i is active (0 freezes its vectors).  mode 0 starts from z = 0 with
 * residual rhs, mode 1 takes one step given q = M p. */
int pamd_bt_cg(int64_t nprob, const pamd_ls_desc* descs, int mode, double* state,
               const double* rhs, double* r, double* p, const double* q,
               double* z, void* stream);

/* x_i += z_i for active problems; v_i = a_i for all problems. */
int pamd_bt_axpy(int64_t nprob, const pamd_ls_desc* descs, const double* state,
                 double* x, const double* z, void* stream);
int pamd_bt_copy(int64_t nprob, const pamd_ls_desc* descs, const double* a,
                 double* v, void* stream);

/* ---- batched sketch and factorisation ------------------------------------
 * The setup half of the batched solve: every problem's sketch Y_i = S_i A_i
 * with its own SJLT (the column draws of pamd_sjlt_draw with seed_i), and the
 * shifted CholeskyQR3 of Y_i, for the whole batch per launch.  Together they
 * replace the per-problem sketch_op_gen + la.qr of drivers/least_squares.py:
 * 302-311.  Work items are numbered problem after problem; pamd_sk_desc
 * carries each problem's first item of every kind and its pool offsets
 * (item counts: ceil(m_i / chunk_rows) chunks, d_i sketch rows, T(T+1)/2
 * upper tiles with T = ceil(n_i / tile), ceil(d_i / strip_rows) strips,
 * ceil(n_i / strip_rows) diagonal blocks; pamd_bt_layout returns the
 * constants).  The caller builds the descriptors; the kernels trust them. */
typedef struct pamd_sk_desc {
  int64_t d;        /* sketch rows, vec_nnz <= d_i <= max_d (or d_i < vec_nnz: with replacement) */
  int64_t y_off;    /* FP64 pool offset of Y_i (d_i x n_i, ld n_i) */
  int64_t rp_off;   /* int64 row-pointer pool offset (d_i + 1 entries) */
  int64_t cnt_off;  /* int32 chunk-histogram pool offset (chunks_i x d_i) */
  int64_t chunk0, row0, tile0, strip0, dblk0;  /* first work item of each kind */
  uint64_t seed;    /* SJLT seed of problem i */
} pamd_sk_desc;

int pamd_bt_layout(int* chunk_rows, int* tile, int* strip_rows, int* max_d, int* max_vec_nnz);

/* CSR of every S_i (d_i x m_i, vec_nnz = k per column, entries +-1/sqrt(k)),
 * each row in column order; problem i's entries start at k * mv_off_i of
 * cols/vals and its row pointers (relative to that start) at rp_off_i.
 * Replaces utils/sketching.py:34-80 for a batch. */
int pamd_bt_sjlt_csr(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                     int64_t nchunk, int k, int32_t* cnt, int64_t* rowptr, int32_t* cols,
                     double* vals, void* stream);
/* Y_i = S_i A_i (FP32 A pool, FP64 sums in column order): the sparse
 * S @ A of drivers/least_squares.py:303, batched. */
int pamd_bt_sjlt_left(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                      int64_t nrows, int k, const int64_t* rowptr, const int32_t* cols,
                      const double* vals, const float* A, double* Y, void* stream);
/* G_i = Y_i^T Y_i into an n_i x n_i pool at r_off_i (upper tiles; the
 * diagonal tiles whole), plus 11 (d n + n (n + 1)) u trace(G_i) on the
 * diagonal when shifted = 1.  Factor G_i with pamd_potrf_batched_f64. */
int pamd_bt_gram(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                 int64_t ntiles, int shifted, const double* Y, double* G, void* stream);
/* Y_i <- Y_i R_i^{-1} in place (R pool at r_off_i, upper, n_i <= nmax <= 512);
 * dinv: 256 doubles per diagonal block (ndblk blocks). */
int pamd_bt_trsm(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                 int64_t nstrips, int64_t ndblk, int nmax, const double* R, double* dinv,
                 double* Y, void* stream);
/* C_i = A_i B_i for upper-triangular A_i, B_i (entries below the diagonal
 * are not read); writes C's upper tiles, zeros below the diagonal.  Builds
 * R = R_2 R_1 R_0 of CholeskyQR3. */
int pamd_bt_trmm_upper(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                       int64_t ntiles, const double* A, const double* B, double* C,
                       void* stream);
/* out[i] = ||A_i||_F^2 (FP32 pool, FP64 sums, fixed order); part: nchunk
 * doubles.  The norm in the backward error, batched. */
int pamd_bt_fro2(int64_t nprob, const pamd_ls_desc* descs, const pamd_sk_desc* sk,
                 int64_t nchunk, const float* A, double* part, double* out, void* stream);
/* out[i] = a_i . b_i over the n-vector pool. */
int pamd_bt_dot(int64_t nprob, const pamd_ls_desc* descs, const double* a, const double* b,
                double* out, void* stream);

/* ---- split-precision GEMM (BF16 pieces on BF16 MFMA) ----------------------
 * An FP64 operand is held as npieces (2 or 3) BF16 pieces x = x_0 + x_1 (+ x_2),
 * x_p the BF16 rounding of what x_0 .. x_{p-1} leave; C = L R^T sums the
 * piece products L_p R_q^T with p + q < npieces in FP32 on
 * v_mfma_f32_16x16x32_bf16 (split-K partials summed in FP64).  16 (24)
 * significant bits per element for 2 (3) pieces.  It runs the products of
 * the power iteration and of B = Q^T A (comps/sketchers/aware.py:175,179;
 * comps/rangefinders.py:186; comps/qb.py:351) in BF16 arithmetic.
 * Piece arrays: npieces arrays (stride pstride elements) of rows x kp BF16
 * (uint16 bits) in 1 KiB blocks of 16 rows x 32 columns, element (i, k) at
 *   ((i/16)(kp/32) + k/32) 512 + (i%16) 32 + 8 (c ^ ((i%16)/4)) + k%8,
 * c = (k%32)/8; zero beyond the matrix.  kp is a multiple of 64. */

/* kp, the rows of L's (mpad) and R's (npad) piece arrays, and the workspace
 * of pamd_split_gemm_f64, for C (M x N) with K reduced. */
int pamd_split_layout(int npieces, int64_t M, int64_t N, int64_t K, int64_t* kp,
                      int64_t* mpad, int64_t* npad, int64_t* workspace_bytes);
/* Pieces of X (rows x cols, row-major FP64) into rows_out x kp arrays
 * (rows_out a multiple of 16), or of X^T (cols x rows) with transpose = 1. */
int pamd_split_bf16(int transpose, int npieces, int64_t rows, int64_t cols, const double* X,
                    int64_t ldx, int64_t rows_out, int64_t kp, uint16_t* pieces,
                    int64_t pstride, void* stream);
/* C[i * lci + j * lcj] = sum_{p+q<npieces} sum_k L_p[i][k] R_q[j][k] for
 * i < M, j < N (lcj = 1: row-major C; lci = 1: C stored transposed); L has
 * mpad rows, R npad rows, both kp columns (pamd_split_layout). */
int pamd_split_gemm_f64(int npieces, int64_t M, int64_t N, int64_t K, const uint16_t* L,
                        int64_t lps, const uint16_t* R, int64_t rps, double* C, int64_t lci,
                        int64_t lcj, void* workspace, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PARLA_AMD_H */
