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
 *    re-entrant (no hidden global state besides the HIP runtime).
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

/* In-place triangular solve with the upper-triangular n x n R:
 * trans = 0: x <- R^{-1} x;  trans = 1: x <- R^{-T} x.
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

/* Rinv (n x n, zeros below the diagonal) = R^{-1} for upper-triangular R
 * (n <= 2048); one triangular solve per column, all columns in parallel. */
int pamd_trtri_upper_f64(int64_t n, const double* R, int64_t ldr, double* Rinv,
                         int64_t ldi, void* stream);

/* In place B (m x n) <- B R^{-1} with R upper triangular n x n (n <= 2048):
 * the Q = Y R^{-1} step of CholeskyQR. */
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
 * iterations. */
int pamd_lsqr_run_f64(const pamd_lsqr_problem* prob, double* x, double* x_orig,
                      double* var, double* arnorms, pamd_lsqr_result* result,
                      void* workspace, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PARLA_AMD_H */
