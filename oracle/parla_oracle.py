"""numpy/scipy restatement of PARLA's randomized-NLA hot path (the oracle).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py``; never by the GPU product path.

Every function cites the reference file:line it restates (paths relative to
the PARLA source tree).  Random draws are made through numpy's public
``Generator`` API in the same order as the reference, so sketch indices, signs
and Gaussian values are bit-identical to what the reference draws for the same
seed (numpy 2.2.6; numpy does not promise stream stability across versions).

Parity status of each function (how it is pinned, see tests/test_oracle_*.py):

* gaussian_operator, sjlt_operator, sjlt_draw -- golden vectors written by the
  reference itself (tests/golden/make_golden.py) + distributional invariants.
* orthonormal_operator -- Q^T Q = I, diag(R) > 0 makes Q unique.
* lsqr, sym_ortho -- the three known-answer cases of the reference docstring
  (lsqr.py:259-305), golden trajectories written by the reference's lsqr,
  scipy.sparse.linalg.lsqr on unpreconditioned problems, closed-form LS.
* spo -- closed-form solutions of the reference's own test problems,
  presolve exactness on consistent problems, convergence-rate fits.
* rs1/rf1/qb1/svd1/evd1 -- exact low-rank recovery, orthonormality,
  B = Q^T A, power-method convergence (the reference's test families).
* sampling_operator -- brute force (index gather == dense selection matrix).
* quantize_q, ozaki_split, ozaki_product -- exact integer arithmetic: the
  pieces reassemble A to within 2^-9P of each chunk-column maximum and the
  product equals an int64 evaluation (tests/test_oracle.py).
* residual_exact -- correctly rounded (exact products + math.fsum), checked
  against exact rational arithmetic; residual_extended against it.
* spo_mixed -- the FP64 least-squares solution (scipy lstsq) and the target
  backward error on the reference's test families.
* mlp_forward, mlp_loss_grad -- central finite differences of the loss;
  bf16_round -- exhaustive comparison with torch's bfloat16 rounding;
  leverage_scores -- sum = n, invariance under right scaling, closed forms.
* bf16_pieces, split_product -- the reconstruction bound of the pieces, and
  exact rational evaluation of the piece products on small cases.
* philox4x32_10 -- the known-answer vectors published with Random123 (Salmon
  et al., SC'11); u53_open0 -- its range and exactness at the end points;
  device_gaussian_fill, device_sjlt_draw, device_sample_rows (the device
  draws, which have no counterpart in the reference) -- moments,
  distinctness, offset invariance and the frequencies of the inverse CDF.
"""
import time
import warnings

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp

EPS = np.finfo(np.float64).eps

__all__ = [
    "orthonormal_operator", "gaussian_operator", "sjlt_operator", "sjlt_draw",
    "sampling_operator", "dim_checks", "sym_ortho", "lsqr", "DenseOp",
    "a_lift_precond", "pcss2", "SketchAndPrecondLog", "spo", "orth",
    "lu_stabilize", "rs1", "rf1", "qb1", "svd1", "evd1", "loglinear_fit",
    "loglog_fit", "backward_error", "EPS", "quantize_q", "ozaki_split", "ozaki_product",
    "residual_extended", "residual_exact", "backward_error_extended", "spo_mixed", "bf16_round",
    "mlp_unpack", "mlp_effective", "mlp_forward", "mlp_loss_grad", "leverage_scores", "coherent_matrix",
    "MLP_NPARAM", "bf16_pieces", "bf16_values", "split_product",
    "svd_right_precond", "pcss2_under", "spu1", "sso1", "pcg", "pcss1", "pcss2_saddle", "sps1", "sps2",
    "project_out", "qb2", "qb3", "evd2", "sparse_sign_operator", "generate_srct", "apply_srct",
    "srct_matrix", "philox4x32_10", "u53_open0", "device_gaussian_fill", "device_sjlt_draw",
    "device_sample_rows",
]


# --------------------------------------------------------------------------
# Sketching operators  (parla/utils/sketching.py)
# --------------------------------------------------------------------------

def gaussian_operator(n_rows, n_cols, rng, normalize=True):
    """utils/sketching.py:20-31.  N(0, 1/min(r, c)) entries, or N(0, 1)."""
    rng = np.random.default_rng(rng)
    if not normalize:
        return rng.standard_normal((n_rows, n_cols))
    sigma = np.sqrt(1.0 / min(n_rows, n_cols))
    return rng.normal(0.0, sigma, (n_rows, n_cols))


def orthonormal_operator(n_rows, n_cols, rng):
    """utils/sketching.py:9-17.  Haar factor: QR of a Gaussian, signs of diag(R)."""
    if n_rows < n_cols:
        return orthonormal_operator(n_cols, n_rows, rng).T
    rng = np.random.default_rng(rng)
    G = gaussian_operator(n_rows, n_cols, rng)
    Q, R = la.qr(G, overwrite_a=True, pivoting=False, mode="economic")
    return Q * np.sign(np.diag(R))


def sjlt_draw(short, long, rng, vec_nnz=8):
    """Raw draws of a wide SJLT (``short`` x ``long``), utils/sketching.py:51-71.

    Returns ``rows`` (long, k) int64 -- the row indices of column j -- and
    ``vals`` (long, k) float64 = +-1/sqrt(k).  The draw order is the
    reference's: one ``choice`` per column (without replacement unless the
    column is shorter than k, :56-65), then one ``random`` block for the
    signs, ``<= 0.5`` giving -1 (:69-70).
    """
    rng = np.random.default_rng(rng)
    k = min(long, vec_nnz)
    with_repl = short < k
    if with_repl:
        warnings.warn(f"Can't set {k} nonzeros per column for columns of length "
                      f"{short}. Sampling indices with replacement instead.")
    rows = np.empty((long, k), dtype=np.int64)
    for j in range(long):
        rows[j] = rng.choice(short, k, replace=with_repl)
    vals = np.ones(long * k)
    vals[rng.random(long * k) <= 0.5] = -1.0
    vals /= np.sqrt(k)
    return rows, vals.reshape(long, k)


def sjlt_operator(n_rows, n_cols, rng, vec_nnz=8):
    """utils/sketching.py:34-80.  Wide: CSC; tall: transpose of the wide draw, CSR."""
    if n_cols >= n_rows:
        rows, vals = sjlt_draw(n_rows, n_cols, rng, vec_nnz)
        k = rows.shape[1]
        cols = np.repeat(np.arange(n_cols), k)
        S = sp.coo_matrix((vals.ravel(), (rows.ravel(), cols)),
                          shape=(n_rows, n_cols))
        return S.tocsc()
    return sjlt_operator(n_cols, n_rows, rng, vec_nnz).T.tocsr()


def sparse_sign_operator(n_rows, n_cols, rng, density=0.05):
    """utils/sketching.py:83-103 as a dense array (same numpy draws)."""
    rng = np.random.default_rng(rng)
    nz = rng.random(n_rows * n_cols) < density
    attempt = 0
    while np.all(~nz):
        if attempt == 10:
            raise RuntimeError('Density too low.')
        nz = rng.random(n_rows * n_cols) < density
        attempt += 1
    rows, cols = np.unravel_index(np.where(nz)[0], (n_rows, n_cols))
    vals = np.ones(rows.size)
    vals[rng.random(vals.size) < 0.5] = -1
    vals /= np.sqrt(min(n_rows, n_cols) * density)
    S = np.zeros((n_rows, n_cols))
    S[rows, cols] = vals
    return S


def generate_srct(n_rows, n_cols, rng):
    """utils/sketching.py:106-116: (r, e, perm) in the reference's draw order."""
    big, small = max(n_rows, n_cols), min(n_rows, n_cols)
    r = rng.choice(big, size=small, replace=False)
    e = rng.random(big)
    e[e > 0.5] = 1.0
    e[e != 1] = -1.0
    e *= np.sqrt(big / small)
    perm = rng.permutation(big)
    return r, e, perm


def apply_srct(r, e, mat, perm=None, forward=True):
    """utils/sketching.py:119-176 with scipy.fft's orthonormal DCT-II / its inverse."""
    from scipy.fft import dct, idct
    mat = np.asarray(mat, dtype=np.float64)
    if forward:
        if perm is not None:
            mat = mat[perm]
        mat = mat * (e[:, None] if mat.ndim > 1 else e)
        return dct(mat, axis=0, norm="ortho")[r]
    shape = (e.size,) if mat.ndim == 1 else (e.size, mat.shape[1])
    out = np.zeros(shape)
    out[r] = mat
    out = idct(out, axis=0, norm="ortho")
    out = out * (e[:, None] if out.ndim > 1 else e)
    if perm is not None:
        out = out[np.argsort(perm)]
    return out


def srct_matrix(n_rows, n_cols, rng):
    """utils/sketching.py:179-201 as a dense array: the SRCT applied to the
    identity (wide), or the transpose of a wide operator (tall)."""
    rng = np.random.default_rng(rng)
    r, e, perm = generate_srct(n_rows, n_cols, rng)
    if n_cols < n_rows:
        # the reference draws (r, e, perm) before it looks at the orientation and
        # the tall operator transposes a second, wide draw (sketching.py:185-200)
        return srct_matrix(n_cols, n_rows, rng).T
    return apply_srct(r, e, np.eye(n_cols), perm)


def sampling_operator(n_rows, n_cols, rng, indices=None):
    """utils/sketching.py:204-232, as a dense index array.

    Returns ``(indices, wide)``: the sorted sample of ``max(r, c)`` of size
    ``min(r, c)`` and whether the operator is wide.  The reference's
    ``rmatmat`` (:222-224) and tall branch (:230) are broken; this restates
    the intended selection ``S @ X == X[indices]`` for wide S.
    """
    rng = np.random.default_rng(rng)
    if indices is None:
        indices = rng.choice(max(n_rows, n_cols), min(n_rows, n_cols),
                             replace=False)
        indices.sort()
    else:
        assert indices.size == min(n_rows, n_cols)
    return indices, n_cols >= n_rows


# --------------------------------------------------------------------------
# LSQR  (parla/comps/determiter/lsqr.py)
# --------------------------------------------------------------------------

def sym_ortho(a, b):
    """Stable Givens rotation (c, s, r), comps/determiter/lsqr.py:63-95."""
    if b == 0:
        return np.sign(a), 0, abs(a)
    if a == 0:
        return 0, np.sign(b), abs(b)
    if abs(b) > abs(a):
        t = a / b
        s = np.sign(b) / np.sqrt(1 + t * t)
        return s * t, s, b / s
    t = b / a
    c = np.sign(a) / np.sqrt(1 + t * t)
    return c, c * t, a / c


class DenseOp:
    """Minimal operator with matvec / rmatvec over a dense array."""

    def __init__(self, A):
        self.A = np.asarray(A, dtype=np.float64)
        self.shape = self.A.shape

    def matvec(self, x):
        return self.A @ x

    def rmatvec(self, y):
        return self.A.T @ y


def lsqr(A, b, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8, iter_lim=None,
         calc_var=False, x0=None):
    """Paige-Saunders LSQR as modified by PARLA, comps/determiter/lsqr.py:98-574.

    ``A`` has ``shape``, ``matvec`` and ``rmatvec``.  Returns the 10-tuple
    (x, istop, itn, r1norm, r2norm, anorm, acond, arnorms, xnorm, var) whose
    8th entry is the per-iteration log of ||A^T r|| estimates, recorded at
    the top of each iteration (:410-413, :572).  On the zero-residual early
    exit (:392-395) that entry is the scalar arnorm, as in the reference.
    """
    if not hasattr(A, "matvec"):
        A = DenseOp(A)
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    if b.ndim > 1:
        b = b.squeeze()
    m, n = A.shape
    if iter_lim is None:
        iter_lim = 2 * n
    var = np.zeros(n)
    ctol = 1.0 / conlim if conlim > 0 else 0.0
    dampsq = damp * damp
    # recurrence state (lsqr.py:342-356)
    itn, istop = 0, 0
    anorm = acond = ddnorm = res2 = xnorm = xxnorm = z = 0.0
    cs2, sn2 = -1.0, 0.0

    bnorm = np.linalg.norm(b)
    if x0 is None:
        x = np.zeros(n)
        u = b
        beta = bnorm
    else:
        x = np.asarray(x0, dtype=np.float64)
        u = b - A.matvec(x)
        beta = np.linalg.norm(u)
    if beta > 0:
        u = (1 / beta) * u
        v = A.rmatvec(u)
        alfa = np.linalg.norm(v)
    else:
        v = x.copy()
        alfa = 0.0
    if alfa > 0:
        v = (1 / alfa) * v
    w = v.copy()
    rhobar, phibar = alfa, beta
    rnorm = r1norm = r2norm = beta
    arnorm = alfa * beta
    if arnorm == 0:
        return x, istop, itn, r1norm, r2norm, anorm, acond, arnorm, xnorm, var

    arnorms = -np.ones(iter_lim)
    while itn < iter_lim:
        arnorms[itn] = arnorm
        itn += 1
        # bidiagonalisation step (lsqr.py:421-430)
        u = A.matvec(v) - alfa * u
        beta = np.linalg.norm(u)
        if beta > 0:
            u = (1 / beta) * u
            anorm = np.sqrt(anorm ** 2 + alfa ** 2 + beta ** 2 + damp ** 2)
            v = A.rmatvec(u) - beta * v
            alfa = np.linalg.norm(v)
            if alfa > 0:
                v = (1 / alfa) * v
        # damping rotation, then the bidiagonal QR step (lsqr.py:434-448)
        rhobar1 = np.sqrt(rhobar ** 2 + damp ** 2)
        cs1, sn1 = rhobar / rhobar1, damp / rhobar1
        psi = sn1 * phibar
        phibar = cs1 * phibar
        cs, sn, rho = sym_ortho(rhobar1, beta)
        theta = sn * alfa
        rhobar = -cs * alfa
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        # solution / search-direction update (lsqr.py:451-460)
        t1, t2 = phi / rho, -theta / rho
        dk = (1 / rho) * w
        x = x + t1 * w
        w = v + t2 * w
        ddnorm = ddnorm + np.linalg.norm(dk) ** 2
        if calc_var:
            var = var + dk ** 2
        # norm(x) estimate (lsqr.py:465-474)
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = np.sqrt(xxnorm + zbar ** 2)
        gamma = np.sqrt(gambar ** 2 + theta ** 2)
        cs2, sn2 = gambar / gamma, theta / gamma
        z = rhs / gamma
        xxnorm = xxnorm + z ** 2
        # convergence tests (lsqr.py:479-526)
        acond = anorm * np.sqrt(ddnorm)
        res1 = phibar ** 2
        res2 = res2 + psi ** 2
        rnorm = np.sqrt(res1 + res2)
        arnorm = alfa * abs(tau)
        r1sq = rnorm ** 2 - dampsq * xxnorm
        r1norm = np.sqrt(abs(r1sq))
        if r1sq < 0:
            r1norm = -r1norm
        r2norm = rnorm
        test1 = rnorm / bnorm
        test2 = arnorm / (anorm * rnorm + EPS)
        test3 = 1.0 / (acond + EPS)
        t1 = test1 / (1 + anorm * xnorm / bnorm)
        rtol = btol + atol * anorm * xnorm / bnorm
        if itn >= iter_lim:
            istop = 7
        if 1 + test3 <= 1:
            istop = 6
        if 1 + test2 <= 1:
            istop = 5
        if 1 + t1 <= 1:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        if istop != 0:
            break
    arnorms = arnorms[arnorms > -1]
    return x, istop, itn, r1norm, r2norm, anorm, acond, arnorms, xnorm, var


# --------------------------------------------------------------------------
# Preconditioned operator and PcSS2  (comps/preconditioning.py, saddle.py)
# --------------------------------------------------------------------------

class _PrecondOp:
    """A_pc = [A; sqrt(delta) I] M with M = R^{-1} (upper_tri) or M = R."""

    def __init__(self, A, delta, R, upper_tri):
        self.A = A
        self.R = R
        self.m = A.shape[0]
        self.sd = np.sqrt(delta)
        self.upper_tri = upper_tri
        lift = A.shape[0] + (A.shape[1] if delta > 0 else 0)
        self.shape = (lift, R.shape[1])

    def m_fwd(self, z):
        if self.upper_tri:
            return la.solve_triangular(self.R, z, lower=False)
        return self.R @ z

    def m_adj(self, w):
        if self.upper_tri:
            return la.solve_triangular(self.R, w, "T", lower=False)
        return self.R.T @ w

    def matvec(self, arg):
        """preconditioning.py:26-31 (tri) / :45-48 (general)."""
        work = self.m_fwd(arg)
        out = self.A @ work
        if self.sd > 0:
            out = np.concatenate((out, self.sd * work))
        return out

    def rmatvec(self, arg):
        """preconditioning.py:33-38 (tri) / :50-54 (general)."""
        work = self.A.T @ arg[:self.m]
        if self.sd > 0:
            work = work + self.sd * arg[self.m:]
        return self.m_adj(work)


def a_lift_precond(A, delta, R, upper_tri=False, k=1):
    """comps/preconditioning.py:16-67 -> (A_pc, M_fwd, M_adj)."""
    if k != 1:
        raise NotImplementedError()
    op = _PrecondOp(A, delta, R, upper_tri)
    return op, op.m_fwd, op.m_adj


def pcss2(A, b, delta, tol, iter_lim, R, upper_tri, z0):
    """Overdetermined branch of PcSS2.__call__, comps/determiter/saddle.py:187-201."""
    m, n = A.shape
    A_pc, M_fwd, _ = a_lift_precond(A, delta, R, upper_tri)
    if delta > 0:
        b = np.concatenate((b, np.zeros(n)))
    res = lsqr(A_pc, b, atol=tol, btol=tol, iter_lim=iter_lim, x0=z0)
    x = M_fwd(res[0])
    y = b[:m] - A @ x
    return x, y, res[7], res


# --------------------------------------------------------------------------
# SPO driver and its log  (drivers/least_squares.py, determiter/logging.py)
# --------------------------------------------------------------------------

class SketchAndPrecondLog:
    """comps/determiter/logging.py:4-94 (times and the error log)."""

    def __init__(self):
        self.time_sketch = 0.0
        self.time_factor = 0.0
        self.time_presolve = 0.0
        self.time_convert = 0.0
        self.time_iterate = 0.0
        self.times = None
        self.errors = None

    @property
    def time_setup(self):
        return self.time_sketch + self.time_factor + self.time_convert

    def wrap_up(self, iter_errors, init_error):
        """logging.py:72-94: amortise iterate time linearly; prepend error at 0."""
        iter_errors = np.atleast_1d(iter_errors)
        setup = self.time_setup
        ramp = np.linspace(0, self.time_iterate, iter_errors.size, endpoint=True)
        self.times = np.concatenate(([setup], setup + self.time_presolve + ramp))
        self.errors = np.concatenate(([init_error], iter_errors))


def dim_checks(sampling_factor, n_rows, n_cols):
    """drivers/least_squares.py:89-104."""
    assert n_rows >= n_cols
    d = int(sampling_factor * n_cols)
    if d > n_rows:
        warnings.warn(f"The embedding dimension d={d} exceeds the {n_rows} rows "
                      f"of the data matrix; proceeding with d={n_rows}.")
        d = n_rows
    assert d >= n_cols
    return d


def _sketch_gen(kind, vec_nnz=8, normalize=True):
    if callable(kind):
        return kind
    if kind == "gaussian":
        return lambda r, c, rng: gaussian_operator(r, c, rng, normalize)
    if kind == "sjlt":
        return lambda r, c, rng: sjlt_operator(r, c, rng, vec_nnz)
    if kind == "srct":
        return srct_matrix
    raise ValueError(kind)


def spo(A, b, delta, tol, iter_lim, rng, sketch="gaussian", sampling_factor=4,
        mode="qr", logging=True, vec_nnz=8, return_info=False):
    """SPO.__call__, drivers/least_squares.py:291-369 (modes 'qr', 'chol', 'svd').

    Returns ``(x, log)``; with ``return_info`` also a dict holding the
    presolve decision, R, z_ske and the raw LSQR tuple.
    """
    m, n = A.shape
    sqrt_delta = np.sqrt(delta)
    d = dim_checks(sampling_factor, m, n)
    rng = np.random.default_rng(rng)
    clock = time.time if logging else (lambda: 0.0)
    log = SketchAndPrecondLog()
    gen = _sketch_gen(sketch, vec_nnz)

    tic = clock()
    S = gen(d, m, rng)
    A_ske = S @ A
    log.time_sketch = clock() - tic

    if mode == "qr":
        tic = clock()
        if delta > 0:
            A_ske = np.vstack((A_ske, sqrt_delta * np.eye(n)))
        Q, R = la.qr(A_ske, overwrite_a=True, mode="economic")
        log.time_factor = clock() - tic
        tic = clock()
        b_ske = S @ b
        z_ske = Q[:d, :].T @ b_ske
        x_ske = la.solve_triangular(R, z_ske, lower=False)
    elif mode == "chol":
        tic = clock()
        G = A_ske.T @ A_ske + delta * np.eye(n)
        R = la.cholesky(G, lower=False)
        log.time_factor = clock() - tic
        tic = clock()
        b_ske = S @ b
        z_ske = la.solve_triangular(R, A_ske.T @ b_ske, lower=False, trans="T")
        x_ske = la.solve_triangular(R, z_ske, lower=False)
    elif mode == "svd":
        # least_squares.py:330-339: R is the n x rank right preconditioner V / sigma
        tic = clock()
        if delta > 0:
            A_ske = np.vstack((A_ske, sqrt_delta * np.eye(n)))
        R, U, sigma, Vh = svd_right_precond(A_ske)
        log.time_factor = clock() - tic
        tic = clock()
        b_ske = S @ b
        z_ske = U[:d, :].T @ b_ske
        x_ske = R @ z_ske
    else:
        raise ValueError()

    # presolve decision (least_squares.py:344-351)
    r = A @ x_ske - b
    if delta > 0:
        r = np.concatenate((r, sqrt_delta * x_ske))
    rel_err = la.norm(r) / la.norm(b)
    used_presolve = not (rel_err >= 1 or (rel_err > 1e-15 and R.shape[0] != R.shape[1]))
    if not used_presolve:
        z_ske = None
    log.time_presolve = clock() - tic

    tic = clock()
    tri = mode in ("qr", "chol")
    x, y, arnorms, raw = pcss2(A, b, delta, tol, iter_lim, R, tri, z_ske)
    log.time_iterate = clock() - tic

    if logging:
        if tri:
            ar0 = la.solve_triangular(R, A.T @ b, "T", lower=False)
        else:
            ar0 = R.T @ (A.T @ b)
        log.wrap_up(arnorms, la.norm(ar0))
    if return_info:
        info = dict(R=R, z_ske=z_ske, x_ske=x_ske, rel_err=rel_err,
                    used_presolve=used_presolve, lsqr=raw, d=d)
        return x, log, info
    return x, log


def backward_error(A, b, x):
    """eta(x) = ||A^T r|| / (||A||_F ||r||), r = b - A x, all in float64."""
    A = np.asarray(A, dtype=np.float64)
    r = b - A @ x
    rn = la.norm(r)
    if rn == 0:
        return 0.0
    return la.norm(A.T @ r) / (la.norm(A, "fro") * rn)


# --------------------------------------------------------------------------
# QB chain  (sketchers/aware.py, rangefinders.py, qb.py, svd.py, evd.py)
# --------------------------------------------------------------------------

def orth(S):
    """utils/linalg_wrappers.py:6-7."""
    return la.qr(S, mode="economic")[0]


def lu_stabilize(S):
    """utils/linalg_wrappers.py:10-12: P L with partial pivoting."""
    return la.lu(S, permute_l=True)[0]


def rs1(A, k, rng, sketch_op_gen=None, num_pass=2, stabilizer=orth,
        passes_per_stab=1):
    """RS1.__call__, comps/sketchers/aware.py:160-184."""
    if sketch_op_gen is None:
        sketch_op_gen = _sketch_gen("gaussian")
    assert num_pass >= 0
    rng = np.random.default_rng(rng)
    done = 0
    if num_pass % 2 == 0:
        S = sketch_op_gen(A.shape[1], k, rng)
    else:
        S = A.T @ sketch_op_gen(A.shape[0], k, rng)
        done += 1
        if passes_per_stab == 1:
            S = stabilizer(S)
    q = (num_pass - done) // 2
    while q > 0:
        S = A @ S
        done += 1
        if done % passes_per_stab == 0:
            S = stabilizer(S)
        S = A.T @ S
        done += 1
        if done % passes_per_stab == 0:
            S = stabilizer(S)
        q -= 1
    return S


def rf1(A, k, tol, rng, rso):
    """RF1.__call__, comps/rangefinders.py:176-188.  ``rso(A, k, rng)``."""
    assert 0 < k <= min(A.shape)
    if not np.isnan(tol):
        warnings.warn("This RangeFinder implementation cannot directly control "
                      "approximation error. Parameter \"tol\" is being ignored.")
    rng = np.random.default_rng(rng)
    S = rso(A, k, rng)
    Y = A @ S
    return la.qr(Y, mode="economic")[0]


def qb1(A, k, tol, rng, rf):
    """QB1.__call__, comps/qb.py:344-352.  ``rf(A, k, tol, rng)``."""
    assert 0 < k <= min(A.shape)
    if not np.isnan(tol):
        assert 0 <= tol < 1
    rng = np.random.default_rng(rng)
    Q = rf(A, k, tol, rng)
    return Q, Q.T @ A


def svd1(A, k, tol, over, rng, qb):
    """SVD1.__call__, drivers/svd.py:161-176.  ``qb(A, k, tol, rng)``."""
    rng = np.random.default_rng(rng)
    Q, B = qb(A, k + over, tol, rng)
    U, s, Vh = la.svd(B, full_matrices=False)
    if over > 0:
        cut = min(k, s.size)
        U, s, Vh = U[:, :cut], s[:cut], Vh[:cut, :]
    keep = ~(s < 10 * EPS)
    if not np.all(keep):
        U, s, Vh = U[:, keep], s[keep], Vh[keep, :]
    return Q @ U, s, Vh


def evd1(A, k, tol, over, rng, qb):
    """EVD1.__call__, drivers/evd.py:270-289.  ``qb(A, k, tol, rng)``."""
    assert 0 < k <= min(A.shape)
    if not np.isnan(tol):
        assert 0 <= tol < np.inf
    rng = np.random.default_rng(rng)
    Q, B = qb(A, k + over, tol / 2, rng)
    C = B @ Q
    lamb, U = la.eigh(C)
    alamb = np.abs(lamb)
    r = min(k, Q.shape[1], np.count_nonzero(alamb > 10 * EPS))
    order = np.argsort(-alamb)[:r]
    return Q @ U[:, order], lamb[order]


# --------------------------------------------------------------------------
# mixed precision: exact BF16 sketch and FP64 iterative refinement (no
# counterpart in the reference; pinned by exact integer arithmetic, extended
# precision residuals and the FP64 least-squares solution)
# --------------------------------------------------------------------------

OZ_CHUNK = 256


def quantize_q(Z):
    """Integer-grid Gaussian: clip(rint(32 Z), -256, 256)."""
    return np.clip(np.rint(32.0 * np.asarray(Z, dtype=np.float64)), -256.0, 256.0)


def ozaki_split(A, b=None, npieces=4):
    """Split [A | b] per 256-row chunk and column into integer pieces:
    A[k, j] = 2^(e_cj - 8) sum_p 2^(-9p) a_p[k, j] + remainder.
    Returns (pieces (P, m, ncols) float64 integers, exps (nchunks, ncols) int)."""
    M = np.asarray(A, dtype=np.float64)
    if b is not None:
        M = np.column_stack([M, np.asarray(b, dtype=np.float64)])
    m, nc = M.shape
    nch = -(-m // OZ_CHUNK)
    exps = np.zeros((nch, nc), dtype=np.int64)
    pieces = np.zeros((npieces, m, nc))
    for c in range(nch):
        blk = M[c * OZ_CHUNK:(c + 1) * OZ_CHUNK]
        mx = np.max(np.abs(blk), axis=0)
        e = np.where(mx > 0, np.frexp(mx)[1], 0)
        exps[c] = e
        y = np.ldexp(blk, (8 - e)[None, :])
        for p in range(npieces):
            a = np.rint(y)
            pieces[p, c * OZ_CHUNK:(c + 1) * OZ_CHUNK] = a
            y = (y - a) * 512.0
    return pieces, exps


def ozaki_product(Q, pieces, exps, scale, group_chunks=None):
    """scale * Q @ (truncated [A | b]) evaluated as the device does: exact
    integer chunk sums, added in FP64 in the order chunk-major, piece-minor
    within each group of `group_chunks` consecutive chunks (all chunks when
    None), then the group sums added in group order."""
    Qi = np.asarray(Q, dtype=np.int64)
    P, m, nc = pieces.shape
    d = Qi.shape[0]
    nchunks = exps.shape[0]
    gc = nchunks if group_chunks is None else int(group_chunks)
    total = np.zeros((d, nc))
    for c0 in range(0, nchunks, gc):
        acc = np.zeros((d, nc))
        for c in range(c0, min(c0 + gc, nchunks)):
            sl = slice(c * OZ_CHUNK, min((c + 1) * OZ_CHUNK, m))
            for p in range(P):
                s = Qi[:, sl] @ pieces[p, sl].astype(np.int64)  # exact
                acc += s.astype(np.float64) * np.ldexp(1.0, exps[c] - 8 - 9 * p)[None, :]
        total += acc
    return scale * total


def _split26(v):
    """Veltkamp split: v = hi + lo with hi, lo of at most 26 significant bits."""
    c = v * 134217729.0  # 2^27 + 1
    hi = c - (c - v)
    return hi, v - hi


def _exact_products(a, x):
    """The four partial products whose sum is exactly a * x (elementwise)."""
    ah, al = _split26(a)
    xh, xl = _split26(x)
    return ah * xh, ah * xl, al * xh, al * xl


def residual_exact(A, x, b):
    """(r, A^T r, ||r||^2) with r_i and (A^T r)_j correctly rounded: exact
    products (Veltkamp/Dekker splitting) summed by math.fsum.  A^T r is the
    correctly rounded product of A^T with the rounded r."""
    import math
    A = np.asarray(A, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m, n = A.shape
    parts = _exact_products(A, x[None, :])
    r = np.empty(m)
    for i in range(m):
        r[i] = math.fsum([b[i]] + [-v for p in parts for v in p[i]])
    parts = _exact_products(A, r[:, None])
    g = np.array([math.fsum(v for p in parts for v in p[:, j]) for j in range(n)])
    rss = math.fsum((r * r).tolist())
    return r, g, rss


def residual_extended(A, x, b):
    """(r, A^T r, ||r||^2) in x87 extended precision (64-bit significand);
    a quick checker, 2^11 times finer than FP64 but not exact."""
    Al = np.asarray(A, dtype=np.longdouble)
    r = np.asarray(b, dtype=np.longdouble) - Al @ np.asarray(x, dtype=np.longdouble)
    g = Al.T @ r
    return r.astype(np.float64), g.astype(np.float64), float(np.sum(r * r))


def backward_error_extended(A, b, x):
    """eta = ||A^T r|| / (||A||_F ||r||) with r and A^T r in extended precision."""
    r, g, rss = residual_extended(A, x, b)
    return float(np.linalg.norm(g) / (np.linalg.norm(A) * np.sqrt(rss)))


def spo_mixed(A, b, tol, rng, sampling_factor=4, npieces=4, inner_tol=1e-6, max_inner=60,
              max_refine=8):
    """Restatement of the mixed-precision SPO: exact BF16-grid Gaussian sketch
    of [A | b], Householder QR, B = A R^{-1} stored in FP32, then FP64
    iterative refinement x <- x + R^{-1} CG(B^T B, R^{-T} A^T r(x)) with the
    residuals in extended precision, until eta(x) <= tol or a step no longer
    halves eta.  Each CG solve stops at the relative residual
    min(0.1, max(inner_tol, 0.1 tol / eta)).  Returns the x with the smallest eta and the eta history."""
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m, n = A.shape
    d = dim_checks(sampling_factor, m, n)
    rng = np.random.default_rng(rng)
    Q = quantize_q(rng.standard_normal((d, m)))
    pieces, exps = ozaki_split(A, b, npieces)
    Mx = ozaki_product(Q, pieces, exps, 2.0 ** -5 / np.sqrt(min(d, m)))
    Qf, R = la.qr(Mx[:, :n], mode='economic')
    z = Qf.T @ Mx[:, n]
    x = la.solve_triangular(R, z)
    B32 = la.solve_triangular(R, A.T, trans='T').T.astype(np.float32)
    B = B32.astype(np.float64)
    anorm = np.linalg.norm(A)
    etas = []
    best = (np.inf, x)
    for step in range(max_refine + 1):
        r, g, rss = residual_extended(A, x, b)
        eta = float(np.linalg.norm(g) / (anorm * np.sqrt(rss)))
        etas.append(eta)
        if eta < best[0]:
            best = (eta, x)
        # stop at the target, or once a step no longer halves eta (the FP64
        # floor of eta for this problem: x itself is rounded to FP64)
        if eta <= tol or (step > 0 and eta > 0.5 * etas[-2]) or step == max_refine:
            break
        rhs = la.solve_triangular(R, g, trans='T')
        # the inner solve only has to take eta down to tol: a tenth of that
        # ratio, between inner_tol and 0.1
        itol = min(0.1, max(inner_tol, 0.1 * tol / eta)) if eta > 0 else inner_tol
        dz = np.zeros(n)
        rv = rhs.copy()
        pv = rv.copy()
        rr = rr0 = rv @ rv
        for _ in range(max_inner):
            q = B.T @ (B @ pv)
            alpha = rr / (pv @ q)
            dz += alpha * pv
            rv -= alpha * q
            rr_new = rv @ rv
            if rr_new <= itol ** 2 * rr0:
                break
            pv = rv + (rr_new / rr) * pv
            rr = rr_new
        x = x + la.solve_triangular(R, dz)
    x = best[1]
    return x, np.array(etas)


# --------------------------------------------------------------------------
# learned leverage-score sketch (no counterpart in the reference; the MLP is
# pinned by finite differences, the labels by scipy's QR, the sampling
# operator by the reference's sampling_operator semantics, sketching.py:204-232)
# --------------------------------------------------------------------------

MLP_IN, MLP_H = 16, 64
MLP_SIZES = [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,))]
MLP_NPARAM = sum(int(np.prod(sh)) for _, sh in MLP_SIZES)


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def bf16_pieces(X, npieces):
    """The BF16 pieces of FP64 values, as uint16 bits, shape (npieces,) + X.shape:
    piece p is the BF16 rounding (through FP32, ties to even) of X minus the
    earlier pieces, the difference taken exactly in FP64.  No reference
    counterpart (the split-precision QB products of parla_amd).  Pinned by
    the reconstruction bound |X - sum| <= 2^-(8 npieces) |X| and by
    bf16_round (tests/test_oracle.py)."""
    r = np.array(X, dtype=np.float64, copy=True)
    out = []
    for _ in range(npieces):
        h = bf16_round(r.astype(np.float32))
        out.append((h.view(np.uint32) >> 16).astype(np.uint16))
        r = r - h.astype(np.float64)
    return np.stack(out)


def bf16_values(bits):
    """FP64 values of BF16 bit patterns."""
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def split_product(Lbits, Rbits):
    """sum over p + q < P of L_p R_q^T in FP64 from piece bit arrays
    (P x M x K and P x N x K): the product the split GEMM forms (it
    accumulates in FP32, so the two differ by rounding).  Pinned against
    exact rational arithmetic on small cases (tests/test_oracle.py)."""
    P = Lbits.shape[0]
    C = np.zeros((Lbits.shape[1], Rbits.shape[1]))
    for p in range(P):
        for q in range(P - p):
            C += bf16_values(Lbits[p]) @ bf16_values(Rbits[q]).T
    return C


def mlp_unpack(theta):
    out, o = {}, 0
    for name, sh in MLP_SIZES:
        k = int(np.prod(sh))
        out[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(sh)
        o += k
    return out


def mlp_effective(theta):
    """Parameters as the device uses them: weights BF16-rounded, biases FP32."""
    p = mlp_unpack(theta)
    for k in ("W1", "W2", "w3"):
        p[k] = bf16_round(p[k]).astype(np.float64)
    for k in ("b1", "b2", "b3"):
        p[k] = p[k].astype(np.float32).astype(np.float64)
    return p


def mlp_forward(p, X):
    """y = w3 . relu(W2^T relu(W1^T asinh(x) + b1) + b2) + b3 for each row x of X."""
    Xt = np.arcsinh(X)
    h1 = np.maximum(Xt @ p["W1"] + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["W2"] + p["b2"], 0.0)
    return h2 @ p["w3"] + p["b3"][0], (h1, h2)


def mlp_loss_grad(p, X, t, rows_total=None):
    """(1/N) sum (y - t)^2 and its gradient, flat in the layout W1|b1|W2|b2|w3|b3."""
    N = X.shape[0] if rows_total is None else rows_total
    y, (h1, h2) = mlp_forward(p, X)
    dy = 2.0 * (y - t) / N
    dh2 = np.outer(dy, p["w3"]) * (h2 > 0)
    dh1 = (dh2 @ p["W2"].T) * (h1 > 0)
    g = {"W1": np.arcsinh(X).T @ dh1, "b1": dh1.sum(0), "W2": h1.T @ dh2, "b2": dh2.sum(0), "w3": h2.T @ dy,
         "b3": np.array([dy.sum()])}
    flat = np.concatenate([g[name].reshape(-1) for name, _ in MLP_SIZES])
    return float(np.sum((y - t) ** 2) / N), flat


def leverage_scores(A):
    """Exact leverage scores ||Q[i, :]||^2 from scipy's FP64 Householder QR."""
    Q = la.qr(np.asarray(A, dtype=np.float64), mode="economic")[0]
    return np.sum(Q * Q, axis=1)


def coherent_matrix(m, n, rng, heavy_frac=0.01, heavy_scale=100.0, cond=1e6):
    """The learned-sketch test family: Gaussian rows, a fraction of rows scaled
    by heavy_scale, right-scaled by diag(logspace(0, -log10 cond)) V^T."""
    rng = np.random.default_rng(rng)
    G = rng.standard_normal((m, n))
    heavy = rng.choice(m, max(1, int(heavy_frac * m)), replace=False)
    G[heavy] *= heavy_scale
    V = la.qr(rng.standard_normal((n, n)))[0]
    return (G * np.logspace(0, -np.log10(cond), n)) @ V.T, heavy


# --------------------------------------------------------------------------
# Fits used by the convergence tests  (parla/utils/stats.py)
# --------------------------------------------------------------------------

def loglinear_fit(x, y):
    """utils/stats.py:6-28: fit log(y) = a + b x; return ([a, b], R^2)."""
    x = np.asarray(x, dtype=float).ravel()
    y = np.asarray(y, dtype=float).ravel()
    keep = y > 0
    x, y = x[keep], y[keep]
    ly = np.log(y)
    X = np.column_stack([np.ones(x.size), x])
    fit = la.lstsq(X, ly)[0]
    ss_tot = np.sum((ly - ly.mean()) ** 2)
    ss_res = np.sum((ly - X @ fit) ** 2)
    return fit, 1 - ss_res / ss_tot


def loglog_fit(x, y):
    """utils/stats.py:31-56: fit log(y) = a + b log(x)."""
    return loglinear_fit(np.log(np.asarray(x, dtype=float)), y)


# --------------------------------------------------------------------------
# The drivers next to the hot path: SVD preconditioner, the underdetermined
# branch, sketch-and-solve, PCG, saddle-point systems, blocked QB, Nystrom
# EVD.  Pinned by the reference-written fixtures (SPO 'chol'/'svd', PcSS2
# with c != 0 and delta > 0, SSO1, PcSS1, SPS1/SPS2, QB3, EVD2) and by closed
# forms where the reference no longer runs under numpy 2 (QB2, SPU1).
# --------------------------------------------------------------------------

def svd_right_precond(A_ske):
    """comps/preconditioning.py:70-79: M = V / sigma over the numerical rank
    (sigma > sigma_1 * ncols * eps); returns (M, U, sigma, Vh)."""
    U, sigma, Vh = la.svd(A_ske, full_matrices=False)
    eps = np.finfo(float).eps
    rank = np.count_nonzero(sigma > sigma[0] * A_ske.shape[1] * eps)
    Vh = Vh[:rank, :]
    U = U[:, :rank]
    sigma = sigma[:rank]
    return Vh.T / sigma, U, sigma, Vh


class _Transposed:
    """The adjoint of an operator with matvec / rmatvec (A_pc.T)."""

    def __init__(self, op):
        self.op = op
        self.shape = (op.shape[1], op.shape[0])

    def matvec(self, x):
        return self.op.rmatvec(x)

    def rmatvec(self, y):
        return self.op.matvec(y)


def pcss2_under(A, c, delta, tol, iter_lim, R, upper_tri):
    """Underdetermined branch of PcSS2.__call__, saddle.py:203-214: LSQR on
    A_pc^T with right-hand side M^T c.  With delta = 0 the reference writes
    ``np.NaN * np.empty(n)`` for x (an attribute numpy 2 removed); NaN is
    the evident value."""
    m, n = A.shape
    A_pc, _, M_adj = a_lift_precond(A, delta, R, upper_tri)
    c_pc = M_adj(c)
    res = lsqr(_Transposed(A_pc), c_pc, atol=tol, btol=tol, iter_lim=iter_lim)
    y = res[0]
    if delta > 0:
        y = y[:m]
        x = (A.T @ y - c) / delta
    else:
        x = np.full(n, np.nan)
    return x, y, res[7], res


def spu1(A, c, tol, iter_lim, rng, sketch="sjlt", sampling_factor=3, vec_nnz=8):
    """SPU1.__call__, drivers/least_squares.py:464-494: min ||y|| s.t. A^T y = c
    from the SVD preconditioner of S A and LSQR on (A M)^T."""
    m, n = A.shape
    d = dim_checks(sampling_factor, m, n)
    rng = np.random.default_rng(rng)
    log = SketchAndPrecondLog()
    S = _sketch_gen(sketch, vec_nnz)(d, m, rng)
    A_ske = S @ A
    M = svd_right_precond(A_ske)[0]
    x, y, errors, _ = pcss2_under(A, c, 0.0, tol, iter_lim, M, False)
    log.wrap_up(errors, la.norm(A @ (M @ (M.T @ c))))
    return y, log


def sso1(A, b, delta, rng, sketch="sjlt", sampling_factor=3, vec_nnz=8, lapack_driver="gelsd"):
    """SSO1.__call__, drivers/least_squares.py:162-189: solve the sketched
    problem [S A; sqrt(delta) I] x = [S b; 0] by LAPACK least squares."""
    m, n = A.shape
    d = dim_checks(sampling_factor, m, n)
    rng = np.random.default_rng(rng)
    S = _sketch_gen(sketch, vec_nnz)(d, m, rng)
    A_ske = S @ A
    b_ske = S @ b
    if delta > 0:
        A_ske = np.vstack((A_ske, (delta ** 0.5) * np.eye(n)))
        b_ske = np.concatenate((b_ske, np.zeros(n)))
    return la.lstsq(A_ske, b_ske, cond=None, lapack_driver=lapack_driver)[0]


def pcg(mv_mat, rhs, mv_pre, iter_lim, tol, x0):
    """comps/determiter/pcg.py:5-47: preconditioned CG on mat x = rhs with
    mv_pre = M M^T; recomputes the residual every 10 steps; residuals[i] =
    ||rhs - mat x_i|| (the recurrence's value), stop at tol * ||r_0||."""
    x = x0.copy()
    residuals = -np.ones(iter_lim)
    r = rhs - mv_mat(x)
    d = mv_pre(r)
    delta1_new = np.dot(r, d)
    cur_err = la.norm(r)
    rel_tol = tol * cur_err
    i = 0
    while i < iter_lim and cur_err > rel_tol:
        residuals[i] = cur_err
        q = mv_mat(d)
        alpha = delta1_new / np.dot(d, q)
        x += alpha * d
        if i % 10 == 0:
            r = rhs - mv_mat(x)
        else:
            r -= alpha * q
        cur_err = la.norm(r)
        s_ = mv_pre(r)
        delta1_old = delta1_new
        delta1_new = np.dot(r, s_)
        d = s_ + (delta1_new / delta1_old) * d
        i += 1
    return x, residuals[:i]


def pcss1(A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
    """PcSS1.__call__, comps/determiter/saddle.py:88-176: PCG on the normal
    equations (A^T A + delta I) x = A^T b - c with the preconditioner
    R R^T (+ I - V V^T when R has fewer columns than A, with R rescaled by
    the regularised singular values)."""
    m, n = A.shape
    if b is None:
        b = np.zeros(m)
    if upper_tri:
        raise NotImplementedError()
    R = np.array(R, dtype=np.float64)
    pc_dim = R.shape[1]
    full = pc_dim == n
    if not full:
        sv = 1.0 / la.norm(R, axis=0)
        V = R * sv
        sv = np.sqrt(sv ** 2 + delta)
        R = V / (sv / sv[-1])

    def mv_pre(vec):
        res = R @ (R.T @ vec)
        if not full:
            res = res + vec - V @ (V.T @ vec)
        return res

    def mv_gram(vec):
        return A.T @ (A @ vec) + delta * vec

    rhs = A.T @ b
    if c is not None:
        rhs = rhs - c
    x0 = np.zeros(n) if (z0 is None or not full) else R @ z0
    x, residuals = pcg(mv_gram, rhs, mv_pre, iter_lim, tol, x0)
    return x, b - A @ x, residuals


def pcss2_saddle(A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
    """PcSS2.__call__ dispatch, saddle.py:186-217."""
    if c is None or la.norm(c) == 0:
        x, y, errors, _ = pcss2(A, b, delta, tol, iter_lim, R, upper_tri, z0)
        return x, y, errors
    if b is None or la.norm(b) == 0:
        x, y, errors, _ = pcss2_under(A, c, delta, tol, iter_lim, R, upper_tri)
        return x, y, errors
    raise ValueError('One of "b" or "c" must be zero.')


def sps1(A, b, c, delta, tol, iter_lim, rng, solver=None, sampling_factor=3, sketch="gaussian"):
    """SPS1.__call__, drivers/saddlesys.py:124-231 (Nystrom strategy 'left'
    when d < n).  ``solver(A, b, c, delta, tol, iter_lim, M, upper_tri, z0)``."""
    solver = solver or pcss1
    m, n = A.shape
    sqrt_delta = np.sqrt(delta)
    d = int(sampling_factor * n)
    rng = np.random.default_rng(rng)
    if b is None:
        b = np.zeros(m)
    log = SketchAndPrecondLog()
    gen = _sketch_gen(sketch)
    nystrom_like = d < n
    if not nystrom_like:
        A_ske = gen(d, m, rng) @ A
        if sqrt_delta > 0:
            A_ske = np.vstack((A_ske, sqrt_delta * np.eye(n)))
        M, U, sigma, Vh = svd_right_precond(A_ske)
    else:
        A_ske = gen(d, m, rng) @ A
        V = la.qr(A_ske.T, mode="economic")[0]
        A_sample = A @ V
        _, sigma, Wt = la.svd(A_sample, full_matrices=False)
        M = V @ (Wt.T / sigma)
    rhs = A.T @ b
    if c is not None:
        rhs = rhs - c
    z_ske = None
    if not nystrom_like:
        z_ske = (Vh @ rhs) / sigma
        x_ske = M @ z_ske
        rhs_pc = M.T @ rhs
        lhs_ske_pc = M.T @ (A.T @ (A @ x_ske) + delta * x_ske)
        if la.norm(lhs_ske_pc - rhs_pc) >= la.norm(rhs_pc):
            z_ske = None
    x, y, errors = solver(A, b, c, delta, tol, iter_lim, M, False, z_ske)
    log.wrap_up(errors, la.norm(rhs))
    return x, y, log


def sps2(A, b, c, delta, tol, iter_lim, rng, sampling_factor=3, sketch="gaussian"):
    """SPS2.__call__, drivers/saddlesys.py:280-327: the saddle-point system as
    an overdetermined LS problem on [A; sqrt(delta) I] with the right-hand
    side shifted by the sketched solution of the c-part, then PcSS2."""
    m, n = A.shape
    sqrt_delta = np.sqrt(delta)
    d = dim_checks(sampling_factor, m, n)
    rng = np.random.default_rng(rng)
    if b is None:
        b = np.zeros(m)
    log = SketchAndPrecondLog()
    S = _sketch_gen(sketch)(d, m, rng)
    A_ske = S @ A
    if sqrt_delta > 0:
        A_ske = np.vstack((A_ske, sqrt_delta * np.eye(n)))
    M, U, sigma, Vh = svd_right_precond(A_ske)
    A_aug = np.vstack((A, sqrt_delta * np.eye(n))) if sqrt_delta > 0 else A
    b_aug = np.concatenate((b, np.zeros(n))) if delta > 0 else b.copy()
    if c is not None and la.norm(c) > 0:
        v = U @ ((1 / sigma) * (Vh @ c))
        b_aug[:m] -= S.T @ v[:d]
        if delta > 0:
            b_aug[m:] -= v[d:]
    z_ske = U[:d, :].T @ (S @ b_aug[:m])
    if delta > 0:
        z_ske += U[d:, :].T @ b_aug[m:]
    x_ske = M @ z_ske
    if la.norm(b_aug - A_aug @ x_ske) >= la.norm(b_aug):
        z_ske = None
    x, _, errors, _ = pcss2(A_aug, b_aug, 0.0, tol, iter_lim, M, False, z_ske)
    y = b - A @ x
    log.wrap_up(errors, la.norm(M.T @ (A_aug.T @ b_aug)))
    return x, y, log


def project_out(Qi, Q):
    """comps/qb.py:598-613 (as_list=False): Qi - Q (Q^T Qi)."""
    return Qi - Q @ (Q.T @ Qi)


def qb2(A, k, tol, rng, rf, blk):
    """QB2.__call__, comps/qb.py:437-482 (overwrite_a=False): blocks of
    rangefinder columns projected out of the previous ones, A deflated by
    each block; stops at k columns or when ||A - Q B||_F <= tol ||A||_F.
    The reference passes the removed ``np.NaN`` as the rangefinder's tol
    (:468); NaN is meant."""
    A = np.array(A, dtype=np.float64)
    assert k > 0
    k = min(k, min(A.shape))
    use_tol = not np.isnan(tol) and tol > 0
    if use_tol:
        sq_norm_A = la.norm(A, "fro") ** 2
        abs_sq_tol = sq_norm_A * tol ** 2
    rng = np.random.default_rng(rng)
    Q = np.empty((A.shape[0], 0))
    B = np.empty((0, A.shape[1]))
    while True:
        if B.shape[0] + blk > k:
            blk = k - B.shape[0]
        Qi = rf(A, blk, np.nan, rng)
        Qi = project_out(Qi, Q)
        Qi = la.qr(Qi, mode="economic")[0]
        Bi = Qi.T @ A
        Q = np.column_stack((Q, Qi))
        B = np.vstack((B, Bi))
        A -= Qi @ Bi
        if use_tol:
            sq_norm_A = sq_norm_A - la.norm(Bi, "fro") ** 2
            if sq_norm_A <= abs_sq_tol:
                break
        if B.shape[0] >= k:
            break
    return Q, B


def qb3(A, k, tol, rng, sk_op, blk):
    """QB3.__call__, comps/qb.py:552-595: G = A S, H = A^T G once, then blocks
    of Q and B from them without further passes over A (early stopping on
    the Frobenius-norm estimate).  ``sk_op(A, k, rng)``."""
    assert 0 < k < min(A.shape)
    use_tol = not np.isnan(tol) and tol > 0
    if use_tol:
        sq_norm_A = la.norm(A, "fro") ** 2
        abs_sq_tol = sq_norm_A * tol ** 2
    Q = np.empty((A.shape[0], 0))
    B = np.empty((0, A.shape[1]))
    rng = np.random.default_rng(rng)
    S = sk_op(A, k, rng)
    G = A @ S
    H = A.T @ G
    for i in range(int(np.ceil(k / blk))):
        b0, b1 = i * blk, min((i + 1) * blk, k)
        Si = S[:, b0:b1]
        BSi = B @ Si
        Yi = G[:, b0:b1] - Q @ BSi
        Qi, Ri = la.qr(Yi, mode="economic")
        Qi = project_out(Qi, Q)
        Qi, Rihat = la.qr(Qi, mode="economic")
        Ri = Rihat @ Ri
        Bi = H[:, b0:b1].T - (Yi.T @ Q) @ B - BSi.T @ B
        Bi = la.solve_triangular(Ri, Bi, trans="T")
        Q = np.column_stack((Q, Qi))
        B = np.vstack((B, Bi))
        if use_tol:
            sq_norm_A = sq_norm_A - la.norm(Bi, "fro") ** 2
            if sq_norm_A <= abs_sq_tol:
                break
    return Q, B


def evd2(A, k, tol, over, rng, sk_op):
    """EVD2.__call__, drivers/evd.py:349-384: Nystrom approximation of a PSD
    A from S = sk_op(A, k + over, rng), regularised by nu = sqrt(n) eps ||A S||,
    truncated where sigma_{i+1}^2 <= nu."""
    n = A.shape[0]
    assert 0 < k < n
    rng = np.random.default_rng(rng)
    S = sk_op(A, k + over, rng)
    Y = A @ S
    nu = np.sqrt(n) * np.finfo(float).eps * la.norm(Y)
    Y = Y + nu * S
    R = la.cholesky(S.T @ Y, lower=False)
    B = la.solve_triangular(R, Y.T, lower=False, trans=True).T
    V, sigma, _ = la.svd(B)
    comp = [k]
    for i in range(k - 1):
        if sigma[i + 1] ** 2 <= nu:
            comp.append(i)
    r = min(comp)
    return V[:, :r], (sigma ** 2)[:r] - nu


# --------------------------------------------------------------------------
# Device draws: counter-based streams of the GPU library (no counterpart in
# the reference, which draws with numpy's Generator).  Restated so that the
# device's Gaussian sketch, SJLT draws and weighted row sampling are checked
# value for value.
# --------------------------------------------------------------------------

_M32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
    as easy as 1, 2, 3", SC'11): ten rounds of two 32 x 32 -> 64-bit
    multiplications (0xD2511F53, 0xCD9E8D57) with the Weyl key schedule
    (0x9E3779B9, 0xBB67AE85).  Counter words and keys broadcast; returns four
    uint32 arrays."""
    c = [np.asarray(x, dtype=np.uint64) & _M32 for x in (c0, c1, c2, c3)]
    c0, c1, c2, c3 = np.broadcast_arrays(*c)
    k0 = np.asarray(k0, dtype=np.uint64) & _M32
    k1 = np.asarray(k1, dtype=np.uint64) & _M32
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        c0, c1, c2, c3 = (p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & _M32, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & _M32
        k0 = (k0 + np.uint64(0x9E3779B9)) & _M32
        k1 = (k1 + np.uint64(0xBB67AE85)) & _M32
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def u53_open0(hi, lo):
    """The uniform in (0, 1] of two 32-bit words: (top 53 bits + 1) * 2^-53."""
    b = ((np.asarray(hi, dtype=np.uint64) << np.uint64(32)) | np.asarray(lo, dtype=np.uint64)) >> np.uint64(11)
    return (b.astype(np.float64) + 1.0) * 2.0 ** -53


def _seed_words(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def device_gaussian_fill(rows, cols, scale, seed, offset=0):
    """The device Gaussian sketch (pamd_gaussian_fill_f64): entry (i, j) is
    draw e = offset + i * cols + j; draws 2p and 2p + 1 are the Box-Muller pair
    of Philox(counter (p, p >> 32, 0x6A09E667, 0), key seed):
    r = sqrt(-2 ln u1), z = (r cos 2 pi u2, r sin 2 pi u2)."""
    k0, k1 = _seed_words(seed)
    e = np.uint64(offset) + np.arange(rows * cols, dtype=np.uint64)
    pair = e >> np.uint64(1)
    x, y, z, w = philox4x32_10(pair & _M32, pair >> np.uint64(32), 0x6A09E667, 0, k0, k1)
    u1, u2 = u53_open0(x, y), u53_open0(z, w)
    rad = np.sqrt(-2.0 * np.log(u1))
    odd = (e & np.uint64(1)).astype(bool)
    val = np.where(odd, rad * np.sin(2.0 * np.pi * u2), rad * np.cos(2.0 * np.pi * u2))
    return (scale * val).reshape(rows, cols)


def device_sjlt_draw(n_short, n_long, k, seed):
    """The device SJLT draw (pamd_sjlt_draw): column j's entry l is row
    floor(x * n_short / 2^32) and sign bit y & 1 of Philox(counter (j,
    j >> 32, l, 0x510E527F + attempt), key seed), redrawn (attempt + 1) while
    the row repeats an earlier one of the column (with replacement when
    n_short < k; after 64 repeats the first unused row).  Returns rows (n_long
    x k, int32) and values +-1/sqrt(k) (negative where the sign bit is set)."""
    k0, k1 = _seed_words(seed)
    rows = np.zeros((n_long, k), dtype=np.int32)
    neg = np.zeros((n_long, k), dtype=bool)
    with_repl = n_short < k
    j = np.arange(n_long, dtype=np.uint64)
    for l in range(k):
        todo = np.ones(n_long, dtype=bool)
        for attempt in range(64):
            x, y, _, _ = philox4x32_10(j & _M32, j >> np.uint64(32), l, 0x510E527F + attempt, k0, k1)
            r = ((x.astype(np.uint64) * np.uint64(n_short)) >> np.uint64(32)).astype(np.int32)
            dup = np.zeros(n_long, dtype=bool) if with_repl else np.any(rows[:, :l] == r[:, None], axis=1)
            take = todo & ~dup
            rows[take, l] = r[take]
            neg[take, l] = (y[take] & 1).astype(bool)
            todo &= dup
            if not todo.any():
                break
        for jj in np.nonzero(todo)[0]:  # 64 straight repeats: first unused row
            used = set(rows[jj, :l].tolist())
            rows[jj, l] = next(r for r in range(n_short) if r not in used)
            neg[jj, l] = bool(y[jj] & 1)
    vals = np.where(neg, -1.0, 1.0) / np.sqrt(k)
    return rows, vals


def device_sample_rows(p, d, seed):
    """The device weighted row sampling (pamd_sample_rows_f64): the CDF of the
    weights summed in 1024 chunks of ceil(m / 1024) rows (chunk totals, then
    their running sum, then within each chunk, in order); draw i takes
    u = u53_open0(x, y) of Philox(counter (i, i >> 32, 0x3C6EF372,
    0x510E527F), key seed), the first row j with cdf[j] >= u * total, and the
    scale 1 / sqrt(d p_j / total)."""
    p = np.asarray(p, dtype=np.float64)
    m = p.size
    chunk = -(-m // 1024)
    cdf = np.empty(m)
    tots = np.array([sum_in_order(p[min(m, t * chunk):min(m, (t + 1) * chunk)]) for t in range(1024)])
    run = 0.0
    starts = np.empty(1024)
    for t in range(1024):
        starts[t] = run
        run += tots[t]
    for t in range(1024):
        i0, i1 = min(m, t * chunk), min(m, (t + 1) * chunk)
        r = starts[t]
        for i in range(i0, i1):
            r += p[i]
            cdf[i] = r
    total = cdf[m - 1]
    k0, k1 = _seed_words(seed)
    i = np.arange(d, dtype=np.uint64)
    x, y, _, _ = philox4x32_10(i & _M32, i >> np.uint64(32), 0x3C6EF372, 0x510E527F, k0, k1)
    target = u53_open0(x, y) * total
    idx = np.searchsorted(cdf, target, side="left")
    idx = np.minimum(idx, m - 1)
    scale = 1.0 / np.sqrt(float(d) * p[idx] / total)
    return idx.astype(np.int64), scale


def sum_in_order(v):
    """Left-to-right FP64 sum (the order of a sequential loop)."""
    s = 0.0
    for x in np.asarray(v, dtype=np.float64).tolist():
        s += x
    return s
