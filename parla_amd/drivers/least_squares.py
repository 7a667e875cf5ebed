"""Over- and underdetermined least squares by sketching (drop-in for
parla/drivers/least_squares.py: OverLstsqSolver, dim_checks, SSO1, sso1, SPO,
spo1, spo3, UnderLstsqSolver, SPU1, spu1).

SPO (qr mode) on the device:
  1. S = sketch_op_gen(d, m, rng); [S A | S b] is formed in one buffer
     (dense Gaussian: MFMA GEMM; SJLT: gather kernel).  With delta > 0 the
     buffer gets the rows [sqrt(delta) I | 0] (least_squares.py:309-310).
  2. Householder QR of the first n columns; the same reflectors turn the
     last column into z = Q^T [S b; 0] = Q[:d].T @ S b (least_squares.py:311-315).
  3. x_ske = R^{-1} z and the presolve test ||A x_ske - b|| / ||b|| < 1
     (least_squares.py:316, 343-351) -- the one host synchronisation.
  4. PcSS2 -> device LSQR on [A; sqrt(delta) I] R^{-1} from z (or from 0).
  5. With logging: errors = [||R^{-T} A^T b||, arnorms...] (least_squares.py:360-367).
"""
import time
import warnings

import numpy as np
import torch

import parla_amd.comps.determiter.saddle as dsad
import parla_amd.comps.sketchers.oblivious as sko
import parla_amd.utils.sketching as usk
from parla_amd import ops
from parla_amd.comps.determiter.logging import SketchAndPrecondLog


class OverLstsqSolver:
    """least_squares.py:16-86: ``x, log = solver(A, b, delta, tol, iter_lim, rng)``."""

    def __call__(self, A, b, delta, tol, iter_lim, rng):
        raise NotImplementedError()


def dim_checks(sampling_factor, n_rows, n_cols):
    """least_squares.py:89-104."""
    assert n_rows >= n_cols
    d = int(sampling_factor * n_cols)
    if d > n_rows:
        msg = f"""
        The embedding dimension "d" should not be larger than the
        number of rows of the data matrix. Here, an embedding dimension
        of d={d} has been requested for a matrix with only {n_rows} rows.
        We will proceed by setting d={n_rows}. This parameter choice will
        result in a very inefficient algorithm!
        """
        warnings.warn(msg)
        d = n_rows
    assert d >= n_cols
    return d


def sso1(A, b, delta, rng, sampling_factor=3, vec_nnz=8, lapack_driver='gelsd'):
    """least_squares.py:107-111 (the reference passes the removed np.NaN as tol)."""
    alg = SSO1(sko.SkOpSJ(vec_nnz), sampling_factor, lapack_driver, overwrite_sketch=True)
    return alg(A, b, delta, np.nan, 1, rng, logging=True)


class SSO1(OverLstsqSolver):
    """Sketch-and-solve for overdetermined least squares (least_squares.py:114-189).

    [S A; sqrt(delta) I] x = [S b; 0] is solved on the device.  The SVD-based
    LAPACK drivers ('gelsd', the default None, and 'gelss') become the
    minimum-norm solution over singular values above eps * sigma_1 (scipy's
    lstsq cutoff for cond=None), from the thin SVD of the sketch; 'gelsy'
    becomes Householder QR with the right-hand side appended (full rank)."""

    def __init__(self, sketch_op_gen, sampling_factor, lapack_driver=None, overwrite_sketch=True):
        self.sketch_op_gen = sketch_op_gen
        self.sampling_factor = sampling_factor
        self.lapack_driver = lapack_driver
        self.overwrite_sketch = overwrite_sketch

    def __call__(self, A, b, delta, tol, iter_lim, rng, logging=True):
        if not np.isnan(tol):
            warnings.warn("""
            This OverLstsqSolver implementation cannot directly control
            approximation error. Parameter "tol" is being ignored.
            """)
        if iter_lim > 1:
            warnings.warn("""
            This OverLstsqSolver implementation is not iterative.
            Parameter "iter_lim" is being ignored.
            """)
        if self.lapack_driver not in (None, 'gelsd', 'gelss', 'gelsy'):
            raise ValueError(f"unknown LAPACK driver {self.lapack_driver}")
        At = ops.as_device(A)
        bt = ops.as_device(b).reshape(-1)
        n_rows, n_cols = At.shape
        d = dim_checks(self.sampling_factor, n_rows, n_cols)
        rng = np.random.default_rng(rng)
        dev = At.device
        log = {'time_sketch': -1.0, 'time_solve': -1.0}

        def clock():
            if logging:
                torch.cuda.synchronize(dev)
                return time.time()
            return 0.0

        tic = clock()
        S = self.sketch_op_gen(d, n_rows, rng)
        rows = d + (n_cols if delta > 0 else 0)
        Mx = torch.empty((rows, n_cols + 1), dtype=ops.F64, device=dev)
        _sketch_into(S, At, Mx[:d, :n_cols])
        b_ske = torch.empty(d, dtype=ops.F64, device=dev)
        _sketch_into(S, bt, b_ske)
        Mx[:d, n_cols].copy_(b_ske)
        log['time_sketch'] = clock() - tic

        tic = clock()
        if delta > 0:
            Mx[d:].zero_()
            Mx[d:, :n_cols].diagonal().fill_(delta ** 0.5)
        if self.lapack_driver == 'gelsy':
            ops.qr_r_(Mx, n_cols)
            x = ops.trsv_(Mx[:n_cols, :n_cols], Mx[:n_cols, n_cols].clone())
        else:
            U, sig, Vh = ops.thin_svd(Mx[:, :n_cols])
            sh = sig.cpu().numpy()
            rank = int(np.count_nonzero(sh > np.finfo(float).eps * sh[0])) if sh.size else 0
            w = ops.gemv(U[:, :rank], Mx[:, n_cols].contiguous(), trans=True)
            w = ops.scale_cols(w.reshape(1, -1), sig[:rank], invert=True).reshape(-1)
            x = ops.gemv(Vh[:rank], w, trans=True)
        log['time_solve'] = clock() - tic
        return ops.like_input(x, A), log


def spo1(A, b, delta, tol, iter_lim, rng, sampling_factor=3, vec_nnz=8):
    """least_squares.py:192-196 (SPO in 'svd' mode with an SJLT)."""
    alg = SPO(sko.SkOpSJ(vec_nnz), sampling_factor, mode='svd')
    return alg(A, b, delta, tol, iter_lim, rng, logging=True)


def spo3(A, b, delta, tol, iter_lim, rng, sampling_factor=3, vec_nnz=8, mode='qr'):
    """least_squares.py:200-203."""
    alg = SPO(sko.SkOpSJ(vec_nnz), sampling_factor, mode)
    return alg(A, b, delta, tol, iter_lim, rng, logging=True)


def _sketch_into(S, X, out):
    """out <- S @ X for a sketch operator S and device X (matrix or vector)."""
    if isinstance(S, usk.DenseSketch):
        if X.dim() == 1:
            ops.gemv(S.S, X, y=out)
        else:
            ops.gemm(S.S, X, out=out)
        return
    if isinstance(S, usk.SJLTSketch) and not S.transposed:
        rp, cl, vl = S.csr_w()
        Xm = X.reshape(-1, 1) if X.dim() == 1 else X
        o = out.reshape(-1, 1) if out.dim() == 1 else out
        ops.sjlt_left(rp, cl, vl, S.n_short, Xm, out=o)
        return
    res = usk.apply(S, X)
    out.copy_(res.reshape(out.shape))


class SPO(OverLstsqSolver):
    """Sketch-and-precondition for overdetermined least squares (least_squares.py:206-369).

    mode 'qr' (Householder QR of the sketch, least_squares.py:306-316),
    'chol' (Cholesky of the sketched Gram matrix plus delta I, :317-329) or
    'svd' (the SVD right preconditioner M = V / sigma over the numerical rank,
    :330-339; rank-deficient A is handled as in the reference)."""

    def __init__(self, sketch_op_gen, sampling_factor: int, mode='qr'):
        self.sketch_op_gen = sketch_op_gen
        self.sampling_factor = sampling_factor
        self.mode = mode
        self.iterative_solver = dsad.PcSS2()

    def __call__(self, A, b, delta, tol, iter_lim, rng, logging=True):
        if self.mode not in ('qr', 'chol', 'svd'):
            raise ValueError()
        At = ops.as_device(A)
        bt = ops.as_device(b).reshape(-1)
        n_rows, n_cols = At.shape
        sqrt_delta = np.sqrt(delta)
        d = dim_checks(self.sampling_factor, n_rows, n_cols)
        rng = np.random.default_rng(rng)
        dev = At.device

        def clock():
            if logging:
                torch.cuda.synchronize(dev)
                return time.time()
            return 0.0

        log = SketchAndPrecondLog()
        tic = clock()
        S = self.sketch_op_gen(d, n_rows, rng)
        rows = d + (n_cols if (delta > 0 and self.mode != 'chol') else 0)
        Mx = torch.empty((rows, n_cols + 1), dtype=ops.F64, device=dev)
        _sketch_into(S, At, Mx[:d, :n_cols])
        log.time_sketch = clock() - tic

        tic = clock()
        b_ske = torch.empty(d, dtype=ops.F64, device=dev)
        if self.mode == 'qr':
            if delta > 0:
                Mx[d:].zero_()
                Mx[d:, :n_cols].diagonal().fill_(sqrt_delta)
            _sketch_into(S, bt, b_ske)
            Mx[:d, n_cols].copy_(b_ske)
            fac = ops.qr_r_(Mx, n_cols)
            # only the upper triangle of this view is R (scratch sits below); every
            # consumer (triangular solves, LSQR) reads the upper triangle only
            R = fac.R_view
            log.time_factor = clock() - tic
            tic = clock()
            z_ske = Mx[:n_cols, n_cols].clone()
            x_ske = ops.trsv_(R, z_ske.clone())
        elif self.mode == 'chol':
            A_ske = Mx[:d, :n_cols]
            G = ops.gemm(A_ske.t(), A_ske)
            if delta > 0:
                G.diagonal().add_(delta)
            R = ops.cholesky_upper(G)
            log.time_factor = clock() - tic
            tic = clock()
            _sketch_into(S, bt, b_ske)
            z_ske = ops.trsv_(R, ops.gemv(A_ske, b_ske, trans=True), trans=True)
            x_ske = ops.trsv_(R, z_ske.clone())
        else:  # 'svd'
            if delta > 0:
                Mx[d:].zero_()
                Mx[d:, :n_cols].diagonal().fill_(sqrt_delta)
            R, U, sigma, Vh = ops.svd_right_precond(Mx[:, :n_cols])
            log.time_factor = clock() - tic
            tic = clock()
            _sketch_into(S, bt, b_ske)
            z_ske = ops.gemv(U[:d, :], b_ske, trans=True)
            x_ske = ops.gemv(R, z_ske)

        # presolve decision (least_squares.py:343-352)
        r = ops.gemv(At, x_ske, y=bt.clone(), alpha=1.0, beta=-1.0)
        rn = ops.nrm2(r)
        bn = ops.nrm2(bt)
        if delta > 0:
            xn = ops.nrm2(x_ske)
            num = torch.sqrt(rn * rn + delta * xn * xn)
        else:
            num = rn
        rel_err = float((num / bn).item())
        if rel_err >= 1 or (rel_err > 1e-15 and R.shape[0] != R.shape[1]):
            z_ske = None
        log.time_presolve = clock() - tic

        tic = clock()
        tri = self.mode in ('qr', 'chol')
        res = self.iterative_solver(At, bt, None, delta, tol, iter_lim, R, tri, z_ske, want_y=False)
        log.time_iterate = clock() - tic

        if logging:
            ar0 = ops.gemv(At, bt, trans=True)
            if tri:
                ops.trsv_(R, ar0, trans=True)
            else:
                ar0 = ops.gemv(R, ar0, trans=True)
            log.wrap_up(res[2], float(ops.nrm2(ar0).item()))
            log.error_desc = self.iterative_solver.ERROR_METRIC_INFO
        return ops.like_input(ops.as_device(res[0]), A), log


class UnderLstsqSolver:
    """least_squares.py:372-419: ``y, log = solver(A, c, tol, iter_lim, rng)``
    for min ||y|| subject to A^T y = c, A tall."""

    def __call__(self, A, c, tol, iter_lim, rng, logging=False):
        raise NotImplementedError()


def spu1(A, c, tol, iter_lim, rng, sampling_factor=3, vec_nnz=8):
    """least_squares.py:422-426."""
    alg = SPU1(sko.SkOpSJ(vec_nnz), sampling_factor)
    return alg(A, c, tol, iter_lim, rng, logging=True)


class SPU1(UnderLstsqSolver):
    """SVD-based sketch-and-precondition for underdetermined least squares
    (least_squares.py:429-494): M = V / sigma from the SVD of S A, then LSQR
    on (A M)^T with right-hand side M^T c (PcSS2's underdetermined branch);
    the minimum-norm y solves A^T y = c."""

    def __init__(self, sketch_op_gen, sampling_factor: int):
        self.sketch_op_gen = sketch_op_gen
        self.sampling_factor = sampling_factor
        self.iterative_solver = dsad.PcSS2()

    def __call__(self, A, c, tol, iter_lim, rng, logging=True):
        At = ops.as_device(A)
        ct = ops.as_device(c).reshape(-1)
        n_rows, n_cols = At.shape
        d = dim_checks(self.sampling_factor, n_rows, n_cols)
        rng = np.random.default_rng(rng)
        dev = At.device

        def clock():
            if logging:
                torch.cuda.synchronize(dev)
                return time.time()
            return 0.0

        log = SketchAndPrecondLog()
        tic = clock()
        S = self.sketch_op_gen(d, n_rows, rng)
        A_ske = torch.empty((d, n_cols), dtype=ops.F64, device=dev)
        _sketch_into(S, At, A_ske)
        log.time_sketch = clock() - tic

        tic = clock()
        M = ops.svd_right_precond(A_ske)[0]
        log.time_factor = clock() - tic

        tic = clock()
        res = self.iterative_solver(At, None, ct, 0.0, tol, iter_lim, M, False, None)
        log.time_iterate = clock() - tic
        y_star = res[1]

        if logging:
            w = ops.gemv(M, ops.gemv(M, ct, trans=True))
            log.wrap_up(res[2], float(ops.nrm2(ops.gemv(At, w)).item()))
            log.error_desc = """
            The logs produced by this algorithm measure error as\n
                || (A M) (A M)' y - (A M) (M' c) ||_2,\n
            where "M" is a right-preconditioner for A. Under typical
            parameter settings, the condition number of A M is <= 10.
            """
        return ops.like_input(ops.as_device(y_star), A), log
