"""Overdetermined least squares by sketch-and-precondition (drop-in for
parla/drivers/least_squares.py: OverLstsqSolver, dim_checks, SPO, spo3).

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
    """Sketch-and-precondition for overdetermined least squares (least_squares.py:206-369)."""

    def __init__(self, sketch_op_gen, sampling_factor: int, mode='qr'):
        self.sketch_op_gen = sketch_op_gen
        self.sampling_factor = sampling_factor
        self.mode = mode
        self.iterative_solver = dsad.PcSS2()

    def __call__(self, A, b, delta, tol, iter_lim, rng, logging=True):
        if self.mode not in ('qr',):
            if self.mode in ('chol', 'svd'):
                raise NotImplementedError(f"SPO mode '{self.mode}' is not part of this library (qr only)")
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
        rows = d + (n_cols if delta > 0 else 0)
        Mx = torch.empty((rows, n_cols + 1), dtype=ops.F64, device=dev)
        _sketch_into(S, At, Mx[:d, :n_cols])
        log.time_sketch = clock() - tic

        tic = clock()
        if delta > 0:
            Mx[d:].zero_()
            Mx[d:, :n_cols].diagonal().fill_(sqrt_delta)
        b_ske = torch.empty(d, dtype=ops.F64, device=dev)
        _sketch_into(S, bt, b_ske)
        Mx[:d, n_cols].copy_(b_ske)
        fac = ops.geqrf_(Mx, n_cols)
        R = fac.R_view  # upper triangle of the factored buffer (ld = n + 1)
        log.time_factor = clock() - tic

        tic = clock()
        z_ske = Mx[:n_cols, n_cols].clone()
        x_ske = ops.trsv_(R, z_ske.clone())
        # presolve decision (least_squares.py:344-351)
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
        Rc = torch.triu(R)
        res = self.iterative_solver(At, bt, None, delta, tol, iter_lim, Rc, True, z_ske, want_y=False)
        log.time_iterate = clock() - tic

        if logging:
            ar0 = ops.gemv(At, bt, trans=True)
            ops.trsv_(Rc, ar0, trans=True)
            log.wrap_up(res[2], float(ops.nrm2(ar0).item()))
            log.error_desc = self.iterative_solver.ERROR_METRIC_INFO
        return ops.like_input(ops.as_device(res[0]), A), log
