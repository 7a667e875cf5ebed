"""Preconditioned saddle-point solvers (drop-in for
parla/comps/determiter/saddle.py: PcSS2 with both its branches)."""
import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.determiter.lsqr import lsqr_with_solution
from parla_amd.comps.preconditioning import a_lift_precond


class PrecondSaddleSolver:
    """saddle.py:20-85: ``x, y, errors = solver(A, b, c, delta, tol, iter_lim, R, upper_tri, z0)``."""

    def __call__(self, A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
        raise NotImplementedError()


def pcss1(A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
    """saddle.py:8-11."""
    return PcSS1()(A, b, c, delta, tol, iter_lim, R, upper_tri, z0)


def pcss2(A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
    """saddle.py:14-17."""
    return PcSS2()(A, b, c, delta, tol, iter_lim, R, upper_tri, z0)


class PcSS1(PrecondSaddleSolver):
    """PCG on the normal equations (A^T A + delta I) x = A^T b - c
    (saddle.py:88-176) with the preconditioner R R^T, or, when R has fewer
    columns than A (a low-rank SVD preconditioner V / s), R R^T + I - V V^T
    with R rescaled by the regularised singular values.  A^T A v is one read
    of A (the library's normal pass); the preconditioner is two or four
    GEMVs.  The reference overwrites the caller's R in the low-rank case
    (saddle.py:142-143); this one works on a copy."""

    ERROR_METRIC_INFO = """
        2-norm of the residual from the normal equations
    """

    def __call__(self, A, b, c, delta, tol, iter_lim, R, upper_tri, z0):
        from parla_amd.comps.determiter.pcg import pcg
        At = ops.as_device(A).contiguous()
        m, n = At.shape
        bt = torch.zeros(m, dtype=ops.F64, device=At.device) if b is None else ops.as_device(b).reshape(-1)
        if bt.dim() != 1:
            raise NotImplementedError()
        if upper_tri:
            raise NotImplementedError()
        Rt = ops.as_device(R).clone().contiguous()
        pc_dim = Rt.shape[1]
        full = pc_dim == n
        V = None
        if not full:
            sv = 1.0 / torch.linalg.vector_norm(Rt, dim=0)
            V = ops.scale_cols(Rt, sv)
            sv = torch.sqrt(sv * sv + delta)
            Rt = ops.scale_cols(V, sv / sv[-1], invert=True)
        normal = ops.NormalOp(At) if n <= 1024 else (lambda v: ops.gemv(At, ops.gemv(At, v), trans=True))

        def mv_pre(vec):
            res = ops.gemv(Rt, ops.gemv(Rt, vec, trans=True))
            if not full:
                res = res + vec - ops.gemv(V, ops.gemv(V, vec, trans=True))
            return res

        def mv_gram(vec):
            out = normal(vec.contiguous())
            if delta != 0:
                out = out + delta * vec
            return out

        rhs = ops.gemv(At, bt, trans=True)
        if c is not None:
            rhs = rhs - ops.as_device(c).reshape(-1)
        if z0 is None or not full:
            x0 = torch.zeros(n, dtype=ops.F64, device=At.device)
        else:
            x0 = ops.gemv(Rt, ops.as_device(z0).reshape(-1))
        x, residuals = pcg(mv_gram, rhs, mv_pre, iter_lim, tol, x0)
        y = ops.gemv(At, x, y=bt.clone(), alpha=-1.0, beta=1.0)
        ref = b if b is not None else c
        return ops.like_input(x, ref), ops.like_input(y, ref), residuals


class PcSS2(PrecondSaddleSolver):
    """LSQR on the preconditioned problem (saddle.py:180-217)."""

    ERROR_METRIC_INFO = """
        2-norm of the residual from the preconditioned normal equations
        (preconditioning on the left and right).
    """

    def __call__(self, A, b, c, delta, tol, iter_lim, R, upper_tri, z0, want_y=True):
        At = ops.as_device(A)
        m, n = At.shape
        k = 1 if (b is None or b.ndim == 1) else b.shape[1]
        A_pc, M_fwd, M_adj = a_lift_precond(At, delta, R, upper_tri, k)
        c_zero = c is None or float(ops.nrm2(ops.as_device(c)).item()) == 0.0
        if c_zero:
            bt = ops.as_device(b)
            if delta > 0:
                bt = torch.cat((bt, torch.zeros(n, dtype=bt.dtype, device=bt.device)))
            z0t = None if z0 is None else ops.as_device(z0)
            result, x = lsqr_with_solution(A_pc, bt, atol=tol, btol=tol, iter_lim=iter_lim, x0=z0t)
            if x is None:  # M applied by the library's triangular solve / GEMV
                x = ops.as_device(M_fwd(ops.as_device(result[0])))
            y = None
            if want_y:
                y = ops.gemv(At, x, y=bt[:m].clone(), alpha=-1.0, beta=1.0)
            return (ops.like_input(x, b), None if y is None else ops.like_input(y, b), result[7])
        b_zero = b is None or float(ops.nrm2(ops.as_device(b)).item()) == 0.0
        if b_zero:
            return self._underdetermined(At, c, delta, tol, iter_lim, R, upper_tri, k)
        raise ValueError('One of "b" or "c" must be zero.')

    @staticmethod
    def _underdetermined(At, c, delta, tol, iter_lim, R, upper_tri, k):
        """saddle.py:203-214: LSQR on A_pc^T = M^T [A; sqrt(delta) I]^T with
        right-hand side M^T c.  A_pc is formed explicitly on the device ([A;
        sqrt(delta) I] R^{-1} by the recursive MFMA triangular solve, or
        [A; sqrt(delta) I] M by GEMM) and transposed, and the device LSQR
        runs on that wide matrix.  y = the LSQR solution (its first m entries
        when delta > 0); x = (A^T y - c) / delta, or NaN when delta = 0 (the
        reference writes the removed ``np.NaN`` there, saddle.py:212)."""
        m, n = At.shape
        ct = ops.as_device(c).reshape(-1)
        Rt = ops.as_device(R)
        sd = float(np.sqrt(delta))
        rows = m + (n if delta > 0 else 0)
        if upper_tri:
            B = torch.empty((rows, n), dtype=ops.F64, device=At.device)
            B[:m].copy_(At)
            if delta > 0:
                B[m:].zero_()
                B[m:].diagonal().fill_(sd)
            ops.trsm_rec_(B, Rt)
            c_pc = ops.trsv_(Rt, ct.clone(), trans=True)
        else:
            B = ops.gemm(At, Rt)
            if delta > 0:
                B = torch.cat((B, sd * Rt), 0)
            c_pc = ops.gemv(Rt, ct, trans=True)
        W = ops.transpose(B)
        result, _ = lsqr_with_solution(W, c_pc, atol=tol, btol=tol, iter_lim=iter_lim)
        y = ops.as_device(result[0])
        if delta > 0:
            y = y[:m].contiguous()
            x = (ops.gemv(At, y, trans=True) - ct) / delta
        else:
            x = torch.full((n,), float('nan'), dtype=ops.F64, device=At.device)
        return (ops.like_input(x, c), ops.like_input(y, c), result[7])
