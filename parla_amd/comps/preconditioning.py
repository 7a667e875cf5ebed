"""Preconditioned operators for LSQR (drop-in for parla/comps/preconditioning.py)."""
import numpy as np
import torch

from parla_amd import ops


def a_lift(A, scale):
    """preconditioning.py:6-13: [A; scale * I] (materialised)."""
    At = ops.as_device(A)
    if scale == 0:
        return ops.like_input(At, A)
    n = At.shape[1]
    eye = torch.eye(n, dtype=ops.F64, device=At.device) * scale
    return ops.like_input(torch.cat((At, eye), 0), A)


class PrecondOperator:
    """A_pc = [A; sqrt(delta) I] M with M = R^{-1} (upper_tri) or M = R.

    Held on the device; ``lsqr`` recognises it and runs the fused device
    iteration.  ``matvec`` / ``rmatvec`` apply it directly (two kernels each),
    preconditioning.py:26-38 and :45-54.
    """

    def __init__(self, A, delta, R, upper_tri):
        self.A = ops.as_device(A)
        self.R = ops.as_device(R)
        self.delta = float(delta)
        self.sqrt_delta = float(np.sqrt(delta))
        self.upper_tri = bool(upper_tri)
        self.m = self.A.shape[0]
        lift = self.m + (self.A.shape[1] if delta > 0 else 0)
        self.shape = (lift, self.R.shape[1])

    def m_fwd(self, z):
        zt = ops.as_device(z).clone()
        if self.upper_tri:
            return ops.like_input(ops.trsv_(self.R, zt), z)
        return ops.like_input(ops.gemv(self.R, zt), z)

    def m_adj(self, w):
        wt = ops.as_device(w).clone()
        if self.upper_tri:
            return ops.like_input(ops.trsv_(self.R, wt, trans=True), w)
        return ops.like_input(ops.gemv(self.R, wt, trans=True), w)

    def matvec(self, arg):
        work = ops.as_device(self.m_fwd(ops.as_device(arg)))
        out = ops.gemv(self.A, work)
        if self.sqrt_delta > 0:
            out = torch.cat((out, self.sqrt_delta * work))
        return ops.like_input(out, arg)

    def rmatvec(self, arg):
        a = ops.as_device(arg)
        work = ops.gemv(self.A, a[: self.m].contiguous(), trans=True)
        if self.sqrt_delta > 0:
            work = work + self.sqrt_delta * a[self.m:]
        return ops.like_input(ops.as_device(self.m_adj(work)), arg)


def a_lift_precond(A, delta, R, upper_tri=False, k=1):
    """preconditioning.py:16-67 -> (A_precond, M_fwd, M_adj)."""
    if k != 1:
        raise NotImplementedError()
    op = PrecondOperator(A, delta, R, upper_tri)
    return op, op.m_fwd, op.m_adj


def svd_right_precond(A_ske):
    """preconditioning.py:70-79: (M, U, sigma, Vh) with M = V / sigma over the
    numerical rank of the sketch (sigma > sigma_1 * ncols * eps), computed on
    the device; results have the kind of the input."""
    t = ops.as_device(A_ske)
    M, U, sig, Vh = ops.svd_right_precond(t)
    return tuple(ops.like_input(v, A_ske) for v in (M, U, sig, Vh))
