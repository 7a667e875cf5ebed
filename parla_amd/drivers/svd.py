"""Randomized SVD backed by a QB decomposition (drop-in for parla/drivers/svd.py; SVD1)."""
import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.qb import QBDecomposer

EPS = np.finfo(float).eps


class SVDecomposer:
    """svd.py:10-55: ``U, s, Vh = alg(A, k, tol, over, rng)``."""

    def __call__(self, A, k, tol, over, rng):
        raise NotImplementedError()


def small_svd(B):
    """Thin SVD of a short-wide (or small) device matrix B (kb x n).

    B^T = Q2 R2 by the library's Householder QR, so B = R2^T Q2^T; the kb x kb
    SVD R2^T = Ur diag(s) Vr^T (rocSOLVER's dgesdd) gives U = Ur,
    Vh = (Q2 Vr)^T.  The large factor never leaves the device.
    """
    kb, n = B.shape
    if kb <= n:
        f = ops.geqrf_(ops.transpose(B))
        Q2 = f.Q()
        R2 = torch.triu(f.R_view)
        Ur, s, Vrh = ops.svd_small(R2.t())
        Vh = ops.transpose(ops.gemm(Q2, Vrh.t()))
        return Ur.contiguous(), s, Vh
    f = ops.geqrf_(B.contiguous().clone())
    Q3 = f.Q()
    R3 = torch.triu(f.R_view)
    Ur, s, Vh = torch.linalg.svd(R3, full_matrices=False)
    return ops.gemm(Q3, Ur), s, Vh.contiguous()


class SVD1(SVDecomposer):
    """svd.py:126-176: Q, B = qb(A, k + over); SVD of B; truncate; U = Q U."""

    def __init__(self, qb: QBDecomposer):
        self.qb = qb

    def __call__(self, A, k, tol, over, rng):
        rng = np.random.default_rng(rng)
        At = ops.as_device(A)
        Q, B = self.qb(At, k + over, tol, rng)
        U, s, Vh = small_svd(B)
        if over > 0:
            cutoff = min(k, s.numel())
            U, s, Vh = U[:, :cutoff], s[:cutoff], Vh[:cutoff, :]
        keep = ~(s < 10 * EPS)
        if not bool(keep.all()):
            idx = torch.nonzero(keep).reshape(-1)
            U, s, Vh = U.index_select(1, idx), s.index_select(0, idx), Vh.index_select(0, idx)
        U = ops.gemm(Q, U)
        return ops.like_input(U, A), ops.like_input(s.contiguous(), A), ops.like_input(Vh.contiguous(), A)
