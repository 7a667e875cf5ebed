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
    """Thin SVD of the k x n factor B on the device (ops.thin_svd: Householder
    QR of the taller orientation, then the k x k SVD of its triangle)."""
    return ops.thin_svd(B)


def svd1(A, k, over, tol, inner_num_pass, block_size, rng):
    """svd.py:58-123: SVD1 on QB2(RF1(RS1(Gaussian, inner_num_pass - 2)), block_size)."""
    from parla_amd.comps.qb import QB2, RF1, RS1
    from parla_amd.comps.sketchers.oblivious import SkOpGA
    import parla_amd.utils.linalg_wrappers as ulaw
    rng = np.random.default_rng(rng)
    rso_ = RS1(SkOpGA(), inner_num_pass - 2, ulaw.orth, 1)
    qb_ = QB2(RF1(rso_), block_size, overwrite_a=False)
    return SVD1(qb_)(A, k, tol, over, rng)


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
