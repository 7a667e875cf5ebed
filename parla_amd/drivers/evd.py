"""Randomized symmetric EVD backed by a QB decomposition (drop-in for
parla/drivers/evd.py; EVD1)."""
import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.qb import QBDecomposer

EPS = np.finfo(float).eps


class EVDecomposer:
    """evd.py:171-208: ``V, lamb = alg(A, k, tol, over, rng)``."""

    TOL_CONTROL = 'none'

    def __call__(self, A, k, tol, over, rng):
        raise NotImplementedError()


class EVD1(EVDecomposer):
    """evd.py:211-289: QB with tol/2, C = B Q, eigh(C), keep the r largest |lambda|, V = Q U."""

    TOL_CONTROL = 'unknown'

    def __init__(self, qb: QBDecomposer):
        self.qb = qb

    def __call__(self, A, k, tol, over, rng):
        assert k > 0
        assert k <= min(A.shape)
        if not np.isnan(tol):
            assert tol >= 0
            assert tol < np.inf
        rng = np.random.default_rng(rng)
        At = ops.as_device(A)
        Q, B = self.qb(At, k + over, tol / 2, rng)
        C = ops.gemm(B, Q)
        lamb, U = torch.linalg.eigh(C, UPLO='L')
        alamb = np.abs(lamb.detach().cpu().numpy())
        d = Q.shape[1]
        r = min(k, d, int(np.count_nonzero(alamb > 10 * EPS)))
        order = np.argsort(-1 * np.abs(alamb))[:r]
        idx = torch.as_tensor(order, dtype=torch.int64, device=At.device)
        U = U.index_select(1, idx).contiguous()
        lamb = lamb.index_select(0, idx).contiguous()
        V = ops.gemm(Q, U)
        return ops.like_input(V, A), ops.like_input(lamb, A)
