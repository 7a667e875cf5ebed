"""Randomized symmetric EVD (drop-in for parla/drivers/evd.py: evd1, evd2,
EVD1 backed by a QB decomposition, EVD2 the Nystrom approximation of a PSD
matrix)."""
import warnings

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
        if C.shape[0] <= ops.JACOBI_MAX:
            lamb, U = ops.sym_eig(C)    # the library's Jacobi, lower triangle (as eigh)
        else:
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


def evd1(A, k, tol, over, inner_num_pass, block_size, rng):
    """evd.py:16-96: EVD1 on QB2(RF1(RS1(Gaussian, inner_num_pass - 2)),
    block_size), called with tol / 2 (EVD1 halves it again, as in the
    reference)."""
    from parla_amd.comps.qb import QB2, RF1, RS1
    from parla_amd.comps.sketchers.oblivious import SkOpGA
    import parla_amd.utils.linalg_wrappers as ulaw
    assert inner_num_pass >= 2
    rng = np.random.default_rng(rng)
    rso_ = RS1(SkOpGA(), inner_num_pass - 2, ulaw.orth, 1)
    qb_ = QB2(RF1(rso_), block_size, overwrite_a=False)
    return EVD1(qb_)(A, k, tol / 2, over, rng)


def evd2(A, k, over, num_passes, rng):
    """evd.py:99-163: EVD2 on RS1(Gaussian, num_passes - 1) (the reference
    passes the removed np.NaN as tol; NaN is meant)."""
    from parla_amd.comps.sketchers.aware import RS1
    from parla_amd.comps.sketchers.oblivious import SkOpGA
    import parla_amd.utils.linalg_wrappers as ulaw
    assert num_passes >= 1
    rng = np.random.default_rng(rng)
    return EVD2(RS1(SkOpGA(), num_passes - 1, ulaw.orth, 1))(A, k, np.nan, over, rng)


class EVD2(EVDecomposer):
    """evd.py:292-384: the rank-(k + over) Nystrom approximation
    (A S)(S^T A S)^+(A S)^T of a PSD A, regularised by
    nu = sqrt(n) eps ||A S||_F, truncated to rank k and where
    sigma_{i+1}^2 <= nu.  On the device: Y = A S + nu S in one GEMM, the
    Cholesky factor of S^T Y, B = Y R^{-1} by the MFMA strip solve, the thin
    SVD of B."""

    TOL_CONTROL = 'none'

    def __init__(self, sk_op):
        self.sk_op = sk_op

    def __call__(self, A, k, tol, over, rng):
        At = ops.as_device(A)
        assert k > 0
        n = At.shape[0]
        assert k < n
        if not np.isnan(tol):
            warnings.warn("""
            This EVDecomposer implementation cannot directly control
            approximation error. Parameter "tol" is being ignored.
            """)
        rng = np.random.default_rng(rng)
        S = self.sk_op(At, k + over, rng)
        if hasattr(S, "S") and not isinstance(S, torch.Tensor):
            S = S.S
        S = ops.as_device(S).contiguous()
        Y = ops.gemm(At, S)
        nu = np.sqrt(n) * EPS * float(ops.nrm2(Y.reshape(-1)).item())
        Y = ops.gemm(At, S, out=S.clone(), alpha=1.0, beta=nu)   # A S + nu S
        R = ops.cholesky_upper(ops.gemm(S.t(), Y))
        B = ops.trsm_mfma(Y, R) if R.shape[0] <= ops.CHOLQR_MAX_K else ops.trsm_rec_(Y.clone(), R)
        V, sigma, _ = ops.thin_svd(B)
        sh = sigma.cpu().numpy()
        comp = [k]
        for i in range(k - 1):
            if sh[i + 1] ** 2 <= nu:
                comp.append(i)
        r = min(comp)
        lamb = sigma[:r] ** 2 - nu
        return ops.like_input(V[:, :r].contiguous(), A), ops.like_input(lamb.contiguous(), A)
