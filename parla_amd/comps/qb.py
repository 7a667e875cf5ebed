"""QB decompositions (drop-in for parla/comps/qb.py: qb, QB1, QB2, QB3,
project_out)."""
import warnings

import numpy as np
import torch

from parla_amd import ops
import parla_amd.utils.sketching as usk
from parla_amd.comps.rangefinders import RangeFinder, RF1  # noqa: F401
from parla_amd.comps.sketchers.aware import RS1, RowSketcher  # noqa: F401
from parla_amd.comps.sketchers.oblivious import SkOpGA
import parla_amd.utils.linalg_wrappers as ulaw


def qb(num_passes, A, k, rng):
    """qb.py:16-82: QB1 on a Gaussian RS1 with num_passes - 2 power passes
    (the reference passes the removed np.NaN as tol; NaN is meant)."""
    rng = np.random.default_rng(rng)
    rso_ = RS1(SkOpGA(), num_passes - 2, ulaw.orth, 1)
    return QB1(RF1(rso_))(A, k, np.nan, rng)


def project_out(Qi, Q, as_list=False):
    """qb.py:598-613: Qi - Q (Q^T Qi), two device GEMMs."""
    if as_list:
        raise NotImplementedError()
    if Q.shape[1] == 0:
        return Qi
    return ops.gemm(Q, ops.gemm(Q.t(), Qi), out=Qi.clone(), alpha=-1.0, beta=1.0)


class QBDecomposer:
    """qb.py:241-283: ``Q, B = qb(A, k, tol, rng)`` with B = Q^T A."""

    TOL_CONTROL = 'none'

    def __call__(self, A, k, tol, rng):
        raise NotImplementedError()


class QB1(QBDecomposer):
    """qb.py:286-352: Q from a rangefinder, then B = Q^T A (one device GEMM)."""

    TOL_CONTROL = 'unknown'

    def __init__(self, rf: RangeFinder):
        self.rangefinder = rf

    def __call__(self, A, k, tol, rng):
        assert k > 0
        assert k <= min(A.shape)
        if not np.isnan(tol):
            assert tol >= 0
            assert tol < 1
        rng = np.random.default_rng(rng)
        At = ops.as_device(A)
        Q = self.rangefinder(At, k, tol, rng)
        B = usk.apply(Q.t(), At)
        return ops.like_input(Q, A), ops.like_input(B, A)


class QB2(QBDecomposer):
    """qb.py:355-482: blocks of `blk` rangefinder columns, each projected out of
    the previous ones and re-orthonormalised, with A deflated by Qi Bi after
    every block (on a device copy unless overwrite_a, and then only for a
    device tensor A); stops at k columns or once ||A - Q B||_F <= tol ||A||_F
    (the Frobenius norms tracked as the reference does, one host read per
    block when tol is active)."""

    TOL_CONTROL = 'full'

    def __init__(self, rf: RangeFinder, blk: int, overwrite_a: bool):
        self.rangefinder = rf
        self.blk = blk
        self.overwrite_a = overwrite_a

    def __call__(self, A, k, tol, rng):
        At = ops.as_device(A)
        if not (self.overwrite_a and At is A):
            At = At.clone()
        At = At.contiguous()
        assert k > 0
        small = min(At.shape)
        if not k <= small:
            warnings.warn(f"""
            The target rank k = {k} is larger than min({tuple(At.shape)}).
            We will proceed with target rank k = {small}.
            """)
            k = small
        use_tol = not np.isnan(tol) and tol > 0
        if use_tol:
            sq_norm = float(ops.nrm2(At.reshape(-1)).item()) ** 2
            abs_sq_tol = sq_norm * tol ** 2
        rng = np.random.default_rng(rng)
        m, n = At.shape
        Qall = torch.empty((m, k), dtype=ops.F64, device=At.device)
        Ball = torch.empty((k, n), dtype=ops.F64, device=At.device)
        done, blk = 0, self.blk
        while True:
            if done + blk > k:
                blk = k - done
            Qi = ops.as_device(self.rangefinder(At, blk, np.nan, rng))
            Qi = project_out(Qi, Qall[:, :done])
            Qi = ops.orth_q(Qi)
            Bi = ops.gemm(Qi.t(), At)
            Qall[:, done:done + blk].copy_(Qi)
            Ball[done:done + blk].copy_(Bi)
            done += blk
            ops.gemm(Qi, Bi, out=At, alpha=-1.0, beta=1.0)   # A -= Qi Bi
            if use_tol:
                sq_norm -= float(ops.nrm2(Bi.reshape(-1)).item()) ** 2
                if sq_norm <= abs_sq_tol:
                    break
            if done >= k:
                break
        Q, B = Qall[:, :done], Ball[:done]
        return ops.like_input(Q, A), ops.like_input(B, A)


class QB3(QBDecomposer):
    """qb.py:485-595: S = sk_op(A, k) once, G = A S and H = A^T G (the only
    passes over A), then blocks of Q and B from (G, H, S) with QR
    re-orthogonalisation and a transposed triangular solve; early stopping on
    the Frobenius-norm estimate.  Householder QR (LAPACK geqrf semantics)
    supplies each block's R."""

    TOL_CONTROL = 'early stopping'

    def __init__(self, sk_op: RowSketcher, blk: int):
        self.sk_op = sk_op
        self.blk = blk

    def __call__(self, A, k, tol, rng):
        At = ops.as_device(A)
        assert k > 0
        assert k < min(At.shape)
        use_tol = not np.isnan(tol) and tol > 0
        if use_tol:
            sq_norm = float(ops.nrm2(At.contiguous().reshape(-1)).item()) ** 2
            abs_sq_tol = sq_norm * tol ** 2
        rng = np.random.default_rng(rng)
        S = self.sk_op(At, k, rng)
        if isinstance(S, usk.DenseSketch):
            S = S.S
        if not isinstance(S, (torch.Tensor, np.ndarray)):
            raise RuntimeError("""
            This implementation requires the sketching routine to return a
            dense matrix, as represented by a numpy ndarray. We received a
            matrix of type %s
            """ % str(type(S)))
        S = ops.as_device(S)
        G = ops.gemm(At, S)
        H = ops.gemm(At.t(), G)
        m, n = At.shape
        Qall = torch.empty((m, k), dtype=ops.F64, device=At.device)
        Ball = torch.empty((k, n), dtype=ops.F64, device=At.device)
        done = 0
        for i in range(int(np.ceil(k / self.blk))):
            b0, b1 = i * self.blk, min((i + 1) * self.blk, k)
            Q, B = Qall[:, :done], Ball[:done]
            Si = S[:, b0:b1]
            BSi = ops.gemm(B, Si) if done else torch.zeros((0, b1 - b0), dtype=ops.F64, device=At.device)
            Yi = G[:, b0:b1].contiguous()
            if done:
                ops.gemm(Q, BSi, out=Yi, alpha=-1.0, beta=1.0)
            f1 = ops.geqrf_(Yi.clone())
            Qi, Ri = f1.Q(), f1.R()
            Qi = project_out(Qi, Q)
            f2 = ops.geqrf_(Qi.contiguous())
            Qi = f2.Q()
            Ri = ops.gemm(f2.R(), Ri)
            Bi = H[:, b0:b1].t().contiguous()
            if done:
                ops.gemm(ops.gemm(Yi.t(), Q), B, out=Bi, alpha=-1.0, beta=1.0)
                ops.gemm(BSi.t(), B, out=Bi, alpha=-1.0, beta=1.0)
            # Bi <- Ri^{-T} Bi  (Bi^T Ri^{-1} transposed)
            Bi = ops.transpose(ops.trsm_mfma(Bi.t(), Ri))
            Qall[:, done:done + (b1 - b0)].copy_(Qi)
            Ball[done:done + (b1 - b0)].copy_(Bi)
            done += b1 - b0
            if use_tol:
                sq_norm -= float(ops.nrm2(Bi.reshape(-1)).item()) ** 2
                if sq_norm <= abs_sq_tol:
                    break
        Q, B = Qall[:, :done], Ball[:done]
        return ops.like_input(Q, A), ops.like_input(B, A)
