"""Stabilizers for power iteration (drop-in for parla/utils/linalg_wrappers.py:6-12)."""
from .. import ops
from ..utils.sketching import DenseSketch


def _dev(S):
    if isinstance(S, DenseSketch):
        return S.S, S
    return ops.as_device(S), S


def orth(S):
    """utils/linalg_wrappers.py:6-7: orthonormal basis of range(S).

    Tall, well-conditioned S (the power-iteration case) goes through
    CholeskyQR2 -- GEMM-shaped work on the MFMA units -- and otherwise through
    Householder QR.  Q equals la.qr(S, mode='economic')[0] up to the sign of
    each column, which no consumer (RS1, RF1, QB1, SVD1, EVD1) depends on.
    """
    St, ref = _dev(S)
    return ops.like_input(ops.orth_q(St), ref)


def orth_intermediate(S):
    """A well-conditioned basis of range(S) for a stabilisation that is not the
    last of a power iteration: one CholeskyQR round, S R^{-1} = Q T with T
    upper triangular and within kappa(S)^2 u of the identity.  The next
    product A (Q T) has the QR factor Q' of A Q up to that same T, so the
    basis the iteration finally returns (stabilised by `orth`) is the one the
    Householder stabiliser yields, up to column signs."""
    St, ref = _dev(S)
    return ops.like_input(ops.orth_q(St, rounds=1), ref)


def lu_stabilize(S):
    """utils/linalg_wrappers.py:10-12: P @ L of LU with partial pivoting."""
    St, ref = _dev(S)
    return ops.like_input(ops.lu_pl(St), ref)
