"""Stabilizers for power iteration (drop-in for parla/utils/linalg_wrappers.py:6-12)."""
from .. import ops
from ..utils.sketching import DenseSketch


def _dev(S):
    if isinstance(S, DenseSketch):
        return S.S, S
    return ops.as_device(S), S


def orth(S):
    """utils/linalg_wrappers.py:6-7: Q factor of the economic Householder QR."""
    St, ref = _dev(S)
    return ops.like_input(ops.qr_q(St), ref)


def lu_stabilize(S):
    """utils/linalg_wrappers.py:10-12: P @ L of LU with partial pivoting."""
    St, ref = _dev(S)
    return ops.like_input(ops.lu_pl(St), ref)
