"""Orthogonal rangefinders (drop-in for parla/comps/rangefinders.py)."""
import warnings

import numpy as np

import parla_amd.utils.linalg_wrappers as ulaw
import parla_amd.utils.sketching as usk
from parla_amd import ops
from parla_amd.comps.sketchers.aware import RS1, RowSketcher
from parla_amd.comps.sketchers.oblivious import SkOpGA


def rf1(A, k, num_pass, rng):
    """rangefinders.py:22-71."""
    rng = np.random.default_rng(rng)
    rso_ = RS1(sketch_op_gen=SkOpGA(), num_pass=num_pass - 1, stabilizer=ulaw.orth,
               passes_per_stab=1)
    return RF1(rso_)(A, k, np.nan, rng)


class RangeFinder:
    """rangefinders.py:78-123: ``Q = rf(A, k, tol, rng)``."""

    def __call__(self, A, k, tol, rng):
        raise NotImplementedError()


class RF1(RangeFinder):
    """rangefinders.py:126-188: S = rso(A, k); Y = A @ S; Q = qr(Y)[0] (Q up to column
    signs, see parla_amd.utils.linalg_wrappers.orth)."""

    def __init__(self, rso: RowSketcher):
        self.rso = rso

    def __call__(self, A, k, tol, rng):
        assert k > 0
        assert k <= min(A.shape)
        if not np.isnan(tol):
            warnings.warn("This RangeFinder implementation cannot directly control approximation "
                          "error. Parameter \"tol\" is being ignored.")
        rng = np.random.default_rng(rng)
        At = ops.as_device(A)
        S = self.rso(At, k, rng)
        Y = usk.apply(At, S)
        Q = ops.orth_q(Y)
        return ops.like_input(Q, A)
