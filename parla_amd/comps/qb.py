"""QB decompositions (drop-in for parla/comps/qb.py; QB1)."""
import numpy as np

from parla_amd import ops
import parla_amd.utils.sketching as usk
from parla_amd.comps.rangefinders import RangeFinder, RF1  # noqa: F401
from parla_amd.comps.sketchers.aware import RS1  # noqa: F401


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
