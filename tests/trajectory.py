"""Rounding envelopes for comparing iterative trajectories (LSQR error logs,
PCG residual sequences) between the device and the oracle.

An iterative method's later iterates depend chaotically on rounding once its
Krylov space is exhausted, so a fixed relative tolerance on the whole
sequence is either too loose early or too tight late.  The envelope is the
largest relative change that ONE extra rounding in the input makes in the
oracle's own sequence, times a safety factor: the device, which rounds
differently throughout, must stay inside it.
"""
import numpy as np


class Rounded:
    """Wraps an operator so A^T y is computed as (A^T (c y)) / c and S @ X as
    ((c S) @ X) / c: the same algebra with one extra rounding per product."""

    def __init__(self, op, c=3.0):
        self.op = op
        self.shape = op.shape
        self.c = c

    def matvec(self, x):
        return self.op @ x if not hasattr(self.op, "matvec") else self.op.matvec(x)

    def rmatvec(self, y):
        if hasattr(self.op, "rmatvec"):
            return self.op.rmatvec(self.c * y) / self.c
        return (self.op.T @ (self.c * y)) / self.c

    def __matmul__(self, X):
        return ((self.c * self.op) @ X) / self.c


ROUNDING_CONSTANTS = (3.0, 5.0, 7.0, 11.0, 13.0)


def rounded_matrix(A, c):
    """(c A) / c elementwise: A with one extra rounding per entry."""
    return (c * A) / c


def perturbed_change(ref, run):
    """Largest relative change, per entry, that one extra rounding makes in
    the oracle's sequence: `run(c)` reruns it with the constant c.  A single
    perturbation is not enough -- past the point where LSQR's trajectory
    turns chaotic its effect on one entry varies by three orders of
    magnitude between constants."""
    d = np.zeros(ref.size)
    for c in ROUNDING_CONSTANTS:
        p = np.atleast_1d(run(c))
        k = min(p.size, ref.size)
        d[:k] = np.maximum(d[:k], np.abs(p[:k] - ref[:k]) / np.abs(ref[:k]))
        d[k:] = np.inf
    return d


def envelope(change, floor=1e-6, factor=1e3):
    """Per-entry tolerance: `factor` times the relative change one rounding
    perturbation makes in the oracle's own sequence (never below `floor`)."""
    return np.maximum(floor, factor * change)


def assert_within_envelope(got, ref, env):
    k = min(got.size, ref.size, env.size)
    d = np.abs(got[:k] - ref[:k]) / np.abs(ref[:k])
    bad = np.nonzero(d > env[:k])[0]
    assert bad.size == 0, f"entries {bad[:5]} differ by {d[bad[:5]]} > envelope {env[bad[:5]]}"
