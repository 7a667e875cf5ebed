"""Data-aware row sketching by power iteration (drop-in for
parla/comps/sketchers/aware.py)."""
import numpy as np

import parla_amd.utils.linalg_wrappers as ulaw
import parla_amd.utils.sketching as usk
from parla_amd import ops
from parla_amd.comps.sketchers.oblivious import SkOpGA


def rs1(A, k, num_pass, rng):
    """aware.py:9-30."""
    assert num_pass >= 0
    assert k >= 1
    assert k <= min(A.shape)
    return RS1(sketch_op_gen=SkOpGA(), num_pass=num_pass, stabilizer=ulaw.orth,
               passes_per_stab=1)(A, k, rng)


class RowSketcher:
    """aware.py:33-73: ``S = sketcher(A, k, rng)`` with S of shape (n, k)."""

    def __call__(self, A, k, rng):
        raise NotImplementedError()


class RS1(RowSketcher):
    """Powered row sketching operator, aware.py:76-184.

    The passes ``A @ S`` and ``A.T @ S`` run as device GEMMs (or SJLT kernels
    for the first pass when the generator is sparse); the stabilizer receives
    and returns device tensors.  In exact arithmetic range(S) equals that of
    (A^T A)^(q) G for even ``num_pass`` and (A^T A)^q A^T G for odd
    (aware.py:149-158).
    """

    def __init__(self, sketch_op_gen, num_pass, stabilizer, passes_per_stab):
        self.sketch_op_gen = sketch_op_gen
        self.num_pass = num_pass
        self.stabilizer = stabilizer
        self.passes_per_stab = passes_per_stab

    def _stabilize(self, S, done):
        # with the library's orth, stabilisations before the last pass take one
        # CholeskyQR round (orth_intermediate): the returned basis is unchanged
        # up to column signs, at half the cost
        if self.stabilizer is ulaw.orth and done < self.num_pass:
            return ulaw.orth_intermediate(S)
        return self.stabilizer(S)

    def __call__(self, A, k, rng):
        assert self.num_pass >= 0
        rng = np.random.default_rng(rng)
        At = ops.as_device(A)
        done = 0
        if self.num_pass % 2 == 0:
            S = self.sketch_op_gen(At.shape[1], k, rng)
        else:
            S = usk.apply(At.t(), self.sketch_op_gen(At.shape[0], k, rng))
            done += 1
            if self.passes_per_stab == 1:
                S = self._stabilize(S, done)
        q = (self.num_pass - done) // 2
        while q > 0:
            S = usk.apply(At, S)
            done += 1
            if done % self.passes_per_stab == 0:
                S = self._stabilize(S, done)
            S = usk.apply(At.t(), S)
            done += 1
            if done % self.passes_per_stab == 0:
                S = self._stabilize(S, done)
            q -= 1
        if isinstance(S, ops.torch.Tensor):
            return ops.like_input(S, A)
        return S
