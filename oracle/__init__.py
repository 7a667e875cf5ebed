"""CPU oracle for the randomized-NLA hot path (test infrastructure only).

This package restates, in numpy/scipy, the algorithms of PARLA 0.1.4 that the
GPU library implements: the sketching operators, sketch-and-precondition least
squares (SPO, qr mode) with preconditioned LSQR, and the RS1 -> RF1 -> QB1 ->
SVD1 / EVD1 chain.  It is the checker the GPU results are compared against and
the CPU baseline that the benchmark times beside the GPU.

Only the test-suite, the smoke check and the benchmark's CPU-baseline leg may
import this package.  The GPU product path never calls into it.
"""
from .parla_oracle import *  # noqa: F401,F403
