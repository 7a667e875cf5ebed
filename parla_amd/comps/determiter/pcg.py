"""Preconditioned conjugate gradients (drop-in for parla/comps/determiter/pcg.py).

Same recurrences as the reference (pcg.py:5-47): the residual is recomputed
from x every ten steps, the log holds ||r|| before each step, and the loop
stops at tol * ||r_0||.  The operator and preconditioner are callables on
device vectors -- in PcSS1 they are one fused pass over A (A^T A v in a single
read, the library's normal pass) and two GEMVs with the preconditioner; the
n-vector updates and dot products are small device tensor operations, and the
host reads one scalar (||r||) per step for the stopping test, as the
reference's loop does.
"""
import numpy as np
import torch

from parla_amd import ops


def pcg(mv_mat, rhs, mv_pre, iter_lim, tol, x0):
    """pcg.py:5-47 -> (x, residuals) with residuals[i] = ||rhs - mat x_i|| (recurrence value)."""
    rt = ops.as_device(rhs)
    x = ops.as_device(x0).clone()
    residuals = []
    r = rt - mv_mat(x)
    d = mv_pre(r)
    delta1_new = torch.dot(r, d)
    cur_err = float(ops.nrm2(r).item())
    rel_tol = tol * cur_err
    i = 0
    while i < iter_lim and cur_err > rel_tol:
        residuals.append(cur_err)
        q = mv_mat(d)
        alpha = delta1_new / torch.dot(d, q)
        x.add_(alpha * d)
        if i % 10 == 0:
            r = rt - mv_mat(x)
        else:
            r.sub_(alpha * q)
        cur_err = float(ops.nrm2(r).item())
        s = mv_pre(r)
        delta1_old = delta1_new
        delta1_new = torch.dot(r, s)
        d = s + (delta1_new / delta1_old) * d
        i += 1
    return ops.like_input(x, rhs), np.array(residuals, dtype=np.float64)
