"""LSQR with device-resident recurrences (drop-in for
parla/comps/determiter/lsqr.py:98-574).

Same arguments, same 10-tuple, same stopping tests and istop codes, and the
PARLA modification that returns the per-iteration ``arnorms`` log at index 7
(the scalar arnorm on the zero-residual early exit, lsqr.py:392-395).  A may
be a dense matrix (numpy or CUDA tensor) or the operator built by
``a_lift_precond``; the whole iteration then runs in libparla_amd.so and only
the final state is read back.
"""
import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.preconditioning import PrecondOperator

_MSG = ('The exact solution is  x = 0                              ',
        'Ax - b is small enough, given atol, btol                  ',
        'The least-squares solution is good enough, given atol     ',
        'The estimate of cond(Abar) has exceeded conlim            ',
        'Ax - b is small enough for this machine                   ',
        'The least-squares solution is good enough for this machine',
        'Cond(Abar) seems to be too large for this machine         ',
        'The iteration limit has been reached                      ')


def _operands(A):
    """(matrix, R or None, sqrt_delta) for the device iteration."""
    if isinstance(A, PrecondOperator):
        if A.upper_tri:
            return A.A, A.R, A.sqrt_delta
        # general right preconditioner M: iterate on the explicit [A M; sd M]
        AM = ops.gemm(A.A, A.R)
        if A.sqrt_delta > 0:
            AM = torch.cat((AM, A.sqrt_delta * A.R), 0)
        return AM, None, 0.0
    if hasattr(A, "matvec") and not isinstance(A, (np.ndarray, torch.Tensor)):
        raise TypeError("parla_amd.lsqr runs on the device: pass a dense matrix or the operator "
                        "returned by parla_amd.comps.preconditioning.a_lift_precond")
    return ops.as_device(A), None, 0.0


def lsqr(A, b, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8, iter_lim=None, show=False,
         calc_var=False, x0=None):
    return lsqr_with_solution(A, b, damp, atol, btol, conlim, iter_lim, show, calc_var, x0)[0]


def lsqr_with_solution(A, b, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8, iter_lim=None, show=False,
                       calc_var=False, x0=None):
    """``(lsqr 10-tuple, x_orig)``: x_orig = M @ result[0] for a preconditioned
    operator, computed with the explicit M the device iteration used (None for
    a plain matrix or a non-triangular M)."""
    M, R, sd = _operands(A)
    m, n = M.shape
    bt = ops.as_device(b)
    if bt.dim() > 1:
        bt = bt.squeeze()
    bt = bt.reshape(-1)
    if sd > 0 and bt.numel() == m + n:
        if bool(torch.any(bt[m:] != 0)):
            raise NotImplementedError("lifted right-hand sides must end in n zeros")
        bt = bt[:m]
    if bt.numel() != m:
        raise ValueError(f"b has {bt.numel()} entries, A has {m} rows")
    if iter_lim is None:
        iter_lim = 2 * n
    x0t = ops.as_device(x0).reshape(-1) if x0 is not None else None
    if show:
        print(' ')
        print('LSQR            Least-squares solution of  Ax = b')
        print(f'The matrix A has {m + (n if sd > 0 else 0)} rows and {n} columns')
        print('damp = %20.14e   calc_var = %8g' % (damp, calc_var))
        print('atol = %8.2e                 conlim = %8.2e' % (atol, conlim))
        print('btol = %8.2e               iter_lim = %8g' % (btol, iter_lim))
    x, xo, var, arn, res = ops.lsqr_run(M, bt, R=R, sqrt_delta=sd, x0=x0t, damp=damp, atol=atol,
                                        btol=btol, conlim=conlim, iter_lim=iter_lim, calc_var=calc_var,
                                        want_orig=True)
    if res.n_log < 0:
        if show:
            print(_MSG[0])
        arnorms = res.arnorm
    else:
        arnorms = arn[: res.n_log].detach().cpu().numpy().copy()
    if show:
        print(' ')
        print('LSQR finished')
        print(_MSG[res.istop])
        print(' ')
        print('istop =%8g   r1norm =%8.1e   anorm =%8.1e   arnorm =%8.1e'
              % (res.istop, res.r1norm, res.anorm, res.arnorm))
        print('itn   =%8g   r2norm =%8.1e   acond =%8.1e   xnorm  =%8.1e'
              % (res.itn, res.r2norm, res.acond, res.xnorm))
        print(' ')
    ref = b
    out = (ops.like_input(x, ref), int(res.istop), int(res.itn), res.r1norm, res.r2norm, res.anorm,
           res.acond, arnorms, res.xnorm, ops.like_input(var, ref))
    return out, (xo if R is not None else None)
