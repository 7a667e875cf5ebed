"""Mixed-precision sketch-and-precondition least squares with FP64 refinement.

Same call contract as SPO (drivers/least_squares.py:206-369 of the reference):
``x, log = SPOMixed(...)(A, b, delta, tol, iter_lim, rng)``, with ``tol`` read
as the target backward error eta(x) = ||A^T r|| / (||A||_F ||r||).

  1. Sketch.  S = 2^-5 / sqrt(d) * Q with Q an integer-valued Gaussian
     (|Q| <= 256, exact in BF16).  [A | b] is split into `npieces` BF16 integer
     pieces per 256-row chunk, and [S A | S b] is formed on BF16 MFMA with
     exact FP32 partial sums and FP64 accumulation: it equals S times A
     truncated to 9 * npieces bits below each chunk-column maximum.
  2. Factor.  QR of S A in FP64; z = Q^T S b.  Sketches of at most 1024 rows
     by CholeskyQR2 (Householder when S A is too ill-conditioned for it),
     taller ones by the communication-avoiding Householder QR.
  3. Precondition explicitly.  B = A R^{-1} in FP64 by a recursive MFMA TRSM,
     stored in FP32 (kappa(B) is about 3, so FP32 storage costs nothing in
     the refinement's contraction rate).
  4. Refine in FP64.  x0 = R^{-1} z; then, until eta(x) <= tol or a step no
     longer halves eta (the FP64 floor: x itself is rounded to FP64),
         g  = A^T (b - A x)                 compensated FP64, one read of A
         dz = CG(B^T B, R^{-T} g)            FP32 B, FP64 vectors, one read of B
         x <- x + R^{-1} dz                   FP64 triangular solves.
     The fixed point is A^T r = 0 evaluated with error-free transformations,
     so the FP32 operator only sets the contraction rate, not the accuracy.

The log's ``errors`` are the backward errors after each refinement step.
"""
import time
import warnings

import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.determiter.logging import SketchAndPrecondLog
from parla_amd.drivers.least_squares import OverLstsqSolver, dim_checks


def cg_normal(op, rhs, max_iter, tol, check_every=4):
    """z ~= (B^T B)^{-1} rhs by conjugate gradients.  The recurrences (dots,
    alpha, beta, the z / r / p updates) run in one device kernel per step
    (pamd_bt_cg on a one-problem descriptor).  The host reads ||r|| after
    `check_every` steps; from the drop measured there it predicts the step at
    which ||r|| <= tol ||r_0|| (CG on a well-conditioned B contracts at a
    near-constant rate) and reads again there, then every `check_every`
    steps.  Each read is a synchronisation, so fewer reads keep the GPU fed."""
    n = rhs.numel()
    dev = rhs.device
    desc = torch.tensor([[1, n, 0, 0, 0, 0, 0, 0]], dtype=torch.int64, device=dev)
    state = torch.tensor([0.0, 1.0, 0.0, 0.0], dtype=torch.float64, device=dev)
    z, r, p, q = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4))
    L, st = ops.lib(), ops.stream(dev)

    def step(mode):
        ops.check(L.pamd_bt_cg(1, ops._p(desc), mode, ops._p(state), ops._p(rhs), ops._p(r), ops._p(p), ops._p(q),
                               ops._p(z), st), "pamd_bt_cg")

    step(0)
    rr0 = state[0:1].clone()
    it, check_at, first = 0, check_every, True
    while it < max_iter:
        op(p, q)
        step(1)
        it += 1
        if it == check_at:
            rr, rr_0 = torch.cat((state[0:1], rr0)).cpu().numpy()  # one read
            if rr <= (tol * tol) * rr_0:
                break
            check_at = it + check_every
            if first and 0.0 < rr < rr_0:
                # iterations to go at the measured rate (rr falls by rr / rr_0 every `it` steps)
                need = int(np.ceil(it * np.log((tol * tol) * rr_0 / rr) / np.log(rr / rr_0)))
                check_at = it + max(1, min(need, max_iter - it))
            first = False
    return z, it


class SPOMixed(OverLstsqSolver):
    """Sketch-and-precondition with an exact BF16 sketch, an FP32-stored
    preconditioned operator and FP64 iterative refinement to ``tol``."""

    def __init__(self, sampling_factor=4, npieces=4, inner_tol=1e-6, max_inner=64, max_refine=10,
                 backend="numpy"):
        if not 1 <= npieces <= 5:
            raise ValueError("npieces must be in 1..5")
        self.sampling_factor = sampling_factor
        self.npieces = npieces
        self.inner_tol = inner_tol
        self.max_inner = max_inner
        self.max_refine = max_refine
        self.backend = backend

    def __call__(self, A, b, delta, tol, iter_lim, rng, logging=True):
        if delta != 0:
            raise NotImplementedError("SPOMixed solves the unregularised problem (delta = 0)")
        At = ops.as_device(A).contiguous()
        bt = ops.as_device(b).reshape(-1).contiguous()
        m, n = At.shape
        if n > 1024:
            raise NotImplementedError("SPOMixed supports n <= 1024")
        d = dim_checks(self.sampling_factor, m, n)
        rng = np.random.default_rng(rng)
        dev = At.device
        log = SketchAndPrecondLog()

        def clock():
            if logging:
                torch.cuda.synchronize(dev)
                return time.time()
            return 0.0

        tic = clock()
        if self.backend == "numpy":
            Q = ops.quantize_q(rng.standard_normal((d, m)))
        else:
            Q = ops.gauss_q(d, m, int(rng.integers(0, 2 ** 63)), dev)
        split = ops.OzakiSplit(At, bt, self.npieces)
        Mx = ops.ozaki_sketch(Q, d, split, 2.0 ** -5 / np.sqrt(min(d, m)))
        del split, Q
        log.time_sketch = clock() - tic

        tic = clock()
        # sketches below the communication-avoiding QR's range: CholeskyQR2
        # of S A when S A is well enough conditioned for it, else Householder
        rz = ops.sketch_rz_cholqr2(Mx, n) if d < ops.QR_R_MIN_ROWS else None
        if rz is not None:
            R, z = rz
        else:
            fac = ops.qr_r_(Mx, n)
            R = torch.triu(fac.R_view).contiguous()
            z = Mx[:n, n].clone()
        B32 = ops.trsm_rec_f32(At, R)
        log.time_factor = clock() - tic

        tic = clock()
        x = ops.trsv_(R, z)
        res = ops.Residual(At)
        normal = ops.NormalOp(B32)
        anorm_d = ops.nrm2(At.reshape(-1))  # read with the first step's norms
        anorm = None
        log.time_presolve = clock() - tic

        tic = clock()
        etas = []
        inner_total = 0
        best_eta, best_x = np.inf, None
        for step in range(self.max_refine + 1):
            g, rss, _ = res(x, bt)
            gn = ops.nrm2(g)
            parts = [gn.reshape(1), rss.reshape(1)] + ([anorm_d.reshape(1)] if anorm is None else [])
            vals = torch.cat(parts).cpu().numpy()  # one host read per step
            if anorm is None:
                anorm = float(vals[2])
            rn = np.sqrt(vals[1])
            eta = float(vals[0] / (anorm * rn)) if rn > 0 else 0.0
            etas.append(eta)
            if eta < best_eta:
                best_eta, best_x = eta, x.clone()
            # stop at the target, or once a step no longer halves eta: x is
            # then at the FP64 floor of this problem's backward error
            if eta <= tol or (step > 0 and eta > 0.5 * etas[-2]) or step == self.max_refine:
                break
            rhs = ops.trsv_(R, g, trans=True)
            # the inner solve only has to take eta down to tol (or to the FP64
            # floor): ask CG for a tenth of that ratio, between inner_tol and 0.1
            inner = min(0.1, max(self.inner_tol, 0.1 * tol / eta)) if eta > 0 else self.inner_tol
            dz, its = cg_normal(normal, rhs, self.max_inner, inner)
            inner_total += its
            x.add_(ops.trsv_(R, dz))
        x = best_x
        if best_eta > tol:
            warnings.warn(f"SPOMixed: backward error {best_eta:.3e} stays above tol {tol:.1e} "
                          f"(the FP64 floor of this problem, or max_refine = {self.max_refine} reached)")
        log.time_iterate = clock() - tic
        log.errors = np.array(etas)
        log.inner_iterations = inner_total
        log.refinement_steps = len(etas) - 1
        log.error_desc = "backward error ||A^T r|| / (||A||_F ||r||) after each FP64 refinement step"
        return ops.like_input(x, A), log
