"""Batched sketch-and-precondition least squares for many problems of uneven size.

No counterpart in the reference, whose drivers take one problem per call
(drivers/least_squares.py:206-369).  Each problem i (A_i FP32 m_i x n_i with
n_i <= 512, b_i FP64) is solved the SPO way, all problems at once:

  setup   per problem: SJLT sketch d_i = sampling_factor * n_i (device draw),
          shifted CholeskyQR3 of S A_i (Fukaya, Kannan, Nakatsukasa, Yamamoto,
          Yanagisawa 2020: a shifted Cholesky first, then CholeskyQR2, stable
          up to kappa ~ 1/u), with every potrf / trtri of the batch in one
          launch;
  solve   FP64 iterative refinement on every problem at once:
              g_i  = A_i^T (b_i - A_i x_i)            compensated, one read of A
              z_i  = CG_K(R_i^{-T} A_i^T A_i R_i^{-1}, R_i^{-T} g_i)
              x_i += R_i^{-1} z_i
          One refinement step (residual, K CG iterations, update) is a fixed
          sequence of batched kernels -- row-block passes over all problems
          plus one-workgroup-per-problem solves and recurrences -- captured
          once in a HIP graph and replayed; the host reads the backward errors
          between replays and freezes converged problems.
  tuning  the launch shape of the row-block passes (workgroup size, blocks
          per wavefront) is chosen by BatchTuner, which times the candidates
          on the batch itself; the choice never changes a result bit.
"""
import time

import numpy as np
import torch

from parla_amd import ops

BLOCK_ROWS = 64
NPL_CLASSES = (1, 2, 4, 8)
MAX_N = 512
U = np.finfo(np.float64).eps / 2


def _npl(n):
    k = -(-n // 64)
    return 1 if k <= 1 else 2 if k <= 2 else 4 if k <= 4 else 8


class BatchTuner:
    """Times each launch-shape candidate of a row-block pass on the batch and
    keeps the fastest per (kernel, npl, number of blocks)."""

    CANDIDATES = [(w, b) for w in (1, 2, 4) for b in (1, 2, 4, 8)]

    def __init__(self, reps=3):
        self.reps = reps
        self.cache = {}

    def choose(self, key, launch):
        if key in self.cache:
            return self.cache[key]
        best, best_t = None, np.inf
        for cand in self.CANDIDATES:
            launch(*cand)  # warm-up
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(self.reps):
                launch(*cand)
            b.record()
            b.synchronize()
            t = a.elapsed_time(b) / self.reps
            if t < best_t:
                best, best_t = cand, t
        self.cache[key] = best
        return best


class Batch:
    """Device pools and descriptors of a batch of problems."""

    def __init__(self, As, bs, dev):
        self.dev = dev
        P = len(As)
        self.P = P
        ms = [int(A.shape[0]) for A in As]
        ns = [int(A.shape[1]) for A in As]
        if max(ns) > MAX_N:
            raise NotImplementedError("batched problems need n <= 512")
        if any(m < n for m, n in zip(ms, ns)):
            raise ValueError("every problem must have m >= n")
        self.ms, self.ns = ms, ns
        a_off = np.cumsum([0] + [m * n for m, n in zip(ms, ns)])
        mv_off = np.cumsum([0] + ms)
        nv_off = np.cumsum([0] + ns)
        r_off = np.cumsum([0] + [n * n for n in ns])
        nb = [-(-m // BLOCK_ROWS) for m in ms]
        blk0 = np.cumsum([0] + nb)
        part_off = np.cumsum([0] + [k * (2 * n + 2) for k, n in zip(nb, ns)])
        desc = np.stack([ms, ns, a_off[:-1], mv_off[:-1], nv_off[:-1], r_off[:-1], blk0[:-1], part_off[:-1]],
                        axis=1).astype(np.int64)
        self.desc = torch.from_numpy(desc).to(dev)
        self.nv_off, self.r_off = nv_off, r_off
        self.n_dev = torch.tensor(ns, dtype=torch.int64, device=dev)
        self.roff_dev = torch.from_numpy(r_off[:-1].astype(np.int64)).to(dev)
        # pools
        self.A = torch.empty(int(a_off[-1]), dtype=torch.float32, device=dev)
        self.b = torch.empty(int(mv_off[-1]), dtype=ops.F64, device=dev)
        for i, (A, b) in enumerate(zip(As, bs)):
            self.A[a_off[i]:a_off[i + 1]].copy_(torch.as_tensor(A).reshape(-1).to(dev, torch.float32))
            self.b[mv_off[i]:mv_off[i + 1]].copy_(torch.as_tensor(b).reshape(-1).to(dev, ops.F64))
        self.a_off, self.mv_off = a_off, mv_off
        NV = int(nv_off[-1])
        f64 = dict(dtype=ops.F64, device=dev)
        self.x, self.z, self.r, self.p, self.q, self.t, self.g, self.rhs = (torch.zeros(NV, **f64) for _ in range(8))
        self.R = torch.zeros(int(r_off[-1]), **f64)
        self.part = torch.zeros(int(part_off[-1]), **f64)
        self.rss = torch.zeros(P, **f64)
        self.state = torch.zeros(4 * P, **f64)
        self.state[1::4] = 1.0
        # 64-row blocks grouped by row length class
        self.blocks = {}
        for c in NPL_CLASSES:
            lst = [(i, r0) for i in range(P) if _npl(ns[i]) == c for r0 in range(0, ms[i], BLOCK_ROWS)]
            if lst:
                self.blocks[c] = torch.tensor(lst, dtype=torch.int32, device=dev).contiguous()
        self.anorm = torch.stack([torch.linalg.vector_norm(self.A_view(i).double()) for i in range(P)])

    def A_view(self, i):
        return self.A[self.a_off[i]:self.a_off[i + 1]].view(self.ms[i], self.ns[i])

    def R_view(self, i):
        n = self.ns[i]
        return self.R[self.r_off[i]:self.r_off[i + 1]].view(n, n)

    def vec(self, pool, i):
        return pool[self.nv_off[i]:self.nv_off[i + 1]]


class BatchedSPO:
    """x_list, log = BatchedSPO(...)(As, bs, tol, rng)."""

    def __init__(self, sampling_factor=4, vec_nnz=8, cg_iters=24, max_refine=8, tuner=None):
        self.sampling_factor = sampling_factor
        self.vec_nnz = vec_nnz
        self.cg_iters = cg_iters
        self.max_refine = max_refine
        self.tuner = tuner or BatchTuner()

    # -- setup ---------------------------------------------------------------
    def factor(self, B, rng):
        """R_i of S_i A_i for every problem (SJLT sketch, shifted CholeskyQR3)."""
        dev = B.dev
        Ys = []
        for i in range(B.P):
            m, n = B.ms[i], B.ns[i]
            d = min(m, int(self.sampling_factor * n))
            rows, vals = ops.sjlt_draw(d, m, self.vec_nnz, int(rng.integers(0, 2 ** 63)), dev)
            rp, cl, cv = ops.sjlt_csr(d, m, self.vec_nnz, rows, vals)
            Ys.append(ops.sjlt_left(rp, cl, cv, d, B.A_view(i).double()))
        Rs = [None] * B.P
        info = torch.zeros(B.P, dtype=torch.int32, device=dev)
        # round 0: shifted Cholesky of the Gram matrix; rounds 1, 2: CholeskyQR2
        for rnd in range(3):
            for i in range(B.P):
                G = ops.gemm(Ys[i].t(), Ys[i])
                if rnd == 0:
                    m_, n_ = Ys[i].shape
                    shift = 11.0 * (m_ * n_ + n_ * (n_ + 1)) * U * float(torch.sum(Ys[i] * Ys[i]).item())
                    G.diagonal().add_(shift)
                B.R_view(i).copy_(G)
            ops.check(ops.lib().pamd_potrf_batched_f64(B.P, ops._p(B.n_dev), ops._p(B.roff_dev), ops._p(B.R),
                                                       ops._p(info), ops.stream(dev)), "pamd_potrf_batched_f64")
            Rinv = torch.empty_like(B.R)
            ops.check(ops.lib().pamd_trtri_batched_f64(B.P, max(B.ns), ops._p(B.n_dev), ops._p(B.roff_dev),
                                                       ops._p(B.R), ops._p(Rinv), ops.stream(dev)),
                      "pamd_trtri_batched_f64")
            for i in range(B.P):
                n = B.ns[i]
                Rk = torch.triu(B.R_view(i))
                Ys[i] = ops.gemm(Ys[i], Rinv[B.r_off[i]:B.r_off[i + 1]].view(n, n))
                Rs[i] = Rk if Rs[i] is None else ops.gemm(Rk, Rs[i])
        bad = info.cpu().numpy()
        if bad.any():
            raise ArithmeticError(f"batched CholeskyQR3 failed for problems {np.nonzero(bad)[0][:8].tolist()}")
        for i in range(B.P):
            B.R_view(i).copy_(torch.triu(Rs[i]))

    # -- batched passes ------------------------------------------------------
    def _pass(self, B, kind):
        L = ops.lib()
        st = ops.stream(B.dev)
        for c, blks in B.blocks.items():
            nb = blks.shape[0]
            if kind == "normal":
                def launch(w, k, c=c, blks=blks, nb=nb):
                    ops.check(L.pamd_bt_normal(c, ops._p(B.desc), ops._p(blks), nb, w, k, ops._p(B.A),
                                               ops._p(B.t), ops._p(B.part), st), "pamd_bt_normal")
            else:
                def launch(w, k, c=c, blks=blks, nb=nb):
                    ops.check(L.pamd_bt_residual(c, ops._p(B.desc), ops._p(blks), nb, w, k, ops._p(B.A),
                                                 ops._p(B.b), ops._p(B.x), ops._p(B.part), st), "pamd_bt_residual")
            w, k = self.tuner.choose((kind, c, nb), launch)
            launch(w, k)

    def normal(self, B, out):
        """out_i = A_i^T A_i t_i for every problem."""
        self._pass(B, "normal")
        ops.check(ops.lib().pamd_bt_reduce(B.P, ops._p(B.desc), ops._p(B.part), 0, ops._p(out), None,
                                           ops.stream(B.dev)), "pamd_bt_reduce")

    def residual(self, B):
        """g_i = A_i^T (b_i - A_i x_i) (compensated) and rss_i = ||r_i||^2."""
        self._pass(B, "residual")
        ops.check(ops.lib().pamd_bt_reduce(B.P, ops._p(B.desc), ops._p(B.part), 1, ops._p(B.g), ops._p(B.rss),
                                           ops.stream(B.dev)), "pamd_bt_reduce")

    def _trsv(self, B, v, trans):
        ops.check(ops.lib().pamd_bt_trsv(B.P, ops._p(B.desc), 1 if trans else 0, ops._p(B.R), ops._p(v),
                                         ops.stream(B.dev)), "pamd_bt_trsv")

    def _copy(self, B, a, v):
        ops.check(ops.lib().pamd_bt_copy(B.P, ops._p(B.desc), ops._p(a), ops._p(v), ops.stream(B.dev)),
                  "pamd_bt_copy")

    def _cg(self, B, mode):
        ops.check(ops.lib().pamd_bt_cg(B.P, ops._p(B.desc), mode, ops._p(B.state), ops._p(B.rhs), ops._p(B.r),
                                       ops._p(B.p), ops._p(B.q), ops._p(B.z), ops.stream(B.dev)), "pamd_bt_cg")

    def refine_step(self, B):
        """x += R^{-1} CG_K(...) on every active problem, then the new residual."""
        self._copy(B, B.g, B.rhs)
        self._trsv(B, B.rhs, True)
        self._cg(B, 0)
        for _ in range(self.cg_iters):
            self._copy(B, B.p, B.t)
            self._trsv(B, B.t, False)
            self.normal(B, B.q)
            self._trsv(B, B.q, True)
            self._cg(B, 1)
        self._copy(B, B.z, B.t)
        self._trsv(B, B.t, False)
        ops.check(ops.lib().pamd_bt_axpy(B.P, ops._p(B.desc), ops._p(B.state), ops._p(B.x), ops._p(B.t),
                                         ops.stream(B.dev)), "pamd_bt_axpy")
        self.residual(B)

    def backward_errors(self, B):
        gn = torch.stack([torch.linalg.vector_norm(B.vec(B.g, i)) for i in range(B.P)])
        eta = gn / (B.anorm * torch.sqrt(B.rss))
        return torch.nan_to_num(eta, nan=0.0).cpu().numpy()

    # -- driver --------------------------------------------------------------
    def __call__(self, As, bs, tol, rng, use_graph=True, logging=True):
        dev = ops.default_device()
        rng = np.random.default_rng(rng)
        tic = time.perf_counter()
        B = Batch(As, bs, dev)
        self.factor(B, rng)
        self.residual(B)  # x = 0; also tunes the residual pass
        torch.cuda.synchronize(dev)
        t_setup = time.perf_counter() - tic
        graph = None
        if use_graph:
            # tune the normal pass outside the capture, then capture one step
            self._copy(B, B.p, B.t)
            self.normal(B, B.q)
            saved = [B.x.clone(), B.g.clone(), B.rss.clone(), B.state.clone()]
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self.refine_step(B)  # warm the stream before capture
            torch.cuda.current_stream().wait_stream(s)
            B.x.copy_(saved[0]), B.g.copy_(saved[1]), B.rss.copy_(saved[2]), B.state.copy_(saved[3])
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self.refine_step(B)
        tic = time.perf_counter()
        etas = [self.backward_errors(B)]
        steps = 0
        while steps < self.max_refine:
            active = etas[-1] > tol
            if not active.any():
                break
            B.state[1::4] = torch.from_numpy(active.astype(np.float64)).to(dev)
            if graph is not None:
                graph.replay()
            else:
                self.refine_step(B)
            steps += 1
            etas.append(self.backward_errors(B))
        torch.cuda.synchronize(dev)
        t_solve = time.perf_counter() - tic
        xs = [B.vec(B.x, i).clone() for i in range(B.P)]
        log = {"setup_s": t_setup, "solve_s": t_solve, "refine_steps": steps, "etas": np.array(etas),
               "tuned": dict(self.tuner.cache)}
        return xs, log


def uneven_batch(P, seed, dev, m_range=(2000, 50000), n_range=(32, 512), ratio=8, cond_range=(1e2, 1e10),
                 noise=1e-8):
    """The batched workload: P problems with m_i, n_i uniform (rejected until
    m_i >= ratio n_i), cond_i log-uniform, A_i = U_i diag(s_i) V_i^T with Haar
    factors (CholeskyQR2 of device Gaussians) and s_i geometric, stored FP32;
    b_i = A_i x_i + noise N(0, 1) in FP64.  Returns (As, bs)."""
    rng = np.random.default_rng(seed)
    As, bs = [], []
    for _ in range(P):
        while True:
            m = int(rng.integers(m_range[0], m_range[1] + 1))
            n = int(rng.integers(n_range[0], n_range[1] + 1))
            if m >= ratio * n:
                break
        cond = 10 ** rng.uniform(np.log10(cond_range[0]), np.log10(cond_range[1]))
        U = ops.orth_q(ops.gaussian_fill(m, n, 1.0, int(rng.integers(0, 2 ** 63)), dev))
        V = ops.orth_q(ops.gaussian_fill(n, n, 1.0, int(rng.integers(0, 2 ** 63)), dev))
        s = torch.from_numpy(np.geomspace(1.0, 1.0 / cond, n)).to(dev)
        A = ops.gemm(U * s, V.t()).to(torch.float32)
        x = torch.from_numpy(rng.standard_normal(n)).to(dev)
        b = A.double() @ x + noise * torch.from_numpy(rng.standard_normal(m)).to(dev)
        As.append(A)
        bs.append(b)
    return As, bs

