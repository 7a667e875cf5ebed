"""Batched sketch-and-precondition least squares for many problems of uneven size.

No counterpart in the reference, whose drivers take one problem per call
(drivers/least_squares.py:206-369).  Each problem i (A_i FP32 m_i x n_i with
n_i <= 512, b_i FP64) is solved the SPO way, all problems at once:

  setup   SJLT sketch Y_i = S_i A_i with d_i = sampling_factor * n_i (device
          draw, CSR transpose and gather product for all problems at once),
          then the shifted CholeskyQR3 of Y_i (Fukaya, Kannan, Nakatsukasa,
          Yamamoto, Yanagisawa 2020: a shifted Cholesky first, then
          CholeskyQR2, stable up to kappa ~ 1/u): every Gram matrix, potrf,
          triangular solve and R product of the batch is one launch;
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
import ctypes
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


def _layout():
    vals = [ctypes.c_int() for _ in range(5)]
    ops.check(ops.lib().pamd_bt_layout(*[ctypes.byref(v) for v in vals]), "pamd_bt_layout")
    return tuple(v.value for v in vals)


def _concat(arrays, dtype, dev):
    """One flat device pool holding the arrays back to back."""
    if all(isinstance(a, torch.Tensor) and a.device == dev and a.dtype == dtype for a in arrays):
        return torch.cat([a.reshape(-1) for a in arrays])
    np_dtype = np.float32 if dtype == torch.float32 else np.float64
    host = np.concatenate([np.asarray(a.cpu() if isinstance(a, torch.Tensor) else a, dtype=np_dtype).reshape(-1)
                           for a in arrays])
    return torch.from_numpy(host).to(dev)


def batch_plan(ms, ns, sampling_factor, layout):
    """Descriptors and pool sizes of a batch (host arithmetic only).

    layout = (chunk_rows, tile, strip_rows, max_d, max_vec_nnz) as returned
    by pamd_bt_layout.  Returns a dict: `desc` (P x 8 int64, pamd_ls_desc),
    `sk` (P x 10 int64, pamd_sk_desc with seed 0), the offset prefixes, the
    64-row block lists per row-length class and the work-item / pool totals."""
    ms = np.asarray(ms, dtype=np.int64)
    ns = np.asarray(ns, dtype=np.int64)
    P = len(ms)
    if ns.max() > MAX_N:
        raise NotImplementedError("batched problems need n <= 512")
    if (ms < ns).any():
        raise ValueError("every problem must have m >= n")
    chunk_rows, tile, strip_rows, max_d, _ = layout
    ds = np.minimum(ms, (sampling_factor * ns).astype(np.int64))
    if ds.max() > max_d:
        raise NotImplementedError(f"batched sketches need d <= {max_d}")

    def prefix(x):
        return np.concatenate([[0], np.cumsum(x)]).astype(np.int64)

    a_off, mv_off, nv_off, r_off = prefix(ms * ns), prefix(ms), prefix(ns), prefix(ns * ns)
    nb = -(-ms // BLOCK_ROWS)
    blk0 = prefix(nb)
    part_off = prefix(nb * (2 * ns + 2))
    desc = np.stack([ms, ns, a_off[:-1], mv_off[:-1], nv_off[:-1], r_off[:-1], blk0[:-1], part_off[:-1]],
                    axis=1).astype(np.int64)
    # sketch descriptors: work items per problem and pool offsets
    nchunk = -(-ms // chunk_rows)
    T = -(-ns // tile)
    ntile = T * (T + 1) // 2
    nstrip = -(-ds // strip_rows)
    ndblk = -(-ns // strip_rows)
    chunk0, row0, tile0, strip0, dblk0 = prefix(nchunk), prefix(ds), prefix(ntile), prefix(nstrip), prefix(ndblk)
    sk = np.stack([ds, prefix(ds * ns)[:-1], prefix(ds + 1)[:-1], prefix(nchunk * ds)[:-1],
                   chunk0[:-1], row0[:-1], tile0[:-1], strip0[:-1], dblk0[:-1],
                   np.zeros(P, dtype=np.int64)], axis=1).astype(np.int64)
    totals = {"chunks": int(chunk0[-1]), "rows": int(row0[-1]), "tiles": int(tile0[-1]),
              "strips": int(strip0[-1]), "dblk": int(dblk0[-1]), "y": int((ds * ns).sum()),
              "rp": int((ds + 1).sum()), "cnt": int((nchunk * ds).sum()), "a": int(a_off[-1]),
              "mv": int(mv_off[-1]), "nv": int(nv_off[-1]), "r": int(r_off[-1]), "part": int(part_off[-1])}
    # 64-row blocks grouped by row length class, problem after problem
    npl = np.array([_npl(int(n)) for n in ns])
    blocks = {}
    for c in NPL_CLASSES:
        sel = np.nonzero(npl == c)[0]
        if len(sel) == 0:
            continue
        cnt = nb[sel]
        prob = np.repeat(sel, cnt)
        local = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        blocks[c] = np.stack([prob, local * BLOCK_ROWS], axis=1).astype(np.int32)
    return {"desc": desc, "sk": sk, "ds": ds, "a_off": a_off, "mv_off": mv_off, "nv_off": nv_off,
            "r_off": r_off, "blocks": blocks, "totals": totals}


class Batch:
    """Device pools and descriptors of a batch of problems.

    Problem i: A_i (FP32, row-major) in one pool, b_i and the n-vectors in
    FP64 pools, R_i in an n_i x n_i pool; `desc` (pamd_ls_desc) and `sk`
    (pamd_sk_desc, the sketch d_i = min(m_i, sampling_factor n_i)) describe
    them to the kernels, and `blocks` lists the 64-row blocks per row-length
    class.  ||A_i||_F is computed on the device."""

    def __init__(self, As, bs, dev, sampling_factor=4):
        self.dev = dev
        P = len(As)
        self.P = P
        if len(bs) != P or any(int(b.shape[0]) != int(A.shape[0]) for A, b in zip(As, bs)):
            raise ValueError("every b_i needs m_i entries")
        plan = batch_plan([int(A.shape[0]) for A in As], [int(A.shape[1]) for A in As], sampling_factor,
                          _layout())
        ms, ns = plan["desc"][:, 0], plan["desc"][:, 1]
        self.ms, self.ns, self.ds = ms.tolist(), ns.tolist(), plan["ds"].tolist()
        self.desc = torch.from_numpy(plan["desc"]).to(dev)
        self.sk_host = plan["sk"]
        self.sk = torch.from_numpy(self.sk_host).to(dev)
        self.totals = plan["totals"]
        a_off, mv_off, nv_off, r_off = plan["a_off"], plan["mv_off"], plan["nv_off"], plan["r_off"]
        self.nv_off, self.r_off = nv_off, r_off
        self.n_dev = torch.from_numpy(ns.copy()).to(dev)
        self.roff_dev = torch.from_numpy(r_off[:-1].copy()).to(dev)
        # pools
        self.A = _concat(As, torch.float32, dev)
        self.b = _concat(bs, ops.F64, dev)
        self.a_off, self.mv_off = a_off, mv_off
        NV = int(nv_off[-1])
        f64 = dict(dtype=ops.F64, device=dev)
        self.x, self.z, self.r, self.p, self.q, self.t, self.g, self.rhs = (torch.zeros(NV, **f64) for _ in range(8))
        self.R = torch.zeros(int(r_off[-1]), **f64)
        self.Rinv = torch.zeros(int(r_off[-1]), **f64)
        self.RinvT = torch.zeros(int(r_off[-1]), **f64)
        self.part = torch.zeros(self.totals["part"], **f64)
        self.rss = torch.zeros(P, **f64)
        self.gg = torch.zeros(P, **f64)
        self.state = torch.zeros(4 * P, **f64)
        self.state[1::4] = 1.0
        self.blocks = {c: torch.from_numpy(lst).to(dev) for c, lst in plan["blocks"].items()}
        fro2 = torch.empty(P, **f64)
        part = torch.empty(max(1, self.totals["chunks"]), **f64)
        ops.check(ops.lib().pamd_bt_fro2(P, ops._p(self.desc), ops._p(self.sk), self.totals["chunks"],
                                         ops._p(self.A), ops._p(part), ops._p(fro2), ops.stream(dev)),
                  "pamd_bt_fro2")
        self.anorm = torch.sqrt(fro2)

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
        self._warm = False

    # -- setup ---------------------------------------------------------------
    def sketch(self, B, rng):
        """Y_i = S_i A_i for every problem (FP64 pool, d_i x n_i each): the
        SJLT of problem i is pamd_sjlt_draw's with seed B.seeds[i]."""
        L, dev, st = ops.lib(), B.dev, ops.stream(B.dev)
        k = self.vec_nnz
        B.seeds = rng.integers(0, 2 ** 63, size=B.P, dtype=np.int64)
        B.sk_host[:, 9] = B.seeds
        B.sk.copy_(torch.from_numpy(B.sk_host))
        nnz = k * int(B.mv_off[-1])
        cnt = torch.empty(max(1, B.totals["cnt"]), dtype=torch.int32, device=dev)
        rowptr = torch.empty(B.totals["rp"], dtype=torch.int64, device=dev)
        cols = torch.empty(nnz, dtype=torch.int32, device=dev)
        vals = torch.empty(nnz, dtype=ops.F64, device=dev)
        ops.check(L.pamd_bt_sjlt_csr(B.P, ops._p(B.desc), ops._p(B.sk), B.totals["chunks"], k, ops._p(cnt),
                                     ops._p(rowptr), ops._p(cols), ops._p(vals), st), "pamd_bt_sjlt_csr")
        Y = torch.empty(B.totals["y"], dtype=ops.F64, device=dev)
        ops.check(L.pamd_bt_sjlt_left(B.P, ops._p(B.desc), ops._p(B.sk), B.totals["rows"], k, ops._p(rowptr),
                                      ops._p(cols), ops._p(vals), ops._p(B.A), ops._p(Y), st), "pamd_bt_sjlt_left")
        return Y

    def factor(self, B, rng):
        """R_i of S_i A_i for every problem: shifted CholeskyQR3 of Y = S A,
        R = R_2 R_1 R_0, each step one launch for the whole batch."""
        L, dev, st = ops.lib(), B.dev, ops.stream(B.dev)
        Y = self.sketch(B, rng)
        P, tot = B.P, B.totals
        D, S = ops._p(B.desc), ops._p(B.sk)
        info = torch.zeros(3 * P, dtype=torch.int32, device=dev)
        dinv = torch.empty(256 * tot["dblk"], dtype=ops.F64, device=dev)
        nmax = max(B.ns)

        def round_(G, r, shifted, update):
            ops.check(L.pamd_bt_gram(P, D, S, tot["tiles"], shifted, ops._p(Y), ops._p(G), st), "pamd_bt_gram")
            ops.check(L.pamd_potrf_batched_f64(P, ops._p(B.n_dev), ops._p(B.roff_dev), ops._p(G),
                                               ops._p(info[r * P:(r + 1) * P]), st), "pamd_potrf_batched_f64")
            if update:
                ops.check(L.pamd_bt_trsm(P, D, S, tot["strips"], tot["dblk"], nmax, ops._p(G), ops._p(dinv),
                                         ops._p(Y), st), "pamd_bt_trsm")

        R0 = torch.empty_like(B.R)
        Rk = torch.empty_like(B.R)
        T1 = torch.zeros_like(B.R)
        round_(R0, 0, 1, True)
        round_(Rk, 1, 0, True)
        ops.check(L.pamd_bt_trmm_upper(P, D, S, tot["tiles"], ops._p(Rk), ops._p(R0), ops._p(T1), st),
                  "pamd_bt_trmm_upper")
        round_(Rk, 2, 0, False)
        B.R.zero_()
        ops.check(L.pamd_bt_trmm_upper(P, D, S, tot["tiles"], ops._p(Rk), ops._p(T1), ops._p(B.R), st),
                  "pamd_bt_trmm_upper")
        bad = info.view(3, P).ne(0).any(0).cpu().numpy()
        if bad.any():
            raise ArithmeticError(f"batched CholeskyQR3 failed for problems {np.nonzero(bad)[0][:8].tolist()}")
        # R_i^{-1} and its transpose once: the refinement applies the
        # preconditioner as products with them (two per CG step)
        ops.check(L.pamd_trtri_batched_f64(P, nmax, ops._p(B.n_dev), ops._p(B.roff_dev), ops._p(B.R),
                                           ops._p(B.Rinv), ops._p(B.RinvT), st), "pamd_trtri_batched_f64")

    # -- batched passes ------------------------------------------------------
    def _pass(self, B, kind):
        L = ops.lib()
        st = ops.stream(B.dev)
        for c, blks in B.blocks.items():
            nb = blks.shape[0]
            if kind == "normal":
                def launch(w, k, c=c, blks=blks, nb=nb):
                    ops.check(L.pamd_bt_normal(c, ops._p(B.desc), ops._p(blks), nb, w, k, ops._p(B.A),
                                               ops._p(B.t), ops._p(B.state), ops._p(B.part), st),
                              "pamd_bt_normal")
            else:
                def launch(w, k, c=c, blks=blks, nb=nb):
                    ops.check(L.pamd_bt_residual(c, ops._p(B.desc), ops._p(blks), nb, w, k, ops._p(B.A),
                                                 ops._p(B.b), ops._p(B.x), ops._p(B.state), ops._p(B.part), st),
                              "pamd_bt_residual")
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
        """v_i <- R_i^{-1} v_i (R_i^{-T} v_i), as products with R_i^{-1} (formed in factor)."""
        ops.check(ops.lib().pamd_bt_trmv(B.P, ops._p(B.desc), 1 if trans else 0, ops._p(B.Rinv), ops._p(B.RinvT),
                                         ops._p(v), ops.stream(B.dev)), "pamd_bt_trmv")

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
        """eta_i = ||A_i^T r_i|| / (||A_i||_F ||r_i||) from the last residual pass."""
        ops.check(ops.lib().pamd_bt_dot(B.P, ops._p(B.desc), ops._p(B.g), ops._p(B.g), ops._p(B.gg),
                                        ops.stream(B.dev)), "pamd_bt_dot")
        eta = torch.sqrt(B.gg) / (B.anorm * torch.sqrt(B.rss))
        return torch.nan_to_num(eta, nan=0.0).cpu().numpy()

    # -- driver --------------------------------------------------------------
    def __call__(self, As, bs, tol, rng, use_graph=True, logging=True):
        dev = ops.default_device()
        rng = np.random.default_rng(rng)
        tic = time.perf_counter()
        B = Batch(As, bs, dev, self.sampling_factor)
        self.factor(B, rng)
        self.residual(B)  # x = 0; also tunes the residual pass
        graph = None
        if use_graph:
            # tune the normal pass outside the capture; nothing in a step
            # allocates or synchronises.  The first call also runs one step
            # with every problem frozen (every pass skips them, x stays) so
            # that each kernel has been launched once before the capture.
            self._copy(B, B.p, B.t)
            self.normal(B, B.q)
            if not self._warm:
                B.state[1::4] = 0.0
                self.refine_step(B)
                B.state[1::4] = 1.0
                self._warm = True
            # the partials of frozen problems must be their residual's: a
            # problem skipped by later passes keeps g_i and ||r_i|| this way
            self.residual(B)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                self.refine_step(B)
        torch.cuda.synchronize(dev)
        t_setup = time.perf_counter() - tic
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
        x = B.x.clone()
        torch.cuda.synchronize(dev)
        t_solve = time.perf_counter() - tic
        xs = [x[B.nv_off[i]:B.nv_off[i + 1]] for i in range(B.P)]
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

