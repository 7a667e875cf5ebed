"""Batched uneven least squares on the device: the batched row-block passes
against numpy per problem, bit-invariance of results under every autotuner
launch shape, graph replays equal to eager execution, and the solver against
the FP64 least-squares solution of each problem."""
import numpy as np
import pytest
import torch

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-batched] {msg}", flush=True)


@pytest.fixture(scope="module")
def bl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops
    ops.lib()
    import parla_amd.drivers.batched_ls as bl_
    return bl_


def _problems(P, seed, mrange=(300, 3000), nrange=(5, 200), cond=(1e2, 1e8)):
    rng = np.random.default_rng(seed)
    As, bs, xs = [], [], []
    for _ in range(P):
        while True:
            m = int(rng.integers(*mrange))
            n = int(rng.integers(*nrange))
            if m >= 8 * n:
                break
        c = 10 ** rng.uniform(np.log10(cond[0]), np.log10(cond[1]))
        U = np.linalg.qr(rng.standard_normal((m, n)))[0]
        V = np.linalg.qr(rng.standard_normal((n, n)))[0]
        A = ((U * np.geomspace(1, 1 / c, n)) @ V.T).astype(np.float32)
        x = rng.standard_normal(n)
        b = A.astype(np.float64) @ x + 1e-6 * rng.standard_normal(m)
        As.append(A)
        bs.append(b)
        xs.append(x)
    return As, bs, xs


def test_batched_passes_match_numpy_and_ignore_launch_shape(bl):
    _say("batched passes")
    from parla_amd import ops
    As, bs, _ = _problems(13, 1, nrange=(1, 513), mrange=(600, 5000))
    dev = torch.device("cuda")
    B = bl.Batch(As, bs, dev)
    rng = np.random.default_rng(2)
    B.t.copy_(torch.from_numpy(rng.standard_normal(B.t.numel())).to(dev))
    B.x.copy_(torch.from_numpy(rng.standard_normal(B.x.numel())).to(dev))
    solver = bl.BatchedSPO()
    outs = {}
    for cand in bl.BatchTuner.CANDIDATES:
        solver.tuner.cache = {}
        solver.tuner.CANDIDATES = [cand]
        solver.normal(B, B.q)
        solver.residual(B)
        outs[cand] = (B.q.clone(), B.g.clone(), B.rss.clone())
    ref = outs[bl.BatchTuner.CANDIDATES[0]]
    for cand, o in outs.items():
        for a, b_ in zip(o, ref):
            assert torch.equal(a, b_), cand  # bit-equal for every launch shape
    q, g, rss = ref
    for i, (A, b) in enumerate(zip(As, bs)):
        Ad = A.astype(np.float64)
        t = B.vec(B.t, i).cpu().numpy()
        x = B.vec(B.x, i).cpu().numpy()
        np.testing.assert_allclose(B.vec(q, i).cpu().numpy(), Ad.T @ (Ad @ t), rtol=1e-11,
                                   atol=1e-11 * np.abs(Ad.T @ (Ad @ t)).max())
        r_ex, g_ex, rss_ex = orc.residual_exact(Ad, x, b)
        np.testing.assert_allclose(B.vec(g, i).cpu().numpy(), g_ex, rtol=0, atol=1e-13 * np.abs(g_ex).max())
        assert abs(rss[i].item() - rss_ex) <= 1e-12 * rss_ex


def test_batched_trsv_and_factors(bl):
    _say("batched factor + preconditioner products + trsv")
    import scipy.linalg as la
    from parla_amd import ops
    As, bs, _ = _problems(9, 3)
    dev = torch.device("cuda")
    B = bl.Batch(As, bs, dev)
    solver = bl.BatchedSPO()
    solver.factor(B, np.random.default_rng(4))
    rng = np.random.default_rng(5)
    v = torch.from_numpy(rng.standard_normal(B.t.numel())).to(dev)
    u = 2.0 ** -53
    for trans in (False, True):
        B.t.copy_(v)
        solver._trsv(B, B.t, trans)  # products with the formed R_i^{-1} (or its transpose)
        for i, A in enumerate(As):
            R = B.R_view(i).cpu().numpy()
            assert np.allclose(R, np.triu(R))
            if not trans:  # R preconditions A: kappa(A R^-1) small
                s = np.linalg.svd(A.astype(np.float64) @ np.linalg.inv(R), compute_uv=False)
                assert s[0] / s[-1] < 10, (i, s[0] / s[-1])
            # to the accuracy of the inverse: ||x - R^{-1} v|| <= c n u kappa(R) ||R^{-1} v||
            x = B.vec(B.t, i).cpu().numpy()
            xr = la.solve_triangular(R, B.vec(v, i).cpu().numpy(), trans=1 if trans else 0)
            n = R.shape[0]
            assert np.linalg.norm(x - xr) <= 10 * n * u * np.linalg.cond(R) * np.linalg.norm(xr), (i, trans)
    # the batched triangular solve entry point: componentwise backward stable
    B.t.copy_(v)
    ops.check(ops.lib().pamd_bt_trsv(B.P, ops._p(B.desc), 0, ops._p(B.R), ops._p(B.t), ops.stream(dev)),
              "pamd_bt_trsv")
    for i, A in enumerate(As):
        R = B.R_view(i).cpu().numpy()
        x = B.vec(B.t, i).cpu().numpy()
        res = np.abs(R @ x - B.vec(v, i).cpu().numpy())
        assert (res <= 1e3 * u * (np.abs(np.triu(R)) @ np.abs(x)) + 1e-300).all(), i


def test_graph_replays_equal_eager(bl):
    """A captured refinement step, replayed several times with eager work in
    between, gives the eager step's result bit for bit each time."""
    _say("graph replays")
    As, bs, _ = _problems(7, 6)
    dev = torch.device("cuda")
    solver = bl.BatchedSPO(cg_iters=6)
    B = bl.Batch(As, bs, dev)
    solver.factor(B, np.random.default_rng(7))
    solver.residual(B)
    solver._copy(B, B.p, B.t)
    solver.normal(B, B.q)
    x0, g0, rss0, st0 = B.x.clone(), B.g.clone(), B.rss.clone(), B.state.clone()
    solver.refine_step(B)
    eager = (B.x.clone(), B.g.clone(), B.rss.clone())
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        solver.refine_step(B)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    B.x.copy_(x0), B.g.copy_(g0), B.rss.copy_(rss0), B.state.copy_(st0)
    with torch.cuda.graph(graph):
        solver.refine_step(B)
    for rep in range(3):
        B.x.copy_(x0), B.g.copy_(g0), B.rss.copy_(rss0), B.state.copy_(st0)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=dev).sum().item()  # unrelated eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        for a, b_ in zip((B.x, B.g, B.rss), eager):
            assert torch.equal(a, b_), rep


def test_batched_spo_solves_every_problem(bl):
    _say("batched solve")
    As, bs, _ = _problems(24, 8, mrange=(2000, 8000), nrange=(32, 257), cond=(1e2, 1e10))
    solver = bl.BatchedSPO(cg_iters=20)
    xs, log = solver(As, bs, 1e-10, 9)
    _say(f"refinement steps {log['refine_steps']}, final etas max {log['etas'][-1].max():.2e}")
    for i, (A, b) in enumerate(zip(As, bs)):
        Ad = A.astype(np.float64)
        x = xs[i].cpu().numpy()
        eta = orc.backward_error_extended(Ad, b, x)
        xl = np.linalg.lstsq(Ad, b, rcond=None)[0]
        eta_l = orc.backward_error_extended(Ad, b, xl)
        assert eta <= max(1e-10, 2 * eta_l), (i, eta, eta_l)
        s = np.linalg.svd(Ad, compute_uv=False)
        assert np.linalg.norm(x - xl) <= 1e-15 * (s[0] / s[-1]) * 1e3 * np.linalg.norm(xl), i


def test_batched_sketch_matches_single_problem_path(bl):
    """Y_i = S_i A_i from the batched CSR build and gather equals the
    single-problem draw -> CSR -> left product with seed_i (same column draws,
    same summation order)."""
    _say("batched sketch vs single-problem sketch")
    from parla_amd import ops
    As, bs, _ = _problems(11, 21, mrange=(40, 9000), nrange=(1, 300))
    As[0] = As[0][:6, :3].copy()  # d = min(m, 4 n) = 6 < vec_nnz: drawn with replacement
    bs[0] = bs[0][:6].copy()
    dev = torch.device("cuda")
    B = bl.Batch(As, bs, dev)
    solver = bl.BatchedSPO()
    Y = solver.sketch(B, np.random.default_rng(22))
    for i, A in enumerate(As):
        m, n = A.shape
        d = B.ds[i]
        rows, vals = ops.sjlt_draw(d, m, 8, int(B.seeds[i]), dev)
        rp, cl, cv = ops.sjlt_csr(d, m, 8, rows, vals)
        ref = ops.sjlt_left(rp, cl, cv, d, torch.from_numpy(A).to(dev).double())
        got = Y[B.sk_host[i, 1]:B.sk_host[i, 1] + d * n].view(d, n)
        scale = torch.abs(ref).max().item()
        assert torch.allclose(got, ref, rtol=0, atol=1e-15 * scale), (i, (got - ref).abs().max().item())


def test_batched_factor_kernels_against_numpy(bl):
    """Gram, triangular solve, triangular product, Frobenius norms and dots of
    the batch against numpy per problem."""
    _say("batched gram / trsm / trmm / fro2 / dot")
    from parla_amd import ops
    import ctypes
    As, bs, _ = _problems(9, 23, mrange=(100, 3000), nrange=(1, 400))
    dev = torch.device("cuda")
    B = bl.Batch(As, bs, dev)
    L, st = ops.lib(), ops.stream(dev)
    D, S, tot, P = ops._p(B.desc), ops._p(B.sk), B.totals, B.P
    rng = np.random.default_rng(24)
    Y = torch.from_numpy(rng.standard_normal(tot["y"])).to(dev)
    G = torch.full((tot["r"],), np.nan, dtype=torch.float64, device=dev)
    ops.check(L.pamd_bt_gram(P, D, S, tot["tiles"], 0, ops._p(Y), ops._p(G), st), "gram")
    # well-conditioned upper triangles for the solve and the product
    Ra = torch.from_numpy(rng.standard_normal(tot["r"])).to(dev)
    Rb = torch.from_numpy(rng.standard_normal(tot["r"])).to(dev)
    for i, n in enumerate(B.ns):
        for Rp in (Ra, Rb):
            v = Rp[B.r_off[i]:B.r_off[i + 1]].view(n, n)
            v.diagonal().copy_(n + torch.rand(n, dtype=torch.float64, device=dev))
    C = torch.zeros(tot["r"], dtype=torch.float64, device=dev)
    ops.check(L.pamd_bt_trmm_upper(P, D, S, tot["tiles"], ops._p(Ra), ops._p(Rb), ops._p(C), st), "trmm")
    X = Y.clone()
    dinv = torch.empty(256 * tot["dblk"], dtype=torch.float64, device=dev)
    ops.check(L.pamd_bt_trsm(P, D, S, tot["strips"], tot["dblk"], max(B.ns), ops._p(Ra), ops._p(dinv),
                             ops._p(X), st), "trsm")
    dots = torch.empty(P, dtype=torch.float64, device=dev)
    ops.check(L.pamd_bt_dot(P, D, ops._p(B.x.copy_(torch.from_numpy(rng.standard_normal(B.x.numel())))),
                            ops._p(B.t.copy_(torch.from_numpy(rng.standard_normal(B.t.numel())))), ops._p(dots),
                            st), "dot")
    import scipy.linalg as sla
    for i, A in enumerate(As):
        m, n = A.shape
        d = B.ds[i]
        yo = B.sk_host[i, 1]
        Yi = Y[yo:yo + d * n].view(d, n).cpu().numpy()
        Gi = G[B.r_off[i]:B.r_off[i + 1]].view(n, n).cpu().numpy()
        Gref = Yi.T @ Yi
        iu = np.triu_indices(n)
        np.testing.assert_allclose(Gi[iu], Gref[iu], rtol=0, atol=1e-13 * np.abs(Gref).max())
        Ai = np.triu(Ra[B.r_off[i]:B.r_off[i + 1]].view(n, n).cpu().numpy())
        Bi = np.triu(Rb[B.r_off[i]:B.r_off[i + 1]].view(n, n).cpu().numpy())
        Ci = C[B.r_off[i]:B.r_off[i + 1]].view(n, n).cpu().numpy()
        np.testing.assert_allclose(Ci, Ai @ Bi, rtol=0, atol=1e-13 * np.abs(Ai @ Bi).max())
        Xi = X[yo:yo + d * n].view(d, n).cpu().numpy()
        Xref = sla.solve_triangular(Ai, Yi.T, trans="T").T
        np.testing.assert_allclose(Xi, Xref, rtol=0, atol=1e-13 * np.abs(Xref).max())
        Ad = A.astype(np.float64)
        assert abs(B.anorm[i].item() - np.linalg.norm(Ad)) <= 1e-14 * np.linalg.norm(Ad)
        xv = B.vec(B.x, i).cpu().numpy()
        tv = B.vec(B.t, i).cpu().numpy()
        assert abs(dots[i].item() - xv @ tv) <= 1e-13 * (np.abs(xv) @ np.abs(tv))


def test_frozen_problems_are_untouched(bl):
    """A refinement step with some problems frozen leaves their x, g and
    ||r||^2 bit for bit as they were (their row blocks are skipped), and
    moves the active ones."""
    _say("frozen problems")
    As, bs, _ = _problems(8, 31)
    dev = torch.device("cuda")
    solver = bl.BatchedSPO(cg_iters=5)
    B = bl.Batch(As, bs, dev)
    solver.factor(B, np.random.default_rng(32))
    solver.residual(B)
    solver.refine_step(B)  # all active: x moves off zero
    x0, g0, rss0 = B.x.clone(), B.g.clone(), B.rss.clone()
    active = np.arange(B.P) % 2 == 0
    B.state[1::4] = torch.from_numpy(active.astype(np.float64)).to(dev)
    solver.refine_step(B)
    for i in range(B.P):
        same = torch.equal(B.vec(B.x, i), B.vec(x0, i)) and torch.equal(B.vec(B.g, i), B.vec(g0, i))
        if active[i]:
            assert not torch.equal(B.vec(B.x, i), B.vec(x0, i)), i
        else:
            assert same and B.rss[i].item() == rss0[i].item(), i
