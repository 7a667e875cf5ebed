"""GPU parity of the least-squares variants next to the hot path: SPO in the
'chol' and 'svd' modes (and on rank-deficient data), the SVD right
preconditioner, PcSS2's underdetermined branch and SPU1, sketch-and-solve
SSO1, and the SRCT and sparse sign operators -- against the fixtures the
reference wrote (tests/golden/make_golden.py), the oracle on the same draws,
and closed forms.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

from trajectory import ROUNDING_CONSTANTS, assert_within_envelope, envelope, perturbed_change, rounded_matrix

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-ls-modes] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pa():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import parla_amd
    from parla_amd import ops
    ops.lib()
    return parla_amd


def _sign_fix(X, ref):
    return X * np.sign(np.sum(X * ref, axis=0))


@pytest.mark.parametrize("tag,gen,mode", [("ga_chol", "ga", "chol"), ("sj_chol_ridge", "sj", "chol"),
                                          ("ga_svd", "ga", "svd"), ("sj_svd_ridge", "sj", "svd")])
def test_spo_modes_match_reference_runs(pa, golden, tag, gen, mode):
    _say(f"SPO {tag}")
    A, b = golden["spo_A"], golden["spo_b"]
    sf, delta = golden[f"spo_{tag}_meta"]
    g = pa.SkOpGA() if gen == "ga" else pa.SkOpSJ(8)
    x, log = pa.SPO(g, float(sf), mode=mode)(A, b, float(delta), 1e-12, 60, np.random.default_rng(5))
    np.testing.assert_allclose(x, golden[f"spo_{tag}_x"], rtol=1e-8)
    ref = golden[f"spo_{tag}_errors"]
    assert abs(log.errors.size - ref.size) <= 1
    k = min(log.errors.size, ref.size)
    np.testing.assert_allclose(log.errors[:k], ref[:k], rtol=1e-6, atol=1e-11 * ref[0])


def test_spo_svd_rank_deficient_matches_reference(pa, golden):
    _say("SPO svd, rank deficient")
    A, b = golden["spo_rd_A"], golden["spo_rd_b"]
    x, log = pa.SPO(pa.SkOpGA(), 3, mode="svd")(A, b, 0.0, 1e-12, 60, np.random.default_rng(6))
    np.testing.assert_allclose(x, golden["spo_rd_svd_x"], rtol=1e-8)
    np.testing.assert_allclose(x, la.lstsq(A, b)[0], rtol=1e-8)   # the minimum-norm solution
    ref = golden["spo_rd_svd_errors"]
    assert abs(log.errors.size - ref.size) <= 1


def test_spo_modes_on_plain_config_match_oracle(pa):
    _say("SPO chol / svd on A 20000 x 200")
    import oracle as orc
    from parla_amd import matmakers as mm
    A, b, _ = mm.config_ls_plain()
    for mode in ("chol", "svd"):
        x, log = pa.SPO(pa.SkOpGA(), 4, mode=mode)(A, b, 0.0, 1e-12, 100, np.random.default_rng(7))
        xo, logo = orc.spo(A, b, 0.0, 1e-12, 100, np.random.default_rng(7), sketch="gaussian",
                           sampling_factor=4, mode=mode)
        assert la.norm(x - xo) <= 1e-8 * la.norm(xo), mode
        assert abs(log.errors.size - logo.errors.size) <= 2, mode
        assert orc.backward_error(A, b, x) <= 2 * max(orc.backward_error(A, b, xo), 1e-11), mode


def test_cholesky_blocked_beyond_one_launch(pa):
    _say("blocked Cholesky n = 700")
    from parla_amd import ops
    rng = np.random.default_rng(3)
    X = rng.standard_normal((900, 700))
    G = X.T @ X
    R = ops.cholesky_upper(torch.from_numpy(G).cuda()).cpu().numpy()
    Rr = la.cholesky(G, lower=False)
    np.testing.assert_allclose(np.triu(R), Rr, rtol=1e-9, atol=1e-9 * np.abs(Rr).max())
    with pytest.raises(np.linalg.LinAlgError):
        ops.cholesky_upper(torch.from_numpy(-np.eye(600)).cuda())


def test_svd_right_precond_matches_reference(pa, golden):
    _say("svd_right_precond")
    import oracle as orc
    from parla_amd.comps.preconditioning import svd_right_precond
    SA = orc.gaussian_operator(80, 300, np.random.default_rng(8)) @ golden["spo_A"]
    M, U, sig, Vh = svd_right_precond(SA)
    np.testing.assert_allclose(sig, golden["svdpc_sigma"], rtol=1e-12)
    # M = V / sigma has entries up to 5.4e3 (cond(SA) = 1.3e4); one extra rounding per entry of SA
    # moves LAPACK's own M by 1.5e-9 in its small entries, so those are held to 10x that
    np.testing.assert_allclose(_sign_fix(M, golden["svdpc_M"]), golden["svdpc_M"], rtol=1e-10, atol=1.5e-8)
    np.testing.assert_allclose(U * sig @ Vh, SA, atol=1e-12 * sig[0])
    # a rank-deficient sketch keeps only the numerical rank
    SAr = SA[:, :8] @ np.random.default_rng(1).standard_normal((8, 20))
    M, U, sig, Vh = svd_right_precond(SAr)
    assert M.shape == (20, 8) and U.shape == (80, 8)


def test_pcss2_underdetermined_matches_reference(pa, golden):
    _say("PcSS2 underdetermined (delta > 0)")
    A, R, c = golden["spo_A"], golden["pcss2u_R"], golden["pcss2u_c"]
    x, y, errors = pa.PcSS2()(A, None, c, 0.5, 1e-12, 60, R, True, None)
    np.testing.assert_allclose(x, golden["pcss2u_x"], rtol=1e-9)
    np.testing.assert_allclose(y, golden["pcss2u_y"], rtol=1e-9, atol=1e-13)
    ref = golden["pcss2u_errors"]
    assert abs(np.size(errors) - ref.size) <= 1
    # the general (non-triangular) preconditioner and delta = 0: x is NaN, y the min-norm solution.
    # R came from the delta = 0.5 sketch, so this run needs 64 iterations (the oracle stops there
    # with istop 1, 1.8e-10 from lstsq); at 60 it stops unconverged, where one extra rounding in A
    # already moves the oracle's y by 1.6e-7.
    M = la.inv(R)
    x0, y0, _ = pa.PcSS2()(A, np.zeros(A.shape[0]), c, 0.0, 1e-12, 100, M, False, None)
    assert np.all(np.isnan(x0))
    yl = la.lstsq(A.T, c)[0]
    assert la.norm(y0 - yl) <= 1e-8 * la.norm(yl)   # the oracle lands 1.8e-10 from it


@pytest.mark.parametrize("m,n,rank", [(2000, 50, 50), (1500, 40, 25)])
def test_spu1_matches_oracle_and_min_norm(pa, m, n, rank):
    _say(f"SPU1 {m}x{n} rank {rank}")
    import oracle as orc
    rng = np.random.default_rng(m + n)
    U = orc.orthonormal_operator(m, rank, rng)
    Vt = orc.orthonormal_operator(rank, n, rng)
    s = np.geomspace(1, 1e-5, rank)
    A = (U * s) @ Vt
    c = A.T @ rng.standard_normal(m)
    y_opt = U @ ((1 / s) * (Vt @ c))
    y, log = pa.SPU1(pa.SkOpSJ(8), 3)(A, c, 1e-12, 100, np.random.default_rng(3))
    yo, logo = orc.spu1(A, c, 1e-12, 100, np.random.default_rng(3))
    assert la.norm(y - y_opt) / (1 + la.norm(y_opt)) <= 1e-8
    assert la.norm(y - yo) <= 1e-8 * la.norm(yo)
    assert abs(log.errors.size - logo.errors.size) <= 2
    np.testing.assert_allclose(log.errors[0], logo.errors[0], rtol=1e-10)
    from parla_amd.drivers.least_squares import spu1
    y3, _ = spu1(A, c, 1e-12, 100, np.random.default_rng(3))
    np.testing.assert_allclose(y3, y, rtol=0, atol=1e-14 * la.norm(y))


def test_sso1_matches_reference(pa, golden):
    _say("SSO1")
    A, b = golden["spo_A"], golden["spo_b"]
    for tag, delta in (("", 0.0), ("_ridge", 0.3)):
        for driver in ("gelsd", "gelsy"):
            x, log = pa.SSO1(pa.SkOpSJ(8), 3, driver)(A, b, delta, np.nan, 1, np.random.default_rng(12))
            np.testing.assert_allclose(x, golden[f"sso1{tag}_x"], rtol=1e-9, err_msg=driver + tag)
            assert log["time_sketch"] >= 0 and log["time_solve"] >= 0
    from parla_amd.drivers.least_squares import sso1
    x, _ = sso1(A, b, 0.0, np.random.default_rng(12))
    np.testing.assert_allclose(x, golden["sso1_x"], rtol=1e-9)


def test_srct_operator_matches_reference(pa, golden):
    _say("SRCT")
    import oracle as orc
    r, e, perm = golden["srct_r"], golden["srct_e"], golden["srct_perm"]
    X = golden["srct_X"]
    np.testing.assert_allclose(pa.apply_srct(r, e, X, perm), golden["srct_fwd"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(pa.apply_srct(r, e, golden["srct_fwd"], perm, forward=False), golden["srct_adj"],
                               rtol=1e-12, atol=1e-13)
    S = pa.SkOpTC()(40, 9, np.random.default_rng(5))   # tall
    np.testing.assert_allclose(S @ golden["srct_tall_X"], golden["srct_tall_out"], rtol=1e-12, atol=1e-13)
    # a larger wide operator against the oracle's DCT (scipy.fft) on the same draws
    Sw = pa.srct_operator(300, 5000, np.random.default_rng(9))
    W = orc.srct_matrix(300, 5000, np.random.default_rng(9))
    Y = np.random.default_rng(10).standard_normal((5000, 4))
    np.testing.assert_allclose(Sw @ Y, W @ Y, rtol=1e-11, atol=1e-11 * la.norm(Y))
    np.testing.assert_allclose(np.asarray(Sw), W, atol=1e-13)


def test_spo_with_srct_and_sparse_sign(pa, golden):
    """test_overdet_least_squares.py:265-270 (SPO with SkOpTC) and the sparse
    sign operator on the same problem; both against the oracle's draws."""
    _say("SPO with SRCT / sparse sign")
    import oracle as orc
    from parla_amd import matmakers as mm
    for tag, (rr, cc, seed, dens) in {"ss20x60": (20, 60, 7, 0.1), "ss50x8": (50, 8, 8, 0.3)}.items():
        S = pa.SkOpSS(dens)(rr, cc, np.random.default_rng(seed))
        np.testing.assert_array_equal(np.asarray(S), golden[f"sps_{tag}"])
        X = np.random.default_rng(1).standard_normal((cc, 3))
        np.testing.assert_allclose(S @ X, golden[f"sps_{tag}"] @ X, rtol=1e-13, atol=1e-14)
        Y = np.random.default_rng(2).standard_normal((4, rr))
        np.testing.assert_allclose(Y @ S, Y @ golden[f"sps_{tag}"], rtol=1e-13, atol=1e-14)
    A, b, x_opt, U, s, Vt = mm.inconsistent_gen()
    for gen, kind in ((pa.SkOpTC(), "srct"),):
        x, log = pa.SPO(gen, 2)(A, b, 0.0, 1e-12, 100, np.random.default_rng(4))
        xo, logo = orc.spo(A, b, 0.0, 1e-12, 100, np.random.default_rng(4), sketch=kind, sampling_factor=2)
        assert la.norm(x - xo) <= 1e-8 * la.norm(xo)
        r = A @ x - b
        assert la.norm(U @ (U.T @ r)) / la.norm(r) <= 1e-6
    x, log = pa.SPO(pa.SkOpSS(0.05), 3)(A, b, 0.0, 1e-12, 100, np.random.default_rng(4))
    assert la.norm(x - la.lstsq(A, b)[0]) <= 1e-8 * la.norm(x)


# --------------------------------------------------------------------------
# blocked QB and the Nystrom EVD (qb.py:355-595, evd.py:292-384)
# --------------------------------------------------------------------------

def test_qb3_and_evd2_match_reference(pa, golden):
    _say("QB3 / EVD2 vs reference")
    from parla_amd.utils import linalg_wrappers as ulaw
    A = golden["qb_A"]
    for tag, blk, tol in (("blk5", 5, 0.0), ("blk7tol", 7, 1e-3)):
        Q, B = pa.QB3(pa.RS1(pa.SkOpGA(), 2, ulaw.orth, 1), blk)(A, 20, tol, np.random.default_rng(31))
        Qr, Br = golden[f"qb3_{tag}_Q"], golden[f"qb3_{tag}_B"]
        assert Q.shape == Qr.shape, tag
        np.testing.assert_allclose(_sign_fix(Q, Qr), Qr, atol=1e-9)
        np.testing.assert_allclose(Q @ B, Qr @ Br, atol=1e-10)
    U = golden["qb_U"]
    Apsd = (U * np.geomspace(1, 1e-4, 30)) @ U.T
    for tag, npass in (("p1", 1), ("p3", 3)):
        V, lamb = pa.EVD2(pa.RS1(pa.SkOpGA(), npass - 1, ulaw.orth, 1))(Apsd, 12, np.nan, 6,
                                                                       np.random.default_rng(41))
        np.testing.assert_allclose(lamb, golden[f"evd2_{tag}_lamb"], rtol=1e-9)
        Vr = golden[f"evd2_{tag}_V"]
        np.testing.assert_allclose(_sign_fix(V, Vr), Vr, atol=1e-8)
    from parla_amd.drivers.evd import evd2
    V, lamb = evd2(Apsd, 12, 6, 3, np.random.default_rng(41))
    np.testing.assert_allclose(lamb, golden["evd2_p3_lamb"], rtol=1e-9)


@pytest.mark.parametrize("maker", ["tall_low_exact_rank", "fast_decay_low_exact_rank", "wide_full_exact_rank",
                                   "s_shaped_decay_low_exact_rank"])
def test_qb2_matches_oracle_and_closed_forms(pa, maker):
    """test_qb.py's QB2 checks: exact-rank recovery, B = Q^T A on the original
    A, the tolerance guarantee -- and the oracle on the same draws."""
    _say(f"QB2 {maker}")
    import oracle as orc
    from parla_amd import matmakers as mm
    from parla_amd.utils import linalg_wrappers as ulaw
    A, U, s, Vt = getattr(mm, maker)()
    for blk in (2, 5):
        qb2 = pa.QB2(pa.RF1(pa.RS1(pa.SkOpGA(), 2, ulaw.orth, 1)), blk, False)
        Q, B = qb2(A, s.size, 0.0, np.random.default_rng(7))
        assert la.norm(Q.T @ Q - np.eye(Q.shape[1])) <= 1e-10
        assert la.norm(A - Q @ B) <= 1e-8
        assert la.norm(B - Q.T @ A) <= 1e-8
        rf = lambda M, k, tol, rng: orc.rf1(M, k, tol, rng, lambda MM, kk, g: orc.rs1(MM, kk, g, None, 2, orc.orth, 1))
        Qo, Bo = orc.qb2(A, s.size, 0.0, np.random.default_rng(7), rf, blk)
        assert la.norm(Q @ Q.T - Qo @ Qo.T) <= 1e-8
        tol = 1e-2
        Q, B = qb2(A, min(A.shape), tol, np.random.default_rng(8))
        Qo, Bo = orc.qb2(A, min(A.shape), tol, np.random.default_rng(8), rf, blk)
        assert Q.shape == Qo.shape
        assert la.norm(A - Q @ B) <= tol * la.norm(A) * (1 + 1e-8)


def test_qb2_overwrite_and_procedural_drivers(pa):
    _say("QB2 overwrite_a, qb / svd1 / evd1")
    from parla_amd import matmakers as mm
    from parla_amd.utils import linalg_wrappers as ulaw
    from parla_amd.comps.qb import qb
    from parla_amd.drivers.svd import svd1
    from parla_amd.drivers.evd import evd1
    A, U, s, Vt = mm.tall_low_exact_rank()
    At = torch.from_numpy(A).cuda()
    qb2 = pa.QB2(pa.RF1(pa.RS1(pa.SkOpGA(), 2, ulaw.orth, 1)), 4, True)
    Q, B = qb2(At, s.size, 0.0, 3)
    assert la.norm(A - (Q @ B).cpu().numpy()) <= 1e-8
    assert torch.linalg.norm(At).item() <= 1e-8 * la.norm(A)   # deflated in place
    Q, B = qb(4, A, s.size, 5)
    assert la.norm(A - Q @ B) <= 1e-8
    # k + over = 15 reaches the exact rank 15: the leading singular values are recovered
    Uk, sk, Vk = svd1(A, 5, 10, np.nan, 3, 2, 6)
    np.testing.assert_allclose(sk, s[:5], rtol=1e-8)
    # below the rank (k + over = 8 of 15) the result is an approximation: the oracle's svd1 on
    # QB2 with the same draws gives the same one
    import oracle as orc
    Uk, sk, Vk = svd1(A, 5, 3, np.nan, 3, 2, 6)
    rf = lambda M, kk, tol, rng: orc.rf1(M, kk, tol, rng, lambda MM, k2, g: orc.rs1(MM, k2, g, None, 1, orc.orth, 1))
    qbo = lambda M, kk, tol, rng: orc.qb2(M, kk, tol, rng, rf, 2)
    Uo, so, Vho = orc.svd1(A, 5, np.nan, 3, np.random.default_rng(6), qbo)
    np.testing.assert_allclose(sk, so, rtol=1e-10)
    assert np.all(sk <= s[:5] * (1 + 1e-12))   # singular values of a projection of A
    S = (U * s) @ U.T
    V, lamb = evd1(S, 5, np.nan, 10, 3, 2, 7)     # k + over = 15: exact
    np.testing.assert_allclose(lamb, s[:5], rtol=1e-8)
    V, lamb = evd1(S, 5, np.nan, 3, 3, 2, 7)      # k + over = 8: the oracle's approximation
    Vo, lo = orc.evd1(S, 5, np.nan, 3, np.random.default_rng(7), qbo)
    np.testing.assert_allclose(lamb, lo, rtol=1e-10)


# --------------------------------------------------------------------------
# saddle-point systems: PcSS1 / pcg, SPS1, SPS2 (saddle.py:88-176,
# pcg.py:5-47, saddlesys.py:87-327)
# --------------------------------------------------------------------------

def test_pcss1_matches_reference(pa, golden):
    _say("PcSS1")
    import oracle as orc
    A, b, c, M = golden["sad_A"], golden["sad_b"], golden["sad_c"], golden["sad_M"]
    for tag, delta in (("", 0.0), ("_ridge", 0.2)):
        M0 = M.copy()
        x, y, res = pa.PcSS1()(A, b, c, delta, 1e-10, 50, M, False, None)
        np.testing.assert_array_equal(M, M0)   # the caller's preconditioner is left alone
        xr, yr, ref = golden[f"pcss1{tag}_x"], golden[f"pcss1{tag}_y"], golden[f"pcss1{tag}_res"]
        assert abs(res.size - ref.size) <= 1
        reruns = [orc.pcss1(rounded_matrix(A, cc), b, c, delta, 1e-10, 50, M, False, None)
                  for cc in ROUNDING_CONSTANTS]
        if delta == 0.0:   # converged in 26 steps: the solution to 1e-8
            np.testing.assert_allclose(x, xr, rtol=1e-8)
            np.testing.assert_allclose(y, yr, rtol=1e-8, atol=1e-12)
        else:
            # this preconditioner leaves kappa = 2.2e5 at delta = 0.2: PCG stops at iter_lim = 50,
            # 0.33% from the solution, where one extra rounding per entry of A moves the oracle's x by
            # up to 0.36%; x is held to 10x that change and must be as close to the solution
            env_x = max(la.norm(r[0] - xr) for r in reruns) / la.norm(xr)
            assert la.norm(x - xr) <= 10 * env_x * la.norm(xr)
            x_opt = la.solve(A.T @ A + delta * np.eye(A.shape[1]), A.T @ b - c)
            assert la.norm(x - x_opt) <= 2 * la.norm(xr - x_opt)
            np.testing.assert_allclose(y, b - A @ x, rtol=1e-12, atol=1e-12 * la.norm(b))
        # CG's residual sequence is rounding-sensitive: one extra rounding per entry of A moves the
        # oracle's own sequence by up to 60% (no ridge) and 110% (ridge) in its tail, so the sequence
        # is held to the envelope such a rounding makes
        change = perturbed_change(ref, lambda cc: reruns[ROUNDING_CONSTANTS.index(cc)][2])
        assert_within_envelope(res, ref, envelope(change, floor=1e-9))


@pytest.mark.parametrize("drv", ["sps1", "sps2"])
def test_sps_drivers_match_reference(pa, golden, drv):
    _say(f"{drv}")
    A, b, c = golden["sad_A"], golden["sad_b"], golden["sad_c"]
    cls = pa.SPS1 if drv == "sps1" else pa.SPS2
    solver = pa.PcSS1() if drv == "sps1" else pa.PcSS2()
    for tag, delta in (("", 0.0), ("_ridge", 0.2)):
        x, y, log = cls(pa.SkOpGA(), 4, solver)(A, b, c, delta, 1e-10, 50, np.random.default_rng(21), logging=True)
        np.testing.assert_allclose(x, golden[f"{drv}{tag}_x"], rtol=1e-8, err_msg=tag)
        np.testing.assert_allclose(y, golden[f"{drv}{tag}_y"], rtol=1e-8, atol=1e-12, err_msg=tag)
        ref = golden[f"{drv}{tag}_errors"]
        assert abs(log.errors.size - ref.size) <= 1, tag
        np.testing.assert_allclose(log.errors[0], ref[0], rtol=1e-10)
        n = A.shape[1]
        x_opt = la.solve(A.T @ A + delta * np.eye(n), A.T @ b - c)
        if drv == "sps2" or delta == 0:
            assert la.norm(x - x_opt) <= 1e-6 * la.norm(x_opt), tag
    if drv == "sps1":   # the Nystrom-style preconditioner (d < n)
        x, y, log = pa.SPS1(pa.SkOpGA(), 0.6, pa.PcSS1())(A, b, c, 0.1, 1e-10, 50, np.random.default_rng(22))
        np.testing.assert_allclose(x, golden["sps1nys_x"], rtol=1e-8)
        ref = golden["sps1nys_errors"]
        assert abs(log.errors.size - ref.size) <= 1
        from parla_amd.drivers.saddlesys import sps
        x2, y2, _ = sps(A, b, c, 0.0, 1e-10, 50, np.random.default_rng(3))
        x_opt = la.solve(A.T @ A, A.T @ b - c)
        assert la.norm(x2 - x_opt) <= 1e-6 * la.norm(x_opt)
