"""GPU parity of the drop-in drivers against the oracle and the reference's
golden vectors: lsqr, PcSS2, SPO, RS1/RF1/QB1, SVD1, EVD1.

Integer outcomes (istop, itn, presolve branch) are compared exactly where the
trajectory is well conditioned; floating results at the tolerances the
reference's own tests use (1e-6 ... 1e-12) or tighter.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

from trajectory import Rounded as _Rounded, assert_within_envelope, envelope, perturbed_change

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-drivers] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pa():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import parla_amd
    from parla_amd import ops
    ops.lib()
    return parla_amd


# --------------------------------------------------------------------------
# lsqr
# --------------------------------------------------------------------------

def test_lsqr_docstring_known_answers(pa):
    _say("lsqr KATs")
    A = np.array([[1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    res = pa.lsqr(A, np.zeros(3))
    assert res[1] == 0 and np.all(res[0] == 0) and np.isscalar(res[7])
    res = pa.lsqr(A, np.array([1.0, 0.0, -1.0]))
    assert res[1] == 1 and res[2] == 1
    np.testing.assert_allclose(res[0], [1.0, -1.0], atol=1e-14)
    res = pa.lsqr(A, np.array([1.0, 0.01, -1.0]))
    assert res[1] == 2
    np.testing.assert_allclose(res[0], [1.00333333, -0.99666667], atol=1e-8)
    np.testing.assert_allclose(res[3], 0.005773502691896255, rtol=1e-10)


@pytest.mark.parametrize("tag,kw", [
    ("plain", dict(atol=1e-10, btol=1e-10)),
    ("x0", dict(atol=1e-10, btol=1e-10, x0=True)),
    ("damp", dict(damp=0.3, atol=1e-10, btol=1e-10, calc_var=True)),
    ("lim", dict(iter_lim=3)),
])
def test_lsqr_matches_reference_trajectory(pa, golden, tag, kw):
    """Unpreconditioned LSQR on a cond-1e3, 60 x 8 matrix.  Once the
    8-dimensional Krylov space is exhausted the Golub-Kahan vectors lose
    orthogonality and the trajectory is driven by rounding: the oracle itself,
    with one extra rounding in A^T y, leaves the reference's log at iteration 7
    and changes itn by one.  So: istop exact, itn within 2, the log within the
    envelope that such a perturbation produces, the solution to 1e-6."""
    import oracle as orc
    _say(f"lsqr golden {tag}")
    kw = dict(kw)
    A, b = golden["lsqr_A60"], golden["lsqr_b60"]
    if kw.pop("x0", False):
        kw["x0"] = golden["lsqr_x060"]
    res = pa.lsqr(A, b, **kw)
    gi = golden[f"lsqr_{tag}_ints"]
    assert res[1] == gi[0]
    assert abs(res[2] - gi[1]) <= 2
    garn = golden[f"lsqr_{tag}_arnorms"]
    change = perturbed_change(garn, lambda c: orc.lsqr(_Rounded(orc.DenseOp(A), c), b, **kw)[7])
    assert_within_envelope(np.atleast_1d(res[7]), garn, envelope(change))
    np.testing.assert_allclose(res[0], golden[f"lsqr_{tag}_x"], rtol=1e-6, atol=1e-9)
    r1, r2, an, ac, xn = golden[f"lsqr_{tag}_scal"]
    np.testing.assert_allclose([res[3], res[4], res[8]], [r1, r2, xn], rtol=1e-6)
    np.testing.assert_allclose(res[5], an, rtol=1e-3)
    if kw.get("calc_var"):
        np.testing.assert_allclose(res[9], golden[f"lsqr_{tag}_var"], rtol=0.05)


def test_lsqr_preconditioned_matches_reference(pa, golden):
    _say("lsqr golden preconditioned")
    from parla_amd.comps.preconditioning import a_lift_precond
    A, b = golden["lsqr_A60"], golden["lsqr_b60"]
    for tag, Rkey, delta in [("pc", "lsqr_R60", 0.0), ("pcridge", "lsqr_Rd60", 0.25)]:
        A_pc = a_lift_precond(A, delta, golden[Rkey], True)[0]
        rhs = b if delta == 0 else np.concatenate((b, np.zeros(8)))
        res = pa.lsqr(A_pc, rhs, atol=1e-12, btol=1e-12, iter_lim=50)
        np.testing.assert_array_equal([res[1], res[2]], golden[f"lsqr_{tag}_ints"])
        np.testing.assert_allclose(res[0], golden[f"lsqr_{tag}_x"], rtol=1e-8)
        np.testing.assert_allclose(res[7], golden[f"lsqr_{tag}_arnorms"], rtol=1e-6)


@pytest.mark.parametrize("m,n", [(3000, 40), (20000, 200), (5000, 700), (2000, 1100)])
def test_lsqr_matches_oracle_at_size(pa, m, n):
    """Well-conditioned A (singular values in [0.5, 1]): the trajectory is
    stable, so the device recurrences must track the oracle's closely."""
    _say(f"lsqr vs oracle m={m} n={n}")
    import oracle as orc
    rng = np.random.default_rng(m + n)
    Q = la.qr(rng.standard_normal((m, n)), mode="economic")[0]
    V = la.qr(rng.standard_normal((n, n)))[0]
    A = (Q * np.linspace(1.0, 0.5, n)) @ V.T
    b = A @ rng.standard_normal(n) + 0.1 * rng.standard_normal(m)
    mine = pa.lsqr(A, b, atol=1e-12, btol=1e-12, iter_lim=80)
    ref = orc.lsqr(A, b, atol=1e-12, btol=1e-12, iter_lim=80)
    assert mine[1] == ref[1] and abs(mine[2] - ref[2]) <= 1
    k = min(mine[7].size, ref[7].size)
    np.testing.assert_allclose(mine[7][:k], ref[7][:k], rtol=1e-6)
    x_ls = la.lstsq(A, b)[0]
    assert la.norm(mine[0] - x_ls) <= 1e-10 * la.norm(x_ls)


# --------------------------------------------------------------------------
# SPO
# --------------------------------------------------------------------------

def test_spo_matches_reference_golden_runs(pa, golden):
    """SPO's error log against the reference's, within the envelope that one
    extra rounding in S @ X produces in the oracle's own log (the SJLT d = 3n
    case amplifies such a perturbation to 50% near the 1e-7 level)."""
    import oracle as orc
    _say("SPO golden runs")
    A, b = golden["spo_A"], golden["spo_b"]
    for tag, gen, kind in [("ga", pa.SkOpGA(), "gaussian"), ("sj", pa.SkOpSJ(8), "sjlt"),
                           ("sjridge", pa.SkOpSJ(8), "sjlt")]:
        sf, delta = golden[f"spo_{tag}_meta"]
        alg = pa.SPO(gen, float(sf), mode="qr")
        x, log = alg(A, b, float(delta), 1e-12, 60, np.random.default_rng(5), logging=True)
        np.testing.assert_allclose(x, golden[f"spo_{tag}_x"], rtol=1e-8)
        ref = golden[f"spo_{tag}_errors"]
        assert abs(log.errors.size - ref.size) <= 2, (tag, log.errors.size, ref.size)

        def rerun(const, kind=kind, sf=sf, delta=delta):
            def pgen(r, c, rng):
                S = orc.sjlt_operator(r, c, rng, 8) if kind == "sjlt" else orc.gaussian_operator(r, c, rng)
                return _Rounded(S, const)

            return orc.spo(A, b, float(delta), 1e-12, 60, np.random.default_rng(5), sketch=pgen,
                           sampling_factor=float(sf))[1].errors

        assert_within_envelope(log.errors, ref, envelope(perturbed_change(ref, rerun), floor=1e-5))


def test_spo_closed_form_reference_problems(pa):
    _say("SPO closed-form families")
    from parla_amd import matmakers as mm
    for gen in (mm.inconsistent_orthog, mm.inconsistent_gen):
        A, b, x_opt, U, s, Vt = gen()
        for seed in (1, 4, 15):
            x, log = pa.SPO(pa.SkOpGA(), 3)(A, b, 0.0, 1e-12, 100, np.random.default_rng(seed))
            r = A @ x - b
            assert la.norm(U @ (U.T @ r)) / la.norm(r) <= 1e-6
            y, y_opt = Vt @ x, Vt @ x_opt
            if la.norm(y_opt) > 1e-8:
                assert y @ y_opt / (la.norm(y) * la.norm(y_opt)) >= 1 - 1e-6
                assert (1 - 1e-6) * la.norm(y_opt) <= la.norm(y) <= (1 + 1e-6) * la.norm(y_opt) + 1e-6


def test_spo_presolve_is_exact_on_consistent(pa):
    _say("SPO consistent presolve")
    from parla_amd import matmakers as mm
    for gen in (mm.consistent_tall, mm.consistent_square):
        A, b, x_opt, *_ = gen()
        for seed in (1, 4, 15, 31, 42):
            x, _ = pa.SPO(pa.SkOpGA(), 1)(A, b, 0.0, 0.0, 1, np.random.default_rng(seed))
            assert la.norm(A @ x - b) <= la.norm(A @ x_opt - b) + 1e-10


@pytest.mark.parametrize("kind", ["gaussian", "sjlt"])
def test_spo_convergence_rate_families(pa, kind):
    _say(f"SPO convergence {kind}")
    from parla_amd import matmakers as mm
    from parla_amd.utils.stats import loglinear_fit
    gen = pa.SkOpGA() if kind == "gaussian" else pa.SkOpSJ(8)
    for maker, ridges in ((mm.inconsistent_gen, (False, True)), (mm.inconsistent_stackid, (False,))):
        A, b, *_ = maker()
        for ridge in ridges:
            delta = 0.25 if ridge else 0.0
            x, log = pa.SPO(gen, 2)(A, b, delta, 1e-12, 100, np.random.default_rng(34998751340))
            fit, r2 = loglinear_fit(np.arange(log.errors.size - 1), log.errors[1:])
            assert r2 >= 0.95 and fit[1] < -0.3 and log.errors[-1] <= 1e-6
            if ridge:
                n = A.shape[1]
                x_opt = la.lstsq(np.vstack((A, np.sqrt(delta) * np.eye(n))),
                                 np.hstack((b, np.zeros(n))))[0]
                assert la.norm(x - x_opt) / (1 + min(la.norm(x_opt), la.norm(x))) <= 1e-6


def test_spo_plain_config_matches_oracle(pa):
    _say("SPO plain LS 20000x200 vs oracle")
    import oracle as orc
    from parla_amd import matmakers as mm
    A, b, x_star = mm.config_ls_plain()
    x, log = pa.SPO(pa.SkOpGA(), 4)(A, b, 0.0, 1e-12, 100, np.random.default_rng(7))
    xo, logo = orc.spo(A, b, 0.0, 1e-12, 100, np.random.default_rng(7), sketch="gaussian",
                       sampling_factor=4)
    # same sketch draws: same preconditioner up to roundoff, same trajectory
    assert abs(log.errors.size - logo.errors.size) <= 1
    k = min(log.errors.size, logo.errors.size)
    np.testing.assert_allclose(log.errors[:k], logo.errors[:k], rtol=1e-4)
    assert la.norm(x - xo) <= 1e-8 * la.norm(xo)
    # ||r|| / ||b|| is 7e-7 here, so evaluating eta = ||A^T r|| / (||A||_F ||r||)
    # in FP64 has a floor near u ||b|| / ||r|| = 1e-10 (LAPACK gelsd: 1.4e-9)
    assert orc.backward_error(A, b, x) <= 1.5 * orc.backward_error(A, b, xo)


# --------------------------------------------------------------------------
# QB chain
# --------------------------------------------------------------------------

def _qb(pa, num_pass=1, stab=None, gen=None):
    from parla_amd.utils import linalg_wrappers as ulaw
    rs = pa.RS1(sketch_op_gen=gen or pa.SkOpGA(), num_pass=num_pass,
                stabilizer=stab or ulaw.orth, passes_per_stab=1)
    return pa.QB1(pa.RF1(rs))


@pytest.mark.parametrize("maker", ["tall_low_exact_rank", "wide_low_exact_rank", "fast_decay_low_exact_rank",
                                   "slow_decay_low_exact_rank", "s_shaped_decay_low_exact_rank",
                                   "tall_full_exact_rank", "wide_full_exact_rank"])
def test_qb1_exact_rank_families(pa, maker):
    _say(f"QB1 {maker}")
    from parla_amd import matmakers as mm
    A, U, s, Vt = getattr(mm, maker)()
    for seed in (1, 4, 15, 31, 42):
        Q, B = _qb(pa)(A, s.size, np.nan, np.random.default_rng(seed))
        assert la.norm(A - Q @ B) <= 1e-8
        assert la.norm(Q.T @ Q - np.eye(Q.shape[1])) <= 1e-8
        assert la.norm(B - Q.T @ A) <= 1e-8
        assert np.linalg.matrix_rank(B) == min(B.shape)


def test_qb1_matches_oracle_with_same_draws(pa):
    _say("QB1 vs oracle, SJLT + 4 passes")
    import oracle as orc
    from parla_amd import matmakers as mm
    A, U, s, Vt = mm.rand_low_rank(600, 300, np.geomspace(1, 1e-3, 60), np.random.default_rng(3), factors=True)
    k = 40
    Q, B = _qb(pa, num_pass=4, gen=pa.SkOpSJ(8))(A, k, np.nan, np.random.default_rng(9))
    rso = lambda M, kk, rng: orc.rs1(M, kk, rng, lambda r, c, g: orc.sjlt_operator(r, c, g, 8), 4, orc.orth, 1)
    Qo, Bo = orc.qb1(A, k, np.nan, np.random.default_rng(9),
                     lambda M, kk, tol, rng: orc.rf1(M, kk, tol, rng, rso))
    # Q is determined up to column signs; compare projectors and errors
    assert la.norm(Q @ Q.T - Qo @ Qo.T) <= 1e-8
    assert abs(la.norm(A - Q @ B) - la.norm(A - Qo @ Bo)) <= 1e-10


def test_rs1_lu_stabilizer_and_power_method(pa):
    _say("RS1 lu + power method")
    from parla_amd import matmakers as mm
    from parla_amd.utils import linalg_wrappers as ulaw
    from parla_amd.utils.stats import loglinear_fit
    rng = np.random.default_rng(0)
    for stab in (ulaw.orth, ulaw.lu_stabilize):
        A = mm.rand_low_rank(40, 16, np.exp(-np.arange(0, 10)), rng)
        gaps = []
        passes = np.array([1, 4, 7, 10, 13])
        for num_pass in passes:
            rs = pa.RS1(pa.SkOpGA(), int(num_pass), stab, 1)
            vec = rs(A, 1, np.random.default_rng(1))[:, 0]
            vec = vec / la.norm(vec)
            gaps.append(la.norm(A.T @ (A @ vec) - vec))
        fit, r2 = loglinear_fit(passes, np.array(gaps))
        assert r2 >= 0.7 and fit[1] <= -0.5


def test_svd1_and_evd1_families(pa):
    _say("SVD1 / EVD1")
    from parla_amd import matmakers as mm
    A, U, s, Vt = mm.tall_low_exact_rank()
    for k, over in ((5, 0), (5, 5), (12, 5), (15, 0)):
        for seed in (38972, 653, 1222):
            Uk, sk, Vk = pa.SVD1(_qb(pa))(A, k, np.nan, over, np.random.default_rng(seed))
            assert sk.size == Uk.shape[1] == Vk.shape[0] <= k
            assert np.all(np.diff(sk[::-1]) >= 0) and sk.min() >= 0
            assert la.norm(Uk.T @ Uk - np.eye(sk.size)) <= 1e-8
            assert la.norm(Vk @ Vk.T - np.eye(sk.size)) <= 1e-8
            if k + over >= s.size:
                np.testing.assert_allclose(sk, s[:k], rtol=1e-8)
    Aw, Uw, sw, Vw = mm.wide_low_exact_rank()
    signs = np.where(np.random.default_rng(0).random(sw.size) < 0.5, -1.0, 1.0)
    lam = sw * signs
    S = (Uw * lam) @ Uw.T
    for seed in (1, 2, 3):
        V, lamb = pa.EVD1(_qb(pa))(S, 12, np.nan, 5, np.random.default_rng(seed))
        assert la.norm(V.T @ V - np.eye(lamb.size)) <= 1e-8
        assert np.all(np.diff(np.abs(lamb)[::-1]) >= 0)
        np.testing.assert_allclose(np.abs(lamb), np.sort(np.abs(lam))[::-1][:12], rtol=1e-8)


def test_device_tensors_stay_on_device(pa):
    _say("tensor in / tensor out")
    rng = np.random.default_rng(0)
    A = torch.from_numpy(rng.standard_normal((500, 20))).cuda()
    b = torch.from_numpy(rng.standard_normal(500)).cuda()
    x, log = pa.SPO(pa.SkOpSJ(8, backend="device"), 4)(A, b, 0.0, 1e-12, 50, 0)
    assert isinstance(x, torch.Tensor) and x.is_cuda
    xo = la.lstsq(A.cpu().numpy(), b.cpu().numpy())[0]
    assert la.norm(x.cpu().numpy() - xo) <= 1e-9 * la.norm(xo)
    Q, B = _qb(pa, 2, gen=pa.SkOpGA(backend="device"))(A, 10, np.nan, 1)
    assert Q.is_cuda and B.is_cuda
