"""GPU parity of the operator-level drop-ins against the oracle: the Haar
operator (SkOpON / orthonormal_operator), the row-selection operator (SkOpIN /
sampling_operator, on the library's gather and scatter kernels), the
preconditioned operator's matvec / rmatvec, the procedural rs1 / rf1 / spo3,
and the triangular solve and preconditioned LSQR past the LDS-resident sizes.

Every comparison uses the same numpy draws on both sides (the library's
default ``backend='numpy'``), so the results agree to rounding.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-operators] {msg}", flush=True)
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
# SkOpON / orthonormal_operator (utils/sketching.py:9-17, oblivious.py:31-35)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(300, 40), (40, 300), (64, 64), (1000, 1)])
def test_orthonormal_operator_matches_oracle(pa, shape):
    _say(f"SkOpON {shape}")
    import oracle as orc
    r, c = shape
    S = pa.SkOpON()(r, c, np.random.default_rng(17))
    got = np.asarray(S)
    ref = orc.orthonormal_operator(r, c, np.random.default_rng(17))
    assert got.shape == ref.shape == shape
    # same Gaussian draws, diag(R) > 0 fixes Q uniquely: equal to rounding
    assert la.norm(got - ref) <= 1e-12 * np.sqrt(min(shape))
    Q = got if r >= c else got.T
    assert la.norm(Q.T @ Q - np.eye(Q.shape[1])) <= 1e-13 * Q.shape[1]
    # the sign convention: R = Q^T G has a positive diagonal
    G = orc.gaussian_operator(max(r, c), min(r, c), np.random.default_rng(17))
    assert np.all(np.diag(Q.T @ G) > 0)
    # the operator applies like its matrix
    X = np.random.default_rng(1).standard_normal((c, 3))
    np.testing.assert_allclose(S @ X, ref @ X, rtol=0, atol=1e-12 * la.norm(X))


# --------------------------------------------------------------------------
# SkOpIN / sampling_operator (utils/sketching.py:204-232, oblivious.py:75-82)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(50, 400), (400, 50), (7, 7)])
@pytest.mark.parametrize("given", [False, True])
def test_sampling_operator_matches_oracle(pa, shape, given):
    _say(f"SkOpIN {shape} given={given}")
    import oracle as orc
    r, c = shape
    big, small = max(r, c), min(r, c)
    indices = None
    if given:
        indices = np.sort(np.random.default_rng(5).choice(big, small, replace=False))
    S = pa.SkOpIN(indices)(r, c, np.random.default_rng(23))
    idx, wide = orc.sampling_operator(r, c, np.random.default_rng(23), indices)
    assert wide == (c >= r)
    np.testing.assert_array_equal(np.asarray(S.indices), idx)
    # brute force: the dense selection matrix of the oracle's indices
    W = np.zeros((small, big))
    W[np.arange(small), idx] = 1.0
    D = W if wide else W.T
    assert D.shape == shape
    rng = np.random.default_rng(8)
    X = rng.standard_normal((c, 6))
    Y = rng.standard_normal((5, r))
    Z = rng.standard_normal((r, 4))
    V = rng.standard_normal((3, c))
    # gathers and scatters are copies: bit-exact
    np.testing.assert_array_equal(S @ X, D @ X)
    np.testing.assert_array_equal(Y @ S, Y @ D)
    np.testing.assert_array_equal(S.T @ Z, D.T @ Z)
    np.testing.assert_array_equal(V @ S.T, V @ D.T)
    np.testing.assert_array_equal(S @ X[:, 0], D @ X[:, 0])
    np.testing.assert_array_equal(Y[0] @ S, Y[0] @ D)
    # device tensors stay on the device; transposed views take the gather directly
    Xt = torch.from_numpy(X).cuda()
    out = S @ Xt
    assert isinstance(out, torch.Tensor) and out.is_cuda
    np.testing.assert_array_equal(out.cpu().numpy(), D @ X)
    Yt = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda().t()
    np.testing.assert_array_equal((Yt @ S).cpu().numpy(), Y @ D)


def test_sampling_sketch_in_spo(pa):
    """A row-selection sketch drives SPO (a coherence-free problem, so uniform
    rows give a usable preconditioner); the same indices in the oracle give
    the same solution."""
    _say("SPO with SkOpIN")
    import oracle as orc
    from parla_amd import matmakers as mm
    A, b, _ = mm.ls_problem(3000, 20, 1e3, 1e-6, np.random.default_rng(4))
    x, log = pa.SPO(pa.SkOpIN(), 5)(A, b, 0.0, 1e-12, 100, np.random.default_rng(6))

    def gen(r, c, rng):
        idx, wide = orc.sampling_operator(r, c, rng)
        W = np.zeros((min(r, c), max(r, c)))
        W[np.arange(idx.size), idx] = 1.0
        return W if wide else W.T

    xo, logo = orc.spo(A, b, 0.0, 1e-12, 100, np.random.default_rng(6), sketch=gen, sampling_factor=5)
    assert la.norm(x - xo) <= 1e-8 * la.norm(xo)
    assert abs(log.errors.size - logo.errors.size) <= 1


# --------------------------------------------------------------------------
# PrecondOperator.matvec / rmatvec (comps/preconditioning.py:16-67)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("upper_tri", [True, False])
@pytest.mark.parametrize("delta", [0.0, 0.25])
def test_precond_operator_matches_oracle(pa, upper_tri, delta):
    _say(f"a_lift_precond upper_tri={upper_tri} delta={delta}")
    import oracle as orc
    from parla_amd.comps.preconditioning import a_lift_precond
    rng = np.random.default_rng(11)
    m, n = 500, 30
    A = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -4, n))
    Ssk = orc.gaussian_operator(4 * n, m, rng)
    SA = Ssk @ A
    if delta > 0:
        SA = np.vstack((SA, np.sqrt(delta) * np.eye(n)))
    if upper_tri:
        R = la.qr(SA, mode="economic")[1]
    else:  # an SVD-style right preconditioner M = V / sigma
        _, sig, Vh = la.svd(SA, full_matrices=False)
        R = Vh.T / sig
    op, mf, ma = a_lift_precond(A, delta, R, upper_tri)
    ref, rmf, rma = orc.a_lift_precond(A, delta, R, upper_tri)
    assert op.shape == ref.shape
    z = rng.standard_normal(n)
    u = rng.standard_normal(ref.shape[0])
    for got, want in ((op.matvec(z), ref.matvec(z)), (op.rmatvec(u), ref.rmatvec(u)),
                      (mf(z), rmf(z)), (ma(z), rma(z))):
        assert isinstance(got, np.ndarray)
        assert la.norm(got - want) <= 1e-12 * la.norm(want)
    # device in, device out
    zt = torch.from_numpy(z).cuda()
    out = op.matvec(zt)
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert la.norm(out.cpu().numpy() - ref.matvec(z)) <= 1e-12 * la.norm(ref.matvec(z))


# --------------------------------------------------------------------------
# procedural entry points: rs1 (aware.py:9-30), rf1 (rangefinders.py:22-71),
# spo3 (least_squares.py:200-203)
# --------------------------------------------------------------------------

def test_procedural_rs1_rf1_match_oracle(pa):
    _say("rs1 / rf1")
    import oracle as orc
    from parla_amd import matmakers as mm
    from parla_amd.comps.sketchers.aware import rs1
    from parla_amd.comps.rangefinders import rf1
    A = mm.rand_low_rank(500, 200, np.geomspace(1, 1e-3, 60), np.random.default_rng(2))
    k = 25
    for num_pass in (1, 2, 3, 4):
        S = rs1(A, k, num_pass, np.random.default_rng(num_pass))
        So = orc.rs1(A, k, np.random.default_rng(num_pass), None, num_pass, orc.orth, 1)
        assert S.shape == So.shape
        if num_pass == 0:
            continue
        # stabilised bases: equal up to column signs, so compare projectors
        P, Po = orc.orth(S), orc.orth(So)
        assert la.norm(P @ P.T - Po @ Po.T) <= 1e-8, num_pass
    for num_pass in (1, 2, 3):
        Q = rf1(A, k, num_pass, np.random.default_rng(40 + num_pass))
        rso = lambda M, kk, rng, q=num_pass - 1: orc.rs1(M, kk, rng, None, q, orc.orth, 1)
        Qo = orc.rf1(A, k, np.nan, np.random.default_rng(40 + num_pass), rso)
        assert la.norm(Q.T @ Q - np.eye(k)) <= 1e-12
        assert la.norm(Q @ Q.T - Qo @ Qo.T) <= 1e-8, num_pass


def test_procedural_spo3_matches_oracle(pa):
    _say("spo3")
    import oracle as orc
    from parla_amd import matmakers as mm
    from parla_amd.drivers.least_squares import spo3
    A, b, _ = mm.ls_problem(4000, 60, 1e5, 1e-6, np.random.default_rng(12))
    for delta in (0.0, 0.1):
        x, log = spo3(A, b, delta, 1e-12, 80, np.random.default_rng(3))
        xo, logo = orc.spo(A, b, delta, 1e-12, 80, np.random.default_rng(3), sketch="sjlt", sampling_factor=3)
        assert la.norm(x - xo) <= 1e-8 * la.norm(xo), delta
        assert abs(log.errors.size - logo.errors.size) <= 1
        k = min(log.errors.size, logo.errors.size, 8)
        np.testing.assert_allclose(log.errors[:k], logo.errors[:k], rtol=1e-6)


# --------------------------------------------------------------------------
# sizes past the LDS-resident triangular solves
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [2049, 3000])
def test_trsv_large_n(pa, n):
    _say(f"trsv n={n}")
    from parla_amd import ops
    rng = np.random.default_rng(n)
    R = np.triu(rng.standard_normal((n, n))) / np.sqrt(n) + 4.0 * np.eye(n)
    b = rng.standard_normal(n)
    Rt = torch.from_numpy(R).cuda()
    for trans in (False, True):
        x = ops.trsv_(Rt, torch.from_numpy(b).cuda(), trans=trans).cpu().numpy()
        xr = la.solve_triangular(R, b, trans="T" if trans else "N", lower=False)
        assert la.norm(x - xr) <= 1e-13 * la.norm(xr) * np.sqrt(n), trans


@pytest.mark.parametrize("n", [1901, 2049])
def test_spo_wide_problems(pa, n):
    """SPO on problems wider than the LDS-resident substitution vector: the
    preconditioned LSQR keeps it in the workspace and reaches the same
    solution as LAPACK."""
    _say(f"SPO n={n}")
    import oracle as orc
    rng = np.random.default_rng(n)
    m = 3 * n + 500
    A = rng.standard_normal((m, n)) @ np.diag(np.geomspace(1, 1e-3, n))
    x_star = rng.standard_normal(n)
    b = A @ x_star + 1e-6 * rng.standard_normal(m)
    x, log = pa.SPO(pa.SkOpSJ(8), 3)(A, b, 0.0, 1e-12, 60, np.random.default_rng(1))
    xl = la.lstsq(A, b)[0]
    assert la.norm(x - xl) <= 1e-8 * la.norm(xl)
    # the oracle on the same draws: same iteration count (to rounding) and a
    # backward error as small as its own (4.4e-10 at n = 1901; LAPACK's gelsd
    # lands at 4.3e-9 on this problem)
    xo, logo = orc.spo(A, b, 0.0, 1e-12, 60, np.random.default_rng(1), sketch="sjlt", sampling_factor=3)
    assert la.norm(x - xo) <= 1e-8 * la.norm(xo)
    assert abs(log.errors.size - logo.errors.size) <= 2
    assert orc.backward_error(A, b, x) <= 2 * orc.backward_error(A, b, xo)
