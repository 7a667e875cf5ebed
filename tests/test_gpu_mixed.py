"""Mixed-precision kernels and the refined solver on the device, against the
oracle's exact integer arithmetic, extended-precision residuals and the FP64
least-squares solution."""
import numpy as np
import pytest
import torch

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-mixed] {msg}", flush=True)


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops as o
    o.lib()
    return o


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_gauss_q_matches_fill_quantised(ops):
    _say("gauss_q vs gaussian_fill")
    rows, m, seed = 37, 1000, 1234
    Q = ops.gauss_q(rows, m, seed, torch.device("cuda"))
    Z = ops.gaussian_fill(rows, m, 1.0, seed, torch.device("cuda"))
    bits = ops.unblock_pieces(Q[None])[0]
    Qh = (bits.astype(np.uint32) << 16).view(np.float32)
    assert Qh.shape == (128, 1024) and np.all(Qh[:, m:] == 0) and np.all(Qh[rows:] == 0)
    np.testing.assert_array_equal(Qh[:rows, :m], orc.quantize_q(Z.cpu().numpy()))
    # the host draw's blocked layout is the same array
    assert torch.equal(ops.quantize_q(Z), Q)


@pytest.mark.parametrize("m,n,with_b,P", [(700, 30, True, 4), (256, 64, False, 3), (1000, 65, True, 5),
                                          (3, 1, False, 1)])
def test_ozaki_split_bit_exact(ops, m, n, with_b, P):
    _say(f"split m={m} n={n} b={with_b} P={P}")
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) * np.exp(rng.uniform(-20, 20, size=(m, 1)))
    b = rng.standard_normal(m) if with_b else None
    sp = ops.OzakiSplit(_dev(A), None if b is None else _dev(b), P)
    pieces, exps = orc.ozaki_split(A, b, P)
    np.testing.assert_array_equal(sp.exps.cpu().numpy(), exps)
    for p in range(P):
        got = sp.piece(p)[:, :m].T
        np.testing.assert_array_equal(got, pieces[p])


@pytest.mark.parametrize("d,m,n,P", [(128, 512, 64, 4), (200, 700, 31, 4), (64, 256, 1, 2), (300, 1300, 130, 5)])
def test_ozaki_sketch_exact(ops, d, m, n, P):
    """The device product equals the int64 evaluation of Q @ pieces summed in
    the same order: bit for bit."""
    _say(f"ozaki sketch d={d} m={m} n={n} P={P}")
    rng = np.random.default_rng(d + m)
    A = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    Z = rng.standard_normal((d, m))
    Q = ops.quantize_q(_dev(Z))
    sp = ops.OzakiSplit(_dev(A), _dev(b), P)
    scale = 2.0 ** -5 / np.sqrt(min(d, m))
    out = ops.ozaki_sketch(Q, d, sp, scale).cpu().numpy()
    pieces, exps = orc.ozaki_split(A, b, P)
    group = ops.ozaki_layout(d, m, n + 1)[3]
    ref = orc.ozaki_product(orc.quantize_q(Z), pieces, exps, scale, group_chunks=group)
    np.testing.assert_array_equal(out, ref)
    # and it is the sketch of [A | b] to the truncation level 2^-9P
    full = scale * orc.quantize_q(Z) @ np.column_stack([A, b])
    bound = scale * np.abs(orc.quantize_q(Z)).sum(axis=1, keepdims=True) * np.abs(np.column_stack([A, b])).max()
    assert np.all(np.abs(out - full) <= 2.0 ** (-9 * P) * bound + 1e-13 * np.abs(full).max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_normal_pass(ops, dtype):
    _say(f"normal pass {dtype}")
    rng = np.random.default_rng(3)
    # 131072 x 200: the grid at its cap (four rows in flight per wavefront)
    for m, n in [(5000, 200), (777, 1024), (64, 3), (131072, 200), (70001, 40)]:
        B = rng.standard_normal((m, n))
        p = rng.standard_normal(n)
        Bt = _dev(B).to(dtype)
        q = ops.NormalOp(Bt)(_dev(p)).cpu().numpy()
        Bh = Bt.double().cpu().numpy()
        ref = Bh.T @ (Bh @ p)
        np.testing.assert_allclose(q, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_compensated_residual(ops):
    """r and A^T r against correctly rounded values, with r 1e-10 of b."""
    _say("compensated residual")
    rng = np.random.default_rng(4)
    m, n = 6000, 200
    A = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -6, n))
    x = rng.standard_normal(n)
    b = A @ x + 1e-10 * rng.standard_normal(m)
    g, rss, r = ops.Residual(_dev(A))(_dev(x), _dev(b), want_r=True)
    r_ref, g_ref, rss_ref = orc.residual_exact(A, x, b)
    r = r.cpu().numpy()
    np.testing.assert_allclose(r, r_ref, rtol=4 * np.finfo(float).eps, atol=0)
    np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=0, atol=1e-14 * np.abs(g_ref).max())
    assert abs(rss.item() - rss_ref) <= 1e-12 * rss_ref
    # the plain FP64 residual is far less accurate here
    r64 = b - A @ x
    assert np.max(np.abs(r64 - r_ref)) > 1e3 * np.max(np.abs(r - r_ref))


@pytest.mark.parametrize("m,n", [(3000, 64), (5000, 200), (2000, 257), (777, 1024), (100, 33)])
def test_trsm_recursive(ops, m, n):
    _say(f"trsm m={m} n={n}")
    rng = np.random.default_rng(m * 7 + n)
    R = np.triu(rng.standard_normal((n, n))) + 4 * np.sqrt(n) * np.eye(n)
    A = rng.standard_normal((m, n))
    X = ops.trsm_rec_(_dev(A), _dev(R)).cpu().numpy()
    ref = A @ np.linalg.inv(R)
    import scipy.linalg as sla
    ref = sla.solve_triangular(R, A.T, trans='T').T
    np.testing.assert_allclose(X, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # out of place with the FP32 copy: A untouched, and bit for bit the FP32
    # rounding of the in-place result (same recursion, same sums)
    Ad = _dev(A)
    X32 = ops.trsm_rec_f32(Ad, _dev(R)).cpu().numpy()
    assert np.array_equal(Ad.cpu().numpy(), A)
    assert X32.dtype == np.float32
    assert np.array_equal(X32, X.astype(np.float32))


def _sketch_like(d, n, cond, seed):
    """[S A | S b]: Haar factors with geometric singular values, b mostly in range."""
    import scipy.linalg as sla
    rng = np.random.default_rng(seed)
    U = sla.qr(rng.standard_normal((d, n)), mode="economic")[0]
    V = sla.qr(rng.standard_normal((n, n)))[0]
    SA = (U * np.geomspace(1.0, 1.0 / cond, n)) @ V.T
    Sb = SA @ rng.standard_normal(n) + 1e-3 * rng.standard_normal(d)
    return np.column_stack([SA, Sb])


@pytest.mark.parametrize("newton", [True, False])
@pytest.mark.parametrize("d,n,cond", [(800, 200, 1e6), (480, 120, 1e4), (300, 257, 1e2), (64, 1, 1.0)])
def test_sketch_rz_cholqr2_matches_householder(ops, monkeypatch, d, n, cond, newton):
    """R and z = Q^T S b from CholeskyQR2 against LAPACK's Householder QR of the
    same sketch, after making R's diagonal positive: both are backward stable,
    so they agree to a small multiple of u kappa(S A).  The second round by
    its Newton form and (threshold 0) by potrf."""
    import scipy.linalg as sla
    _say(f"sketch R, z by CholeskyQR2 d={d} n={n} cond={cond:g} newton={newton}")
    if not newton:
        monkeypatch.setattr(ops, "SKETCH_NEWTON_MAX_E", 0.0)
    M = _sketch_like(d, n, cond, d + n)
    rz = ops.sketch_rz_cholqr2(_dev(M), n)
    assert rz is not None
    R, z = (t.cpu().numpy() for t in rz)
    Ql, Rl = sla.qr(M[:, :n], mode="economic")
    sg = np.sign(np.diag(Rl))
    Rl, zl = sg[:, None] * Rl, sg * (Ql.T @ M[:, n])
    assert np.all(np.diag(R) > 0) and np.all(np.tril(R, -1) == 0)
    tol = 100 * np.finfo(float).eps * cond * np.sqrt(n)
    assert np.linalg.norm(R - Rl) <= tol * np.linalg.norm(Rl)
    assert np.linalg.norm(z - zl) <= tol * np.linalg.norm(M[:, n])
    # the preconditioned sketch S A R^{-1} is as orthonormal as with LAPACK's R
    # (what the refinement's operator B = A R^{-1} relies on); both measured by
    # the same host solve, whose own error is of order u kappa
    def orth_err(Rx):
        Qc = sla.solve_triangular(Rx, M[:, :n].T, trans="T").T
        return np.linalg.norm(Qc.T @ Qc - np.eye(n))
    assert orth_err(R) <= 10 * orth_err(Rl) + 1e-13 * np.sqrt(n)


def test_sketch_rz_cholqr2_declines_ill_conditioned(ops):
    """kappa(S A) = 1e10 is beyond CholeskyQR2: None, and SPOMixed factors by Householder."""
    _say("sketch R, z by CholeskyQR2 declines kappa 1e10")
    assert ops.sketch_rz_cholqr2(_dev(_sketch_like(600, 100, 1e10, 5)), 100) is None


@pytest.mark.parametrize("cond", [1e4, 1e8])
def test_spo_mixed_reaches_fp64_floor(ops, cond):
    """The refined solution's backward error (exactly evaluated) is below
    1e-12 or no worse than LAPACK's FP64 solution of the same problem, and x
    agrees with it to the conditioning bound and with the oracle's
    restatement of the algorithm on the same draw."""
    import parla_amd as pa
    from parla_amd import matmakers as mm
    _say(f"SPOMixed cond={cond:g}")
    rng = np.random.default_rng(11)
    m, n = 6000, 120
    U = mm.haar(m, n, rng)
    V = mm.haar(n, n, rng)
    s = np.logspace(0, -np.log10(cond), n)
    A = (U * s) @ V.T
    xs = rng.standard_normal(n)
    b = A @ xs + 1e-8 * rng.standard_normal(m)
    alg = pa.SPOMixed(4, npieces=4)
    x, log = alg(_dev(A), _dev(b), 0.0, 1e-12, 100, 5)
    x = x.cpu().numpy()
    xl = np.linalg.lstsq(A, b, rcond=None)[0]
    eta = orc.backward_error_extended(A, b, x)
    eta_l = orc.backward_error_extended(A, b, xl)
    assert eta <= max(1e-12, 1.5 * eta_l), (eta, eta_l, log.errors)
    assert log.refinement_steps <= 6
    assert np.linalg.norm(x - xl) <= 1e-15 * cond * 100 * np.linalg.norm(xl)
    xo, eo = orc.spo_mixed(A, b, 1e-12, 5)
    assert np.linalg.norm(x - xo) <= 1e-15 * cond * 100 * np.linalg.norm(xl)


def test_ozaki_sketch_full_size_long_groups(ops):
    """At the 262144 x 1024 shape every workgroup streams 1024 K steps: the
    result is reproducible bit for bit and within the truncation bound of the
    FP64 product everywhere (a stale LDS buffer shows up as a piece-sized
    error in some entries and a run-to-run difference)."""
    d, m, n, P = 4096, 262144, 1024, 4
    _say(f"ozaki sketch full size d={d} m={m} n={n} P={P}")
    dev = torch.device("cuda")
    assert ops.ozaki_layout(d, m, n + 1)[3] == 128  # chunks per K group
    Q = ops.gauss_q(d, m, 99, dev)
    Qf = torch.clamp(torch.round(32.0 * ops.gaussian_fill(d, m, 1.0, 99, dev)), -256.0, 256.0)
    g = torch.Generator(device=dev).manual_seed(3)
    A = torch.randn((m, n), generator=g, dtype=torch.float64, device=dev)
    b = torch.randn((m,), generator=g, dtype=torch.float64, device=dev)
    sp = ops.OzakiSplit(A, b, P)
    scale = 2.0 ** -5 / np.sqrt(d)
    out1 = ops.ozaki_sketch(Q, d, sp, scale)
    out2 = ops.ozaki_sketch(Q, d, sp, scale)
    assert torch.equal(out1, out2)
    Ab = torch.cat([A, b[:, None]], dim=1)
    full = scale * (Qf @ Ab)
    bound = scale * Qf.abs().sum(dim=1, keepdim=True) * Ab.abs().max()
    bad = int(((out1 - full).abs() > 2.0 ** (-9 * P) * bound + 1e-13 * full.abs().max()).sum().item())
    assert bad == 0, bad


def test_gauss_q_full_size_matches_quantised_fill(ops):
    """The FP32 draw with its FP64 fallback gives, bit for bit, the quantised
    FP64 fill at the 4096 x 262144 sketch size (zero as +0 on both sides)."""
    rows, m, seed = 4096, 262144, 12345
    _say(f"gauss_q full size {rows}x{m}")
    dev = torch.device("cuda")
    Q = ops.gauss_q(rows, m, seed, dev)
    Z = ops.gaussian_fill(rows, m, 1.0, seed, dev)
    Qf = ops.quantize_q(Z)
    del Z
    assert torch.equal(Q, Qf)
