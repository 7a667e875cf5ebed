"""GPU parity of the individual kernels against numpy / scipy / the oracle.

Tolerances are stated per test: FP64 products are compared at a few hundred
units of roundoff scaled by the operand norms (summation orders differ), and
everything integer (CSR structure, pivots, iteration counts) bit-exactly.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def _say(msg):
    print(f"[gpu-kernels] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops as _ops
    _ops.lib()
    return _ops


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel(a, b):
    return la.norm(a - b) / max(la.norm(b), 1e-300)


# --------------------------------------------------------------------------
# GEMM / GEMV / norms / TRSV
# --------------------------------------------------------------------------

@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (37, 53, 29), (64, 64, 16), (130, 70, 200),
                                   (200, 45, 20000), (5, 300, 7), (266, 266, 266), (100, 40, 1000)])
def test_gemm_matches_numpy(ops, ta, tb, M, N, K):
    _say(f"gemm ta={ta} tb={tb} M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    At, Bt = dev(A), dev(B)
    C = ops.gemm(At.t() if ta else At, Bt.t() if tb else Bt)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    tol = 50 * EPS * np.sqrt(K)
    assert rel(host(C), ref) < tol


def test_gemm_alpha_beta_and_strided_output(ops):
    _say("gemm alpha/beta/ld")
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((33, 17)), rng.standard_normal((17, 21))
    big = rng.standard_normal((40, 30))
    out = dev(big.copy())
    ops.gemm(dev(A), dev(B), out=out[:33, 2:23], alpha=-2.0, beta=0.5)
    ref = big.copy()
    ref[:33, 2:23] = -2.0 * A @ B + 0.5 * big[:33, 2:23]
    assert np.allclose(host(out), ref, rtol=1e-13, atol=1e-13)


def test_mfma_f64_lane_layout(ops):
    # a 16 x 16 x 4 product exercises exactly one MFMA per wave: any layout
    # mistake permutes rows/columns and fails the bit-for-bit comparison of
    # an integer-valued product
    A = np.arange(16 * 4, dtype=np.float64).reshape(16, 4)
    B = np.arange(4 * 16, dtype=np.float64).reshape(4, 16) * 3 + 1
    C = host(ops.gemm(dev(A), dev(B)))
    np.testing.assert_array_equal(C, A @ B)


@pytest.mark.parametrize("m,n", [(1, 1), (100, 7), (2000, 300), (20000, 200), (50, 1500)])
def test_gemv_both_ways_and_nrm2(ops, m, n):
    _say(f"gemv m={m} n={n}")
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    Ad = dev(A)
    assert rel(host(ops.gemv(Ad, dev(x))), A @ x) < 50 * EPS * np.sqrt(n)
    assert rel(host(ops.gemv(Ad, dev(y), trans=True)), A.T @ y) < 50 * EPS * np.sqrt(m)
    yd = dev(y.copy())
    ops.gemv(Ad, dev(x), y=yd, alpha=2.0, beta=-1.0)
    assert rel(host(yd), 2 * A @ x - y) < 100 * EPS * np.sqrt(n)
    nr = float(host(ops.nrm2(dev(y)))[0])
    assert abs(nr - la.norm(y)) <= 10 * EPS * np.sqrt(m) * la.norm(y)


@pytest.mark.parametrize("n", [1, 5, 64, 65, 200, 513, 1024])
def test_trsv_matches_scipy(ops, n):
    _say(f"trsv n={n}")
    rng = np.random.default_rng(n)
    R = np.triu(rng.standard_normal((n, n))) + np.diag(3 + rng.random(n))
    b = rng.standard_normal(n)
    for trans in (False, True):
        x = host(ops.trsv_(dev(R), dev(b.copy()), trans=trans))
        ref = la.solve_triangular(R, b, trans='T' if trans else 'N', lower=False)
        # backward error of a triangular solve: |R x - b| <= c n u |R||x|
        Rop = R.T if trans else R
        assert la.norm(Rop @ x - b) <= 10 * n * EPS * la.norm(R) * la.norm(x)
        assert rel(x, ref) < 1e-10


def test_trsv_ignores_strict_lower_part(ops):
    rng = np.random.default_rng(3)
    n = 150
    R = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    junk = R + np.tril(rng.standard_normal((n, n)) * 1e3, -1)
    b = rng.standard_normal(n)
    x = host(ops.trsv_(dev(junk), dev(b.copy())))
    assert rel(x, la.solve_triangular(R, b)) < 1e-12


# --------------------------------------------------------------------------
# Householder QR, LU
# --------------------------------------------------------------------------

@pytest.mark.parametrize("m,n", [(1, 1), (10, 3), (64, 64), (800, 200), (300, 9), (2000, 266), (4100, 40), (4096, 97),
                                 (513, 200), (9000, 20)])
def test_geqrf_orgqr_match_lapack(ops, m, n):
    _say(f"geqrf m={m} n={n}")
    rng = np.random.default_rng(m * 3 + n)
    A = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -6, n))
    M = dev(A.copy())
    f = ops.geqrf_(M)
    R = host(f.R())
    Q = host(f.Q())
    Qr, Rr = la.qr(A, mode="economic")
    # same Householder sign convention as LAPACK: R agrees entrywise
    assert la.norm(R - Rr) <= 1e3 * EPS * la.norm(Rr) * np.sqrt(m)
    assert la.norm(Q.T @ Q - np.eye(n)) <= 1e3 * EPS * np.sqrt(m)
    assert la.norm(Q @ R - A) <= 1e3 * EPS * la.norm(A) * np.sqrt(m)
    assert not f.singular()


def test_geqrf_applies_qt_to_appended_columns(ops):
    _say("geqrf augmented")
    rng = np.random.default_rng(11)
    d, n = 800, 200
    A = rng.standard_normal((d, n))
    b = rng.standard_normal((d, 3))
    M = dev(np.hstack([A, b]))
    f = ops.geqrf_(M, n)
    z = host(M)[:n, n:]
    Q, R = la.qr(A, mode="economic")
    assert la.norm(z - Q.T @ b) <= 1e3 * EPS * la.norm(b) * np.sqrt(d)
    assert la.norm(host(f.R()) - R) <= 1e3 * EPS * la.norm(R) * np.sqrt(d)


def test_geqrf_flags_exact_zero_column(ops):
    A = np.random.default_rng(2).standard_normal((30, 6))
    A[:, 3] = 0.0
    A[:, 3] = A[:, 1]  # dependent column: R[3,3] is roundoff, not flagged
    f = ops.geqrf_(dev(A.copy()))
    R = host(f.R())
    assert abs(R[3, 3]) < 1e-12 * la.norm(A)
    B = np.zeros((20, 4))
    B[:, 0] = 1.0
    f = ops.geqrf_(dev(B))
    assert f.singular()


@pytest.mark.parametrize("m,n", [(7, 3), (200, 15), (4096, 64), (50, 50)])
def test_lu_pl_matches_scipy(ops, m, n):
    _say(f"lu m={m} n={n}")
    rng = np.random.default_rng(m + 13 * n)
    A = rng.standard_normal((m, n))
    PL = host(ops.lu_pl(dev(A)))
    ref = la.lu(A, permute_l=True)[0]
    assert np.allclose(PL, ref, rtol=1e-10, atol=1e-12)


# --------------------------------------------------------------------------
# sketches
# --------------------------------------------------------------------------

def test_gaussian_fill_moments_and_tiling(ops):
    _say("gaussian fill")
    d = torch.device("cuda")
    X = host(ops.gaussian_fill(1000, 1000, 1.0, 1234, d))
    assert abs(X.mean()) < 5e-3
    assert abs(X.std() - 1.0) < 5e-3
    # fourth moment of N(0,1) is 3
    assert abs((X ** 4).mean() - 3.0) < 0.05
    # element (i, j) depends only on its linear index: a sub-block drawn with
    # an offset reproduces the same values
    Y = host(ops.gaussian_fill(10, 1000, 1.0, 1234, d, offset=500 * 1000))
    np.testing.assert_array_equal(Y, X[500:510])
    Z = host(ops.gaussian_fill(1000, 1000, 1.0, 1235, d))
    assert np.corrcoef(X.ravel(), Z.ravel())[0, 1] < 5e-3


def _golden_csc(golden, tag):
    ip = golden[f"sjlt_{tag}_indptr"]
    ix = golden[f"sjlt_{tag}_indices"]
    dt = golden[f"sjlt_{tag}_data"]
    r, c, seed, k, is_csc = golden[f"sjlt_{tag}_meta"]
    return int(r), int(c), int(k), int(seed), bool(is_csc), ip, ix, dt


@pytest.mark.parametrize("tag", ["w16x64_s0", "w40x300_s42", "w800x2000_s7", "w30x50_s9_k3"])
def test_sjlt_csr_and_products_match_scipy(ops, golden, tag):
    _say(f"sjlt {tag}")
    import oracle as orc
    r, c, k, seed, is_csc, ip, ix, dt = _golden_csc(golden, tag)
    rows, vals = orc.sjlt_draw(r, c, np.random.default_rng(seed), k)
    S = sp.csc_matrix((dt, ix, ip), shape=(r, c))
    rows_d = dev(rows.astype(np.int32))
    vals_d = dev(vals)
    rowptr, cols, cv = ops.sjlt_csr(r, c, k, rows_d, vals_d)
    Scsr = S.tocsr()
    Scsr.sort_indices()
    np.testing.assert_array_equal(host(rowptr), Scsr.indptr)
    np.testing.assert_array_equal(host(cols)[: Scsr.nnz], Scsr.indices)
    np.testing.assert_array_equal(host(cv)[: Scsr.nnz], Scsr.data)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((c, 37))
    left = host(ops.sjlt_left(rowptr, cols, cv, r, dev(A)))
    assert rel(left, S @ A) < 100 * EPS
    X = rng.standard_normal((23, c))
    right = host(ops.sjlt_right(rowptr, cols, cv, r, dev(X)))
    assert rel(right, X @ S.T) < 100 * EPS


@pytest.mark.parametrize("r,c,k", [(2048, 131072, 8), (266, 4096, 8), (300, 2049, 8), (7, 70000, 4)])
def test_sjlt_csr_many_column_blocks(ops, r, c, k):
    """Past 32 blocks of 64 columns the CSR offsets come from the grouped
    scan: the result must equal scipy's CSC -> CSR of the same draw."""
    _say(f"sjlt csr {r} x {c}")
    import oracle as orc
    rows, vals = orc.device_sjlt_draw(r, c, k, 1234)
    S = sp.csc_matrix((vals.reshape(-1), rows.reshape(-1), np.arange(0, c * k + 1, k)), shape=(r, c))
    rowptr, cols, cv = ops.sjlt_csr(r, c, k, dev(rows.reshape(-1)), dev(vals.reshape(-1)))
    Scsr = S.tocsr()
    Scsr.sort_indices()
    np.testing.assert_array_equal(host(rowptr), Scsr.indptr)
    np.testing.assert_array_equal(host(cols)[: Scsr.nnz], Scsr.indices)
    np.testing.assert_array_equal(host(cv)[: Scsr.nnz], Scsr.data)


def test_sjlt_right_wide_rows_cross_lds_tiles(ops):
    _say("sjlt right n=4096 k=266")
    import oracle as orc
    k, n = 266, 4096
    rows, vals = orc.sjlt_draw(k, n, np.random.default_rng(11), 8)
    S = orc.sjlt_operator(n, k, np.random.default_rng(11), 8)  # tall: W^T
    rowptr, cols, cv = ops.sjlt_csr(k, n, 8, dev(rows.astype(np.int32)), dev(vals))
    A = np.random.default_rng(0).standard_normal((1000, n))
    out = host(ops.sjlt_right(rowptr, cols, cv, k, dev(A)))
    assert rel(out, A @ S) < 100 * EPS


def _fma_chain(terms):
    """Correctly rounded fma(v, a, acc) chain from acc = 0, in the given order."""
    from fractions import Fraction
    acc = 0.0
    for v, a in terms:
        acc = float(Fraction(v) * Fraction(a) + Fraction(acc))
    return acc


@pytest.mark.parametrize("k,n,m", [(100, 600, 45), (300, 1100, 33), (600, 700, 20)])
def test_sjlt_right_sums_in_column_order_with_fma(ops, k, n, m):
    """Both right-product kernels (lane per row for k <= 512, thread per output
    column beyond) give every output the fma chain over its CSR entries in
    ascending column order, bit for bit."""
    _say(f"sjlt right exact order k={k} n={n} m={m}")
    import oracle as orc
    rows, vals = orc.sjlt_draw(k, n, np.random.default_rng(k + n), 8)
    rowptr, cols, cv = ops.sjlt_csr(k, n, 8, dev(rows.astype(np.int32)), dev(vals))
    A = np.random.default_rng(m).standard_normal((m, n))
    out = host(ops.sjlt_right(rowptr, cols, cv, k, dev(A)))
    rp, cc, vv = host(rowptr), host(cols), host(cv)
    rng = np.random.default_rng(1)
    for r, c in zip(rng.integers(0, m, 60), rng.integers(0, k, 60)):
        p0, p1 = rp[c], rp[c + 1]
        assert np.all(np.diff(cc[p0:p1]) > 0)
        want = _fma_chain([(vv[p], A[r, cc[p]]) for p in range(p0, p1)])
        assert out[r, c] == want, (r, c, out[r, c], want)


def test_device_sjlt_draw_invariants(ops):
    _say("sjlt device draw")
    d = torch.device("cuda")
    rows, vals = ops.sjlt_draw(50, 20000, 8, 99, d)
    R, V = host(rows), host(vals)
    assert R.min() >= 0 and R.max() < 50
    assert all(len(set(r)) == 8 for r in R[:2000])
    np.testing.assert_allclose(np.abs(V), 1 / np.sqrt(8))
    assert abs((V < 0).mean() - 0.5) < 0.01
    counts = np.bincount(R.ravel(), minlength=50)
    expect = 20000 * 8 / 50
    assert np.all(np.abs(counts - expect) < 6 * np.sqrt(expect))
    rows2, _ = ops.sjlt_draw(50, 20000, 8, 99, d)
    np.testing.assert_array_equal(host(rows2), R)  # deterministic


# --------------------------------------------------------------------------
# CholeskyQR2 pieces
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 31, 32, 33, 266, 512])
def test_potrf_matches_scipy(ops, n):
    _say(f"potrf n={n}")
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n + 5, n))
    G = X.T @ X
    Gd = dev(G.copy())
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops.potrf_(Gd, info)
    assert int(host(info)[0]) == 0
    R = np.triu(host(Gd))
    Rr = la.cholesky(G, lower=False)
    assert la.norm(R - Rr) <= 1e3 * EPS * la.norm(Rr) * n


def test_potrf_reports_failed_pivot(ops):
    G = np.eye(40)
    G[17, 17] = -1.0
    Gd = dev(G)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops.potrf_(Gd, info)
    assert int(host(info)[0]) == 18


@pytest.mark.parametrize("m,n", [(100, 7), (8192, 266), (3000, 300), (500, 64)])
def test_trsm_right_and_cholqr2(ops, m, n):
    _say(f"trsm/cholqr2 m={m} n={n}")
    rng = np.random.default_rng(m + n)
    Y = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -4, n))
    R = la.qr(Y, mode="economic")[1]
    B = dev(Y.copy())
    ops.trsm_right_(B, dev(R))
    ref = la.solve_triangular(R, Y.T, trans="T").T
    assert rel(host(B), ref) < 1e-9
    Q = ops.cholqr2(dev(Y))
    assert Q is not None
    Qh = host(Q)
    assert la.norm(Qh.T @ Qh - np.eye(n)) <= 1e3 * EPS * np.sqrt(m)
    # same range as Y: Q Q^T Y = Y
    assert la.norm(Qh @ (Qh.T @ Y) - Y) <= 1e-12 * la.norm(Y)
    # agrees with LAPACK's Q up to column signs
    Ql = la.qr(Y, mode="economic")[0]
    assert la.norm(np.abs(Ql.T @ Qh) - np.eye(n)) <= 1e-8


def test_cholqr2_declines_rank_deficient(ops):
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((300, 10)) @ rng.standard_normal((10, 20))  # rank 10
    assert ops.cholqr2(dev(Y)) is None
    Q = host(ops.orth_q(dev(Y)))  # falls back to Householder
    assert la.norm(Q.T @ Q - np.eye(20)) <= 1e-12


@pytest.mark.parametrize("k", [1, 7, 266, 512])
def test_small_combine_and_cholqr_stats(ops, k):
    """The CholeskyQR2 helpers: A + alpha B + sigma I and its halved upper
    triangle, the Frobenius norm, and the packed acceptance figures."""
    import torch
    rng = np.random.default_rng(k)
    A = rng.standard_normal((k, k))
    B = rng.standard_normal((k, k))
    full = A - 0.5 * B + 2.0 * np.eye(k)
    nrm = torch.empty(1, dtype=torch.float64, device="cuda")
    got = host(ops.small_combine(dev(A), dev(B), alpha=-0.5, sigma=2.0, nrm=nrm))
    np.testing.assert_allclose(got, full, rtol=0, atol=1e-15 * np.abs(full).max())
    assert abs(nrm.item() - la.norm(full)) <= 1e-13 * la.norm(full)
    # without the norm the combine runs over many workgroups: the same values
    np.testing.assert_array_equal(host(ops.small_combine(dev(A), dev(B), alpha=-0.5, sigma=2.0)), got)
    h = host(ops.small_combine(dev(A), sigma=-1.0, upper_half=True))
    E = A - np.eye(k)
    want = np.triu(E)
    want[np.diag_indices(k)] *= 0.5
    np.testing.assert_array_equal(h, want)
    G = dev(np.diag(rng.uniform(0.5, 3.0, k)) * np.sign(rng.standard_normal(k))[:, None])
    info = torch.tensor([3], dtype=torch.int32, device="cuda")
    stats = torch.empty(4, dtype=torch.float64, device="cuda")
    ops.check(ops.lib().pamd_cholqr_stats_f64(k, ops._p(G), G.stride(0), ops._p(info), ops._p(nrm),
                                              ops._p(stats), ops.stream()), "stats")
    d = np.abs(np.diag(host(G)))
    assert host(stats).tolist() == [d.min(), d.max(), nrm.item(), 3.0]


@pytest.mark.parametrize("n", [1, 7, 64, 266, 1024])
def test_trtri_matches_scipy(ops, n):
    rng = np.random.default_rng(n)
    R = np.triu(rng.standard_normal((n, n))) + np.diag(4 + rng.random(n))
    Ri = host(ops.trtri(dev(R)))
    assert np.all(np.tril(Ri, -1) == 0)
    assert la.norm(R @ Ri - np.eye(n)) <= 100 * n * EPS * la.norm(R) * la.norm(Ri)


@pytest.mark.parametrize("m,k", [(8192, 266), (100, 7), (3000, 300), (777, 512), (16, 16), (33, 1)])
def test_trsm_mfma_matches_scipy(ops, m, k):
    """X = Y R^{-1} by the MFMA strip solver, out of place and in place; only
    the upper triangle of R is read (junk below it)."""
    import scipy.linalg as sla
    rng = np.random.default_rng(m + k)
    R = np.triu(rng.standard_normal((k, k))) + 2 * np.sqrt(k) * np.eye(k)
    Rj = R + np.tril(rng.standard_normal((k, k)), -1)
    Y = rng.standard_normal((m, k))
    ref = sla.solve_triangular(R, Y.T, trans='T').T
    Yd = torch.from_numpy(Y).cuda()
    X = ops.trsm_mfma(Yd, torch.from_numpy(Rj).cuda()).cpu().numpy()
    np.testing.assert_allclose(X, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    ops.trsm_mfma(Yd, torch.from_numpy(Rj).cuda(), out=Yd)
    np.testing.assert_allclose(Yd.cpu().numpy(), ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())

