"""GPU parity of the one-sided block Jacobi SVD and the symmetric
eigendecomposition built on it (csrc/jacobi.hip), through the C ABI, against
LAPACK (scipy.linalg.svd / eigh).

Bounds: singular values to a small multiple of m u sigma_1 (LAPACK's own
values carry an error of that order); right singular vectors to
u / (relative gap)^2; reconstruction and orthogonality to a small multiple
of m u.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = np.finfo(float).eps


def _say(msg):
    print(f"[gpu-jacobi] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops
    ops.lib()
    return ops


def _haar(m, n, cond, seed):
    """U diag(s) V^T with Haar U, V and geometric s: every column mixes every
    singular direction (the slow case for unpreconditioned Jacobi)."""
    rng = np.random.default_rng(seed)
    U = la.qr(rng.standard_normal((m, n)), mode="economic")[0]
    V = la.qr(rng.standard_normal((n, n)))[0]
    s = np.geomspace(1.0, 1.0 / cond, n)
    return (U * s) @ V.T, s


def _graded(m, n, cond, seed):
    """Gaussian columns scaled geometrically: A = G D."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, n)) * np.geomspace(1.0, 1.0 / cond, n)


def _check_svd(A, U, s, Vt, rtol_vec=True):
    m, n = A.shape
    sr = la.svd(A, compute_uv=False)
    assert np.all(np.diff(s) <= 0)
    # LAPACK's singular values carry an absolute error of order sigma_1 u, as do ours
    assert np.max(np.abs(s - sr)) <= 50 * m * EPS * sr[0]
    assert la.norm(U * s @ Vt - A) <= 20 * m * EPS * la.norm(A)
    assert la.norm(U.T @ U - np.eye(n)) <= 20 * m * EPS * np.sqrt(n)
    assert la.norm(Vt @ Vt.T - np.eye(n)) <= 20 * m * EPS * np.sqrt(n)
    # right singular vectors: |v_j . v_ref_j| = 1 within u / (relative gap)^2
    Ur, _, Vtr = la.svd(A, full_matrices=False)
    gap = np.minimum(np.r_[np.inf, -np.diff(sr)], np.r_[-np.diff(sr), np.inf]) / sr[0]
    dots = np.abs(np.sum(Vt * Vtr, axis=1))
    assert np.all(1.0 - dots <= 100 * m * EPS / np.minimum(gap, 1.0) ** 2 + 1e-15)


@pytest.mark.parametrize("m,n,cond", [(1, 1, 1.0), (5, 3, 10.0), (80, 20, 1e4), (266, 266, 1e6), (300, 257, 1e3),
                                      (512, 512, 1e2), (600, 300, 1e8), (1024, 700, 1e4)])
def test_svd_jacobi_matches_lapack(ops, m, n, cond):
    """The kernel on its own (no preconditioning), on column-graded matrices."""
    _say(f"svd_jacobi {m}x{n} cond {cond:g}")
    A = _graded(m, n, cond, m + n)
    U, s, Vt = (t.cpu().numpy() for t in ops.svd_jacobi(torch.from_numpy(A).cuda()))
    _check_svd(A, U, s, Vt)


@pytest.mark.parametrize("m,n", [(266, 266), (120, 60)])
def test_svd_jacobi_haar_factors(ops, m, n):
    """Unpreconditioned Jacobi on Haar singular vectors takes ~27 sweeps at
    266 x 266 but converges to the same accuracy."""
    _say(f"svd_jacobi haar {m}x{n}")
    A, _ = _haar(m, n, 1e6, 7 * m + n)
    U, s, Vt = (t.cpu().numpy() for t in ops.svd_jacobi(torch.from_numpy(A).cuda()))
    _check_svd(A, U, s, Vt)


def test_svd_jacobi_rank_deficient_and_zero(ops):
    _say("svd_jacobi rank deficient")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((60, 5)) @ rng.standard_normal((5, 40))
    A[:, 7] = 0.0
    U, s, Vt = (t.cpu().numpy() for t in ops.svd_jacobi(torch.from_numpy(A).cuda()))
    sr = la.svd(A, compute_uv=False)
    np.testing.assert_allclose(s[:5], sr[:5], rtol=1e-13)
    assert np.all(s[5:] <= 1e-13 * s[0])
    assert la.norm(U * s @ Vt - A) <= 1e-13 * la.norm(A)
    Z = np.zeros((30, 10))
    U, s, Vt = (t.cpu().numpy() for t in ops.svd_jacobi(torch.from_numpy(Z).cuda()))
    assert np.all(s == 0) and np.all(U == 0)
    assert la.norm(Vt @ Vt.T - np.eye(10)) == 0.0


def test_svd_jacobi_repeatable_and_view_input(ops):
    _say("svd_jacobi repeat / strided view")
    A = _graded(200, 150, 1e5, 9)
    big = np.zeros((200, 170))
    big[:, 10:160] = A
    At = torch.from_numpy(big).cuda()[:, 10:160]
    r1 = [t.cpu().numpy() for t in ops.svd_jacobi(At)]
    r2 = [t.cpu().numpy() for t in ops.svd_jacobi(torch.from_numpy(A).cuda())]
    for a, b in zip(r1, r2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("d,n", [(80, 20), (20, 80), (3000, 266), (266, 4096)])
def test_thin_svd_both_orientations(ops, d, n):
    """thin_svd: Householder QR of the taller orientation, then the Jacobi SVD
    of R^T."""
    _say(f"thin_svd {d}x{n}")
    A, _ = _haar(max(d, n), min(d, n), 1e6, d * n)
    if d < n:
        A = A.T.copy()
    U, s, Vh = (t.cpu().numpy() for t in ops.thin_svd(torch.from_numpy(A).cuda()))
    k = min(d, n)
    assert U.shape == (d, k) and Vh.shape == (k, n)
    sr = la.svd(A, compute_uv=False)
    # Haar factors: every method determines sigma_j only to about u sigma_1
    assert np.max(np.abs(s - sr)) <= 1e-12 * sr[0]
    assert la.norm(U * s @ Vh - A) <= 1e-13 * la.norm(A) * np.sqrt(k)
    assert la.norm(U.T @ U - np.eye(k)) <= 1e-12 * np.sqrt(k)
    assert la.norm(Vh @ Vh.T - np.eye(k)) <= 1e-12 * np.sqrt(k)


@pytest.mark.parametrize("n,kind", [(30, "psd"), (266, "psd"), (120, "indefinite"), (64, "pm_pair"),
                                    (50, "rank_deficient")])
def test_sym_eig_matches_eigh(ops, n, kind):
    _say(f"sym_eig {n} {kind}")
    rng = np.random.default_rng(n)
    Q = la.qr(rng.standard_normal((n, n)))[0]
    lam = np.geomspace(1.0, 1e-6, n)
    if kind == "indefinite":
        lam = lam * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    elif kind == "pm_pair":        # exactly equal magnitudes of opposite sign: the shifted path
        lam[1] = -lam[0]
        lam[5] = -lam[4]
    elif kind == "rank_deficient":
        lam[20:] = 0.0
    C = (Q * lam) @ Q.T
    C = np.tril(C) + np.triu(rng.standard_normal((n, n)), 1)   # only the lower triangle is read
    lt, Vt = (t.cpu().numpy() for t in ops.sym_eig(torch.from_numpy(C).cuda()))
    Cs = np.tril(C) + np.tril(C, -1).T
    lr, Vr = la.eigh(Cs)
    order = np.argsort(-np.abs(lr), kind="stable")
    assert np.all(np.diff(np.abs(lt)) <= 1e-15)
    np.testing.assert_allclose(np.sort(lt), np.sort(lr), rtol=0, atol=20 * n * EPS * np.abs(lr).max())
    assert la.norm(Cs @ Vt - Vt * lt) <= 20 * n * EPS * la.norm(Cs)
    assert la.norm(Vt.T @ Vt - np.eye(n)) <= 20 * n * EPS * np.sqrt(n)
    if kind in ("psd", "indefinite"):
        dots = np.abs(np.sum(Vt * Vr[:, order], axis=0))
        assert np.all(1.0 - dots[:10] <= 1e-8)
