"""R factor by communication-avoiding Householder QR (pamd_qr_r_f64) against
LAPACK (scipy.linalg.qr), the factorisation of the sketch in
drivers/least_squares.py:311-316.

R is unique up to the signs of its rows for a full-rank matrix; the kernel
returns the non-negative-diagonal one, so it is compared with D R_lapack,
D = sign(diag(R_lapack)), at a few hundred units of roundoff (well-conditioned
inputs).  Ill-conditioned inputs are checked through what the factorisation
must satisfy whatever the algorithm: R^T R = A^T A and R^T z = A^T b
(z = the first n rows of Q^T b) to backward-error size.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def _say(msg):
    print(f"[gpu-caqr] {msg}", flush=True)
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
    return t.cpu().numpy()


def _positive(R):
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return d[:, None] * R, d


# shapes around the kernel's cases: one block (m <= 512), P = ceil(m / 512)
# blocks with a stack of P R factors, partial last panels (n % 32 != 0),
# appended columns, square matrices and the sketches of the LS workloads
@pytest.mark.parametrize("m,n,extra", [(1, 1, 0), (10, 3, 1), (64, 64, 0), (300, 9, 2), (512, 32, 1),
                                       (513, 33, 1), (800, 200, 1), (1024, 100, 0), (1025, 40, 3),
                                       (2000, 266, 1), (4096, 97, 0), (4096, 513, 1), (8192, 64, 1),
                                       (700, 700, 0)])
def test_qr_r_matches_lapack(ops, m, n, extra):
    _say(f"qr_r m={m} n={n} extra={extra}")
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, n))
    B = rng.standard_normal((m, extra))
    M = dev(np.hstack([A, B]))
    f = ops.qr_r_caqr_(M, n)
    R = host(f.R())
    Z = host(M)[:n, n:]
    Q, Rl = la.qr(A, mode="economic")
    Rp, d = _positive(Rl)
    assert np.all(np.diag(R) >= 0)
    kappa = np.linalg.cond(A) if n > 1 else 1.0
    tol = 200 * EPS * np.sqrt(m) * kappa
    assert la.norm(R - Rp) <= tol * la.norm(Rp)
    if extra:
        assert la.norm(Z - d[:, None] * (Q.T @ B)) <= tol * la.norm(B)
    assert not f.singular()


@pytest.mark.parametrize("m,n", [(800, 200), (4096, 513), (3000, 129)])
def test_qr_r_ill_conditioned_backward_error(ops, m, n):
    """Columns scaled over 1e-8 .. 1 (kappa ~ 1e8): R^T R = A^T A and
    R^T z = A^T b hold to a backward error of a few hundred roundoffs."""
    _say(f"qr_r ill-conditioned m={m} n={n}")
    rng = np.random.default_rng(n)
    U, _ = la.qr(rng.standard_normal((m, n)), mode="economic")
    V, _ = la.qr(rng.standard_normal((n, n)))
    A = (U * np.logspace(0, -8, n)) @ V.T
    b = rng.standard_normal(m)
    M = dev(np.hstack([A, b[:, None]]))
    f = ops.qr_r_caqr_(M, n)
    R = host(f.R())
    z = host(M)[:n, n]
    nA = la.norm(A, 2)
    assert la.norm(R.T @ R - A.T @ A) <= 100 * EPS * np.sqrt(m) * nA ** 2
    assert la.norm(R.T @ z - A.T @ b) <= 100 * EPS * np.sqrt(m) * nA * la.norm(b)


def test_qr_r_is_deterministic(ops):
    _say("qr_r determinism")
    rng = np.random.default_rng(5)
    A = rng.standard_normal((4096, 300))
    outs = []
    for _ in range(2):
        M = dev(A)
        ops.qr_r_caqr_(M, 300)
        outs.append(host(M)[:300, :300])
    assert np.array_equal(np.triu(outs[0]), np.triu(outs[1]))


def test_qr_r_flags_exact_zero_column(ops):
    A = np.random.default_rng(2).standard_normal((30, 6))
    A[:, 3] = A[:, 1]  # dependent column: R[3,3] is roundoff, not flagged
    f = ops.qr_r_caqr_(dev(A.copy()))
    assert abs(host(f.R())[3, 3]) < 1e-12 * la.norm(A)
    B = np.zeros((600, 4))
    B[:, 0] = 1.0
    f = ops.qr_r_caqr_(dev(B))
    assert f.singular()


@pytest.mark.parametrize("m", [700, 9000])
def test_qr_r_outside_the_kernel_rows_goes_through_geqrf(ops, m):
    """qr_r_ hands matrices of at most 1024 or more than 8192 rows to the
    blocked Householder QR (LAPACK's signs)."""
    _say(f"qr_r routing m={m}")
    rng = np.random.default_rng(8)
    A = rng.standard_normal((m, 20))
    b = rng.standard_normal(m)
    M = dev(np.hstack([A, b[:, None]]))
    f = ops.qr_r_(M, 20)
    Q, Rl = la.qr(A, mode="economic")
    assert f.householder is not None
    assert la.norm(host(f.R()) - Rl) <= 1e3 * EPS * np.sqrt(m) * la.norm(Rl)
    assert la.norm(host(M)[:20, 20] - Q.T @ b) <= 1e3 * EPS * np.sqrt(m) * la.norm(b)
    assert not f.singular()


def test_qr_r_routes_tall_sketches_to_the_kernel(ops):
    _say("qr_r routing m=2000")
    A = np.random.default_rng(9).standard_normal((2000, 50))
    f = ops.qr_r_(dev(A), 50)
    assert f.householder is None
    Rp, _ = _positive(la.qr(A, mode="r")[0][:50])
    assert la.norm(host(f.R()) - Rp) <= 1e3 * EPS * np.sqrt(2000) * la.norm(Rp)
