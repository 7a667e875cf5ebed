"""Synthetic test matrices with prescribed spectra (host-side data generation).

Restates PARLA's ``parla/tests/matmakers.py:7-41`` and the problem generators of
its least-squares / QB tests, plus the benchmark configurations.  Everything
here runs on the host with numpy and produces ndarrays; it is input data, not
part of any computation that the GPU library performs.

Haar factors follow ``orthonormal_operator`` (parla/utils/sketching.py:9-17):
QR of a Gaussian with the signs of diag(R) folded into Q.
"""
import numpy as np
import scipy.linalg as la


def haar(n_rows, n_cols, rng):
    """Haar-distributed orthonormal factor (utils/sketching.py:9-17)."""
    if n_rows < n_cols:
        return haar(n_cols, n_rows, rng).T
    rng = np.random.default_rng(rng)
    G = rng.normal(0.0, np.sqrt(1.0 / n_cols), (n_rows, n_cols))
    Q, R = la.qr(G, overwrite_a=True, mode="economic")
    return Q * np.sign(np.diag(R))


def rand_low_rank(n_rows, n_cols, spectrum, rng, factors=False):
    """tests/matmakers.py:7-21."""
    rng = np.random.default_rng(rng)
    if isinstance(spectrum, int):
        spectrum = rng.random(size=(spectrum,))
    spectrum = np.sort(spectrum)[::-1]
    spectrum = spectrum / spectrum[0]
    rank = spectrum.size
    U = haar(n_rows, rank, rng)
    V = haar(rank, n_cols, rng)
    M = (U * spectrum) @ V
    return (M, U, spectrum, V) if factors else M


def simple_mat(n_rows, n_cols, scale, rng):
    """tests/matmakers.py:24-31."""
    rng = np.random.default_rng(rng)
    A = rng.normal(0, 1, (n_rows, n_cols))
    QA, RA = la.qr(A)
    RA *= 1 / np.sqrt(1 + scale * np.arange(n_cols))
    return QA @ RA


def exponent_spectrum(n_rows, n_cols, rank, rng, spectrum_param, factors=False):
    """tests/matmakers.py:34-36."""
    spectrum = np.exp((-np.arange(1, rank) + 1) / spectrum_param)
    return rand_low_rank(n_rows, n_cols, spectrum, rng, factors)


def s_shaped_spectrum(n_rows, n_cols, rank, rng, factors=False):
    """tests/matmakers.py:39-41."""
    spectrum = 0.0001 + 1 / (1 + np.exp(np.arange(1, rank) - 29))
    return rand_low_rank(n_rows, n_cols, spectrum, rng, factors)


# ---------------------------------------------------------------------------
# Least-squares test problems (test_overdet_least_squares.py:11-110)
# Each returns (A, b, x_opt, U, s, Vt).
# ---------------------------------------------------------------------------

def consistent_tall():
    rng = np.random.default_rng(190489290)
    A = simple_mat(100, 10, 1, rng)
    U, s, Vt = la.svd(A)
    x = rng.standard_normal(10)
    return A, A @ x, x, U, s, Vt


def consistent_square():
    rng = np.random.default_rng(3278992245)
    n = 10
    U = haar(n, n, rng)
    s = rng.random(n) + 1e-4
    Vt = haar(n, n, rng)
    A = (U * s) @ Vt
    x = rng.standard_normal(n)
    return A, A @ x, x, U, s, Vt


def inconsistent_orthog():
    rng = np.random.default_rng(19837647834763)
    n, m = 1000, 100
    U = haar(n, m, rng)
    Vt = haar(m, m, rng)
    s = rng.random(m) + 1e-4
    A = (U * s) @ Vt
    b = rng.standard_normal(n)
    b = b - U @ (U.T @ b)
    b *= 1e2 / la.norm(b)
    return A, b, np.zeros(m), U, s, Vt


def inconsistent_gen():
    rng = np.random.default_rng(897809809)
    m, n, num_hi = 1000, 100, 30
    hi = 1e5 * np.ones(num_hi) + rng.random(num_hi)
    lo = np.ones(n - num_hi) + rng.random(n - num_hi)
    spec = np.concatenate([hi, lo])
    U = haar(m, n, rng)
    Vt = haar(n, n, rng)
    A = (U * spec) @ Vt
    x = np.concatenate([rng.standard_normal(num_hi) / 1e5,
                        rng.standard_normal(n - num_hi)])
    b_orth = rng.standard_normal(m) * 1e2
    b_orth -= U @ (U.T @ b_orth)
    return A, A @ x + b_orth, x, U, spec, Vt


def inconsistent_stackid():
    rng = np.random.default_rng(2837592038243)
    A = np.tile(np.eye(70), (10, 1))
    m, n = A.shape
    A = (A.T * rng.lognormal(size=(m,))).T
    x0 = rng.standard_normal(size=(n,))
    b = A @ x0 + rng.standard_normal(size=(m,))
    U, spec, Vt = la.svd(A, full_matrices=False)
    x = Vt.T @ (U.T @ b) / spec
    return A, b, x, U, spec, Vt


# ---------------------------------------------------------------------------
# QB test problems (test_qb.py:12-73).  Each returns (A, U, s, Vt).
# ---------------------------------------------------------------------------

_QB_SEED = 89374539423


def tall_low_exact_rank():
    return rand_low_rank(200, 50, 15, np.random.default_rng(_QB_SEED), factors=True)


def wide_low_exact_rank():
    return rand_low_rank(50, 200, 15, np.random.default_rng(_QB_SEED), factors=True)


def tall_full_exact_rank():
    return rand_low_rank(200, 50, 50, np.random.default_rng(_QB_SEED), factors=True)


def wide_full_exact_rank():
    return rand_low_rank(50, 200, 50, np.random.default_rng(_QB_SEED), factors=True)


def fast_decay_low_exact_rank():
    return exponent_spectrum(50, 200, 15, np.random.default_rng(_QB_SEED), 2, factors=True)


def slow_decay_low_exact_rank():
    return exponent_spectrum(50, 200, 15, np.random.default_rng(_QB_SEED), 150, factors=True)


def s_shaped_decay_low_exact_rank():
    return s_shaped_spectrum(50, 200, 45, np.random.default_rng(_QB_SEED), factors=True)


# ---------------------------------------------------------------------------
# Benchmark workloads.  Data seed 20260000 + c, sketch seed 20261000 + c.
# ---------------------------------------------------------------------------

def data_seed(c):
    return 20260000 + c


def sketch_seed(c):
    return 20261000 + c


def ls_problem(m, n, cond, noise, rng, spectrum="geom"):
    """A = U diag(s) V^T with Haar U, V; b = A x* + noise * N(0, 1).

    ``spectrum='geom'``: s geometric from 1 to 1/cond.  Returns (A, b, x_star).
    """
    rng = np.random.default_rng(rng)
    U = haar(m, n, rng)
    V = haar(n, n, rng)
    s = np.geomspace(1.0, 1.0 / cond, n)
    A = (U * s) @ V.T
    x_star = rng.standard_normal(n)
    b = A @ x_star + noise * rng.standard_normal(m)
    return A, b, x_star


def config_ls_plain(m=20000, n=200):
    """Plain FP64 sketch-and-precondition LS: A 20000x200, s from 1 to 1e-6."""
    return ls_problem(m, n, 1e6, 1e-8, np.random.default_rng(data_seed(1)))


def config_ls_mixed(m=262144, n=1024):
    """Mixed-precision LS: A 262144x1024, s log-spaced from 1 to 1e-8 (cond 1e8)."""
    return ls_problem(m, n, 1e8, 1e-8, np.random.default_rng(data_seed(3)))


def config_qb(m=8192, n=4096, rank_hi=512):
    """QB -> SVD input: A = U diag(s) V^T + 1e-8 N(0,1), s_i = i^-2 (i <= 512), else 1e-6."""
    rng = np.random.default_rng(data_seed(2))
    k = min(m, n)
    U = haar(m, k, rng)
    V = haar(n, k, rng)
    i = np.arange(1, k + 1, dtype=np.float64)
    s = np.where(i <= rank_hi, i ** -2.0, 1e-6)
    A = (U * s) @ V.T + 1e-8 * rng.standard_normal((m, n))
    return A, s
