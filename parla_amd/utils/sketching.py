"""Sketching operators on the device (drop-in for parla/utils/sketching.py).

Each generator returns an operator object that lives on the GPU and supports
the products PARLA's consumers use -- ``S @ X``, ``X @ S`` and ``S.T``
(drivers/least_squares.py:303,314; comps/sketchers/aware.py:166,175,179;
comps/rangefinders.py:186) -- with ``X`` a numpy array or a CUDA tensor.  The
result has the kind of ``X`` (numpy in, numpy out; CUDA tensor in, CUDA tensor
out).  All products run in libparla_amd.so.

Random draws.  ``backend='numpy'`` (the default) draws through the numpy
``Generator`` in exactly the reference's call order, so indices, signs and
Gaussian values are bit-identical to PARLA's for the same seed; the draws are
then uploaded.  ``backend='device'`` takes one 64-bit seed from the Generator
and draws on the GPU with Philox-4x32-10 (same distributions, different
stream) -- the form used when generation time matters.
"""
import warnings

import numpy as np
import torch

from .. import ops

DEFAULT_BACKEND = "numpy"


def _seed_from(rng):
    return int(rng.integers(0, 2 ** 63 - 1, dtype=np.int64))


def _prep(X):
    """(2-D float64 CUDA tensor, was_vector, ref-for-output-kind)."""
    t = ops.as_device(X)
    if t.dim() == 1:
        return t.reshape(-1, 1), True, X
    return t, False, X


def _finish(out, was_vec, ref):
    if was_vec:
        out = out.reshape(-1)
    return ops.like_input(out, ref)


class DenseSketch:
    """A dense operator S held on the device."""

    __array_ufunc__ = None  # make numpy defer X @ S to __rmatmul__

    def __init__(self, S):
        self.S = S
        self.shape = tuple(S.shape)

    @property
    def T(self):
        return DenseSketch(self.S.t())

    @property
    def tensor(self):
        return self.S

    def toarray(self):
        return self.S.detach().cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.toarray()
        return a.astype(dtype) if dtype is not None else a

    def __matmul__(self, X):
        if isinstance(X, (DenseSketch, SJLTSketch)):
            return NotImplemented
        Xt, vec, ref = _prep(X)
        if vec:
            out = ops.gemv(self.S, Xt.reshape(-1)).reshape(-1, 1)
        else:
            out = ops.gemm(self.S, Xt)
        return _finish(out, vec, ref)

    def __rmatmul__(self, X):
        Xt, vec, ref = _prep(X)
        if vec:  # x @ S == S^T x
            out = ops.gemv(self.S, Xt.reshape(-1), trans=True).reshape(-1, 1)
        else:
            out = ops.gemm(Xt, self.S)
        return _finish(out, vec, ref)


class SJLTSketch:
    """Sparse sign operator with k nonzeros per column of the wide draw W.

    W is n_short x n_long; column j holds rows ``rows[j, :]`` with values
    ``vals[j, :]`` (the CSC layout of utils/sketching.py:63-74).  The operator
    is W itself (wide) or W^T (tall, utils/sketching.py:75-79).
    """

    __array_ufunc__ = None

    def __init__(self, n_short, n_long, k, rows, vals, transposed=False, _shared=None):
        self.n_short, self.n_long, self.k = int(n_short), int(n_long), int(k)
        self.rows, self.vals = rows, vals
        self.transposed = transposed
        self.shape = (self.n_long, self.n_short) if transposed else (self.n_short, self.n_long)
        self._cache = _shared if _shared is not None else {}

    @property
    def T(self):
        return SJLTSketch(self.n_short, self.n_long, self.k, self.rows, self.vals,
                          not self.transposed, self._cache)

    def csr_w(self):
        """CSR of W (built on the device once, deterministic order)."""
        if "w" not in self._cache:
            self._cache["w"] = ops.sjlt_csr(self.n_short, self.n_long, self.k, self.rows, self.vals)
        return self._cache["w"]

    def csr_wt(self):
        """CSR of W^T: row j of W^T is column j of W, i.e. the draw itself."""
        if "wt" not in self._cache:
            dev = self.rows.device
            rowptr = torch.arange(0, self.n_long * self.k + 1, self.k, dtype=torch.int64, device=dev)
            self._cache["wt"] = (rowptr, self.rows.reshape(-1), self.vals.reshape(-1))
        return self._cache["wt"]

    def toarray(self):
        rows = self.rows.cpu().numpy()
        vals = self.vals.cpu().numpy()
        W = np.zeros((self.n_short, self.n_long))
        for l in range(self.k):
            np.add.at(W, (rows[:, l], np.arange(self.n_long)), vals[:, l])
        return W.T.copy() if self.transposed else W

    def __array__(self, dtype=None, copy=None):
        a = self.toarray()
        return a.astype(dtype) if dtype is not None else a

    # S @ X
    def __matmul__(self, X):
        if isinstance(X, (DenseSketch, SJLTSketch)):
            return NotImplemented
        Xt, vec, ref = _prep(X)
        if Xt.shape[0] != self.shape[1]:
            raise ValueError(f"shape mismatch {self.shape} @ {tuple(Xt.shape)}")
        if not self.transposed:
            rp, cl, vl = self.csr_w()
            out = ops.sjlt_left(rp, cl, vl, self.n_short, Xt)
        else:
            rp, cl, vl = self.csr_wt()
            out = ops.sjlt_left(rp, cl, vl, self.n_long, Xt)
        return _finish(out, vec, ref)

    # X @ S
    def __rmatmul__(self, X):
        Xt, vec, ref = _prep(X)
        if vec:  # x @ S == (S^T x)
            out = (self.T @ Xt)
            return _finish(out, True, ref)
        if Xt.shape[1] != self.shape[0]:
            raise ValueError(f"shape mismatch {tuple(Xt.shape)} @ {self.shape}")
        base, trans, _ = ops.mat_desc(Xt)
        if self.transposed:
            if trans:  # A^T @ W^T == (W @ A)^T
                rp, cl, vl = self.csr_w()
                out = ops.transpose(ops.sjlt_left(rp, cl, vl, self.n_short, base))
            elif self.n_short <= 1024:
                rp, cl, vl = self.csr_w()
                out = ops.sjlt_right(rp, cl, vl, self.n_short, Xt)
            else:  # (X W^T)^T = W X^T
                rp, cl, vl = self.csr_w()
                out = ops.transpose(ops.sjlt_left(rp, cl, vl, self.n_short, ops.transpose(Xt)))
        else:
            # X @ W == (W^T X^T)^T
            rp, cl, vl = self.csr_wt()
            XT = base if trans else ops.transpose(Xt)
            out = ops.transpose(ops.sjlt_left(rp, cl, vl, self.n_long, XT))
        return _finish(out, False, ref)


class SamplingSketch:
    """Row / column selection operator (utils/sketching.py:204-232).

    Wide (n_cols >= n_rows): S @ X = X[indices] and its adjoint S.T @ Y puts
    the rows of Y at `indices` (zeros elsewhere).  The tall operator is the
    transpose of the wide one with the same indices -- the reference's tall
    branch hands `indices` to the generator as a seed (sketching.py:230) and
    its rmatmat builds no result (:222-224); both are the evident intent here.
    Every product is the library's row gather or row scatter (on the
    transposed operand for products from the right).
    """

    __array_ufunc__ = None

    def __init__(self, n_rows, n_cols, indices, _dev_idx=None):
        self.shape = (n_rows, n_cols)
        self.wide = n_cols >= n_rows
        self.indices = indices  # host int64 array
        self._dev = _dev_idx if _dev_idx is not None else {}

    @property
    def T(self):
        return SamplingSketch(self.shape[1], self.shape[0], self.indices, self._dev)

    def _idx(self, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = torch.as_tensor(np.asarray(self.indices, dtype=np.int64), device=dev)
        return self._dev[key]

    def _select(self, Xt):   # X[indices] (rows)
        return ops.row_gather(Xt, self._idx(Xt.device))

    def _spread(self, Xt, rows):   # rows of X placed at indices of a zero matrix
        return ops.row_scatter(Xt, self._idx(Xt.device), rows)

    def __matmul__(self, X):
        if isinstance(X, (DenseSketch, SJLTSketch, SamplingSketch)):
            return NotImplemented
        Xt, vec, ref = _prep(X)
        if Xt.shape[0] != self.shape[1]:
            raise ValueError(f"shape mismatch {self.shape} @ {tuple(Xt.shape)}")
        out = self._select(Xt) if self.wide else self._spread(Xt, self.shape[0])
        return _finish(out, vec, ref)

    def __rmatmul__(self, X):
        Xt, vec, ref = _prep(X)
        if vec:
            return _finish((self.T @ Xt), True, ref)
        if Xt.shape[1] != self.shape[0]:
            raise ValueError(f"shape mismatch {tuple(Xt.shape)} @ {self.shape}")
        # X @ S = (S^T X^T)^T
        base, trans, _ = ops.mat_desc(Xt)
        XT = base if trans else ops.transpose(Xt)
        if self.wide:   # S^T scatters
            out = self._spread(XT, self.shape[1])
        else:           # S^T (wide) selects
            out = self._select(XT)
        return _finish(ops.transpose(out), False, ref)

    def toarray(self):
        d = min(self.shape)
        W = np.zeros((d, max(self.shape)))
        W[np.arange(d), np.asarray(self.indices)] = 1.0
        return W if self.wide else W.T.copy()

    def __array__(self, dtype=None, copy=None):
        a = self.toarray()
        return a.astype(dtype) if dtype is not None else a


class SparseSketch:
    """A general sparse operator (the sparse sign operator of
    utils/sketching.py:83-103) held on the device as the CSR of S and of S^T;
    S @ X and X @ S = (S^T X^T)^T run on the library's CSR product."""

    __array_ufunc__ = None

    def __init__(self, csr_s, csr_st, shape):
        self._s, self._st = csr_s, csr_st
        self.shape = tuple(shape)

    @classmethod
    def from_scipy(cls, S, device):
        import scipy.sparse as sp

        def up(M):
            M = sp.csr_matrix(M)
            M.sort_indices()
            return (torch.from_numpy(M.indptr.astype(np.int64)).to(device),
                    torch.from_numpy(M.indices.astype(np.int32)).to(device),
                    torch.from_numpy(M.data.astype(np.float64)).to(device))
        return cls(up(S), up(S.T), S.shape)

    @property
    def T(self):
        return SparseSketch(self._st, self._s, (self.shape[1], self.shape[0]))

    def toarray(self):
        rp, cl, vl = (t.cpu().numpy() for t in self._s)
        W = np.zeros(self.shape)
        for i in range(self.shape[0]):
            W[i, cl[rp[i]:rp[i + 1]]] += vl[rp[i]:rp[i + 1]]
        return W

    def __array__(self, dtype=None, copy=None):
        a = self.toarray()
        return a.astype(dtype) if dtype is not None else a

    def __matmul__(self, X):
        if isinstance(X, (DenseSketch, SJLTSketch, SamplingSketch, SparseSketch)):
            return NotImplemented
        Xt, vec, ref = _prep(X)
        if Xt.shape[0] != self.shape[1]:
            raise ValueError(f"shape mismatch {self.shape} @ {tuple(Xt.shape)}")
        rp, cl, vl = self._s
        return _finish(ops.sjlt_left(rp, cl, vl, self.shape[0], Xt), vec, ref)

    def __rmatmul__(self, X):
        Xt, vec, ref = _prep(X)
        if vec:
            return _finish(self.T @ Xt, True, ref)
        if Xt.shape[1] != self.shape[0]:
            raise ValueError(f"shape mismatch {tuple(Xt.shape)} @ {self.shape}")
        base, trans, _ = ops.mat_desc(Xt)
        XT = base if trans else ops.transpose(Xt)
        rp, cl, vl = self._st
        return _finish(ops.transpose(ops.sjlt_left(rp, cl, vl, self.shape[1], XT)), False, ref)


# ---------------------------------------------------------------------------
# generators (parla/utils/sketching.py)
# ---------------------------------------------------------------------------

def gaussian_operator(n_rows, n_cols, rng, normalize=True, backend=None, device=None):
    """utils/sketching.py:20-31 -> DenseSketch."""
    backend = backend or DEFAULT_BACKEND
    rng = np.random.default_rng(rng)
    dev = device or ops.default_device()
    scale = np.sqrt(1.0 / min(n_rows, n_cols)) if normalize else 1.0
    if backend == "device":
        return DenseSketch(ops.gaussian_fill(n_rows, n_cols, scale, _seed_from(rng), dev))
    if normalize:
        S = rng.normal(0.0, scale, (n_rows, n_cols))
    else:
        S = rng.standard_normal((n_rows, n_cols))
    return DenseSketch(torch.from_numpy(S).to(dev))


def _sjlt_draw_numpy(short, long, rng, vec_nnz):
    """The reference's draw order (utils/sketching.py:51-71)."""
    k = min(long, vec_nnz)
    with_repl = short < k
    if with_repl:
        warnings.warn(f"Can't set {k} nonzeros per column for columns of length {short}. "
                      "Sampling indices with replacement instead.")
    rows = np.empty((long, k), dtype=np.int32)
    for j in range(long):
        rows[j] = rng.choice(short, k, replace=with_repl)
    vals = np.ones(long * k)
    vals[rng.random(long * k) <= 0.5] = -1.0
    vals /= np.sqrt(k)
    return rows, vals.reshape(long, k), k


def sjlt_operator(n_rows, n_cols, rng, vec_nnz=8, backend=None, device=None):
    """utils/sketching.py:34-80 -> SJLTSketch (wide: W, tall: W^T of the wide draw)."""
    backend = backend or DEFAULT_BACKEND
    rng = np.random.default_rng(rng)
    dev = device or ops.default_device()
    transposed = n_cols < n_rows
    short, long = (n_cols, n_rows) if transposed else (n_rows, n_cols)
    if backend == "device":
        k = min(long, vec_nnz)
        if short < k:
            warnings.warn(f"Can't set {k} nonzeros per column for columns of length {short}. "
                          "Sampling indices with replacement instead.")
        rows, vals = ops.sjlt_draw(short, long, k, _seed_from(rng), dev)
    else:
        r, v, k = _sjlt_draw_numpy(short, long, rng, vec_nnz)
        rows = torch.from_numpy(r).to(dev)
        vals = torch.from_numpy(v).to(dev)
    return SJLTSketch(short, long, k, rows, vals, transposed)


def orthonormal_operator(n_rows, n_cols, rng, device=None):
    """utils/sketching.py:9-17: Haar factor, QR of a Gaussian with diag(R) > 0."""
    if n_rows < n_cols:
        return orthonormal_operator(n_cols, n_rows, rng, device).T
    rng = np.random.default_rng(rng)
    G = gaussian_operator(n_rows, n_cols, rng, device=device).S
    f = ops.geqrf_(G.contiguous().clone())
    Q = f.Q()
    signs = torch.sign(torch.diagonal(f.R_view))
    D = torch.diag(signs)
    return DenseSketch(ops.gemm(Q, D))


def sparse_sign_operator(n_rows, n_cols, rng, density=0.05, device=None):
    """utils/sketching.py:83-103: each entry nonzero with probability
    `density` (redrawn up to ten times if all are zero), values
    +-1/sqrt(min(r, c) density); the draws are the reference's numpy calls."""
    import scipy.sparse as sp
    rng = np.random.default_rng(rng)
    nz = rng.random(n_rows * n_cols) < density
    attempt = 0
    while np.all(~nz):
        if attempt == 10:
            raise RuntimeError('Density too low.')
        nz = rng.random(n_rows * n_cols) < density
        attempt += 1
    rows, cols = np.unravel_index(np.where(nz)[0], (n_rows, n_cols))
    vals = np.ones(rows.size)
    vals[rng.random(vals.size) < 0.5] = -1
    vals /= np.sqrt(min(n_rows, n_cols) * density)
    S = sp.coo_matrix((vals, (rows, cols)), shape=(n_rows, n_cols)).tocsr()
    return SparseSketch.from_scipy(S, device or ops.default_device())


def generate_srct(n_rows, n_cols, rng):
    """utils/sketching.py:106-116: restriction r, signs e (scaled by
    sqrt(big / small)) and permutation perm, drawn in the reference's order."""
    big, small = max(n_rows, n_cols), min(n_rows, n_cols)
    r = rng.choice(big, size=small, replace=False)
    e = rng.random(big)
    e[e > 0.5] = 1.0
    e[e != 1] = -1.0
    e *= np.sqrt(big / small)
    perm = rng.permutation(big)
    return r, e, perm


def _srct_matrix(r, e, perm, dev):
    N, d = e.size, r.size
    pinv = np.argsort(perm)
    rt = torch.from_numpy(np.ascontiguousarray(r, dtype=np.int64)).to(dev)
    et = torch.from_numpy(np.ascontiguousarray(e, dtype=np.float64)).to(dev)
    pt = torch.from_numpy(np.ascontiguousarray(pinv, dtype=np.int64)).to(dev)
    S = torch.empty((d, N), dtype=ops.F64, device=dev)
    ops.check(ops.lib().pamd_srct_fill_f64(d, N, ops._p(rt), ops._p(et), ops._p(pt), ops._p(S), max(N, 1),
                                           ops.stream(dev)), "pamd_srct_fill_f64")
    return S


def apply_srct(r, e, mat, perm=None, forward=True):
    """utils/sketching.py:119-176: the SRCT (r, e, perm) or its adjoint applied
    to the columns of mat, as the device product with the operator's rows."""
    Xt = ops.as_device(mat)
    perm = np.arange(np.asarray(e).size) if perm is None else perm
    S = DenseSketch(_srct_matrix(np.asarray(r), np.asarray(e), np.asarray(perm), Xt.device))
    out = (S @ Xt) if forward else (S.T @ Xt)
    return ops.like_input(out, mat)


def srct_operator(n_rows, n_cols, rng, device=None):
    """utils/sketching.py:179-201: the SRCT as a device operator.  Its d rows
    of the orthonormal DCT-II (composed with the signs and the permutation)
    are generated on the GPU, so S @ X, X @ S and S.T run on the FP64 MFMA
    GEMM; `sketch_data` holds (r, e, perm) as in the reference."""
    rng = np.random.default_rng(rng)
    r, e, perm = generate_srct(n_rows, n_cols, rng)
    if n_cols < n_rows:
        # the reference draws before it looks at the orientation; the tall
        # operator is the transpose of a second, wide draw (sketching.py:185-200)
        return srct_operator(n_cols, n_rows, rng, device).T
    S = DenseSketch(_srct_matrix(r, e, perm, device or ops.default_device()))
    S.sketch_data = (r, e, perm)
    return S


def sampling_operator(n_rows, n_cols, rng, indices=None):
    """utils/sketching.py:204-232 (wide selection; the tall form is its transpose)."""
    rng = np.random.default_rng(rng)
    if indices is None:
        indices = rng.choice(max(n_rows, n_cols), min(n_rows, n_cols), replace=False)
        indices.sort()
    else:
        indices = np.asarray(indices)
        assert indices.size == min(n_rows, n_cols)
    return SamplingSketch(n_rows, n_cols, indices)


def apply(X, Y):
    """X @ Y for device tensors / sketch operators, always on the device.

    Used by the drivers in place of the bare ``@`` so that a device tensor on
    the left of a sketch operator, or two device tensors, go through the
    library's kernels rather than torch's.
    """
    ops_types = (DenseSketch, SJLTSketch, SamplingSketch, SparseSketch)
    if getattr(X, "device_operator", False):  # split operator: sketches see its FP64 matrix
        return Y.__rmatmul__(X.dense()) if isinstance(Y, ops_types) else X @ Y
    if getattr(Y, "device_operator", False):
        return X.__matmul__(Y.dense()) if isinstance(X, ops_types) else Y.__rmatmul__(X)
    if isinstance(Y, ops_types):
        return Y.__rmatmul__(X)
    if isinstance(X, ops_types):
        return X.__matmul__(Y)
    Xt = ops.as_device(X)
    Yt = ops.as_device(Y)
    if Yt.dim() == 1:
        return ops.gemv(Xt, Yt)
    if Xt.dim() == 1:
        return ops.gemv(Yt, Xt, trans=True)
    return ops.gemm(Xt, Yt)


def as_tensor(S):
    """Dense device tensor for an operator or array (operators are densified)."""
    if isinstance(S, DenseSketch):
        return S.S
    if isinstance(S, (SJLTSketch, SamplingSketch, SparseSketch)):
        return ops.as_device(S.toarray())
    return ops.as_device(S)
