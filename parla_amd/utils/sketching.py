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
    """Row / column selection operator (utils/sketching.py:204-232)."""

    __array_ufunc__ = None

    def __init__(self, n_rows, n_cols, indices):
        self.shape = (n_rows, n_cols)
        self.wide = n_cols >= n_rows
        self.indices = indices  # host int64 array, sorted

    @property
    def T(self):
        return SamplingSketch(self.shape[1], self.shape[0], self.indices)

    def _idx(self, dev):
        return torch.as_tensor(self.indices, dtype=torch.int64, device=dev)

    def __matmul__(self, X):
        Xt, vec, ref = _prep(X)
        if self.wide:  # S @ X = X[indices]
            out = Xt.index_select(0, self._idx(Xt.device))
        else:  # S (tall) @ X scatters X into the selected rows
            out = torch.zeros((self.shape[0], Xt.shape[1]), dtype=Xt.dtype, device=Xt.device)
            out.index_copy_(0, self._idx(Xt.device), Xt)
        return _finish(out, vec, ref)

    def __rmatmul__(self, X):
        Xt, vec, ref = _prep(X)
        if vec:
            return _finish((self.T @ Xt), True, ref)
        if self.wide:  # X @ S scatters columns
            out = torch.zeros((Xt.shape[0], self.shape[1]), dtype=Xt.dtype, device=Xt.device)
            out.index_copy_(1, self._idx(Xt.device), Xt)
        else:  # X @ S_tall = X[:, indices]
            out = Xt.index_select(1, self._idx(Xt.device))
        return _finish(out, False, ref)


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
    ops_types = (DenseSketch, SJLTSketch, SamplingSketch)
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
    if isinstance(S, (SJLTSketch, SamplingSketch)):
        return ops.as_device(S.toarray())
    return ops.as_device(S)
