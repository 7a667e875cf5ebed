"""A matrix held as BF16 pieces for the products of the power iteration.

No counterpart in the reference, whose RS1 / RF1 / QB1 multiply in the dtype
of A (comps/sketchers/aware.py:175,179; comps/rangefinders.py:186;
comps/qb.py:351).  ``SplitOperator(A, npieces)`` keeps FP64 A on the device
together with its BF16 pieces in both orientations, and the drivers'
products with it -- A @ X, A.T @ X and X @ A -- run on the split GEMM
(``pamd_split_gemm_f64``): npieces = 2 carries 16 significant bits of every
element and forms three piece products, npieces = 3 carries 24 bits (the
FP32 significand) with six.  Everything else stays FP64: the sparse first
sketch (applied to FP64 A), the orthonormalisations, B's small SVD.

Use: ``QB1(RF1(RS1(...)))(SplitOperator(A, 2), k, tol, rng)``; the drivers
accept it wherever they accept A.
"""
import torch

from parla_amd import ops


class SplitOperator:
    device_operator = True

    def __init__(self, A, npieces=2, _parts=None, _transposed=False):
        if npieces not in (2, 3):
            raise ValueError("npieces must be 2 or 3")
        if _parts is None:
            Ad = ops.as_device(A)
            _parts = (Ad, ops.split_bf16(Ad, npieces), ops.split_bf16(Ad, npieces, transpose=True))
        self._parts = _parts
        self.npieces = npieces
        self._transposed = _transposed
        m, n = _parts[0].shape
        self.shape = (n, m) if _transposed else (m, n)
        self.dtype = torch.float64
        self.device = _parts[0].device

    def t(self):
        return SplitOperator(None, self.npieces, self._parts, not self._transposed)

    @property
    def T(self):
        return self.t()

    def dense(self):
        """The FP64 matrix (a view)."""
        A = self._parts[0]
        return A.t() if self._transposed else A

    def _left(self):
        """Pieces of this operator as the K-contiguous left factor (M = rows)."""
        return self._parts[2] if self._transposed else self._parts[1]

    def _right_of(self, X, K, N, transpose):
        kp, _, npad, _ = ops.split_layout(self.npieces, 1, N, K)
        return ops.split_bf16(X, self.npieces, transpose=transpose, rows_out=npad, kp=kp)

    def __matmul__(self, X):
        """self @ X for a device matrix or vector X (shape[1] rows), FP64 result."""
        Xt = ops.as_device(X)
        vec = Xt.dim() == 1
        if vec:
            Xt = Xt.reshape(-1, 1)
        L = self._left()
        if Xt.shape[0] != L.K:
            raise ValueError("shape mismatch")
        R = self._right_of(Xt, L.K, Xt.shape[1], transpose=True)  # pieces of X^T
        C = ops.split_gemm(L, R)
        return C.reshape(-1) if vec else C

    def __rmatmul__(self, X):
        """X @ self for a device matrix X (shape[0] columns): (self^T X^T)^T."""
        Xt = ops.as_device(X)
        L = self.t()._left()
        if Xt.shape[1] != L.K:
            raise ValueError("shape mismatch")
        if Xt.stride(0) == 1 and Xt.stride(1) != 1:  # a transposed view: split its row-major base
            R = self._right_of(Xt.t(), L.K, Xt.shape[0], transpose=True)
        else:
            R = self._right_of(Xt, L.K, Xt.shape[0], transpose=False)
        return ops.split_gemm(L, R, transpose_out=True)
