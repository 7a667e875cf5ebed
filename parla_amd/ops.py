"""Tensor-level wrappers over the C ABI (device memory and streams from torch).

Every function here takes / returns float64 CUDA (HIP) tensors, enqueues work
on torch's current stream and allocates its workspace through torch's caching
allocator.  torch is used for memory and streams only; all arithmetic happens
in libparla_amd.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check

F64 = torch.float64

# Optional per-launch timing: when a list is installed here, the wrappers
# below bracket their library call with HIP events on the launch stream and
# append (name, algorithmic flops, algorithmic bytes, start, end).
_event_log = None


def set_event_log(log):
    """Install (list) or remove (None) the per-launch event log; returns the old one."""
    global _event_log
    old, _event_log = _event_log, log
    return old


class _Timed:
    def __init__(self, name, flops=0.0, nbytes=0.0):
        self.rec = None
        if _event_log is not None:
            self.rec = (name, float(flops), float(nbytes), torch.cuda.Event(enable_timing=True),
                        torch.cuda.Event(enable_timing=True))

    def __enter__(self):
        if self.rec:
            self.rec[3].record()
        return self

    def __exit__(self, *exc):
        if self.rec:
            self.rec[4].record()
            _event_log.append(self.rec)
        return False


def lib():
    return _lib.load()


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("parla_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream(dev=None):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def as_device(X, dev=None):
    """float64 CUDA tensor view of X (numpy array, CPU or CUDA tensor); an
    operator that holds its matrix on the device in another form (the split
    operator) passes through unchanged."""
    if getattr(X, "device_operator", False):
        return X
    if isinstance(X, torch.Tensor):
        if X.device.type == "cuda":
            return X if X.dtype == F64 else X.to(F64)
        X = X.detach().numpy()
    arr = np.asarray(X, dtype=np.float64)
    t = torch.from_numpy(np.ascontiguousarray(arr))
    return t.to(dev or default_device())


def like_input(t, ref):
    """Return t in the kind of `ref`: CUDA tensor (or device operator) stays,
    anything else -> numpy."""
    if getattr(ref, "device_operator", False) or (isinstance(ref, torch.Tensor) and ref.device.type == "cuda"):
        return t
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return t


def workspace(nbytes, dev):
    return torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=dev)


def _query(fn, *args):
    out = ctypes.c_int64(0)
    check(fn(*args, ctypes.byref(out)), fn.__name__)
    return out.value


def mat_desc(X):
    """(base, trans, ld): X == base (trans 0) or X == base^T (trans 1), base row-major with ld."""
    if X.dim() != 2:
        raise ValueError("expected a matrix")
    r, c = X.shape
    if r == 0 or c == 0:
        return X.contiguous(), 0, max(c, 1)
    s0, s1 = X.stride()
    if s1 == 1 and (r == 1 or s0 >= c):
        return X, 0, (s0 if r > 1 else c)
    if s0 == 1 and (c == 1 or s1 >= r):
        return X.t(), 1, (s1 if c > 1 else r)
    return X.contiguous(), 0, c


# ---------------------------------------------------------------------------
# dense
# ---------------------------------------------------------------------------

def gemm(A, B, out=None, alpha=1.0, beta=0.0):
    """out = alpha * A @ B + beta * out for 2-D float64 CUDA tensors (views allowed)."""
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError(f"matmul shape mismatch {tuple(A.shape)} @ {tuple(B.shape)}")
    dev = A.device
    if out is None:
        out = torch.empty((M, N), dtype=F64, device=dev)
        beta = 0.0
    if out.stride(1) != 1 and N > 1:
        raise ValueError("gemm output must have unit column stride")
    At, ta, lda = mat_desc(A)
    Bt, tb, ldb = mat_desc(B)
    ldc = out.stride(0) if M > 1 else max(N, 1)
    ws_bytes = _query(lib().pamd_gemm_workspace, M, N, K)
    ws = workspace(ws_bytes, dev) if ws_bytes > 0 else None
    with _Timed("gemm_f64", 2.0 * M * N * K, 8.0 * (M * K + K * N + M * N)):
        check(lib().pamd_gemm_f64(ta, tb, M, N, K, alpha, _p(At), lda, _p(Bt), ldb, beta, _p(out),
                                  max(ldc, N, 1), _p(ws), ws_bytes, stream(dev)), "pamd_gemm_f64")
    return out


def gemv(A, x, trans=False, y=None, alpha=1.0, beta=0.0):
    """y = alpha * op(A) x + beta * y, A row-major (unit column stride)."""
    At, ta, lda = mat_desc(A)
    if ta:  # A is a transposed view of At: op flips
        trans = not trans
    m, n = At.shape
    dev = A.device
    ylen = n if trans else m
    if y is None:
        y = torch.empty(ylen, dtype=F64, device=dev)
        beta = 0.0
    x = x.contiguous()
    ws_bytes = _query(lib().pamd_gemv_workspace, int(trans), m, n)
    ws = workspace(ws_bytes, dev) if ws_bytes > 0 else None
    check(lib().pamd_gemv_f64(int(trans), m, n, alpha, _p(At), lda, _p(x), beta, _p(y), _p(ws),
                              ws_bytes, stream(dev)), "pamd_gemv_f64")
    return y


def nrm2(x):
    """||x||_2 as a 1-element device tensor."""
    x = x.contiguous()
    out = torch.empty(1, dtype=F64, device=x.device)
    n = x.numel()
    if n > (1 << 18):  # long vectors: one partial per workgroup
        ws = workspace(_query(lib().pamd_nrm2_workspace, n), x.device)
        check(lib().pamd_nrm2_large_f64(n, _p(x), _p(out), _p(ws), stream(x.device)), "pamd_nrm2_large_f64")
        return out
    check(lib().pamd_nrm2_f64(n, _p(x), _p(out), stream(x.device)), "pamd_nrm2_f64")
    return out


def trsv_(R, x, trans=False):
    """In place: x <- R^{-1} x (trans False) or R^{-T} x; R upper, unit column stride."""
    if R.stride(1) != 1 and R.shape[1] > 1:
        R = R.contiguous()
    n = R.shape[0]
    check(lib().pamd_trsv_upper_f64(int(trans), n, _p(R), max(R.stride(0), n), _p(x),
                                    stream(x.device)), "pamd_trsv_upper_f64")
    return x


def row_gather(A, idx, scale=None, out=None):
    """out[r, :] = scale[r] * A[idx[r], :] (scale None: 1); idx a device int64 tensor."""
    A = A if A.stride(-1) == 1 else A.contiguous()
    m, n = A.shape
    d = idx.numel()
    if out is None:
        out = torch.empty((d, n), dtype=F64, device=A.device)
    check(lib().pamd_row_gather_f64(d, m, n, _p(idx), _p(scale), _p(A), max(A.stride(0), n, 1), _p(out),
                                    max(out.stride(0), n, 1), stream(A.device)), "pamd_row_gather_f64")
    return out


def row_scatter(A, idx, rows, out=None):
    """out (rows x n, zero outside the selected rows) with out[idx[r], :] = A[r, :]."""
    A = A if A.stride(-1) == 1 else A.contiguous()
    d, n = A.shape
    if out is None:
        out = torch.zeros((rows, n), dtype=F64, device=A.device)
    check(lib().pamd_row_scatter_f64(d, rows, n, _p(idx), _p(A), max(A.stride(0), n, 1), _p(out),
                                     max(out.stride(0), n, 1), stream(A.device)), "pamd_row_scatter_f64")
    return out


def transpose(A):
    """Contiguous A^T."""
    A = A.contiguous()
    r, c = A.shape
    out = torch.empty((c, r), dtype=F64, device=A.device)
    check(lib().pamd_transpose_f64(r, c, _p(A), max(c, 1), _p(out), max(r, 1), stream(A.device)),
          "pamd_transpose_f64")
    return out


# ---------------------------------------------------------------------------
# QR
# ---------------------------------------------------------------------------

class QRFactors:
    """Result of geqrf: the factored matrix (R in its upper triangle) + reflectors."""

    def __init__(self, M, n, ws):
        self.M = M
        self.m = M.shape[0]
        self.n = n
        self.ncols = M.shape[1]
        self.ws = ws

    @property
    def R_view(self):
        """n x n view whose upper triangle is R (strictly lower part holds reflectors)."""
        return self.M[: self.n, : self.n]

    def R(self):
        return torch.triu(self.R_view)

    def Q(self):
        Q = torch.empty((self.m, self.n), dtype=F64, device=self.M.device)
        check(lib().pamd_orgqr_f64(self.m, self.n, self.ncols, _p(self.ws), _p(Q), max(self.n, 1),
                                   stream(Q.device)), "pamd_orgqr_f64")
        return Q

    def singular(self):
        flag = ctypes.c_void_p(0)
        check(lib().pamd_qr_singular_flag(_p(self.ws), self.m, self.n, self.ncols, ctypes.byref(flag)),
              "pamd_qr_singular_flag")
        base = self.ws.data_ptr()
        off = flag.value - base
        return bool(self.ws[off:off + 4].view(torch.int32).item())


def geqrf_(M, n=None):
    """Householder QR of the first n columns of M (m x ncols, contiguous), in place."""
    if not M.is_contiguous():
        raise ValueError("geqrf_ needs a contiguous matrix")
    m, ncols = M.shape
    n = ncols if n is None else n
    if n > m:
        raise ValueError("geqrf_ needs m >= n")
    dev = M.device
    ws = workspace(_query(lib().pamd_geqrf_workspace, m, n, ncols), dev)
    check(lib().pamd_geqrf_f64(m, n, ncols, _p(M), max(ncols, 1), _p(ws), stream(dev)), "pamd_geqrf_f64")
    return QRFactors(M, n, ws)


# pamd_qr_r_f64 takes 1024 < m <= 8192; at m <= 1024 the one-workgroup
# Householder panels of geqrf are faster than its two-level block/stack steps
QR_R_MIN_ROWS, QR_R_MAX_ROWS = 1025, 8192


class RFactor:
    """Result of qr_r_: R in the upper triangle of M[:n, :n] and the first n
    rows of Q^T times the appended columns in M[:n, n:].  R has a
    non-negative diagonal when pamd_qr_r_f64 factored M, LAPACK's signs when
    geqrf did (`householder` holds its factors then)."""

    def __init__(self, M, n, ws, householder):
        self.M = M
        self.m = M.shape[0]
        self.n = n
        self.ws = ws
        self.householder = householder

    @property
    def R_view(self):
        """n x n view whose upper triangle is R (the rest is scratch)."""
        return self.M[: self.n, : self.n]

    def R(self):
        return torch.triu(self.R_view)

    def singular(self):
        if self.householder is not None:
            return self.householder.singular()
        flag = ctypes.c_void_p(0)
        check(lib().pamd_qr_r_singular_flag(_p(self.ws), self.m, ctypes.byref(flag)), "pamd_qr_r_singular_flag")
        off = flag.value - self.ws.data_ptr()
        return bool(self.ws[off:off + 4].view(torch.int32).item())


def qr_r_(M, n=None):
    """R factor of the first n columns of M (m x ncols, contiguous), in place,
    with Q^T applied to columns n..; Q is not formed.  Communication-avoiding
    Householder QR (pamd_qr_r_f64) for QR_R_MIN_ROWS <= m <= QR_R_MAX_ROWS,
    blocked Householder (geqrf) otherwise; R's rows may differ from LAPACK's
    in sign."""
    if not M.is_contiguous():
        raise ValueError("qr_r_ needs a contiguous matrix")
    m, ncols = M.shape
    n = ncols if n is None else n
    if n > m:
        raise ValueError("qr_r_ needs m >= n")
    if not QR_R_MIN_ROWS <= m <= QR_R_MAX_ROWS:
        return RFactor(M, n, None, geqrf_(M, n))
    dev = M.device
    ws = workspace(_query(lib().pamd_qr_r_workspace, m, n, ncols), dev)
    check(lib().pamd_qr_r_f64(m, n, ncols, _p(M), max(ncols, 1), _p(ws), stream(dev)), "pamd_qr_r_f64")
    return RFactor(M, n, ws, None)


def qr_r_caqr_(M, n=None):
    """pamd_qr_r_f64 on any m <= QR_R_MAX_ROWS (what qr_r_ routes to for tall
    matrices); for tests and measurements of the kernel itself."""
    m, ncols = M.shape
    n = ncols if n is None else n
    ws = workspace(_query(lib().pamd_qr_r_workspace, m, n, ncols), M.device)
    check(lib().pamd_qr_r_f64(m, n, ncols, _p(M), max(ncols, 1), _p(ws), stream(M.device)), "pamd_qr_r_f64")
    return RFactor(M, n, ws, None)


def qr_q(Y):
    """Economic Q of Y (m x n): la.qr(Y, mode='economic')[0], Q is m x min(m, n)."""
    M = Y.clone(memory_format=torch.contiguous_format)
    return geqrf_(M, min(M.shape)).Q()


CHOLQR_MAX_K = 512
CHOLQR_MAX_COND = 1e7  # CholeskyQR2 is orthogonal to O(u) while kappa(Y) << u^{-1/2}


def potrf_(G, info):
    check(lib().pamd_potrf_upper_f64(G.shape[0], _p(G), max(G.stride(0), G.shape[0], 1), _p(info),
                                     stream(G.device)), "pamd_potrf_upper_f64")


def trtri(R):
    """R^{-1} (upper triangular, zeros below) for upper-triangular R."""
    n = R.shape[0]
    Ri = torch.empty((n, n), dtype=F64, device=R.device)
    check(lib().pamd_trtri_upper_f64(n, _p(R), max(R.stride(0), n, 1), _p(Ri), max(n, 1), stream(R.device)),
          "pamd_trtri_upper_f64")
    return Ri


def trsm_right_(B, R):
    """In place B <- B R^{-1} (R upper)."""
    m, n = B.shape
    check(lib().pamd_trsm_right_upper_f64(m, n, _p(R), max(R.stride(0), n, 1), _p(B), max(B.stride(0), n, 1),
                                          stream(B.device)), "pamd_trsm_right_upper_f64")


NEWTON_MAX_E = 1e-4  # second round by a Newton correction while ||Q1^T Q1 - I||_F stays below this


def trsm_mfma(Y, R, out=None):
    """Y R^{-1} (R upper, k <= 512; only R's upper triangle is read) by the
    MFMA strip solver; out may be Y itself."""
    m, k = Y.shape
    Yc = Y if Y.stride(1) == 1 else Y.contiguous()
    X = torch.empty((m, k), dtype=F64, device=Y.device) if out is None else out
    ws = workspace(_query(lib().pamd_trsm_mfma_workspace, k), Y.device)
    check(lib().pamd_trsm_right_upper_mfma_f64(m, k, _p(R), max(R.stride(0), k, 1), _p(Yc), max(Yc.stride(0), k, 1),
                                               _p(X), max(X.stride(0), k, 1), _p(ws), stream(Y.device)),
          "pamd_trsm_right_upper_mfma_f64")
    return X


def _chol_round(Q, info):
    G = gemm(Q.t(), Q)
    potrf_(G, info)
    return G, trsm_mfma(Q, G)  # Q R^{-1}


def small_combine(A, B=None, alpha=0.0, sigma=0.0, upper_half=False, out=None, nrm=None):
    """out = A + alpha B + sigma I, or its upper triangle with the diagonal
    halved (upper_half); nrm (a one-element device tensor) receives
    ||A + alpha B + sigma I||_F.  One launch (k <= 512)."""
    k = A.shape[0]
    if out is None:
        out = torch.empty((k, k), dtype=F64, device=A.device)
    check(lib().pamd_small_combine_f64(k, _p(A), A.stride(0), _p(B), B.stride(0) if B is not None else k,
                                       float(alpha), float(sigma), 1 if upper_half else 0, _p(out), out.stride(0),
                                       _p(nrm), stream(A.device)), "pamd_small_combine_f64")
    return out


def _newton_round(Q):
    """Second CholeskyQR round for a nearly orthonormal Q (Q^T Q = I + E):
    the Cholesky factor of I + E is I + X with X + X^T = E - X^T X, solved by
    two fixed-point steps X = H(E - X^T X) (H takes the upper triangle and
    halves the diagonal; the error after them is O(||E||^3)), and
    (I + X)^{-1} = I - X + X^2 + O(||X||^3).  Three k x k products and one
    m x k x k GEMM instead of potrf and trtri.  Returns (Q (I + X)^{-1}, ||E||_F)
    with ||E||_F as a one-element device tensor."""
    QtQ = gemm(Q.t(), Q)
    # E, ||E||_F and H(E) by the many-workgroup combine and nrm2: the one-workgroup combine that also forms the norm
    # moves k^2 doubles three times through one CU (the same values, 8 us slower at k = 266, 56 us at k = 512)
    E = small_combine(QtQ, sigma=-1.0)
    enorm = nrm2(E.reshape(-1))
    X = small_combine(E, upper_half=True)                                       # H(E)
    X = small_combine(QtQ, gemm(X.t(), X), alpha=-1.0, sigma=-1.0, upper_half=True)  # H(E - X^T X)
    M = small_combine(gemm(X, X), X, alpha=-1.0, sigma=1.0)                     # I - X + X^2
    return gemm(Q, M), enorm


def cholqr2(Y, rounds=2):
    """Q spanning range(Y) by CholeskyQR2 (two rounds of G = Q^T Q, R = chol(G),
    Q <- Q R^{-1}); None when Y is too ill-conditioned for it (a pivot fails or
    the first R's diagonal spread exceeds CHOLQR_MAX_COND).  The second round,
    whose Gram matrix is I + E with E of order kappa(Y)^2 u, is a Newton
    correction (_newton_round) while ||E||_F <= NEWTON_MAX_E, else a second
    potrf round.  rounds=1 is plain CholeskyQR: Q = Y R^{-1} spans range(Y)
    and equals the two-round Q times an upper-triangular factor within
    kappa(Y)^2 u of the identity."""
    m, k = Y.shape
    if k == 0 or k > CHOLQR_MAX_K or m < k:
        return None
    info = torch.empty(2, dtype=torch.int32, device=Y.device)  # potrf writes info[0] (0 or the failing pivot)
    G, Q1 = _chol_round(Y, info[0:1])
    if rounds == 1:
        Q, enorm = Q1, None
    else:
        Q, enorm = _newton_round(Q1)
    # one synchronisation: pivot, conditioning estimate, size of E
    stats = torch.empty(4, dtype=F64, device=Y.device)
    check(lib().pamd_cholqr_stats_f64(k, _p(G), G.stride(0), _p(info), _p(enorm), _p(stats), stream(Y.device)),
          "pamd_cholqr_stats_f64")
    dd = stats.cpu().numpy()
    if dd[3] != 0 or not np.all(np.isfinite(dd)) or dd[0] <= 0 or dd[1] / dd[0] > CHOLQR_MAX_COND:
        return None
    if rounds == 2 and dd[2] > NEWTON_MAX_E:
        _, Q = _chol_round(Q1, info[1:2])
        if int(info[1].item()) != 0:
            return None
    return Q


SKETCH_NEWTON_MAX_E = 1e-2  # sketch_rz_cholqr2: second round by Newton while ||Q1^T Q1 - I||_F is below this


def sketch_rz_cholqr2(M, n):
    """(R, z) of a sketch M = [S A | S b] (d x (n + 1)) by CholeskyQR2 on
    S A: R1 = chol((S A)^T S A), Q1 = S A R1^{-1}, R2 = chol(Q1^T Q1); then
    R = R2 R1 (positive diagonal) and z = R2^{-T} Q1^T S b = Q^T S b for
    Q = Q1 R2^{-1}, which is never formed.  These are a Householder QR's R
    and Q^T S b up to the signs of R's rows, to O(u kappa) for
    kappa(S A) << u^{-1/2}.  Q1^T Q1 = I + E with ||E|| of order
    kappa^2 u: while ||E||_F <= SKETCH_NEWTON_MAX_E, R2 = I + X comes from
    two fixed-point steps X = H(E - X^T X) (H: the upper triangle with the
    diagonal halved; the error after them is O(||E||^3)), else from potrf.
    None when S A is too ill-conditioned for it (a pivot fails or R1's
    diagonal spread exceeds CHOLQR_MAX_COND); the caller then factors M by
    Householder.  On sketches of a few hundred rows this is much shorter than
    a panel-by-panel Householder QR."""
    d, ncols = M.shape
    if n < 1 or n > CHOLQR_MAX_K or d < n or ncols != n + 1:
        return None
    SA, Sb = M[:, :n], M[:, n:]
    info = torch.zeros(2, dtype=torch.int32, device=M.device)
    G1, Q1 = _chol_round(SA, info[0:1])
    QtQ = gemm(Q1.t(), Q1)
    E = small_combine(QtQ, sigma=-1.0)  # as in _newton_round: E, ||E||_F, H(E)
    enorm = nrm2(E.reshape(-1))
    X = small_combine(E, upper_half=True)
    stats = torch.empty(4, dtype=F64, device=M.device)
    check(lib().pamd_cholqr_stats_f64(n, _p(G1), G1.stride(0), _p(info), _p(enorm), _p(stats), stream(M.device)),
          "pamd_cholqr_stats_f64")
    dd = stats.cpu().numpy()  # one read: R1's diagonal spread, ||E||_F, the first pivot
    if dd[3] != 0 or not np.all(np.isfinite(dd)) or dd[0] <= 0 or dd[1] / dd[0] > CHOLQR_MAX_COND:
        return None
    w = gemm(Q1.t(), Sb).reshape(n)
    if dd[2] <= SKETCH_NEWTON_MAX_E:
        X = small_combine(QtQ, gemm(X.t(), X), alpha=-1.0, sigma=-1.0, upper_half=True)  # H(E - X^T X)
        R2 = small_combine(X, sigma=1.0)  # I + X (upper triangular)
    else:
        potrf_(QtQ, info[1:2])
        if int(info[1].item()) != 0:
            return None
        R2 = torch.triu(QtQ)
    R = gemm(R2, torch.triu(G1))
    z = trsv_(R2, w, trans=True)
    return R, z


def orth_q(Y, rounds=2):
    """Orthonormal basis of range(Y) (m x min(m, n)).  CholeskyQR2 on MFMA when Y
    is tall and well conditioned, else Householder; the two agree up to the sign
    of each column.  rounds=1: one CholeskyQR round (a well-conditioned basis of
    range(Y), orthonormal to kappa(Y)^2 u)."""
    m, n = Y.shape
    if m >= n:
        Q = cholqr2(Y, rounds)
        if Q is not None:
            return Q
    return qr_q(Y)


_ROCSOLVER = None


def _rocsolver():
    """(librocsolver, handle) for the small dense SVD; loaded on first use."""
    global _ROCSOLVER
    if _ROCSOLVER is None:
        rb = ctypes.CDLL("librocblas.so")
        rs = ctypes.CDLL("librocsolver.so")
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        rb.rocblas_create_handle.restype = i32
        rb.rocblas_create_handle.argtypes = [ctypes.POINTER(vp)]
        rb.rocblas_set_stream.restype = i32
        rb.rocblas_set_stream.argtypes = [vp, vp]
        # rocblas_status rocsolver_dgesdd(handle, left_svect, right_svect, m, n, A, lda, S,
        #                                 U, ldu, V, ldv, info)
        rs.rocsolver_dgesdd.restype = i32
        rs.rocsolver_dgesdd.argtypes = [vp, i32, i32, i32, i32, vp, i32, vp, vp, i32, vp, i32, vp]
        h = ctypes.c_void_p()
        if rb.rocblas_create_handle(ctypes.byref(h)) != 0:
            raise RuntimeError("rocblas_create_handle failed")
        _ROCSOLVER = (rb, rs, h)
    return _ROCSOLVER


def svd_small(M):
    """Thin SVD M = U diag(s) Vh of a small square device matrix by rocSOLVER's
    divide-and-conquer dgesdd (13 ms at 266 x 266, against 35 ms for the
    Jacobi SVD torch.linalg.svd picks).  rocSOLVER is column-major, so the
    row-major buffer of M is M^T to it: M^T = U' S V'^T gives U = V', Vh = U'^T,
    which are the row-major views of its V^T and U outputs."""
    n = M.shape[0]
    if M.shape[1] != n:
        raise ValueError("svd_small needs a square matrix")
    rb, rs, h = _rocsolver()
    dev = M.device
    A = M.contiguous().clone()
    s = torch.empty(n, dtype=F64, device=dev)
    Ub = torch.empty((n, n), dtype=F64, device=dev)
    Vb = torch.empty((n, n), dtype=F64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    rb.rocblas_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    st = rs.rocsolver_dgesdd(h, 192, 192, n, n, _p(A), n, _p(s), _p(Ub), n, _p(Vb), n, _p(info))
    if st != 0:
        raise RuntimeError(f"rocsolver_dgesdd failed (status {st})")
    bad = int(info.item())  # scipy's la.svd raises LinAlgError when gesdd does not converge
    if bad != 0:
        raise np.linalg.LinAlgError(f"SVD did not converge (rocsolver_dgesdd info {bad})")
    return Vb, s, Ub


JACOBI_MAX = 1024          # largest dimension the LDS-resident Jacobi pair handles
JACOBI_SWEEPS = 40
# Sweeps enqueued first.  Every sweep is nb - 1 launches, and sweeps after
# convergence still cost a launch each (their rounds return at once): the
# preconditioned matrices of the drivers (R^T of a QR) converge in 8-10
# sweeps, so 12 are enqueued, and only a matrix that needs more is solved
# again from the start with JACOBI_SWEEPS.
JACOBI_SWEEPS_FIRST = 12


def svd_jacobi(X):
    """Thin SVD X = U diag(s) Vt of a device matrix X (m x n, n <= m <=
    1024) by the library's one-sided block Jacobi; s descending.  Raises
    numpy.linalg.LinAlgError when the sweeps do not converge (as la.svd)."""
    m, n = X.shape
    Xc = X if X.stride(-1) == 1 else X.contiguous()
    dev = X.device
    U = torch.empty((m, n), dtype=F64, device=dev)
    s = torch.empty(n, dtype=F64, device=dev)
    Vt = torch.empty((n, n), dtype=F64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = workspace(_query(lib().pamd_svd_jacobi_workspace, m, n), dev)
    with _Timed("svd_jacobi", 0.0, 8.0 * (m * n + n * n)):
        for sweeps in (JACOBI_SWEEPS_FIRST, JACOBI_SWEEPS):
            # the entry point writes info (0: a sweep within the limit rotated nothing)
            check(lib().pamd_svd_jacobi_f64(m, n, _p(Xc), max(Xc.stride(0), n, 1), _p(U), n, _p(s), _p(Vt), n,
                                            sweeps, _p(ws), _p(info), stream(dev)), "pamd_svd_jacobi_f64")
            if int(info.item()) == 0:
                return U, s, Vt
    raise np.linalg.LinAlgError(f"Jacobi SVD did not converge in {JACOBI_SWEEPS} sweeps")


def thin_svd(X):
    """X = U diag(s) Vh (thin: k = min(d, n) singular triplets) of a device
    matrix X (d x n), s descending.

    Preconditioned one-sided Jacobi (Drmac and Veselic): T, the taller
    orientation of X (m x k), is factored T = Q R by Householder QR, and the
    Jacobi SVD (svd_jacobi) runs on R^T, whose columns are far closer to
    orthogonal than T's.  R^T = Uj diag(s) Vtj gives T = (Q Vtj^T) diag(s) Uj^T.
    Beyond k = 1024 the k x k SVD of R is rocSOLVER's dgesdd (svd_small).
    Returns (U, s, Vh)."""
    d, n = X.shape
    T = X.contiguous().clone() if d >= n else transpose(X)
    f = geqrf_(T)
    R = torch.triu(f.R_view)
    k = R.shape[0]
    if k <= JACOBI_MAX:
        Uj, s, Vtj = svd_jacobi(transpose(R))
        UT, VhT = gemm(f.Q(), Vtj.t()), transpose(Uj)
    else:
        Ur, s, VhT = svd_small(R)
        UT, VhT = gemm(f.Q(), Ur), VhT.contiguous()
    if d >= n:
        return UT, s, VhT
    return transpose(VhT), s, transpose(UT)    # X = T^T = VhT^T diag(s) UT^T


def sym_eig(C):
    """Eigendecomposition of the symmetric device matrix C (lower triangle
    read, as scipy.linalg.eigh): (lamb, V), ordered by decreasing |lamb|.
    The eigenvalues of C are its singular values with the signs of
    u_j . v_j, so this is thin_svd (the preconditioned Jacobi SVD) plus one
    small kernel.  When an eigenvalue pair of nearly equal magnitude and
    opposite sign mixes the singular vectors, the decomposition is redone on
    C + ||C||_2 I, which is positive semidefinite (the eigenvalues then carry
    the absolute error u ||C||, as eigh's do)."""
    n = C.shape[0]
    Cc = C if C.stride(-1) == 1 else C.contiguous()
    dev = C.device
    lamb = torch.empty(n, dtype=F64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    Cs = torch.empty((n, n), dtype=F64, device=dev)

    def run(shift):
        check(lib().pamd_sym_shift_f64(n, _p(Cc), max(Cc.stride(0), n, 1), float(shift), _p(Cs), n, stream(dev)),
              "pamd_sym_shift_f64")
        U, s, Vh = thin_svd(Cs)
        info.zero_()
        check(lib().pamd_eig_from_svd_f64(n, _p(U), n, _p(Vh), n, _p(s), float(shift), _p(lamb), _p(info),
                                          stream(dev)), "pamd_eig_from_svd_f64")
        return Vh, s

    Vh, s = run(0.0)
    if int(info.item()) == 2:
        Vh, s = run(float(s[0].item()))
        order = torch.argsort(-lamb.abs(), stable=True)   # decreasing lambda now: by |lambda|
        return lamb[order].contiguous(), transpose(Vh)[:, order].contiguous()
    return lamb, transpose(Vh)


def svd_right_precond(A_ske):
    """comps/preconditioning.py:70-79 on the device: the thin SVD of the
    sketch, its numerical rank (sigma > sigma_1 * ncols * eps, one host read)
    and M = V / sigma (n x rank).  Returns (M, U, sigma, Vh) truncated to the
    rank."""
    U, s, Vh = thin_svd(A_ske)
    sh = s.cpu().numpy()
    eps = np.finfo(float).eps
    rank = int(np.count_nonzero(sh > sh[0] * A_ske.shape[1] * eps)) if sh.size else 0
    U, s, Vh = U[:, :rank], s[:rank].contiguous(), Vh[:rank, :]
    M = transpose(Vh)
    scale_cols(M, s, invert=True, out=M)
    return M, U, s, Vh


def scale_cols(A, s, invert=False, out=None):
    """out = A diag(s) (or A diag(s)^{-1}); A row-major (unit column stride)."""
    if A.stride(-1) != 1:
        A = A.contiguous()
    r, c = A.shape
    if out is None:
        out = torch.empty((r, c), dtype=F64, device=A.device)
    check(lib().pamd_scale_cols_f64(r, c, _p(A), max(A.stride(0), c, 1), _p(s.contiguous()), int(bool(invert)),
                                    _p(out), max(out.stride(0), c, 1), stream(A.device)), "pamd_scale_cols_f64")
    return out


POTRF_BLOCK = 256


def cholesky_upper(G):
    """R with R^T R = G (upper triangle of the result; scipy.linalg.cholesky
    semantics, only G's upper triangle is read).  One potrf launch up to
    n = 512; beyond, right-looking blocked over 256-column panels: potrf of
    the diagonal block, R12 from R12^T = G12^T R11^{-1} (the MFMA strip
    solve) and the trailing G22 -= R12^T R12 (MFMA GEMM).  Raises
    numpy.linalg.LinAlgError when G is not positive definite (as la.cholesky)."""
    n = G.shape[0]
    R = G.contiguous().clone()
    info = torch.zeros(1, dtype=torch.int32, device=G.device)
    if n <= CHOLQR_MAX_K:
        potrf_(R, info)
        bad = int(info.item())
        if bad:
            raise np.linalg.LinAlgError(f"{bad}-th leading minor not positive definite")
        return R
    for j0 in range(0, n, POTRF_BLOCK):
        j1 = min(n, j0 + POTRF_BLOCK)
        potrf_(R[j0:j1, j0:j1], info)
        bad = int(info.item())
        if bad:
            raise np.linalg.LinAlgError(f"{j0 + bad}-th leading minor not positive definite")
        if j1 < n:
            X = trsm_mfma(R[j0:j1, j1:].t(), R[j0:j1, j0:j1])   # R12^T = G12^T R11^{-1}
            R[j0:j1, j1:] = transpose(X)
            gemm(X, R[j0:j1, j1:], out=R[j1:, j1:], alpha=-1.0, beta=1.0)
    return R


def lu_pl(A):
    """P @ L of the partially pivoted LU of A (scipy.linalg.lu(A, permute_l=True)[0])."""
    At = A if A.stride(-1) == 1 else A.contiguous()
    m, n = At.shape
    k = min(m, n)
    dev = At.device
    PL = torch.empty((m, k), dtype=F64, device=dev)
    ws = workspace(_query(lib().pamd_lu_workspace, m, n), dev)
    check(lib().pamd_lu_pl_f64(m, n, _p(At), max(At.stride(0), n, 1), _p(PL), max(k, 1), _p(ws), stream(dev)),
          "pamd_lu_pl_f64")
    return PL


# ---------------------------------------------------------------------------
# sketches
# ---------------------------------------------------------------------------

def gaussian_fill(rows, cols, scale, seed, dev, offset=0):
    out = torch.empty((rows, cols), dtype=F64, device=dev)
    check(lib().pamd_gaussian_fill_f64(rows, cols, scale, seed & (2 ** 64 - 1), offset, _p(out),
                                       max(cols, 1), stream(dev)), "pamd_gaussian_fill_f64")
    return out


# ---------------------------------------------------------------------------
# mixed precision
# ---------------------------------------------------------------------------

OZ_CHUNK = 256
OZ_TN = 64


def _round_up(x, q):
    return ((x + q - 1) // q) * q


def ozaki_layout(d, m, ncols):
    """(kp, dpad, ncp, group_chunks, workspace bytes) of the exact sketch."""
    v = [ctypes.c_int64() for _ in range(5)]
    check(lib().pamd_ozaki_layout(d, m, ncols, *[ctypes.byref(x) for x in v]), "pamd_ozaki_layout")
    return tuple(x.value for x in v)


def gauss_q(rows, m, seed, dev, offset=0):
    """Quantised Gaussian Q = rint(32 z) for the Philox stream of
    gaussian_fill, as a blocked BF16 array (rows rounded up to 128 x kp, kp
    = m rounded up to 256, zeros beyond); `unblock_pieces(Q[None])` gives the
    row-major view."""
    kp = _round_up(max(m, 1), OZ_CHUNK)
    rows_out = _round_up(max(rows, 1), 128)
    Q = torch.empty((rows_out, kp), dtype=torch.uint16, device=dev)
    check(lib().pamd_gauss_q_bf16(rows, rows_out, m, seed & (2 ** 64 - 1), offset, _p(Q), stream(dev)),
          "pamd_gauss_q_bf16")
    return Q


def block_rows(X, rows_out, kp):
    """Blocked layout (blocked.h) of a row-major 2-D device tensor of 16-bit
    values, zero-padded to rows_out x kp: the device-side form of
    block_pieces for one array."""
    R0, K0 = X.shape
    Xp = torch.zeros((rows_out, kp), dtype=torch.int16, device=X.device)  # indexing has no uint16 kernel
    Xp[:R0, :K0] = X.view(torch.int16)
    b = Xp.view(rows_out // 16, 16, kp // 32, 4, 8).permute(0, 2, 1, 3, 4)
    r = torch.arange(16, device=X.device)[:, None]
    logical = torch.arange(4, device=X.device)[None, :] ^ (r >> 2)
    return b[:, :, r, logical, :].contiguous().view(rows_out, kp).view(X.dtype)


def quantize_q(Z, m=None):
    """Q = clip(rint(32 Z), -256, 256) in the blocked BF16 layout of gauss_q
    (the host draw of gauss_q: numpy normals -> the same integer grid)."""
    Z = as_device(Z)
    rows, mm = Z.shape
    kp = _round_up(max(mm, 1), OZ_CHUNK)
    q = (torch.clamp(torch.round(32.0 * Z), -256.0, 256.0) + 0.0).to(torch.bfloat16).view(torch.uint16)  # zero is +0
    return block_rows(q, _round_up(max(rows, 1), 128), kp)


class OzakiSplit:
    """[A | b] cut into `npieces` BF16 integer pieces per 256-row chunk
    (blocked arrays of ncp x kp, one per piece)."""

    def __init__(self, A, b=None, npieces=4):
        m, n = A.shape
        if A.stride(1) != 1:
            A = A.contiguous()
        self.m, self.n, self.npieces = m, n, npieces
        self.ncols = n + (1 if b is not None else 0)
        self.kp = _round_up(max(m, 1), OZ_CHUNK)
        self.ncp = _round_up(max(self.ncols, 1), OZ_TN)
        self.pstride = self.ncp * self.kp
        dev = A.device
        self.pieces = torch.empty((npieces, self.ncp, self.kp), dtype=torch.uint16, device=dev)
        self.exps = torch.empty((self.kp // OZ_CHUNK, self.ncols), dtype=torch.int32, device=dev)
        bb = None if b is None else b.contiguous()
        check(lib().pamd_ozaki_split_bf16(m, n, _p(A), max(A.stride(0), n, 1), _p(bb), npieces, _p(self.pieces),
                                          self.pstride, _p(self.exps), self.ncols, stream(dev)),
              "pamd_ozaki_split_bf16")

    def piece(self, p):
        """Piece p as a row-major numpy array (ncols x kp, BF16 values as float32)."""
        bits = unblock_pieces(self.pieces[p:p + 1])[0, :self.ncols]
        return (bits.astype(np.uint32) << 16).view(np.float32)


def ozaki_sketch(Q, d, split, scale, out=None):
    """out (d x ncols) = scale * Q[:d] @ [A | b] from the exact BF16 products
    (Q from gauss_q / quantize_q, d its logical rows)."""
    kp, dpad, ncp, _, wsb = ozaki_layout(d, split.m, split.ncols)
    if Q.shape[0] < dpad or Q.shape[1] != kp or split.kp != kp:
        raise ValueError("ozaki_sketch: Q does not match the layout")
    if out is None:
        out = torch.empty((d, split.ncols), dtype=F64, device=Q.device)
    ws = workspace(wsb, Q.device)
    with _Timed("ozaki_sketch", 2.0 * dpad * kp * ncp * split.npieces,
                2.0 * dpad * kp + 2.0 * split.npieces * kp * ncp + 8.0 * d * split.ncols):
        check(lib().pamd_ozaki_sketch_f64(d, split.m, split.ncols, split.npieces, _p(Q), _p(split.pieces),
                                          split.pstride, _p(split.exps), split.ncols, float(scale), _p(out),
                                          out.stride(0), _p(ws), stream(Q.device)),
              "pamd_ozaki_sketch_f64")
    return out


class NormalOp:
    """q = B^T (B p) for an FP32- or FP64-stored B, FP64 vectors."""

    def __init__(self, B):
        if B.dtype not in (torch.float32, F64) or B.stride(1) != 1:
            raise ValueError("NormalOp needs a row-major float32 or float64 matrix")
        self.B = B
        m, n = B.shape
        self.ws = workspace(_query(lib().pamd_normal_pass_workspace, m, n), B.device)

    def __call__(self, p, q=None):
        m, n = self.B.shape
        if q is None:
            q = torch.empty(n, dtype=F64, device=p.device)
        check(lib().pamd_normal_pass(1 if self.B.dtype == torch.float32 else 0, m, n, _p(self.B),
                                     self.B.stride(0), _p(p), _p(q), _p(self.ws), stream(p.device)),
              "pamd_normal_pass")
        return q


class Residual:
    """Compensated r = b - A x, g = A^T r and ||r||^2, one read of A."""

    def __init__(self, A):
        self.A = A if A.stride(1) == 1 else A.contiguous()
        m, n = A.shape
        self.ws = workspace(_query(lib().pamd_residual_workspace, m, n), A.device)

    def __call__(self, x, b, want_r=False):
        m, n = self.A.shape
        dev = self.A.device
        g = torch.empty(n, dtype=F64, device=dev)
        rss = torch.empty(1, dtype=F64, device=dev)
        r = torch.empty(m, dtype=F64, device=dev) if want_r else None
        check(lib().pamd_residual_f64(m, n, _p(self.A), self.A.stride(0), _p(x.contiguous()), _p(b.contiguous()),
                                      _p(r), _p(g), _p(rss), _p(self.ws), stream(dev)), "pamd_residual_f64")
        return g, rss, r


def trsm_rec_f32(A, R):
    """FP32 rounding of A R^{-1} (R upper), A unmodified: the recursive TRSM
    out of place with the FP32 copy written by its leaves (no copy of A, no
    separate conversion pass).  Equal bit for bit to trsm_rec_(A.clone(), R)
    followed by .to(torch.float32)."""
    m, n = A.shape
    if A.stride(1) != 1 and n > 1:
        A = A.contiguous()
    X = torch.empty((m, n), dtype=F64, device=A.device)
    X32 = torch.empty((m, n), dtype=torch.float32, device=A.device)
    ws_bytes = _query(lib().pamd_gemm_workspace, m, n, n)
    ws = workspace(ws_bytes, A.device)
    check(lib().pamd_trsm_right_upper_rec_f32out(m, n, _p(R), max(R.stride(0), n, 1), _p(A), max(A.stride(0), n, 1),
                                                 _p(X), max(n, 1), _p(X32), max(n, 1), _p(ws), ws_bytes,
                                                 stream(A.device)), "pamd_trsm_right_upper_rec_f32out")
    del X
    return X32


def trsm_rec_(X, R):
    """In place X <- X R^{-1} (R upper), recursive MFMA TRSM for large m."""
    m, n = X.shape
    ws_bytes = _query(lib().pamd_gemm_workspace, m, n, n)
    ws = workspace(ws_bytes, X.device)
    check(lib().pamd_trsm_right_upper_rec_f64(m, n, _p(R), max(R.stride(0), n, 1), _p(X), max(X.stride(0), n, 1),
                                              _p(ws), ws_bytes, stream(X.device)), "pamd_trsm_right_upper_rec_f64")
    return X


# ---------------------------------------------------------------------------
# Split-precision GEMM: FP64 operands as BF16 pieces on BF16 MFMA
# ---------------------------------------------------------------------------
def split_layout(npieces, M, N, K):
    """(kp, mpad, npad, workspace bytes) of the split GEMM C (M x N) = L R^T
    with K reduced: piece arrays have kp columns, L's mpad rows, R's npad."""
    v = [ctypes.c_int64() for _ in range(4)]
    check(lib().pamd_split_layout(npieces, M, N, K, *[ctypes.byref(x) for x in v]), "pamd_split_layout")
    return tuple(x.value for x in v)


class SplitPieces:
    """BF16 pieces (npieces x rows_out x kp, uint16 bits, 16 x 32 blocked
    layout) of a K-contiguous operand with `rows` logical rows and K
    logical columns."""

    def __init__(self, pieces, rows, K):
        self.pieces = pieces
        self.rows = rows
        self.K = K

    @property
    def npieces(self):
        return self.pieces.shape[0]


def split_bf16(X, npieces, transpose=False, rows_out=None, kp=None):
    """Pieces of X (row-major FP64, rows x cols) or, with transpose, of X^T;
    rows_out defaults to the row count rounded up to 128 (a left operand)."""
    Xc = X if X.stride(-1) == 1 else X.contiguous()
    rows, cols = Xc.shape
    out_rows, K = (cols, rows) if transpose else (rows, cols)
    kp = kp or _round_up(max(K, 1), 64)
    rows_out = rows_out or _round_up(max(out_rows, 1), 128)
    P = torch.empty((npieces, rows_out, kp), dtype=torch.uint16, device=X.device)
    check(lib().pamd_split_bf16(int(transpose), npieces, rows, cols, _p(Xc), max(Xc.stride(0), cols, 1), rows_out,
                                kp, _p(P), rows_out * kp, stream(X.device)), "pamd_split_bf16")
    return SplitPieces(P, out_rows, K)


def split_gemm(L, R, out=None, transpose_out=False):
    """C = sum_{p+q<P} L_p R_q^T (M x N, FP64) from SplitPieces L (M x K) and
    R (N x K, its piece array with the layout's npad rows); with
    transpose_out the result is written as C^T (N x M)."""
    npieces = L.npieces
    M, N, K = L.rows, R.rows, L.K
    kp, mpad, npad, wsb = split_layout(npieces, M, N, K)
    if (L.pieces.shape[2] != kp or R.pieces.shape[2] != kp or L.pieces.shape[1] < mpad
            or R.pieces.shape[1] < npad or R.npieces != npieces or R.K != K):
        raise ValueError("split_gemm: piece arrays do not match the layout")
    dev = L.pieces.device
    if out is None:
        out = torch.empty((N, M) if transpose_out else (M, N), dtype=F64, device=dev)
    lci, lcj = (1, out.stride(0)) if transpose_out else (out.stride(0), 1)
    ws = workspace(wsb, dev)
    Lp, Rp = L.pieces, R.pieces
    npairs = npieces * (npieces + 1) // 2
    # logged flops: the BF16 MFMA work executed (padded tile shapes, every piece pair)
    with _Timed("split_gemm", 2.0 * mpad * npad * kp * npairs, 2.0 * npieces * (M + N) * K + 8.0 * M * N):
        check(lib().pamd_split_gemm_f64(npieces, M, N, K, _p(Lp), Lp.shape[1] * kp, _p(Rp), Rp.shape[1] * kp,
                                        _p(out), lci, lcj, _p(ws), stream(dev)), "pamd_split_gemm_f64")
    return out


def block_pieces(a):
    """Blocked piece array (uint16 tensor) from a row-major numpy one
    (npieces x rows x kp, rows a multiple of 16, kp of 64)."""
    P, R, kp = a.shape
    b = np.ascontiguousarray(a).reshape(P, R // 16, 16, kp // 32, 4, 8).transpose(0, 1, 3, 2, 4, 5)
    r = np.arange(16)[:, None]
    logical = np.arange(4)[None, :] ^ (r >> 2)  # the logical chunk stored at physical position c in row r
    return torch.from_numpy(np.ascontiguousarray(b[:, :, :, r, logical, :]).reshape(P, R, kp))


def unblock_pieces(pieces):
    """Row-major view (numpy, npieces x rows_out x kp) of a blocked piece
    array: undoes the 16 x 32 blocking and the chunk swizzle."""
    a = pieces.cpu().numpy()
    P, R, kp = a.shape
    b = a.reshape(P, R // 16, kp // 32, 16, 4, 8)
    r = np.arange(16)[:, None]
    phys = np.arange(4)[None, :] ^ (r >> 2)  # physical chunk of logical chunk c in row r
    b = b[:, :, :, r, phys, :]               # (P, R/16, kp/32, 16, 4, 8) in logical order
    return b.transpose(0, 1, 3, 2, 4, 5).reshape(P, R, kp)


def sjlt_draw(n_short, n_long, k, seed, dev):
    rows = torch.empty((n_long, k), dtype=torch.int32, device=dev)
    vals = torch.empty((n_long, k), dtype=F64, device=dev)
    check(lib().pamd_sjlt_draw(n_short, n_long, k, seed & (2 ** 64 - 1), _p(rows), _p(vals), stream(dev)),
          "pamd_sjlt_draw")
    return rows, vals


def sjlt_csr(n_short, n_long, k, rows, vals):
    """CSR (rowptr int64, cols int32, vals) of the wide SJLT whose CSC draw is (rows, vals)."""
    dev = rows.device
    rowptr = torch.empty(n_short + 1, dtype=torch.int64, device=dev)
    cols = torch.empty(max(n_long * k, 1), dtype=torch.int32, device=dev)
    cvals = torch.empty(max(n_long * k, 1), dtype=F64, device=dev)
    ws = workspace(_query(lib().pamd_sjlt_csr_workspace, n_short, n_long), dev)
    check(lib().pamd_sjlt_csr(n_short, n_long, k, _p(rows), _p(vals), _p(rowptr), _p(cols), _p(cvals),
                              _p(ws), stream(dev)), "pamd_sjlt_csr")
    return rowptr, cols, cvals


def sjlt_left(rowptr, cols, vals, d, A, out=None):
    """out (d x n) = W @ A with W (d x m) in CSR."""
    A = A if A.stride(-1) == 1 else A.contiguous()
    m, n = A.shape
    if out is None:
        out = torch.empty((d, n), dtype=F64, device=A.device)
    check(lib().pamd_sjlt_left_f64(d, m, n, _p(rowptr), _p(cols), _p(vals), _p(A), max(A.stride(0), n, 1),
                                   _p(out), max(out.stride(0), n, 1), stream(A.device)), "pamd_sjlt_left_f64")
    return out


def sjlt_right(rowptr, cols, vals, k, A, out=None):
    """out (m x k) = A @ W^T with W (k x n) in CSR."""
    A = A if A.stride(-1) == 1 else A.contiguous()
    m, n = A.shape
    if out is None:
        out = torch.empty((m, k), dtype=F64, device=A.device)
    check(lib().pamd_sjlt_right_f64(m, n, k, _p(rowptr), _p(cols), _p(vals), _p(A), max(A.stride(0), n, 1),
                                    _p(out), max(out.stride(0), k, 1), stream(A.device)), "pamd_sjlt_right_f64")
    return out


# ---------------------------------------------------------------------------
# LSQR
# ---------------------------------------------------------------------------

def _lsqr_problem(A, b, R, sqrt_delta, x0, damp, atol, btol, conlim, iter_lim, calc_var):
    """The pamd_lsqr_problem struct and the tensors it points to (kept alive by the caller)."""
    A = A if A.stride(-1) == 1 else A.contiguous()
    m, n = A.shape
    b = b.contiguous()
    prob = _lib.LsqrProblem()
    prob.m, prob.n = m, n
    prob.A, prob.lda = A.data_ptr(), max(A.stride(0), n)
    prob.sqrt_delta = float(sqrt_delta)
    if R is not None:
        if R.stride(1) != 1:
            R = R.contiguous()
        prob.R, prob.ldr = R.data_ptr(), max(R.stride(0), n)
    else:
        prob.R, prob.ldr = None, n
    prob.b = b.data_ptr()
    if x0 is not None:
        x0 = x0.contiguous()
        prob.x0 = x0.data_ptr()
    else:
        prob.x0 = None
    prob.damp, prob.atol, prob.btol, prob.conlim = float(damp), float(atol), float(btol), float(conlim)
    prob.iter_lim = iter_lim
    prob.calc_var = int(bool(calc_var))
    return prob, (A, b, R, x0)


def lsqr_run(A, b, R=None, sqrt_delta=0.0, x0=None, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8,
             iter_lim=None, calc_var=False, want_orig=False):
    """Device LSQR on [A; sqrt_delta I] R^{-1} with rhs [b; 0].

    Returns (x, var, arnorms, result) -- or (x, x_orig, var, arnorms, result)
    with ``want_orig``, x_orig = R^{-1} x computed by the iteration's own
    triangular substitution."""
    m, n = A.shape
    dev = A.device
    iter_lim = int(2 * n if iter_lim is None else iter_lim)
    prob, keep = _lsqr_problem(A, b, R, sqrt_delta, x0, damp, atol, btol, conlim, iter_lim, calc_var)
    x = torch.empty(n, dtype=F64, device=dev)
    xo = torch.empty(n, dtype=F64, device=dev) if want_orig else None
    var = torch.zeros(n, dtype=F64, device=dev)
    arn = torch.empty(max(iter_lim, 1), dtype=F64, device=dev)
    ws = workspace(_query(lib().pamd_lsqr_workspace, m, n, iter_lim), dev)
    res = _lib.LsqrResult()
    check(lib().pamd_lsqr_run_f64(ctypes.byref(prob), _p(x), _p(xo), _p(var), _p(arn), ctypes.byref(res),
                                  _p(ws), stream(dev)), "pamd_lsqr_run_f64")
    del keep
    if want_orig:
        return x, xo, var, arn, res
    return x, var, arn, res


LSQR_RESULT_BYTES = ctypes.sizeof(_lib.LsqrResult)


def lsqr_result_from_device(t):
    """A pamd_lsqr_result read back from the device bytes pamd_lsqr_launch_f64 wrote."""
    raw = bytes(t.cpu().numpy().tobytes())
    return _lib.LsqrResult.from_buffer_copy(raw)


def lsqr_launch(A, b, R=None, sqrt_delta=0.0, x0=None, damp=0.0, atol=1e-8, btol=1e-8, conlim=1e8,
                iter_lim=None, calc_var=False, out=None):
    """LSQR as launches only (graph-capturable): every iteration up to
    iter_lim is launched and the result stays on the device.

    ``out`` = (x, x_orig or None, var, arnorms, result bytes, workspace) from a
    previous call (or from :func:`lsqr_launch_buffers`) is reused, so a
    captured graph owns no allocation.  Returns that tuple."""
    m, n = A.shape
    dev = A.device
    iter_lim = int(2 * n if iter_lim is None else iter_lim)
    if out is None:
        out = lsqr_launch_buffers(m, n, iter_lim, dev)
    x, xo, var, arn, resb, ws = out
    prob, keep = _lsqr_problem(A, b, R, sqrt_delta, x0, damp, atol, btol, conlim, iter_lim, calc_var)
    check(lib().pamd_lsqr_launch_f64(ctypes.byref(prob), _p(x), _p(xo), _p(var), _p(arn), _p(resb), _p(ws),
                                     stream(dev)), "pamd_lsqr_launch_f64")
    del keep
    return out


def lsqr_launch_buffers(m, n, iter_lim, dev, want_orig=False):
    """Outputs and workspace of :func:`lsqr_launch` for an m x n problem."""
    x = torch.empty(n, dtype=F64, device=dev)
    xo = torch.empty(n, dtype=F64, device=dev) if want_orig else None
    var = torch.zeros(n, dtype=F64, device=dev)
    arn = torch.zeros(max(iter_lim, 1), dtype=F64, device=dev)  # entries past the stop stay 0
    resb = torch.zeros(LSQR_RESULT_BYTES, dtype=torch.uint8, device=dev)
    ws = workspace(_query(lib().pamd_lsqr_workspace, m, n, iter_lim), dev)
    return x, xo, var, arn, resb, ws
