"""The allocation-free Jacobi SVD entry point (pamd_svd_jacobi_noalloc_f64)
against the workspace one (pamd_svd_jacobi_f64), through the C ABI: the same
kernels on the library's static per-GPU array must give the same bits, on two
calls in a row (the second smaller, so the array holds the first call's
larger contents), and sizes beyond the array are rejected without a launch.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import _lib
    return _lib.load()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(lib, A, noalloc):
    m, n = A.shape
    dev = A.device
    U = torch.full((m, n), np.nan, dtype=torch.float64, device=dev)
    s = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    Vt = torch.full((n, n), np.nan, dtype=torch.float64, device=dev)
    info = torch.full((1,), -7, dtype=torch.int32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if noalloc:
        rc = lib.pamd_svd_jacobi_noalloc_f64(m, n, _p(A), n, _p(U), n, _p(s), _p(Vt), n, 40, _p(info), st)
    else:
        nb = ctypes.c_int64(0)
        assert lib.pamd_svd_jacobi_workspace(m, n, ctypes.byref(nb)) == 0
        ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        rc = lib.pamd_svd_jacobi_f64(m, n, _p(A), n, _p(U), n, _p(s), _p(Vt), n, 40, _p(ws), _p(info), st)
    assert rc == 0, rc
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    return U.cpu(), s.cpu(), Vt.cpu()


def test_noalloc_svd_matches_workspace_entry(lib):
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    for m, n in ((700, 520), (90, 33)):   # the second call is smaller than the first
        _say(f"noalloc vs workspace SVD {m}x{n}")
        A = torch.from_numpy(rng.standard_normal((m, n)) * np.geomspace(1.0, 1e-6, n)).to(dev)
        ref = _run(lib, A, noalloc=False)
        got = _run(lib, A, noalloc=True)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        Ah = A.cpu().numpy()
        U, s, Vt = (t.numpy() for t in got)
        assert np.linalg.norm(U * s @ Vt - Ah) <= 1e-12 * np.linalg.norm(Ah)


def test_noalloc_svd_rejects_sizes_beyond_its_array(lib):
    _say("noalloc size limit")
    dev = torch.device("cuda", 0)
    A = torch.zeros((1025, 4), dtype=torch.float64, device=dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib.pamd_svd_jacobi_noalloc_f64(1025, 4, _p(A), 4, _p(out), 4, _p(out), _p(out), 4, 40, _p(info), st)
    assert rc == -1
    rc = lib.pamd_svd_jacobi_noalloc_f64(10, 11, _p(A), 11, _p(out), 11, _p(out), _p(out), 11, 40, _p(info), st)
    assert rc == -2
