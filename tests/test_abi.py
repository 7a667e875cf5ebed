"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and the Python binding covers the same set."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parla_amd.h")
LIB = os.path.join(ROOT, "parla_amd", "libparla_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pamd_[a-z0-9_]+)\s*\(", text)))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if " T pamd_" in ln})


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        pytest.fail("libparla_amd.so is not built; run `python -m parla_amd.build`")
    return LIB


def test_header_declares_entry_points():
    names = declared()
    for must in ["pamd_gemm_f64", "pamd_geqrf_f64", "pamd_orgqr_f64", "pamd_lsqr_run_f64",
                 "pamd_sjlt_left_f64", "pamd_sjlt_right_f64", "pamd_trsv_upper_f64",
                 "pamd_gaussian_fill_f64", "pamd_lu_pl_f64"]:
        assert must in names


def test_library_exports_every_declared_symbol(built):
    missing = set(declared()) - set(exported())
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_python_binding_matches_header(built):
    from parla_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared())
    lib = _lib.load()
    for name in _lib.SIGNATURES:
        assert hasattr(lib, name)
    assert lib.pamd_abi_version() == 1
    assert lib.pamd_status_string(-3).decode().startswith("invalid argument")


def test_argument_validation_without_gpu(built):
    """Bad arguments are rejected on the host, before any launch."""
    import ctypes
    from parla_amd import _lib
    lib = _lib.load()
    # transA = 2 -> argument 1 invalid
    rc = lib.pamd_gemm_f64(2, 0, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None, 0, None)
    assert rc == -1
    # lda < K
    rc = lib.pamd_gemm_f64(0, 0, 4, 4, 4, 1.0, ctypes.c_void_p(8), 2, ctypes.c_void_p(8), 4, 0.0,
                           ctypes.c_void_p(8), 4, None, 0, None)
    assert rc == -8
    # trsv: n must stay below 2^30 (beyond 2048 the vector is solved in place in global memory)
    rc = lib.pamd_trsv_upper_f64(0, 1 << 30, ctypes.c_void_p(8), 1 << 30, ctypes.c_void_p(8), None)
    assert rc == -2
    b = ctypes.c_int64(0)
    assert lib.pamd_geqrf_workspace(10, 20, 20, ctypes.byref(b)) == -2  # n > m
    assert lib.pamd_geqrf_workspace(100, 20, 21, ctypes.byref(b)) == 0 and b.value > 0
    assert lib.pamd_sjlt_csr_workspace(20000, 100, ctypes.byref(b)) == -1  # n_short too large
    # batched sketch: vec_nnz above the register-resident draw bound
    assert lib.pamd_bt_sjlt_csr(1, ctypes.c_void_p(8), ctypes.c_void_p(8), 1, 17, *[ctypes.c_void_p(8)] * 4,
                                None) == -5


def test_batched_layout_constants(built):
    """The constants the host-side batch plan is tested with are the library's."""
    import ctypes
    from parla_amd import _lib
    import test_batched_plan as tbp
    lib = _lib.load()
    vals = [ctypes.c_int() for _ in range(5)]
    assert lib.pamd_bt_layout(*[ctypes.byref(v) for v in vals]) == 0
    assert tuple(v.value for v in vals) == tbp.LAYOUT


def test_status_mapping():
    from parla_amd import _lib
    with pytest.raises(ValueError):
        _lib.check(-4, "f")
    with pytest.raises(NotImplementedError):
        _lib.check(-1001, "f")
    with pytest.raises(_lib.PamdError):
        _lib.check(-1000, "f")
    with pytest.raises(ArithmeticError):
        _lib.check(1, "f")
