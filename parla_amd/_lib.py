"""ctypes binding of libparla_amd.so (the C ABI declared in include/parla_amd.h).

The library is loaded after ``import torch`` so that its ``libamdhip64.so.7``
dependency resolves, by SONAME, to the HIP runtime torch already loaded: one
runtime, one device context, and torch's device pointers and streams are valid
in the library.  There is no fallback: if the library is missing the import of
any compute routine raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see above)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparla_amd.so")

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_u64 = ctypes.c_uint64


class LsqrProblem(ctypes.Structure):
    _fields_ = [("m", c_i64), ("n", c_i64), ("A", c_vp), ("lda", c_i64),
                ("sqrt_delta", c_dbl), ("R", c_vp), ("ldr", c_i64), ("b", c_vp),
                ("x0", c_vp), ("damp", c_dbl), ("atol", c_dbl), ("btol", c_dbl),
                ("conlim", c_dbl), ("iter_lim", c_i64), ("calc_var", c_int)]


class LsqrResult(ctypes.Structure):
    _fields_ = [("istop", c_i64), ("itn", c_i64), ("r1norm", c_dbl), ("r2norm", c_dbl),
                ("anorm", c_dbl), ("acond", c_dbl), ("arnorm", c_dbl), ("xnorm", c_dbl),
                ("n_log", c_i64)]


# name -> (restype, argtypes); mirrors include/parla_amd.h
SIGNATURES = {
    "pamd_abi_version": (c_int, []),
    "pamd_status_string": (ctypes.c_char_p, [c_int]),
    "pamd_gemm_f64": (c_int, [c_int, c_int, c_i64, c_i64, c_i64, c_dbl, c_vp, c_i64, c_vp, c_i64,
                              c_dbl, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_gemm_workspace": (c_int, [c_i64, c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_gemv_f64": (c_int, [c_int, c_i64, c_i64, c_dbl, c_vp, c_i64, c_vp, c_dbl, c_vp, c_vp,
                              c_i64, c_vp]),
    "pamd_gemv_workspace": (c_int, [c_int, c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_nrm2_f64": (c_int, [c_i64, c_vp, c_vp, c_vp]),
    "pamd_trsv_upper_f64": (c_int, [c_int, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_gaussian_fill_f64": (c_int, [c_i64, c_i64, c_dbl, c_u64, c_u64, c_vp, c_i64, c_vp]),
    "pamd_sjlt_draw": (c_int, [c_i64, c_i64, c_int, c_u64, c_vp, c_vp, c_vp]),
    "pamd_srct_fill_f64": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "pamd_sjlt_csr_workspace": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_sjlt_csr": (c_int, [c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_sjlt_left_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                                   c_i64, c_vp]),
    "pamd_sjlt_right_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                                    c_i64, c_vp]),
    "pamd_transpose_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_svd_jacobi_workspace": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_svd_jacobi_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_vp,
                                    c_vp]),
    "pamd_svd_jacobi_noalloc_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_vp,
                                            c_vp]),
    "pamd_sym_shift_f64": (c_int, [c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp]),
    "pamd_eig_from_svd_f64": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_dbl, c_vp, c_vp, c_vp]),
    "pamd_scale_cols_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_int, c_vp, c_i64, c_vp]),
    "pamd_geqrf_workspace": (c_int, [c_i64, c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_geqrf_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_orgqr_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp]),
    "pamd_qr_singular_flag": (c_int, [c_vp, c_i64, c_i64, c_i64, ctypes.POINTER(c_vp)]),
    "pamd_qr_r_workspace": (c_int, [c_i64, c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_qr_r_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_qr_r_singular_flag": (c_int, [c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "pamd_lu_workspace": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_lu_pl_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_potrf_upper_f64": (c_int, [c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_trtri_upper_f64": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_small_combine_f64": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_dbl, c_dbl, c_int, c_vp, c_i64, c_vp,
                                       c_vp]),
    "pamd_cholqr_stats_f64": (c_int, [c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_trsm_mfma_workspace": (c_int, [c_i64, ctypes.POINTER(c_i64)]),
    "pamd_trsm_right_upper_mfma_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "pamd_trsm_right_upper_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_ozaki_layout": (c_int, [c_i64, c_i64, c_i64] + [ctypes.POINTER(c_i64)] * 5),
    "pamd_gauss_q_bf16": (c_int, [c_i64, c_i64, c_i64, ctypes.c_uint64, ctypes.c_uint64, c_vp, c_vp]),
    "pamd_ozaki_split_bf16": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_int, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_ozaki_sketch_f64": (c_int, [c_i64, c_i64, c_i64, c_int, c_vp, c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64,
                                      c_vp, c_vp]),
    "pamd_normal_pass_workspace": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_normal_pass": (c_int, [c_int, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_residual_workspace": (c_int, [c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_residual_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_nrm2_workspace": (c_int, [c_i64, ctypes.POINTER(c_i64)]),
    "pamd_potrf_batched_f64": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_trtri_batched_f64": (c_int, [c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_normal": (c_int, [c_int, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_residual": (c_int, [c_int, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_reduce": (c_int, [c_i64, c_vp, c_vp, c_int, c_vp, c_vp, c_vp]),
    "pamd_bt_trsv": (c_int, [c_i64, c_vp, c_int, c_vp, c_vp, c_vp]),
    "pamd_bt_trmv": (c_int, [c_i64, c_vp, c_int, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_cg": (c_int, [c_i64, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_axpy": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_copy": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_layout": (c_int, [ctypes.POINTER(c_int)] * 5),
    "pamd_split_layout": (c_int, [c_int, c_i64, c_i64, c_i64] + [ctypes.POINTER(c_i64)] * 4),
    "pamd_split_bf16": (c_int, [c_int, c_int, c_i64, c_i64, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "pamd_split_gemm_f64": (c_int, [c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp,
                                    c_vp]),
    "pamd_bt_sjlt_csr": (c_int, [c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_sjlt_left": (c_int, [c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_gram": (c_int, [c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp]),
    "pamd_bt_trsm": (c_int, [c_i64, c_vp, c_vp, c_i64, c_i64, c_int, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_trmm_upper": (c_int, [c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_fro2": (c_int, [c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_bt_dot": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_nrm2_large_f64": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_trsm_right_upper_rec_f64": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_trsm_right_upper_rec_f32out": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_mlp_nparams": (c_int, []),
    "pamd_mlp_fwd": (c_int, [c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_mlp_bwd_workspace": (c_int, [c_i64, ctypes.POINTER(c_i64)]),
    "pamd_mlp_bwd": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pamd_mlp_adam": (c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, c_i64, c_vp, c_vp, c_vp]),
    "pamd_row_gather_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_sample_rows_workspace": (c_int, [c_i64, ctypes.POINTER(c_i64)]),
    "pamd_sample_rows_f64": (c_int, [c_i64, c_vp, c_i64, c_u64, c_vp, c_vp, c_vp, c_vp]),
    "pamd_row_scatter_f64": (c_int, [c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "pamd_lsqr_workspace": (c_int, [c_i64, c_i64, c_i64, ctypes.POINTER(c_i64)]),
    "pamd_lsqr_run_f64": (c_int, [ctypes.POINTER(LsqrProblem), c_vp, c_vp, c_vp, c_vp,
                                  ctypes.POINTER(LsqrResult), c_vp, c_vp]),
    "pamd_lsqr_launch_f64": (c_int, [ctypes.POINTER(LsqrProblem), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
}

_lib = None


def load():
    """Load (once) and return the ctypes handle; raise if the build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m parla_amd.build` "
            "(the GPU library has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PamdError(RuntimeError):
    """A HIP runtime failure reported by the library."""


def check(status, fname):
    """Map a library status onto the exception types the reference raises."""
    if status == 0:
        return
    msg = load().pamd_status_string(status).decode()
    if status == -1000:
        raise PamdError(f"{fname}: {msg}")
    if status == -1001:
        raise NotImplementedError(f"{fname}: {msg}")
    if status < 0:
        raise ValueError(f"{fname}: argument {-status} is invalid ({msg})")
    raise ArithmeticError(f"{fname}: {msg} (status {status})")
