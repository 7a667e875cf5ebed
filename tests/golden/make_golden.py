"""Write golden vectors by running PARLA's own code (run once, in the build container).

The fixtures pin the parts of the hot path that no closed form fixes:

* the exact random draws of ``sjlt_operator`` (row indices, signs, CSC/CSR
  layout) and ``gaussian_operator`` for given seeds;
* the exact LSQR trajectory (istop, itn, arnorms log, estimates) of the
  reference's modified ``lsqr`` on small plain and preconditioned problems;
* the SPO error log (``log.errors``) and solution on a small problem.

Usage:  python tests/golden/make_golden.py [/path/to/parla/source]
The output ``tests/golden/parla_golden.npz`` holds plain arrays only.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    sys.path.insert(0, ref_root)
    import scipy.linalg as la
    from parla.utils import sketching as usk
    from parla.comps.determiter.lsqr import lsqr
    from parla.comps.preconditioning import a_lift_precond
    import parla.comps.sketchers.oblivious as sko
    import parla.drivers.least_squares as rlsq

    out = {}

    # --- SJLT draws: wide (CSC) and tall (CSR) ---------------------------
    for tag, (r, c, seed, k) in {
        "w16x64_s0": (16, 64, 0, 8), "w16x64_s1": (16, 64, 1, 8),
        "w40x300_s42": (40, 300, 42, 8), "w800x2000_s7": (800, 2000, 7, 8),
        "w5x30_s3": (5, 30, 3, 8),          # short columns: with replacement
        "t64x16_s5": (64, 16, 5, 8), "t4096x266_s11": (4096, 266, 11, 8),
        "w30x50_s9_k3": (30, 50, 9, 3),
    }.items():
        S = usk.sjlt_operator(r, c, np.random.default_rng(seed), k)
        out[f"sjlt_{tag}_indptr"] = S.indptr.astype(np.int64)
        out[f"sjlt_{tag}_indices"] = S.indices.astype(np.int64)
        out[f"sjlt_{tag}_data"] = S.data.astype(np.float64)
        out[f"sjlt_{tag}_meta"] = np.array([r, c, seed, k, int(S.format == "csc")])

    # --- Gaussian draws ----------------------------------------------------
    for tag, (r, c, seed, norm) in {
        "g8x20_s0": (8, 20, 0, True), "g30x6_s3": (30, 6, 3, True),
        "g5x7_s4_raw": (5, 7, 4, False),
    }.items():
        out[f"gauss_{tag}"] = usk.gaussian_operator(r, c, np.random.default_rng(seed), norm)
        out[f"gauss_{tag}_meta"] = np.array([r, c, seed, int(norm)])

    # --- LSQR trajectories ---------------------------------------------------
    def put_lsqr(tag, A_op, b, **kw):
        res = lsqr(A_op, b, **kw)
        out[f"lsqr_{tag}_x"] = np.asarray(res[0], dtype=np.float64)
        out[f"lsqr_{tag}_ints"] = np.array([res[1], res[2]])
        out[f"lsqr_{tag}_scal"] = np.array([res[3], res[4], res[5], res[6], res[8]],
                                           dtype=np.float64)
        out[f"lsqr_{tag}_arnorms"] = np.atleast_1d(np.asarray(res[7], dtype=np.float64))
        out[f"lsqr_{tag}_var"] = np.asarray(res[9], dtype=np.float64)

    A3 = np.array([[1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    out["lsqr_A3"] = A3
    put_lsqr("kat1", A3, np.array([1.0, 0.0, -1.0]))
    put_lsqr("kat2", A3, np.array([1.0, 0.01, -1.0]))

    rng = np.random.default_rng(2026)
    A = rng.standard_normal((60, 8)) @ np.diag(np.logspace(0, -3, 8))
    b = rng.standard_normal(60)
    x0 = rng.standard_normal(8)
    out["lsqr_A60"], out["lsqr_b60"], out["lsqr_x060"] = A, b, x0
    put_lsqr("plain", A, b, atol=1e-10, btol=1e-10)
    put_lsqr("x0", A, b, atol=1e-10, btol=1e-10, x0=x0)
    put_lsqr("damp", A, b, damp=0.3, atol=1e-10, btol=1e-10, calc_var=True)
    put_lsqr("lim", A, b, iter_lim=3)

    # preconditioned operator, as PcSS2 builds it (tri and ridge)
    S = usk.gaussian_operator(24, 60, np.random.default_rng(77))
    R = la.qr(S @ A, mode="economic")[1]
    out["lsqr_R60"] = R
    A_pc = a_lift_precond(A, 0.0, R, True)[0]
    put_lsqr("pc", A_pc, b, atol=1e-12, btol=1e-12, iter_lim=50)
    Rd = la.qr(np.vstack((S @ A, np.sqrt(0.25) * np.eye(8))), mode="economic")[1]
    out["lsqr_Rd60"] = Rd
    A_pcd = a_lift_precond(A, 0.25, Rd, True)[0]
    put_lsqr("pcridge", A_pcd, np.concatenate((b, np.zeros(8))),
             atol=1e-12, btol=1e-12, iter_lim=50)

    # --- SPO on a small inconsistent problem ------------------------------
    rng = np.random.default_rng(31337)
    m, n = 300, 20
    U = usk.orthonormal_operator(m, n, rng)
    V = usk.orthonormal_operator(n, n, rng)
    s = np.geomspace(1, 1e-4, n)
    As = (U * s) @ V
    bs = As @ rng.standard_normal(n) + 1e-3 * rng.standard_normal(m)
    out["spo_A"], out["spo_b"] = As, bs
    for tag, gen, sf, delta in [("ga", sko.SkOpGA(), 4, 0.0),
                                ("sj", sko.SkOpSJ(8), 3, 0.0),
                                ("sjridge", sko.SkOpSJ(8), 3, 0.25)]:
        alg = rlsq.SPO(gen, sf, mode="qr")
        x, log = alg(As, bs, delta, 1e-12, 60, np.random.default_rng(5), logging=True)
        out[f"spo_{tag}_x"] = x
        out[f"spo_{tag}_errors"] = log.errors
        out[f"spo_{tag}_meta"] = np.array([sf, delta])

    path = os.path.join(HERE, "parla_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
