"""Write golden vectors by running PARLA's own code (run once, in the build container).

The fixtures pin the parts of the hot path that no closed form fixes:

* the exact random draws of ``sjlt_operator`` (row indices, signs, CSC/CSR
  layout) and ``gaussian_operator`` for given seeds;
* the exact LSQR trajectory (istop, itn, arnorms log, estimates) of the
  reference's modified ``lsqr`` on small plain and preconditioned problems;
* the SPO error log (``log.errors``) and solution on a small problem, in
  the modes 'qr', 'chol' and 'svd' (and a rank-deficient 'svd' run);
* the next drivers on the path: PcSS2's underdetermined branch (delta > 0),
  SSO1, PcSS1 / pcg, SPS1 / SPS2, QB3 and EVD2.  Routines that the
  reference can no longer run under numpy 2 (QB2, SPU1 with delta = 0, the
  procedural sso1 / evd2, all through the removed ``np.NaN``) are pinned by
  closed forms in the tests instead.

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
    import parla.comps.preconditioning as rpc
    import parla.comps.determiter.saddle as dsad
    import parla.drivers.saddlesys as sad
    import parla.drivers.evd as evd
    import parla.comps.qb as qbm
    import parla.utils.linalg_wrappers as ulaw
    from parla.comps.sketchers.aware import RS1

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

    # --- SRCT draws and products, sparse sign draws ---------------------------
    r, e, perm = usk.generate_srct(12, 50, np.random.default_rng(3))
    out["srct_r"], out["srct_e"], out["srct_perm"] = r, e, perm
    Xs = np.random.default_rng(4).standard_normal((50, 3))
    out["srct_X"] = Xs
    out["srct_fwd"] = usk.apply_srct(r, e, Xs, perm)
    out["srct_adj"] = usk.apply_srct(r, e, out["srct_fwd"], perm, forward=False)
    St = usk.srct_operator(40, 9, np.random.default_rng(5))   # tall: transpose of the wide draw
    out["srct_tall_X"] = np.random.default_rng(6).standard_normal((9, 2))
    out["srct_tall_out"] = St @ out["srct_tall_X"]
    for tag, (rr, cc, seed, dens) in {"ss20x60": (20, 60, 7, 0.1), "ss50x8": (50, 8, 8, 0.3)}.items():
        S = usk.sparse_sign_operator(rr, cc, np.random.default_rng(seed), dens)
        out[f"sps_{tag}"] = S.toarray()

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

    # --- SPO modes 'chol' and 'svd' (least_squares.py:317-339) -----------
    for tag, gen, sf, mode, delta in [("ga_chol", sko.SkOpGA(), 4, "chol", 0.0),
                                      ("sj_chol_ridge", sko.SkOpSJ(8), 3, "chol", 0.25),
                                      ("ga_svd", sko.SkOpGA(), 4, "svd", 0.0),
                                      ("sj_svd_ridge", sko.SkOpSJ(8), 3, "svd", 0.25)]:
        alg = rlsq.SPO(gen, sf, mode=mode)
        x, log = alg(As, bs, delta, 1e-12, 60, np.random.default_rng(5), logging=True)
        out[f"spo_{tag}_x"] = x
        out[f"spo_{tag}_errors"] = log.errors
        out[f"spo_{tag}_meta"] = np.array([sf, delta])
    # rank-deficient data: the SVD preconditioner drops the null directions
    rng = np.random.default_rng(4242)
    Ur = usk.orthonormal_operator(200, 12, rng)
    Vr = usk.orthonormal_operator(12, 16, rng)
    Ard = (Ur * np.geomspace(1, 1e-2, 12)) @ Vr
    brd = Ard @ rng.standard_normal(16) + 1e-3 * rng.standard_normal(200)
    out["spo_rd_A"], out["spo_rd_b"] = Ard, brd
    x, log = rlsq.SPO(sko.SkOpGA(), 3, mode="svd")(Ard, brd, 0.0, 1e-12, 60, np.random.default_rng(6),
                                                    logging=True)
    out["spo_rd_svd_x"] = x
    out["spo_rd_svd_errors"] = log.errors
    # svd_right_precond on a sketch (M M^T and sigma are sign-free)
    M, U, sig, Vh = rpc.svd_right_precond(usk.gaussian_operator(80, 300, np.random.default_rng(8)) @ As)
    out["svdpc_M"], out["svdpc_sigma"] = M, sig  # M up to column signs

    # --- PcSS2's underdetermined branch with delta > 0 (saddle.py:203-214) ---
    Rs = la.qr(np.vstack((usk.gaussian_operator(80, 300, np.random.default_rng(9)) @ As,
                          np.sqrt(0.5) * np.eye(20))), mode="economic")[1]
    cvec = np.random.default_rng(10).standard_normal(20)
    xu, yu, eu = dsad.PcSS2()(As, None, cvec, 0.5, 1e-12, 60, Rs, True, None)
    out["pcss2u_R"], out["pcss2u_c"] = Rs, cvec
    out["pcss2u_x"], out["pcss2u_y"], out["pcss2u_errors"] = xu, yu, np.asarray(eu)

    # --- sketch-and-solve SSO1 (least_squares.py:114-189) -------------------
    for tag, delta in (("", 0.0), ("_ridge", 0.3)):
        alg = rlsq.SSO1(sko.SkOpSJ(8), 3, "gelsd", overwrite_sketch=True)
        x, _ = alg(As, bs, delta, np.nan, 1, np.random.default_rng(12), logging=True)
        out[f"sso1{tag}_x"] = x

    # --- PcSS1 / pcg and the saddle-point drivers (saddlesys.py) -------------
    rng = np.random.default_rng(555)
    m, n = 400, 25
    Us = usk.orthonormal_operator(m, n, rng)
    Vs = usk.orthonormal_operator(n, n, rng)
    Asd = (Us * np.geomspace(1, 1e-3, n)) @ Vs
    bsd = rng.standard_normal(m)
    csd = rng.standard_normal(n)
    out["sad_A"], out["sad_b"], out["sad_c"] = Asd, bsd, csd
    Msd = rpc.svd_right_precond(usk.gaussian_operator(100, m, np.random.default_rng(13)) @ Asd)[0]
    out["sad_M"] = Msd
    for tag, delta in (("", 0.0), ("_ridge", 0.2)):
        x, y, res = dsad.PcSS1()(Asd, bsd, csd, delta, 1e-10, 50, Msd.copy(), False, None)
        out[f"pcss1{tag}_x"], out[f"pcss1{tag}_y"], out[f"pcss1{tag}_res"] = x, y, np.asarray(res)
    for tag, cls, solver in (("sps1", sad.SPS1, dsad.PcSS1()), ("sps2", sad.SPS2, dsad.PcSS2())):
        for dtag, delta in (("", 0.0), ("_ridge", 0.2)):
            alg = cls(sko.SkOpGA(), 4, solver)
            x, y, log = alg(Asd, bsd, csd, delta, 1e-10, 50, np.random.default_rng(21), logging=True)
            out[f"{tag}{dtag}_x"], out[f"{tag}{dtag}_y"] = x, y
            out[f"{tag}{dtag}_errors"] = log.errors
    # the Nystrom-style SPS1 (sampling factor below 1: d < n)
    alg = sad.SPS1(sko.SkOpGA(), 0.6, dsad.PcSS1())
    x, y, log = alg(Asd, bsd, csd, 0.1, 1e-10, 50, np.random.default_rng(22), logging=True)
    out["sps1nys_x"], out["sps1nys_y"], out["sps1nys_errors"] = x, y, log.errors

    # --- QB3 and EVD2 (qb.py:485-595, evd.py:292-384) ------------------------
    rng = np.random.default_rng(777)
    Uq = usk.orthonormal_operator(150, 30, rng)
    Vq = usk.orthonormal_operator(30, 90, rng)
    Aq = (Uq * np.geomspace(1, 1e-4, 30)) @ Vq
    out["qb_A"] = Aq
    for tag, blk, tol in (("blk5", 5, 0.0), ("blk7tol", 7, 1e-3)):
        Q, B = qbm.QB3(RS1(sko.SkOpGA(), 2, ulaw.orth, 1), blk)(Aq, 20, tol, np.random.default_rng(31))
        out[f"qb3_{tag}_Q"], out[f"qb3_{tag}_B"] = Q, B  # Q up to column signs
    Apsd = (Uq * np.geomspace(1, 1e-4, 30)) @ Uq.T  # rebuilt from qb_U in the tests
    out["qb_U"] = Uq
    for tag, npass in (("p1", 1), ("p3", 3)):
        V, lamb = evd.EVD2(RS1(sko.SkOpGA(), npass - 1, ulaw.orth, 1))(Apsd, 12, np.nan, 6,
                                                                      np.random.default_rng(41))
        out[f"evd2_{tag}_lamb"], out[f"evd2_{tag}_V"] = lamb, V  # V up to column signs

    path = os.path.join(HERE, "parla_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
