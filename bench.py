#!/usr/bin/env python
"""Benchmark of the randomized-NLA hot path on one MI355X.

Headline (``value``): QB GB/s on the QB -> SVD workload -- A 8192 x 4096 FP64
with s_i = i^-2 (i <= 512, then 1e-6) plus 1e-8 Gaussian noise, target rank
256 + 10 oversampling, 2 power iterations with QR re-orthogonalisation
(RS1 num_pass = 4), SJLT start with 8 nonzeros +-1/sqrt(8) per column.  One
step = one QB1 call (RS1 -> RF1 -> B = Q^T A) on A resident in HBM; it reads A
6 times, so QB GB/s = 6 * m * n * 8 B / t_step.

Also reported (``extra``): LS solves/s of plain FP64 sketch-and-precondition
(A 20000 x 200, Gaussian d = 800, Householder QR, LSQR to tol 1e-12) with its
backward error, the SVD1 time, and for each the CPU oracle timed on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LS solves/s at ≤1e-10 backward err (chosen vs plain FP64 GPU); QB GB/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2516.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def roofline_terms(terms, measured_s):
    """Roofline time of a solve from its algorithmic work: terms is a list of
    (name, amount, rate) with amount in flop or bytes and rate in flop/s or
    B/s at the published peak.  Returns the dict the bench reports."""
    parts = {name: amount / rate * 1e3 for name, amount, rate in terms}
    t_roof = sum(parts.values())
    return {"roofline_ms": round(t_roof, 4), "measured_ms": round(measured_s * 1e3, 4),
            "frac": round(t_roof / 1e3 / measured_s, 4),
            "parts_ms": {k: round(v, 4) for k, v in parts.items()}}


def device_eta(ops, torch, A, b, x):
    """eta = ||A^T r|| / (||A||_F ||r||) from the library's compensated
    residual pass (r and A^T r to about working precision)."""
    g, rss, _ = ops.Residual(A)(x, b)
    return float((torch.linalg.norm(g) / (torch.linalg.norm(A) * torch.sqrt(rss))).item())


def cpu_threads():
    try:
        from threadpoolctl import threadpool_info
        n = [d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return len(os.sched_getaffinity(0))


def gemm_traffic():
    """HBM bytes per launch of the QB GEMMs, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (calibrated by a copy of known size)."""
    try:
        with open(os.path.join(ROOT, "profiles", "gemm_f64_qb_traffic.json")) as f:
            return json.load(f)["mean_hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def timed_steps(fn, steps, warmup, torch):
    for i in range(warmup):
        fn(1000 + i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fn(i)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle timings")
    ap.add_argument("--ls-steps", type=int, default=10)
    ap.add_argument("--legs", default="qb,ls,mixed,learned,batched",
                    help="comma list of qb, ls, mixed, learned, batched")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--train-matrices", type=int, default=2000)
    ap.add_argument("--pmc-gemm", action="store_true",
                    help="only run the three QB GEMM shapes once each plus a 512 MiB calibration copy "
                         "(the workload of the rocprofv3 FETCH_SIZE / WRITE_SIZE passes)")
    ap.add_argument("--pmc-bf16", action="store_true",
                    help="only run the two BF16 MFMA GEMMs once each: the split GEMM on the QB1 shape "
                         "A S (8192 x 266, K 4096, 2 pieces) and the exact sketch at A 262144 x 1024 "
                         "(d 4096, 4 pieces) (the workload of the rocprofv3 SQ counter passes)")
    ap.add_argument("--profile-qr", action="store_true",
                    help="only time the sketch factorisations: blocked Householder (geqrf) and the "
                         "communication-avoiding R factor (pamd_qr_r_f64), 5 calls each on the LS sketch shapes")
    ap.add_argument("--profile-ls", action="store_true",
                    help="only run 12 plain FP64 SPO and 12 SPOMixed solves on A 20000 x 200 (the workload of the LS "
                         "kernel traces)")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only run the batched solve of the 512 uneven problems twice (the workload of the batched "
                         "kernel trace)")
    ap.add_argument("--profile-learned", action="store_true",
                    help="only train on --train-matrices matrices and run 4 learned-sketch and 4 plain SJLT solves "
                         "on A 131072 x 512 (the workload of the learned-sketch kernel trace)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print("bench.py measures one GPU: parla_amd runs on a single MI355X and has no "
              "multi-GPU path; run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import parla_amd as pa
    from parla_amd import matmakers as mm
    from parla_amd import ops

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    ops.lib()
    if args.pmc_gemm:
        torch.manual_seed(0)
        A = torch.randn(8192, 4096, dtype=torch.float64, device=dev)
        S = torch.randn(4096, 266, dtype=torch.float64, device=dev)
        Y = ops.gemm(A, S)          # Y = A S
        ops.gemm(A.t(), Y)          # Z = A^T Y
        ops.gemm(Y.t(), A)          # B = Q^T A
        x = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        x.clone()                   # calibration: 512 MiB read, 512 MiB written
        torch.cuda.synchronize()
        return
    if args.pmc_bf16:
        torch.manual_seed(0)
        A = torch.randn(8192, 4096, dtype=torch.float64, device=dev)
        S = torch.randn(4096, 266, dtype=torch.float64, device=dev)
        kp, mpad, npad, _ = ops.split_layout(2, 8192, 266, 4096)
        ops.split_gemm(ops.split_bf16(A, 2), ops.split_bf16(S, 2, transpose=True, rows_out=npad))
        del A, S
        A = torch.randn(262144, 1024, dtype=torch.float64, device=dev)
        b = torch.randn(262144, dtype=torch.float64, device=dev)
        Q = ops.gauss_q(4096, 262144, 1, dev)
        ops.ozaki_sketch(Q, 4096, ops.OzakiSplit(A, b, 4), 1.0)
        torch.cuda.synchronize()
        return

    if args.profile_qr:
        # the sketches [S A | S b] of the LS workloads: C1 (800 x 201), the SJLT
        # and learned sketches of A 131072 x 512 (2048 / 4096 x 513), C3 (4096 x 1025)
        for (mq, nq) in [(800, 200), (2048, 512), (4096, 512), (4096, 1024)]:
            Mq = torch.randn(mq, nq + 1, dtype=torch.float64, device=dev)
            res = {}
            for name, fn in [("geqrf", ops.geqrf_), ("qr_r", ops.qr_r_caqr_)]:
                ts = []
                for it in range(6):
                    M = Mq.clone()
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    fn(M, nq)
                    e1.record()
                    torch.cuda.synchronize()
                    ts.append(e0.elapsed_time(e1))
                res[name] = float(np.median(ts[1:]))
            log(f"profile-qr: {mq} x {nq + 1}: geqrf {res['geqrf']:.3f} ms, qr_r {res['qr_r']:.3f} ms")
        return
    if args.profile_ls:
        Al_h, bl_h, _ = mm.config_ls_plain()
        Al, bl = torch.from_numpy(Al_h).to(dev), torch.from_numpy(bl_h).to(dev)
        spo = pa.SPO(pa.SkOpGA(backend="device"), 4)
        for i in range(12):
            spo(Al, bl, 0.0, 1e-12, 100, i, logging=False)
        torch.cuda.synchronize()
        log("profile-ls: 12 plain FP64 SPO solves done")
        # the chosen path on the same problem, to the metric's threshold
        chosen = pa.SPOMixed(4, npieces=4, backend="device")
        for i in range(12):
            chosen(Al, bl, 0.0, 1e-10, 100, i, logging=False)
        torch.cuda.synchronize()
        log("profile-ls: 12 SPOMixed solves done")
        return
    if args.profile_learned:
        import parla_amd.comps.sketchers.learned as lm
        m5, n5, d5 = 131072, 512, 4096
        data = lm.training_set(args.train_matrices, m5, n5, 50000, dev, log=log)
        net = lm.LeverageMLP(seed=0)
        net.fit(data, epochs=4, lr=2e-3, batch_rows=65536, log=log)
        del data
        A5, _ = lm.family_matrix(m5, n5, 777, dev)
        rng5 = np.random.default_rng(5)
        b5 = A5.double() @ torch.from_numpy(rng5.standard_normal(n5)).to(dev) + \
            1e-6 * torch.from_numpy(rng5.standard_normal(m5)).to(dev)
        learned = lm.SPOLearned(net, d=d5, kappa_max=30.0)
        plain5 = pa.SPO(pa.SkOpSJ(8, backend="device"), 4)
        for i in range(4):
            learned(A5, b5, 0.0, 1e-12, 200, i, logging=False)
            plain5(A5, b5, 0.0, 1e-12, 200, i, logging=False)
        torch.cuda.synchronize()
        log("profile-learned: 4 learned and 4 plain SJLT solves done")
        return
    if args.profile_batched:
        from parla_amd.drivers import batched_ls as bl_
        As4, bs4 = bl_.uneven_batch(args.batch_problems, mm.data_seed(4), dev)
        solver = bl_.BatchedSPO(cg_iters=20)
        for i in range(2):
            solver(As4, bs4, 1e-10, 1 + i)
        torch.cuda.synchronize()
        log("profile-batched: 2 batched solves done")
        return

    # ------------------------------------------------------------------ QB
    m, n, k, over = 8192, 4096, 256, 10
    log("generating the QB workload on the host (Haar factors by numpy QR)")
    A_host, s_true = mm.config_qb(m, n)
    A = torch.from_numpy(A_host).to(dev)
    qb = pa.QB1(pa.RF1(pa.RS1(pa.SkOpSJ(8, backend="device"), 4, pa.orth, 1)))
    out = {}

    def qb_step(seed):
        out["QB"] = qb(A, k + over, np.nan, seed)

    log(f"QB: {args.warmup} warmup + {args.steps} timed steps")
    t_qb = timed_steps(qb_step, args.steps, args.warmup, torch)
    qb_bytes = 6.0 * m * n * 8
    value = qb_bytes / t_qb / 1e9

    # errors on the device (FP64): no host BLAS between the timed legs
    Q, B = out["QB"]
    fro = float(torch.linalg.norm(A))
    qb_err = float(torch.linalg.norm(A - Q @ B)) / fro
    orth_err = float(torch.linalg.norm(Q.t() @ Q - torch.eye(Q.shape[1], dtype=Q.dtype, device=dev)))
    tail = float(np.sqrt(np.sum(s_true[k + over:] ** 2)) / np.sqrt(np.sum(s_true ** 2)))

    # roofline of the dominant kernel (the FP64 MFMA GEMM): per-launch events
    events = []
    ops.set_event_log(events)
    for i in range(3):
        qb_step(500 + i)
    torch.cuda.synchronize()
    ops.set_event_log(None)
    g = [e for e in events if e[0] == "gemm_f64" and e[1] > 1e10]  # the 5 passes over A (1.785e10 flop each)
    g_ms = [e[3].elapsed_time(e[4]) for e in g]
    avg_ms = float(np.mean(g_ms))
    avg_flops = float(np.mean([e[1] for e in g]))
    achieved = avg_flops / (avg_ms / 1e3) / 1e12

    # the chosen QB: the same chain on a split operator (A as BF16 pieces,
    # the five dense passes on BF16 MFMA); the split of A is inside the step
    split_extra = {}
    for npieces, label in ((2, "bf16x3"), (3, "bf16x6")):
        def qbs_step(seed, npieces=npieces):
            out["QBS"] = qb(pa.SplitOperator(A, npieces), k + over, np.nan, seed)

        t_s = timed_steps(qbs_step, args.steps, args.warmup, torch)
        Qs_, Bs_ = out["QBS"]
        err_s = float(torch.linalg.norm(A - Qs_ @ Bs_)) / fro
        ev = []
        ops.set_event_log(ev)
        qbs_step(600)
        torch.cuda.synchronize()
        ops.set_event_log(None)
        sg = [e for e in ev if e[0] == "split_gemm" and e[1] > 1e10]
        sg_ms = float(np.mean([e[3].elapsed_time(e[4]) for e in sg]))
        sg_tf = float(np.mean([e[1] for e in sg])) / (sg_ms / 1e3) / 1e12
        split_extra[f"qb_{label}"] = {
            "pieces": npieces, "ms": round(t_s * 1e3, 3), "speedup_vs_f64": round(t_qb / t_s, 3),
            "gb_per_s_f64_equiv": round(qb_bytes / t_s / 1e9, 2), "rel_err_fro": err_s,
            "split_gemm_avg_us": round(sg_ms * 1e3, 1), "split_gemm_bf16_tflops_executed": round(sg_tf, 1),
            "split_gemm_frac_bf16_peak": round(sg_tf / 2516.6, 4)}

    # SVD1 on the same workload
    svd = pa.SVD1(qb)

    def svd_step(seed):
        out["SVD"] = svd(A, k, np.nan, over, seed)

    t_svd = timed_steps(svd_step, max(3, args.steps // 2), 1, torch)
    s_est = out["SVD"][1].cpu().numpy()
    sv_rel = float(np.max(np.abs(s_est[:k] - s_true[:k]) / s_true[:k]))

    # ------------------------------------------------------------------ LS
    log("LS: plain FP64 SPO on A 20000 x 200")
    Al_h, bl_h, _ = mm.config_ls_plain()
    Al = torch.from_numpy(Al_h).to(dev)
    bl = torch.from_numpy(bl_h).to(dev)
    spo = pa.SPO(pa.SkOpGA(backend="device"), 4)

    def ls_step(seed):
        out["LS"] = spo(Al, bl, 0.0, 1e-12, 100, seed, logging=False)

    t_ls = timed_steps(ls_step, args.ls_steps, 2, torch)
    x_ls = out["LS"][0].cpu().numpy()
    import oracle as orc
    eta = orc.backward_error_extended(Al_h, bl_h, x_ls)
    _, lg = spo(Al, bl, 0.0, 1e-12, 100, 7, logging=True)
    ls_iters = int(lg.errors.size - 1)
    ml, nl, dl = Al_h.shape[0], Al_h.shape[1], 4 * Al_h.shape[1]
    # the algorithmic work of one solve as the survey of the reference counts it: S (device-drawn,
    # written and read once), S A on FP64 MFMA, Householder QR of [SA | Sb], and per LSQR
    # iteration the two products A v and A^T u (66 MB; the fused kernel reads A once instead)
    ls_roof = roofline_terms([
        ("gaussian_S_write_read", 2 * 8.0 * dl * ml, HBM_PEAK_GBS * 1e9),
        ("sketch_SA_f64", 2.0 * dl * ml * nl, FP64_PEAK_TFLOPS * 1e12),
        ("qr_of_sketch", 2.0 * dl * (nl + 1) ** 2 - 2.0 * (nl + 1) ** 3 / 3, FP64_PEAK_TFLOPS * 1e12),
        ("lsqr_two_passes_per_iter", ls_iters * 2 * 8.0 * ml * nl, HBM_PEAK_GBS * 1e9)], t_ls)
    ls_roof["dominant_kernels"] = "lsqr_tail_kernel (one workgroup, latency) and lsqr_pass_kernel (HBM); " \
                                  "see profiles/ls_plain_20000x200_kernel_stats.md"

    # the chosen path on the same problem: exact BF16 sketch + FP64 refinement
    # until eta <= 1e-10 (the metric's threshold)
    chosen = pa.SPOMixed(4, npieces=4, backend="device")

    def ls_chosen_step(seed):
        out["LSC"] = chosen(Al, bl, 0.0, 1e-10, 100, seed, logging=False)

    t_lsc = timed_steps(ls_chosen_step, args.ls_steps, 2, torch)
    eta_c = orc.backward_error_extended(Al_h, bl_h, out["LSC"][0].cpu().numpy())
    lgc = out["LSC"][1]
    # algorithmic work of one chosen solve: the integer Gaussian grid (BF16, written and read), the split of
    # [A | b] into 4 BF16 pieces (A read, pieces written), the 4 piece products on BF16 MFMA, the two Gram
    # matrices and the triangular solve of CholeskyQR2 and B = A R^-1 on FP64 MFMA (A read, B written in FP32),
    # and per refinement step one FP64 read of A (the compensated residual) plus one FP32 read of B per CG step
    res_reads = lgc.refinement_steps + 1
    ls_chosen_roof = roofline_terms([
        ("gaussian_grid_write_read", 2 * 2.0 * dl * ml, HBM_PEAK_GBS * 1e9),
        ("split_A_into_pieces", 8.0 * ml * (nl + 1) + 4 * 2.0 * ml * (nl + 1), HBM_PEAK_GBS * 1e9),
        ("exact_sketch_4_bf16_pieces", 4 * 2.0 * dl * ml * (nl + 1), BF16_PEAK_TFLOPS * 1e12),
        ("cholqr2_of_sketch", 2 * 2.0 * dl * nl * nl + dl * nl * nl, FP64_PEAK_TFLOPS * 1e12),
        ("B_eq_A_Rinv", 1.0 * ml * nl * nl, FP64_PEAK_TFLOPS * 1e12),
        ("B_read_A_write_f32", 8.0 * ml * nl + 4.0 * ml * nl, HBM_PEAK_GBS * 1e9),
        ("residual_passes_f64", res_reads * 8.0 * ml * nl, HBM_PEAK_GBS * 1e9),
        ("cg_normal_passes_f32", lgc.inner_iterations * 4.0 * ml * nl, HBM_PEAK_GBS * 1e9)], t_lsc)
    ls_chosen_roof["refinement_steps"] = int(lgc.refinement_steps)
    ls_chosen_roof["cg_iterations"] = int(lgc.inner_iterations)
    ls_chosen_roof["dominant_kernels"] = "normal_pass_kernel + its partial reduction and CG update per CG step " \
                                         "(three launches), the 200 x 200 triangular solves, potrf_la; see " \
                                         "profiles/ls_chosen_20000x200_kernel_stats.md"

    # ------------------------------------------------------------- mixed LS
    mixed_extra = {}
    if "mixed" in legs:
        import oracle as orc
        log("C3: generating A 262144 x 1024 (cond 1e8) on the host")
        A3h, b3h, _ = mm.config_ls_mixed()
        A3 = torch.from_numpy(A3h).to(dev)
        b3 = torch.from_numpy(b3h).to(dev)
        mixed = pa.SPOMixed(4, npieces=4, backend="device")

        def mixed_step(seed):
            out["MX"] = mixed(A3, b3, 0.0, 1e-12, 100, seed, logging=False)

        log("C3: mixed-precision SPO (exact BF16 sketch, FP32 operator, FP64 refinement)")
        t_mx = timed_steps(mixed_step, 3, 1, torch)
        x_mx = out["MX"][0].cpu().numpy()
        _, lgm = mixed(A3, b3, 0.0, 1e-12, 100, 7, logging=True)
        plain3 = pa.SPO(pa.SkOpGA(backend="device"), 4)

        def plain3_step(seed):
            out["P3"] = plain3(A3, b3, 0.0, 1e-12, 100, seed, logging=False)

        log("C3: plain FP64 SPO on the same problem")
        t_p3 = timed_steps(plain3_step, 2, 1, torch)
        x_p3 = out["P3"][0].cpu().numpy()
        _, lgp3 = plain3(A3, b3, 0.0, 1e-12, 100, 7, logging=True)
        log("C3: backward errors in extended precision on the host")
        eta_mx = orc.backward_error_extended(A3h, b3h, x_mx)
        eta_p3 = orc.backward_error_extended(A3h, b3h, x_p3)
        m3, n3, d3 = A3h.shape[0], A3h.shape[1], 4096
        steps3 = int(len(lgm.errors) - 1)
        qr3 = 2.0 * d3 * (n3 + 1) ** 2 - 2.0 * (n3 + 1) ** 3 / 3
        mx_roof = roofline_terms([
            ("exact_sketch_4_bf16_pieces", 4 * 2.0 * d3 * m3 * (n3 + 1), BF16_PEAK_TFLOPS * 1e12),
            ("qr_of_sketch", qr3, FP64_PEAK_TFLOPS * 1e12),
            ("B_eq_A_Rinv_f64", float(m3) * n3 * n3, FP64_PEAK_TFLOPS * 1e12),
            ("B_written_f32", 4.0 * m3 * n3, HBM_PEAK_GBS * 1e9),
            ("residual_passes_f64", (steps3 + 1) * 8.0 * m3 * n3, HBM_PEAK_GBS * 1e9),
            ("cg_normal_passes_f32", int(lgm.inner_iterations) * 4.0 * m3 * n3, HBM_PEAK_GBS * 1e9)], t_mx)
        mx_roof["dominant_kernels"] = "oz_gemm_kernel (BF16 MFMA), trsm leaves + gemm_f64 (B = A R^-1), " \
                                      "normal_pass_kernel (HBM); see profiles/ls_mixed_262144x1024_kernel_stats.md"
        iters_p3 = int(lgp3.errors.size - 1)
        p3_roof = roofline_terms([
            ("gaussian_S_write_read", 2 * 8.0 * d3 * m3, HBM_PEAK_GBS * 1e9),
            ("sketch_SA_f64", 2.0 * d3 * m3 * n3, FP64_PEAK_TFLOPS * 1e12),
            ("qr_of_sketch", qr3, FP64_PEAK_TFLOPS * 1e12),
            ("lsqr_two_passes_per_iter", iters_p3 * 2 * 8.0 * m3 * n3, HBM_PEAK_GBS * 1e9)], t_p3)
        cpu3 = None
        if not args.no_cpu:
            # bounded CPU sample: the same family at 1/8 of the rows (the sketch, the LSQR passes and
            # the draws scale with m; the 4096 x 1024 QR does not, so the extrapolation is generous)
            m3s = m3 // 8
            log(f"C3: CPU oracle SPO on a {m3s} x {n3} sample of the family")
            A3s, b3s, _ = mm.config_ls_mixed(m=m3s)
            ts3 = []
            for i in range(3):
                t0 = time.perf_counter()
                x3s, _ = orc.spo(A3s, b3s, 0.0, 1e-12, 100, np.random.default_rng(i), sketch="gaussian",
                                 sampling_factor=4, logging=False)
                ts3.append(time.perf_counter() - t0)
            t3s = float(np.median(ts3))
            cpu3 = {"sample": f"A {m3s}x{n3} of the same family (cond 1e8), Gaussian d=4096, LSQR tol 1e-12; "
                              "median of 3", "sample_s": round(t3s, 3),
                    "full_size_solves_per_s_est": round(1.0 / (8 * t3s), 5), "cores": cpu_threads(),
                    "cpu": cpu_model(), "sample_eta": float(orc.backward_error(A3s, b3s, x3s))}
            del A3s, b3s
        mixed_extra = {
            "ls_mixed_workload": "A 262144x1024 f64, s geometric 1..1e-8, b = A x* + 1e-8 N(0,1); exact BF16 "
                                 "Gaussian sketch d=4096 (4 BF16 pieces of A), FP64 QR, B = A R^-1 stored FP32, "
                                 "FP64 refinement with CG on B^T B to eta <= 1e-12 or the FP64 floor",
            "ls_mixed_solves_per_s": round(1.0 / t_mx, 3), "ls_mixed_ms": round(t_mx * 1e3, 3),
            "ls_mixed_eta": eta_mx, "ls_mixed_refinement_etas": [float(e) for e in lgm.errors],
            "ls_mixed_inner_cg_iters": int(lgm.inner_iterations),
            "ls_mixed_phase_ms": {"sketch": round(lgm.time_sketch * 1e3, 3), "factor": round(lgm.time_factor * 1e3, 3),
                                  "presolve": round(lgm.time_presolve * 1e3, 3),
                                  "refine": round(lgm.time_iterate * 1e3, 3)},
            "ls_plain_f64_c3_solves_per_s": round(1.0 / t_p3, 3), "ls_plain_f64_c3_ms": round(t_p3 * 1e3, 3),
            "ls_plain_f64_c3_eta": eta_p3, "ls_plain_f64_c3_lsqr_iters": iters_p3,
            "ls_mixed_speedup_vs_plain_f64": round(t_p3 / t_mx, 3),
            "ls_mixed_roofline": mx_roof, "ls_plain_f64_c3_roofline": p3_roof,
            "ls_c3_cpu_baseline": cpu3,
        }
        del A3, b3

    # ---------------------------------------------------------- learned LS
    learned_extra = {}
    if "learned" in legs:
        import parla_amd.comps.sketchers.learned as lm
        m5, n5, d5 = 131072, 512, 4096
        log(f"C5: training set of {args.train_matrices} matrices {m5} x {n5} (labels from FP64 QR)")
        t0 = time.perf_counter()
        data = lm.training_set(args.train_matrices, m5, n5, 50000, dev, log=log)
        torch.cuda.synchronize()
        t_data = time.perf_counter() - t0
        net = lm.LeverageMLP(seed=0)
        t0 = time.perf_counter()
        losses = net.fit(data, epochs=4, lr=2e-3, batch_rows=65536, log=log)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        rows_trained = sum(F.shape[0] for F, _ in data)
        del data
        A5, DG5 = lm.family_matrix(m5, n5, 777, dev)
        lev5 = torch.exp(lm.leverage_labels(DG5).double()) * n5 / m5
        del DG5
        rng5 = np.random.default_rng(5)
        b5 = A5.double() @ torch.from_numpy(rng5.standard_normal(n5)).to(dev) + \
            1e-6 * torch.from_numpy(rng5.standard_normal(m5)).to(dev)
        G5 = lm.presketch(n5, 4242, dev)
        F5 = lm.features(A5, G5)
        yl = net.predict(F5)
        probs = {"learned": lm.sampling_probabilities(yl, 0.1, n5),
                 "uniform": torch.full((m5,), 1.0 / m5, dtype=torch.float64, device=dev)}
        rn5 = torch.sum(F5.double() ** 2, dim=1)
        probs["rownorm"] = 0.9 * rn5 / rn5.sum() + 0.1 / m5
        probe = lm.SPOLearned(net, d=d5, kappa_max=1e30, lanczos_steps=12)  # closer kappa estimates for the report
        kap, tv = {}, {}
        for name, p5 in probs.items():
            kap[name] = probe.sketch_factor(A5, b5, np.random.default_rng(2), probs=p5)[2]
            tv[name] = float(0.5 * torch.sum(torch.abs(p5 - lev5 / n5)).item())
        learned = lm.SPOLearned(net, d=d5, kappa_max=30.0)
        plain5 = pa.SPO(pa.SkOpSJ(8, backend="device"), 4)

        def learned_step(seed):
            out["L5"] = learned(A5, b5, 0.0, 1e-12, 200, seed, logging=False)

        def plain5_step(seed):
            out["S5"] = plain5(A5, b5, 0.0, 1e-12, 200, seed, logging=False)

        log("C5: learned-sketch SPO vs plain SJLT SPO")
        t_l5 = timed_steps(learned_step, 3, 1, torch)
        t_s5 = timed_steps(plain5_step, 3, 1, torch)
        x_l5, lg5 = out["L5"]
        x_s5 = out["S5"][0]
        _, lgs5 = plain5(A5, b5, 0.0, 1e-12, 200, 7, logging=True)
        learned_extra = {
            "learned_workload": f"A {m5}x{n5} FP32 (Gaussian rows, 1% x100, right-scaled to cond 1e6; the FP64 "
                                f"solvers widen it inside the timed solve), features A G16, "
                                f"MLP 16-64-64-1 BF16 weights, row sampling d={d5}, SJLT fallback above kappa 30",
            "learned_training": {"matrices": args.train_matrices, "rows": int(rows_trained),
                                 "data_s": round(t_data, 2), "train_s": round(t_train, 2),
                                 "epoch_losses": [round(v, 4) for v in losses]},
            "learned_kappa_est": {k: round(v, 3) for k, v in kap.items()},
            "learned_tv_to_leverage": {k: round(v, 4) for k, v in tv.items()},
            "learned_fallback": bool(getattr(lg5, "fallback", False)),
            "learned_ls_ms": round(t_l5 * 1e3, 3), "plain_sjlt_ls_ms": round(t_s5 * 1e3, 3),
            "learned_lsqr_iters": int(lg5.errors.size - 1) if lg5.errors is not None else None,
            "plain_sjlt_lsqr_iters": int(lgs5.errors.size - 1),
            "learned_rel_diff_vs_plain_sjlt": float((torch.linalg.norm(x_l5 - x_s5) / torch.linalg.norm(x_s5)).item()),
        }
        del A5, b5, F5

    # ---------------------------------------------------------- batched LS
    batched_extra = {}
    if "batched" in legs:
        import oracle as orc
        from parla_amd.drivers import batched_ls as bl
        P4 = args.batch_problems
        log(f"C4: generating {P4} uneven problems on the device")
        As4, bs4 = bl.uneven_batch(P4, mm.data_seed(4), dev)
        solver = bl.BatchedSPO(cg_iters=20)
        solver(As4, bs4, 1e-10, 1)  # warm-up: tunes launch shapes, builds the graph once
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs4, lg4 = solver(As4, bs4, 1e-10, 2)
        torch.cuda.synchronize()
        t_b4 = time.perf_counter() - t0
        # the plain FP64 path, one problem at a time, stopped at the metric's threshold: the loosest
        # LSQR tolerance for which every sampled solve reaches eta <= 1e-10 (checked by the
        # compensated residual), then timed at that tolerance
        n_s4 = min(64, P4)
        idx = np.random.default_rng(3).choice(P4, n_s4, replace=False)
        A4d = {int(i): As4[i].double() for i in idx}
        plain4 = pa.SPO(pa.SkOpSJ(8, backend="device"), 4)
        # plain FP64 sketch-and-precondition reaches a backward error of about u kappa(A): on the
        # worst-conditioned problems (kappa up to 1e10) no tolerance gets it to 1e-10, so the
        # tolerance kept is the loosest that reaches the threshold on every sampled problem, else
        # the one that reaches it on the most, and the count is reported
        tol4, etas_p4, best = None, None, -1
        for cand in (1e-10, 1e-11, 1e-12, 1e-13, 1e-14):
            log(f"C4: plain FP64 SPO on {n_s4} sampled problems at LSQR tol {cand:g}")
            e_c = []
            for i in idx:
                xi, _ = plain4(A4d[int(i)], bs4[i], 0.0, cand, 400, int(i), logging=False)
                e_c.append(device_eta(ops, torch, A4d[int(i)], bs4[i], xi))
            meet = int(np.sum(np.array(e_c) <= 1e-10))
            if meet > best:
                tol4, etas_p4, best = cand, e_c, meet
            if meet == n_s4:
                break
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in idx:
            plain4(A4d[int(i)], bs4[i], 0.0, tol4, 400, int(i), logging=False)
        torch.cuda.synchronize()
        t_p4 = (time.perf_counter() - t0) / len(idx)
        del A4d
        etas = []
        for i in idx[:6]:
            etas.append(orc.backward_error_extended(As4[i].double().cpu().numpy(), bs4[i].cpu().numpy(),
                                                    xs4[i].cpu().numpy()))
        mn4 = np.array([A.shape[0] * A.shape[1] for A in As4], dtype=np.float64)
        work = int(mn4.sum())
        # algorithmic HBM bytes of the batched solve: the FP32 A_i read by the sketch, by the
        # residual at x = 0, and per refinement step by 20 normal passes and one residual pass
        # of every problem still active
        et4 = lg4["etas"]
        pass_elems = 2.0 * mn4.sum() + sum(float(mn4[et4[s] > 1e-10].sum()) * 21 for s in range(len(et4) - 1))
        b4_roof = roofline_terms([("fp32_passes_over_A_i", 4.0 * pass_elems, HBM_PEAK_GBS * 1e9)], t_b4)
        b4_roof["dominant_kernels"] = "bt_normal_kernel (HBM); see profiles/ls_batched_512_kernel_stats.md"
        batched_extra = {
            "batched_workload": f"{P4} problems, m_i U[2000,50000], n_i U[32,512], m_i >= 8 n_i, cond log-U[1e2,1e10], "
                                "FP32 A_i; SJLT d_i = 4 n_i, shifted CholeskyQR3, graph-replayed FP64 refinement "
                                "(CG 20 per step) to eta <= 1e-10",
            "batched_solves_per_s": round(P4 / t_b4, 2), "batched_ms_per_batch": round(t_b4 * 1e3, 2),
            "batched_setup_ms": round(lg4["setup_s"] * 1e3, 2), "batched_solve_ms": round(lg4["solve_s"] * 1e3, 2),
            "batched_refine_steps": int(lg4["refine_steps"]),
            "batched_max_eta_device": float(lg4["etas"][-1].max()),
            "batched_eta_extended_sample": [float(e) for e in etas],
            "batched_sum_mn": int(work),
            "plain_f64_one_at_a_time_solves_per_s": round(1.0 / t_p4, 2),
            "plain_f64_sample": f"{n_s4} problems, SJLT SPO at LSQR tol {tol4:g} (the loosest of 1e-10 ... 1e-14 "
                                "at which every sampled eta <= 1e-10, else the one meeting it most often)",
            "plain_f64_lsqr_tol": tol4, "plain_f64_max_eta": float(max(etas_p4)),
            "plain_f64_meet_threshold": f"{best} of {n_s4}",
            "plain_f64_etas": [float(e) for e in etas_p4],
            "batched_speedup_vs_plain_f64": round((P4 / t_b4) * t_p4, 2),
            "batched_roofline": b4_roof,
            "batched_tuned_shapes": {f"{k[0]}/npl{k[1]}": list(v) for k, v in lg4["tuned"].items()},
        }
        del As4, bs4, xs4

    # ------------------------------------------------------------------ CPU
    cpu = None
    extra_cpu = {}
    if not args.no_cpu:
        import oracle as orc
        threads = cpu_threads()
        log(f"CPU oracle QB1 (numpy/scipy, {threads} BLAS threads)")
        rso = lambda M, kk, rng: orc.rs1(M, kk, rng, lambda r_, c_, g_: orc.sjlt_operator(r_, c_, g_, 8),
                                         4, orc.orth, 1)
        rf = lambda M, kk, tol, rng: orc.rf1(M, kk, tol, rng, rso)
        ts = []
        for i in range(4):
            t0 = time.perf_counter()
            Qo, Bo = orc.qb1(A_host, k + over, np.nan, np.random.default_rng(i), rf)
            ts.append(time.perf_counter() - t0)
        t_cpu = float(np.median(ts[1:]))
        cpu = {"value": qb_bytes / t_cpu / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
               "cpu": cpu_model(),
               "sample": "full QB1 workload (8192 x 4096, k = 266, SJLT drawn with numpy), "
                         "1 warm-up + median of 3"}
        extra_cpu["qb_oracle_rel_err"] = float(np.linalg.norm(A_host - Qo @ Bo) / fro)
        # SVD1's singular values from the oracle's QB: the same algorithmic error as the device's
        so = np.linalg.svd(Bo, compute_uv=False)[:k]
        extra_cpu["svd1_oracle_max_rel_err_top256"] = float(np.max(np.abs(so - s_true[:k]) / s_true[:k]))
        log("CPU oracle SPO on A 20000 x 200")
        ts = []
        for i in range(4):
            t0 = time.perf_counter()
            xo, _ = orc.spo(Al_h, bl_h, 0.0, 1e-12, 100, np.random.default_rng(i), sketch="gaussian",
                            sampling_factor=4, logging=False)
            ts.append(time.perf_counter() - t0)
        t_cpu_ls = float(np.median(ts[1:]))
        extra_cpu["ls_cpu_solves_per_s"] = 1.0 / t_cpu_ls
        extra_cpu["ls_cpu_eta"] = float(orc.backward_error(Al_h, bl_h, xo))

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_qb * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Haar factors, prescribed spectrum)",
        "config": {"workload": "QB1 (RS1 num_pass=4 -> RF1 -> B = Q^T A), A 8192x4096 f64, "
                               "k = 256 + 10, SJLT 8 nnz drawn on the device", "m": m, "n": n,
                   "k": k + over, "passes_over_A": 6, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "gemm_f64_kernel (FP64 MFMA)",
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": gemm_traffic(),
                     "avg_launch_ms": round(avg_ms, 4), "flops_per_launch": avg_flops,
                     "launches_timed": len(g)},
        "cpu_baseline": cpu,
        "extra": {
            "qb_rel_err_fro": qb_err, "qb_orth_err": orth_err, "qb_tail_bound": tail,
            **split_extra,
            "qb_split_workload": "same QB1 chain on SplitOperator(A, pieces): the five dense passes as BF16 MFMA "
                                 "piece products (2 pieces: 3 products, 16 bits; 3 pieces: 6 products, 24 bits) "
                                 "with FP32 accumulation; sparse first pass, CholeskyQR and the split of A in "
                                 "FP64 inside the timed step",
            "svd1_ms": round(t_svd * 1e3, 3), "svd1_max_rel_err_top256": sv_rel,
            "ls_plain_f64_solves_per_s": round(1.0 / t_ls, 3), "ls_plain_f64_ms": round(t_ls * 1e3, 3),
            "ls_plain_f64_eta": eta, "ls_plain_f64_iters": ls_iters, "ls_plain_f64_roofline": ls_roof,
            "ls_chosen_solves_per_s": round(1.0 / t_lsc, 3), "ls_chosen_ms": round(t_lsc * 1e3, 3),
            "ls_chosen_eta": eta_c, "ls_chosen_roofline": ls_chosen_roof,
            "ls_chosen_workload": "same problem; SPOMixed: exact BF16 Gaussian sketch d=800 (4 pieces), FP64 QR, "
                                  "B = A R^-1 in FP32, FP64 refinement to eta <= 1e-10",
            "ls_chosen_speedup_vs_plain_f64": round(t_ls / t_lsc, 3),
            "ls_workload": "SPO qr, A 20000x200 f64, s geometric 1..1e-6, Gaussian d=800 drawn on the "
                           "device, LSQR tol 1e-12, logging off",
            **extra_cpu,
            **mixed_extra,
            **learned_extra,
            **batched_extra,
        },
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
