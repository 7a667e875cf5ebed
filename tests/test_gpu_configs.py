"""Parity at the full workload sizes.

* SVD1 on the QB workload (A 8192 x 4096, k = 256 + 10, two power iterations,
  SJLT with 8 nonzeros) with numpy draws, against the oracle's svd1 on the
  same draws: the singular values agree to 1e-8 sigma_1, and both carry the
  same error against the prescribed spectrum (the error is the algorithm's).
* The batched solve of 512 uneven problems (m_i in [2000, 50000], n_i in
  [32, 512], cond up to 1e10): every device backward error <= 1e-10, and an
  extended-precision check of a sample.
"""
import sys
import time

import numpy as np
import pytest
import scipy.linalg as la

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-configs] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pa():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import parla_amd
    from parla_amd import ops
    ops.lib()
    return parla_amd


def test_svd1_qb_workload_matches_oracle(pa):
    import oracle as orc
    from parla_amd import matmakers as mm
    m, n, k, over = 8192, 4096, 256, 10
    t0 = time.time()
    A, s_true = mm.config_qb(m, n)
    _say(f"SVD1 8192x4096: workload generated in {time.time() - t0:.1f} s")
    qb = pa.QB1(pa.RF1(pa.RS1(pa.SkOpSJ(8), 4, pa.orth, 1)))
    U, s, Vh = pa.SVD1(qb)(torch.from_numpy(A).cuda(), k, np.nan, over, np.random.default_rng(2026))
    s = s.cpu().numpy()
    _say("SVD1 on the device done; oracle svd1 on the host")
    rso = lambda M, kk, rng: orc.rs1(M, kk, rng, lambda r_, c_, g_: orc.sjlt_operator(r_, c_, g_, 8), 4, orc.orth, 1)
    rf = lambda M, kk, tol, rng: orc.rf1(M, kk, tol, rng, rso)
    qbo = lambda M, kk, tol, rng: orc.qb1(M, kk, tol, rng, rf)
    Uo, so, Vho = orc.svd1(A, k, np.nan, over, np.random.default_rng(2026), qbo)
    assert s.size == so.size == k
    assert np.max(np.abs(s - so)) <= 1e-8 * so[0]
    err = np.max(np.abs(s - s_true[:k]) / s_true[:k])
    err_o = np.max(np.abs(so - s_true[:k]) / s_true[:k])
    _say(f"top-{k} relative error vs the prescribed spectrum: device {err:.4e}, oracle {err_o:.4e}")
    # 1e-8 sigma_1 is 7e-4 of sigma_256 = 256^-2: the tail agrees far closer
    np.testing.assert_allclose(s, so, rtol=1e-6)
    assert abs(err - err_o) <= 1e-4 * err_o
    # the leading singular vectors agree too (well separated at the top)
    Ut = U.cpu().numpy()
    for j in range(8):
        assert abs(abs(Ut[:, j] @ Uo[:, j]) - 1.0) <= 1e-8, j


def test_batched_full_workload_backward_errors(pa):
    import oracle as orc
    from parla_amd import matmakers as mm
    from parla_amd.drivers import batched_ls as bl
    dev = torch.device("cuda")
    t0 = time.time()
    As, bs = bl.uneven_batch(512, mm.data_seed(4), dev)
    torch.cuda.synchronize()
    _say(f"512 uneven problems generated in {time.time() - t0:.1f} s")
    xs, log = bl.BatchedSPO(cg_iters=20)(As, bs, 1e-10, 1)
    etas = log["etas"][-1]
    _say(f"{log['refine_steps']} refinement steps, max device eta {etas.max():.3e}")
    assert len(xs) == 512 and etas.size == 512
    assert np.all(etas <= 1e-10)
    idx = np.random.default_rng(5).choice(512, 8, replace=False)
    for i in idx:
        A = As[i].double().cpu().numpy()
        b = bs[i].cpu().numpy()
        eta = orc.backward_error_extended(A, b, xs[i].cpu().numpy())
        assert eta <= 1e-10, (int(i), A.shape, eta)
