"""Learned leverage-score sketch on the device: MLP forward/backward/Adam
against the oracle, the gather, and what the trained network gains over the
same row-sampling method without it (uniform or row-norm probabilities)."""
import numpy as np
import pytest
import torch

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-learned] {msg}", flush=True)


@pytest.fixture(scope="module")
def lm():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops
    ops.lib()
    import parla_amd.comps.sketchers.learned as lm_
    return lm_


def test_nparams(lm):
    from parla_amd import ops
    assert ops.lib().pamd_mlp_nparams() == lm.NPARAM == orc.MLP_NPARAM


def test_mlp_forward_matches_oracle(lm):
    _say("mlp forward")
    rng = np.random.default_rng(1)
    theta = rng.standard_normal(lm.NPARAM) * 0.3
    net = lm.LeverageMLP()
    net.set_theta(theta)
    X = rng.standard_normal((1000, 16)).astype(np.float32)
    y = net.predict(torch.from_numpy(X).cuda()).cpu().numpy()
    ref, _ = orc.mlp_forward(orc.mlp_effective(theta), X.astype(np.float64))
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


@pytest.mark.parametrize("rows", [1, 300, 5000])
def test_mlp_backward_matches_oracle(lm, rows):
    _say(f"mlp backward rows={rows}")
    rng = np.random.default_rng(rows)
    theta = rng.standard_normal(lm.NPARAM) * 0.3
    net = lm.LeverageMLP()
    net.set_theta(theta)
    X = rng.standard_normal((rows, 16)).astype(np.float32)
    t = rng.standard_normal(rows).astype(np.float32)
    g, loss = net.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda())
    lref, gref = orc.mlp_loss_grad(orc.mlp_effective(theta), X.astype(np.float64), t.astype(np.float64))
    np.testing.assert_allclose(g.cpu().numpy(), gref, rtol=2e-4, atol=2e-5 * np.abs(gref).max())
    assert abs(loss.item() - lref) <= 1e-4 * lref


def test_adam_step_matches_oracle(lm):
    _say("adam")
    rng = np.random.default_rng(5)
    theta = rng.standard_normal(lm.NPARAM).astype(np.float32)
    net = lm.LeverageMLP()
    net.set_theta(theta)
    m = np.zeros(lm.NPARAM, np.float32)
    v = np.zeros(lm.NPARAM, np.float32)
    w = theta.copy()
    for step in range(1, 4):
        g = rng.standard_normal(lm.NPARAM).astype(np.float32)
        net.adam(torch.from_numpy(g).cuda(), lr=1e-2)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        w = w - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(net.theta(), w, rtol=1e-5, atol=1e-6)
    # the published BF16 weights are the round-to-nearest-even of the masters
    wb = net.wb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    p = orc.mlp_unpack(net.theta())
    masters = np.concatenate([p["W1"].ravel(), p["W2"].ravel(), p["w3"].ravel()]).astype(np.float32)
    np.testing.assert_array_equal(wb.view(np.float32), orc.bf16_round(masters))


def test_row_gather(lm):
    _say("row gather")
    rng = np.random.default_rng(6)
    A = rng.standard_normal((500, 37))
    idx = rng.integers(0, 500, 120)
    sc = rng.uniform(0.5, 2, 120)
    out = torch.empty((120, 37), dtype=torch.float64, device="cuda")
    lm.row_gather(torch.from_numpy(A).cuda(), idx, sc, out)
    np.testing.assert_array_equal(out.cpu().numpy(), sc[:, None] * A[idx])


def test_leverage_labels_match_scipy_qr(lm):
    _say("leverage labels")
    rng = np.random.default_rng(7)
    A, _ = orc.coherent_matrix(4000, 48, rng)
    t = lm.leverage_labels(torch.from_numpy(A).cuda()).cpu().numpy()
    ref = np.log(orc.leverage_scores(A) * 4000 / 48)
    np.testing.assert_allclose(t, ref, rtol=0, atol=1e-4)


def test_trained_network_gains_over_the_method_without_it(lm):
    """Train on a small version of the family (the device's own fwd/bwd/Adam),
    then compare kappa(A R^{-1}) of the row-sampling sketch driven by the
    network, by uniform probabilities (the method without the network) and by
    the pre-sketch row norms; the learned sketch's LS solution is judged by
    the plain SJLT SPO."""
    import parla_amd as pa
    dev = torch.device("cuda")
    m, n, d = 16384, 64, 512
    _say("training set")
    data = lm.training_set(24, m, n, 100, dev)
    net = lm.LeverageMLP(seed=3)
    losses = net.fit(data, epochs=8, lr=3e-3, batch_rows=8192)
    _say(f"losses {losses}")
    assert losses[-1] < 0.5 * losses[0]
    A, DG = lm.family_matrix(m, n, 999, dev)
    lev = torch.exp(lm.leverage_labels(DG).double()) * n / m
    rng = np.random.default_rng(1)
    assert A.dtype == torch.float32  # the family's data type
    b = A.double() @ torch.from_numpy(rng.standard_normal(n)).to(dev) + 1e-6 * torch.from_numpy(rng.standard_normal(m)).to(dev)
    G = lm.presketch(n, 4242, dev)
    F = lm.features(A, G)
    p_learned = lm.sampling_probabilities(net.predict(F), 0.1, n)
    p_uniform = torch.full((m,), 1.0 / m, dtype=torch.float64, device=dev)
    rn = torch.sum(F.double() ** 2, dim=1)
    p_rownorm = 0.9 * rn / rn.sum() + 0.1 / m
    alg = lm.SPOLearned(net, d=d, kappa_max=1e30, lanczos_steps=12)
    kap = {}
    for name, p in [("learned", p_learned), ("uniform", p_uniform), ("rownorm", p_rownorm)]:
        R, z, k = alg.sketch_factor(A, b, np.random.default_rng(2), probs=p)
        kap[name] = k
    # total-variation distance of the sampling distributions to the leverage one
    tv = {k: float(0.5 * torch.sum(torch.abs(p - lev / n)).item())
          for k, p in [("learned", p_learned), ("uniform", p_uniform), ("rownorm", p_rownorm)]}
    _say(f"kappa {kap}  TV to leverage {tv}")
    assert kap["learned"] < kap["uniform"] / 3
    assert tv["learned"] < tv["uniform"]
    # the learned sketch's solution, judged by the plain SJLT method
    alg = lm.SPOLearned(net, d=d, kappa_max=50.0)
    x, log = alg(A, b, 0.0, 1e-12, 200, 5)
    assert not log.fallback
    xs, _ = pa.SPO(pa.SkOpSJ(8), 4)(A, b, 0.0, 1e-12, 200, 5)
    assert torch.linalg.norm(x - xs) <= 1e-6 * torch.linalg.norm(xs)
    # a threshold below the estimate switches to the SJLT fallback
    x2, log2 = lm.SPOLearned(net, d=d, kappa_max=1.0)(A, b, 0.0, 1e-12, 200, 5)
    assert log2.fallback
    assert torch.linalg.norm(x2 - xs) <= 1e-6 * torch.linalg.norm(xs)
