"""The asynchronous LSQR entry point (pamd_lsqr_launch_f64): launches only, no
host read-back, so it can be captured in a HIP graph.  It must give the
blocking entry point's (pamd_lsqr_run_f64) results bit for bit -- the same
kernels run in the same order, the extra launches after the stopping test
return at once -- and a captured graph must reproduce them on every replay.
"""
import sys

import numpy as np
import pytest
import scipy.linalg as la

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-lsqr-launch] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops
    ops.lib()
    return ops


def _problem(m, n, seed, d):
    """A with singular values spread over 1e4 and R from the QR of a d x m
    Gaussian sketch of it (the SPO set-up), on the device."""
    rng = np.random.default_rng(seed)
    U = la.qr(rng.standard_normal((m, n)), mode="economic")[0]
    V = la.qr(rng.standard_normal((n, n)))[0]
    A = (U * np.logspace(0, -4, n)) @ V.T
    b = A @ rng.standard_normal(n) + 1e-3 * rng.standard_normal(m)
    S = rng.standard_normal((d, m)) / np.sqrt(d)
    R = la.qr(S @ A, mode="r")[0][:n]
    dev = torch.device("cuda")
    return (torch.from_numpy(A).to(dev), torch.from_numpy(b).to(dev), torch.from_numpy(np.ascontiguousarray(R)).to(dev))


FIELDS = ("istop", "itn", "r1norm", "r2norm", "anorm", "acond", "arnorm", "xnorm", "n_log")


@pytest.mark.parametrize("m,n,d,it,x0,calc_var", [(20000, 200, 800, 120, False, False),
                                                  (3000, 64, 256, 120, True, True),
                                                  (6000, 1950, 3900, 40, False, False)])
def test_launch_equals_blocking_run(ops, m, n, d, it, x0, calc_var):
    """Bit-identical x, x_orig, var, the arnorms log and every result field,
    for runs that converge and for one stopped by the iteration limit (n =
    1950, past the LDS substitution vector: the workspace path)."""
    _say(f"launch vs run m={m} n={n} x0={x0} calc_var={calc_var}")
    A, b, R = _problem(m, n, m + n, d)
    z0 = torch.full((n,), 0.01, dtype=torch.float64, device=A.device) if x0 else None
    kw = dict(R=R, x0=z0, atol=1e-12, btol=1e-12, iter_lim=it, calc_var=calc_var)
    x, xo, var, arn, res = ops.lsqr_run(A, b, want_orig=True, **kw)
    bufs = ops.lsqr_launch_buffers(m, n, it, A.device, want_orig=True)
    ops.lsqr_launch(A, b, out=bufs, **kw)
    torch.cuda.synchronize()
    x2, xo2, var2, arn2, resb, _ = bufs
    r2 = ops.lsqr_result_from_device(resb)
    _say(f"istop {res.istop} itn {res.itn}")
    for f in FIELDS:
        assert getattr(r2, f) == getattr(res, f), f
    assert torch.equal(x, x2) and torch.equal(xo, xo2)
    if calc_var:
        assert torch.equal(var, var2)
    k = res.itn
    assert torch.equal(arn[:k], arn2[:k])
    if n < 1000:
        # these converge well inside the limit, and R^{-1} x solves the LS problem
        assert res.istop in (1, 2) and res.itn < it
        xl = la.lstsq(A.cpu().numpy(), b.cpu().numpy())[0]
        assert np.linalg.norm(xo.cpu().numpy() - xl) <= 1e-6 * np.linalg.norm(xl)
    else:
        assert res.istop in (1, 2, 7)


def test_captured_graph_replays_equal_eager(ops):
    """Captured once, replayed three times with unrelated eager work and
    synchronisations between the replays: every replay equals the eager
    launch bit for bit."""
    _say("lsqr launch graph replays")
    m, n, it = 20000, 200, 120
    A, b, R = _problem(m, n, 11, 800)
    kw = dict(R=R, atol=1e-12, btol=1e-12, iter_lim=it)
    eager = ops.lsqr_launch(A, b, out=ops.lsqr_launch_buffers(m, n, it, A.device, want_orig=True), **kw)
    torch.cuda.synchronize()
    ref = [t.clone() for t in eager[:5] if t is not None]
    bufs = ops.lsqr_launch_buffers(m, n, it, A.device, want_orig=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ops.lsqr_launch(A, b, out=bufs, **kw)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        ops.lsqr_launch(A, b, out=bufs, **kw)
    for rep in range(3):
        for t in bufs[:5]:
            if t is not None:
                t.fill_(0)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=A.device).sum().item()  # unrelated eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        got = [t for t in bufs[:5] if t is not None]
        for a, r in zip(got, ref):
            assert torch.equal(a, r), rep
