"""The device's counter-based draws against the oracle's restatement of them
(no counterpart in the reference, which draws with numpy): the Gaussian
sketch fill, the SJLT column draws and the weighted row sampling.  Integer
outcomes (rows, signs, sampled indices) must be identical; Gaussian values
agree to the accuracy of the two log / sin / cos implementations.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[gpu-draws] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops
    ops.lib()
    return ops


@pytest.mark.parametrize("rows,cols,seed,offset", [(800, 2000, 12345, 0), (3, 7, 2 ** 63 + 5, 11), (257, 129, 1, 2 ** 33)])
def test_gaussian_fill_matches_oracle(ops, rows, cols, seed, offset):
    import oracle as orc
    _say(f"gaussian fill {rows}x{cols} seed {seed} offset {offset}")
    X = ops.gaussian_fill(rows, cols, 0.5, seed, torch.device("cuda"), offset=offset).cpu().numpy()
    R = orc.device_gaussian_fill(rows, cols, 0.5, seed, offset)
    # |z| < 9; the two libraries' log and sin / cos differ by a few ulp, and
    # numpy rounds 2 pi u before its sin / cos (about 7e-16 absolute)
    np.testing.assert_allclose(X, R, rtol=0, atol=5e-14)


@pytest.mark.parametrize("n_short,n_long,k,seed", [(266, 4096, 8, 7), (30, 5000, 8, 2 ** 40 + 3), (5, 300, 8, 9),
                                                   (1000, 100, 33, 4)])
def test_sjlt_draw_matches_oracle(ops, n_short, n_long, k, seed):
    import oracle as orc
    _say(f"sjlt draw {n_short} x {n_long} k {k}")
    rows, vals = ops.sjlt_draw(n_short, n_long, k, seed, torch.device("cuda"))
    r_o, v_o = orc.device_sjlt_draw(n_short, n_long, k, seed)
    np.testing.assert_array_equal(rows.cpu().numpy().reshape(n_long, k), r_o)
    np.testing.assert_array_equal(vals.cpu().numpy().reshape(n_long, k), v_o)


@pytest.mark.parametrize("m,d,seed", [(3000, 4096, 77), (131072, 4096, 2 ** 50 + 1), (5, 100, 3), (1025, 700, 9)])
def test_sample_rows_matches_oracle(ops, m, d, seed):
    import oracle as orc
    from parla_amd.comps.sketchers.learned import sample_rows
    _say(f"sample rows m {m} d {d}")
    rng = np.random.default_rng(m)
    p = rng.random(m) ** 3
    p[::5] = 0.0
    pt = torch.from_numpy(p).cuda()

    class _Seed:  # hands the exact seed to sample_rows
        def integers(self, lo, hi):
            return seed

    idx, scale = sample_rows(pt, d, _Seed(), backend="device")
    i_o, s_o = orc.device_sample_rows(p, d, seed)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_o)
    np.testing.assert_allclose(scale.cpu().numpy(), s_o, rtol=4e-16, atol=0)
