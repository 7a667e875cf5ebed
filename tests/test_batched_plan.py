"""Host-side layout of the batched solver (no GPU): every work item of every
batched kernel decodes to a problem and a local index inside that problem's
range, by the same binary search and tile decode the kernels use, and every
problem's slice of every pool lies inside the pool."""
import numpy as np
import pytest

from parla_amd.drivers.batched_ls import BLOCK_ROWS, batch_plan, _npl

LAYOUT = (2048, 64, 16, 4096, 16)  # chunk_rows, tile, strip_rows, max_d, max_vec_nnz


def bt_find(first, idx):
    """The kernels' search: the largest p with first[p] <= idx."""
    lo, hi = 0, len(first) - 1
    while lo < hi:
        mid = (lo + hi + 1) >> 1
        if first[mid] <= idx:
            lo = mid
        else:
            hi = mid - 1
    return lo


def upper_tile(t, T):
    ti = 0
    while t >= T - ti:
        t -= T - ti
        ti += 1
    return ti, ti + t


def _sizes(seed, P):
    rng = np.random.default_rng(seed)
    ms, ns = [], []
    while len(ms) < P:
        m, n = int(rng.integers(1, 9000)), int(rng.integers(1, 513))
        if m >= n:
            ms.append(m)
            ns.append(n)
    return ms, ns


@pytest.mark.parametrize("seed,P,factor", [(0, 1, 4), (1, 7, 4), (2, 40, 4), (3, 25, 2), (4, 12, 8)])
def test_every_work_item_decodes_in_range(seed, P, factor):
    ms, ns = _sizes(seed, P)
    if seed == 0:
        ms, ns = [512 * 8], [512]  # largest problem shape of the workload
    plan = batch_plan(ms, ns, factor, LAYOUT)
    chunk_rows, tile, strip_rows, max_d, _ = LAYOUT
    sk, desc, tot = plan["sk"], plan["desc"], plan["totals"]
    ms, ns, ds = desc[:, 0], desc[:, 1], sk[:, 0]
    assert (ds <= max_d).all() and (ds <= ms).all() and (ds == np.minimum(ms, factor * ns)).all()
    counts = {"chunks": -(-ms // chunk_rows), "rows": ds, "tiles": None, "strips": -(-ds // strip_rows),
              "dblk": -(-ns // strip_rows)}
    T = -(-ns // tile)
    counts["tiles"] = T * (T + 1) // 2
    col = {"chunks": 4, "rows": 5, "tiles": 6, "strips": 7, "dblk": 8}
    for kind, c in col.items():
        first = sk[:, c]
        assert tot[kind] == int(counts[kind].sum())
        seen = np.zeros(len(ms), dtype=np.int64)
        for idx in range(tot[kind]):
            p = bt_find(first, idx)
            local = idx - first[p]
            assert 0 <= local < counts[kind][p], (kind, idx, p, local)
            seen[p] += 1
        assert (seen == counts[kind]).all(), kind
    # upper tiles: every (ti, tj) with ti <= tj exactly once
    for p in range(len(ms)):
        pairs = [upper_tile(t, int(T[p])) for t in range(int(counts["tiles"][p]))]
        assert sorted(pairs) == [(i, j) for i in range(T[p]) for j in range(i, T[p])]
    # pool slices
    k = LAYOUT[4]
    for p in range(len(ms)):
        m, n, d = int(ms[p]), int(ns[p]), int(ds[p])
        assert desc[p, 2] + m * n <= tot["a"]
        assert desc[p, 3] + m <= tot["mv"] and k * (desc[p, 3] + m) <= k * tot["mv"]
        assert desc[p, 4] + n <= tot["nv"]
        assert desc[p, 5] + n * n <= tot["r"]
        nb = -(-m // BLOCK_ROWS)
        assert desc[p, 7] + nb * (2 * n + 2) <= tot["part"]
        assert sk[p, 1] + d * n <= tot["y"]
        assert sk[p, 2] + d + 1 <= tot["rp"]
        assert sk[p, 3] + counts["chunks"][p] * d <= tot["cnt"]
        assert 256 * (sk[p, 8] + counts["dblk"][p]) <= 256 * tot["dblk"]
    # 64-row blocks: each problem's rows exactly once, in its row-length class
    covered = {p: [] for p in range(len(ms))}
    for c, lst in plan["blocks"].items():
        for prob, row0 in lst:
            assert _npl(int(ns[prob])) == c
            assert 0 <= row0 < ms[prob] and row0 % BLOCK_ROWS == 0
            covered[int(prob)].append(int(row0))
    for p in range(len(ms)):
        assert covered[p] == list(range(0, int(ms[p]), BLOCK_ROWS))


def test_plan_rejects_unsupported_shapes():
    with pytest.raises(NotImplementedError):
        batch_plan([5000], [513], 4, LAYOUT)
    with pytest.raises(ValueError):
        batch_plan([10], [20], 4, LAYOUT)
    with pytest.raises(NotImplementedError):
        batch_plan([100000], [512], 9, LAYOUT)
