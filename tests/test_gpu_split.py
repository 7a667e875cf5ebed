"""Split-precision (BF16 pieces) products on the device: the split kernels
bit for bit against the oracle's pieces, the GEMM's piece pairing exactly,
the GEMM against the oracle's product at FP32-accumulation accuracy, and QB
with a split operator against the FP64 QB."""
import numpy as np
import pytest
import torch

import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-split] {msg}", flush=True)


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from parla_amd import ops as ops_
    ops_.lib()
    return ops_


@pytest.mark.parametrize("P", [2, 3])
@pytest.mark.parametrize("rows,cols", [(1, 1), (37, 200), (300, 65), (130, 257)])
def test_split_kernels_bit_exact(ops, P, rows, cols):
    _say(f"split P={P} {rows}x{cols}")
    rng = np.random.default_rng(rows * 1000 + cols)
    X = rng.standard_normal((rows, cols)) * np.exp(rng.uniform(-10, 10, (rows, cols)))
    Xd = torch.from_numpy(X).cuda()
    ref = orc.bf16_pieces(X, P)
    a = ops.split_bf16(Xd, P, rows_out=(rows + 5 + 15) // 16 * 16)
    got = ops.unblock_pieces(a.pieces)
    kp = got.shape[2]
    assert kp % 64 == 0 and kp >= cols
    assert np.array_equal(got[:, :rows, :cols], ref)
    assert not got[:, rows:, :].any() and not got[:, :, cols:].any()
    b = ops.split_bf16(Xd, P, transpose=True, rows_out=(cols + 3 + 15) // 16 * 16)
    got = ops.unblock_pieces(b.pieces)
    assert np.array_equal(got[:, :cols, :rows], ref.transpose(0, 2, 1))
    assert not got[:, cols:, :].any() and not got[:, :, rows:].any()


def test_split_gemm_pairs_pieces_exactly(ops):
    """With one nonzero piece on each side the product is exact in FP32
    (small integers): C = L_p R_q^T when p + q < P, zero otherwise."""
    _say("piece pairing")
    rng = np.random.default_rng(5)
    M, N, K = 200, 300, 640
    for P in (2, 3):
        kp, mpad, npad, _ = ops.split_layout(P, M, N, K)
        Lv = rng.integers(-8, 9, (M, K)).astype(np.float32)
        Rv = rng.integers(-8, 9, (N, K)).astype(np.float32)
        Lb = (Lv.view(np.uint32) >> 16).astype(np.uint16)  # small integers are exact BF16
        Rb = (Rv.view(np.uint32) >> 16).astype(np.uint16)
        exact = Lv.astype(np.float64) @ Rv.astype(np.float64).T
        for p in range(P):
            for q in range(P):
                Lp = np.zeros((P, mpad, kp), dtype=np.uint16)
                Rp = np.zeros((P, npad, kp), dtype=np.uint16)
                Lp[p, :M, :K] = Lb
                Rp[q, :N, :K] = Rb
                C = ops.split_gemm(ops.SplitPieces(ops.block_pieces(Lp).cuda(), M, K),
                                   ops.SplitPieces(ops.block_pieces(Rp).cuda(), N, K)).cpu().numpy()
                want = exact if p + q < P else np.zeros_like(exact)
                assert np.array_equal(C, want), (P, p, q)


@pytest.mark.parametrize("P", [2, 3])
@pytest.mark.parametrize("M,N,K,tout", [(300, 70, 1000, False), (129, 289, 64, True), (8192, 266, 4096, False),
                                        (4096, 266, 8192, True)])
def test_split_gemm_matches_oracle(ops, P, M, N, K, tout):
    _say(f"split gemm P={P} M={M} N={N} K={K} transposed={tout}")
    rng = np.random.default_rng(M + N + K + P)
    L = rng.standard_normal((M, K))
    R = rng.standard_normal((N, K))
    kp, mpad, npad, _ = ops.split_layout(P, M, N, K)
    Lp = ops.split_bf16(torch.from_numpy(L).cuda(), P)
    Rp = ops.split_bf16(torch.from_numpy(R).cuda(), P, rows_out=npad)
    C = ops.split_gemm(Lp, Rp, transpose_out=tout).cpu().numpy()
    if tout:
        C = C.T
    ref = orc.split_product(orc.bf16_pieces(L, P), orc.bf16_pieces(R, P))
    # FP32 accumulation inside each K split: error ~ sqrt(K) 2^-24 |C| for random signs
    err = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    assert err <= 4 * np.sqrt(K) * 2.0 ** -24, err
    full = L @ R.T
    err_full = np.linalg.norm(C - full) / np.linalg.norm(full)
    assert err_full <= 4 * np.sqrt(K) * 2.0 ** -24 + 2.0 ** -(8 * P - 4), err_full
    C2 = ops.split_gemm(Lp, Rp, transpose_out=tout).cpu().numpy()
    assert np.array_equal(C2.T if tout else C2, C)  # deterministic


@pytest.mark.parametrize("P", [2, 3])
def test_qb_with_split_operator(ops, P):
    """QB1 on a split operator: Q orthonormal to FP64 accuracy, B = Q^T A to
    the pieces' accuracy, and the approximation error that of the FP64 QB."""
    _say(f"QB with split operator P={P}")
    import parla_amd as pa
    from parla_amd import matmakers as mm
    m, n, k = 4096, 2048, 128 + 10
    A_host, _ = mm.config_qb(m, n)
    A = torch.from_numpy(A_host).cuda()
    qb = pa.QB1(pa.RF1(pa.RS1(pa.SkOpSJ(8, backend="device"), 4, pa.orth, 1)))
    Q64, B64 = qb(A, k, np.nan, 7)
    op = pa.SplitOperator(A, P)
    Q, B = qb(op, k, np.nan, 7)
    assert isinstance(Q, torch.Tensor) and isinstance(B, torch.Tensor)
    fro = torch.linalg.norm(A).item()
    e64 = torch.linalg.norm(A - Q64 @ B64).item() / fro
    e = torch.linalg.norm(A - Q @ B).item() / fro
    _say(f"rel err FP64 {e64:.3e}, split P={P} {e:.3e}")
    assert e <= 1.05 * e64 + 2.0 ** -(8 * P - 2)
    I = torch.eye(k, dtype=torch.float64, device="cuda")
    assert torch.linalg.norm(Q.t() @ Q - I).item() <= 1e-13
    Bex = (Q.t() @ A)
    assert torch.linalg.norm(B - Bex).item() <= 2.0 ** -(8 * P - 3) * fro


@pytest.mark.parametrize("P", [2, 3])
def test_split_gemm_long_k_reproducible(ops, P):
    """32768 K per workgroup (4 K splits of 131072): bit-for-bit reproducible
    and within the FP32-accumulation bound of the FP64 product."""
    M, N, K = 8192, 266, 131072
    _say(f"split gemm long K P={P} M={M} N={N} K={K}")
    g = torch.Generator(device="cuda").manual_seed(P)
    L = torch.randn((M, K), generator=g, dtype=torch.float64, device="cuda")
    R = torch.randn((N, K), generator=g, dtype=torch.float64, device="cuda")
    kp, mpad, npad, _ = ops.split_layout(P, M, N, K)
    Lp = ops.split_bf16(L, P)
    Rp = ops.split_bf16(R, P, rows_out=npad)
    C1 = ops.split_gemm(Lp, Rp)
    C2 = ops.split_gemm(Lp, Rp)
    assert torch.equal(C1, C2)
    ref = L @ R.t()
    err = (torch.linalg.norm(C1 - ref) / torch.linalg.norm(ref)).item()
    assert err <= 4 * np.sqrt(K) * 2.0 ** -24 + 2.0 ** -(8 * P - 4), err
