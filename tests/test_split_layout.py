"""Blocked BF16 piece layout (no GPU): the host helpers that block and
unblock piece arrays agree with the element offset formula the kernels use,
((i/16)(kp/32) + k/32) 512 + (i%16) 32 + 8 (c ^ ((i%16)/4)) + k%8."""
import numpy as np

from parla_amd import ops


def _offset(i, k, kp):
    r, c = i % 16, (k % 32) // 8
    return ((i // 16) * (kp // 32) + k // 32) * 512 + r * 32 + 8 * (c ^ (r // 4)) + k % 8


def test_block_unblock_follow_the_offset_formula():
    rng = np.random.default_rng(0)
    P, R, kp = 2, 48, 128
    a = rng.integers(0, 65535, (P, R, kp)).astype(np.uint16)
    flat = np.zeros((P, R * kp), dtype=np.uint16)
    i, k = np.meshgrid(np.arange(R), np.arange(kp), indexing="ij")
    off = _offset(i, k, kp)
    assert np.array_equal(np.sort(off.ravel()), np.arange(R * kp))  # a permutation
    for p in range(P):
        flat[p, off.ravel()] = a[p].ravel()
    blocked = ops.block_pieces(a).numpy().reshape(P, -1)
    assert np.array_equal(blocked, flat)
    import torch
    assert np.array_equal(ops.unblock_pieces(torch.from_numpy(flat.reshape(P, R, kp))), a)


def test_swizzle_makes_fragment_reads_conflict_free():
    """The 16 lanes that read one chunk column of a block (rows 0..15, same
    logical chunk) hit 16 distinct 4-bank groups of the 64 LDS banks."""
    for c in range(4):
        banks = set()
        for r in range(16):
            byte = r * 64 + 16 * (c ^ (r // 4))
            banks.add((byte // 4) % 64 // 4)
        assert len(banks) == 16


def test_block_rows_pads_and_matches_block_pieces():
    """The tensor form (used for the quantised Gaussian) pads with zeros and
    lays out exactly as the numpy form."""
    import torch
    rng = np.random.default_rng(1)
    X = rng.integers(0, 65535, (37, 200)).astype(np.uint16)
    got = ops.block_rows(torch.from_numpy(X), 128, 256).numpy()
    padded = np.zeros((1, 128, 256), dtype=np.uint16)
    padded[0, :37, :200] = X
    assert np.array_equal(got, ops.block_pieces(padded).numpy()[0])
