"""The branch-free decimation score of the analysis kernels (decimate_mask / decimate_mask64 in
rgh_analyse.hip) against the reference's walk over the set bits (enc.rs:583 decimate_masked,
enc.rs:557 decimate_score with the DECIMATE4 / DECIMATE8 tables of enc.rs:88): every 16-bit mask
and a seeded sample of 64-bit masks, plus every one- and two-bit 64-bit mask. CPU only."""
import numpy as np

DECIMATE4 = [3, 2, 2, 1, 1, 1] + [0] * 10
DECIMATE8 = [3] * 4 + [2] * 8 + [1] * 12 + [0] * 40


def walk(mask, table):
    # The reference's loop: from the top set bit down, the run of zeros below each set bit.
    score, m = 0, mask
    while m:
        p = m.bit_length() - 1
        m ^= 1 << p
        below = m.bit_length() - 1 if m else -1
        score += table[p - below - 1]
    return score


def branch_free16(m):
    # decimate_mask: [r <= 5] + [r <= 2] + [r == 0] as popcounts against the smeared mask.
    M = 0xFFFFFFFF
    t3 = (m | (m << 1) | (m << 2)) & M
    t6 = (t3 | (t3 << 3)) & M
    return (bin(m & (((m << 1) & M) | 1)).count("1") + bin(m & (((t3 << 1) & M) | 7)).count("1")
            + bin(m & (((t6 << 1) & M) | 63)).count("1"))


def branch_free64(m):
    # decimate_mask64: [r <= 23] + [r <= 11] + [r <= 3].
    M = (1 << 64) - 1
    t2 = (m | (m << 1)) & M
    t4 = (t2 | (t2 << 2)) & M
    t12 = (t4 | (t4 << 4) | (t4 << 8)) & M
    t24 = (t12 | (t12 << 12)) & M
    return (bin(m & (((t4 << 1) & M) | 0xF)).count("1") + bin(m & (((t12 << 1) & M) | 0xFFF)).count("1")
            + bin(m & (((t24 << 1) & M) | 0xFFFFFF)).count("1"))


def test_every_16_bit_mask():
    bad = [m for m in range(1 << 16) if branch_free16(m) != walk(m, DECIMATE4)]
    assert not bad, bad[:5]


def test_64_bit_masks():
    rng = np.random.default_rng(11)
    masks = [1 << p for p in range(64)] + [(1 << p) | (1 << q) for p in range(64) for q in range(p)]
    for density in (0.02, 0.05, 0.1, 0.3, 0.6):
        bits = rng.random((4000, 64)) < density
        masks += [sum(1 << int(i) for i in np.nonzero(row)[0]) for row in bits]
    bad = [m for m in masks if branch_free64(m) != walk(m, DECIMATE8)]
    assert not bad, [hex(m) for m in bad[:5]]
