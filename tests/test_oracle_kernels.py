"""Pins of the oracle's kernels by closed forms, brute force and published examples."""
import numpy as np
import pytest

from oracle_lib import lib, ptr

RNG = np.random.default_rng(1234)
H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]])
CF = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]])


def u8(shape, lo=0, hi=256):
    return RNG.integers(lo, hi, shape).astype(np.uint8)


def test_sad16_brute_force():
    for _ in range(50):
        a, b = u8((16, 16)), u8((16, 16))
        assert lib().ora_sad16(ptr(a), 16, ptr(b), 16) == int(np.abs(a.astype(int) - b.astype(int)).sum())


def test_satd_is_halved_hadamard_magnitude():
    # SATD = sum over 4x4 blocks of |H d H^T|, halved once at the end (x264_pixel_satd).
    for _ in range(50):
        a, b = u8((16, 16)), u8((16, 16))
        d = a.astype(int) - b.astype(int)
        tot = sum(np.abs(H4 @ d[y:y + 4, x:x + 4] @ H4.T).sum() for y in range(0, 16, 4) for x in range(0, 16, 4))
        assert lib().ora_satd16(ptr(a), 16, ptr(b), 16) == tot >> 1
        assert lib().ora_satd4(ptr(a), 16, ptr(b), 16) == np.abs(H4 @ d[:4, :4] @ H4.T).sum() >> 1


def test_forward_dct4_is_the_core_transform():
    # ITU-T H.264 8.5 forward core transform: Y = Cf X Cf^T.
    for _ in range(100):
        x = RNG.integers(-255, 256, (4, 4)).astype(np.int16)
        out = np.zeros(16, np.int16)
        lib().ora_fdct4(ptr(np.ascontiguousarray(x)), ptr(out))
        assert (out.reshape(4, 4) == CF @ x.astype(int) @ CF.T).all()


def idct4_spec(c):
    """ITU-T H.264 8.5.12.2, equations 8-338 .. 8-351, written out per row then column."""
    c = c.astype(int)
    f = np.zeros((4, 4), int)
    for i in range(4):
        e0 = c[i, 0] + c[i, 2]
        e1 = c[i, 0] - c[i, 2]
        e2 = (c[i, 1] >> 1) - c[i, 3]
        e3 = c[i, 1] + (c[i, 3] >> 1)
        f[i] = [e0 + e3, e1 + e2, e1 - e2, e0 - e3]
    r = np.zeros((4, 4), int)
    for j in range(4):
        g0 = f[0, j] + f[2, j]
        g1 = f[0, j] - f[2, j]
        g2 = (f[1, j] >> 1) - f[3, j]
        g3 = f[1, j] + (f[3, j] >> 1)
        h = [g0 + g3, g1 + g2, g1 - g2, g0 - g3]
        for i in range(4):
            r[i, j] = (h[i] + 32) >> 6
    return r


def test_inverse_dct4_matches_the_standard():
    for _ in range(200):
        c = RNG.integers(-2000, 2001, (4, 4)).astype(np.int16)
        out = np.zeros(16, np.int16)
        lib().ora_idct4(ptr(np.ascontiguousarray(c)), ptr(out))
        assert (out.reshape(4, 4) == idct4_spec(c)).all()


M8 = np.array([[8, 8, 8, 8, 8, 8, 8, 8], [12, 10, 6, 3, -3, -6, -10, -12], [8, 4, -4, -8, -8, -4, 4, 8],
               [10, -3, -12, -6, 6, 12, 3, -10], [8, -8, -8, 8, 8, -8, -8, 8], [6, -12, 3, 10, -10, -3, 12, -6],
               [4, -8, 8, -4, -4, 8, -8, 4], [3, -6, 10, -12, 12, -10, 6, -3]])


def test_forward_dct8_is_the_8x8_core_transform():
    # The High-profile forward 8x8 transform is (1/8) M8 X M8^T (M8 the integer basis of the
    # standard's 8x8 transform); its >>1 and >>2 steps are exact when every sample is a
    # multiple of 64, so there the integer butterfly must equal the matrix form exactly.
    for _ in range(100):
        x = (RNG.integers(-1, 2, (8, 8)) * 64).astype(np.int16)
        out = np.zeros(64, np.int16)
        lib().ora_fdct8(ptr(np.ascontiguousarray(x)), ptr(out))
        want = M8 @ x.astype(int) @ M8.T // 64
        assert (out.reshape(8, 8) == want).all()
    x = np.full(64, 7, np.int16)
    lib().ora_fdct8(ptr(x), ptr(out))
    assert out[0] == 64 * 7 and (out[1:] == 0).all()


def tap6(a, b, c, d, e, f):
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f


def test_luma_mc_half_pel_samples_follow_the_six_tap_filter():
    P = 8
    src = u8((16 + 2 * P, 16 + 2 * P)).astype(np.uint8)
    o = P * src.shape[1] + P
    flat = np.ascontiguousarray(src).reshape(-1)
    ss = src.shape[1]
    s = src.astype(int)
    for mx, my in [(2, 0), (0, 2), (2, 2), (0, 0)]:
        dst = np.zeros((16, 16), np.uint8)
        base = flat[o:]
        lib().ora_luma_mc(ptr(dst), 16, ptr(base), ss, 16, 16, mx, my)
        for r in range(16):
            for c in range(16):
                R, C = r + P, c + P
                if (mx, my) == (0, 0):
                    want = s[R, C]
                elif (mx, my) == (2, 0):
                    want = np.clip((tap6(*s[R, C - 2:C + 4]) + 16) >> 5, 0, 255)
                elif (mx, my) == (0, 2):
                    want = np.clip((tap6(*s[R - 2:R + 4, C]) + 16) >> 5, 0, 255)
                else:
                    hh = [tap6(*s[R + k, C - 2:C + 4]) for k in range(-2, 4)]
                    want = np.clip((tap6(*hh) + 512) >> 10, 0, 255)
                assert dst[r, c] == want, (mx, my, r, c)


def test_window_table_equals_spec_luma_mc():
    # enc.rs:4030 window_prediction_matches_luma_mc_up_to_the_vector_bounds.
    P = 24
    for _ in range(4):
        plane = np.ascontiguousarray(u8((16 + 2 * P, 16 + 2 * P)))
        ss = plane.shape[1]
        flat = plane.reshape(-1)
        base = flat[P * ss + P:]
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                a = np.zeros((16, 16), np.uint8)
                b = np.zeros((16, 16), np.uint8)
                lib().ora_window_pred(ptr(a), ptr(base), ss, dx, dy)
                off = (dy >> 2) * ss + (dx >> 2)
                lib().ora_luma_mc(ptr(b), 16, ptr(flat[P * ss + P + off:]), ss, 16, 16, dx & 3, dy & 3)
                assert (a == b).all(), (dx, dy)


def test_chroma_mc_is_the_bilinear_eighth_sample_filter():
    src = np.ascontiguousarray(u8((9, 16)))
    s = src.astype(int)
    for fx in range(8):
        for fy in range(8):
            out = np.zeros(64, np.uint8)
            lib().ora_chroma_mc8(ptr(out), ptr(src), 16, fx, fy)
            want = ((8 - fx) * (8 - fy) * s[:8, :8] + fx * (8 - fy) * s[:8, 1:9] + (8 - fx) * fy * s[1:9, :8] +
                    fx * fy * s[1:9, 1:9] + 32) >> 6
            assert (out.reshape(8, 8) == want).all()


def dec_score(v):
    a = np.asarray(v, np.int16)
    return lib().ora_decimate_score(ptr(a), len(a))


def test_decimation_scores_follow_x264():
    # enc.rs:4190 decimation_scores_follow_x264
    one = lambda i, v: [v if k == i else 0 for k in range(16)]
    assert dec_score([0] * 16) == 0
    assert dec_score(one(0, 1)) == 3
    assert dec_score(one(15, -1)) == 0
    assert dec_score(one(1, 2)) == 9
    two = one(0, 1)
    two[1] = -1
    assert dec_score(two) == 6


def cavlc_bits(levels, nc):
    a = np.asarray(levels, np.int16)
    out = np.zeros(64, np.uint8)
    n = lib().ora_cavlc_block(ptr(a), len(a), nc, ptr(out), 64)
    return "".join(format(b, "08b") for b in out[:(n + 7) // 8])[:n]


def test_cavlc_published_examples():
    # Richardson, "The H.264 Advanced Video Compression Standard", CAVLC examples:
    # block 0 3 -1 0 / 0 -1 1 0 / 1 0 0 0 / 0 0 0 0 -> zig-zag 0,3,0,1,-1,-1,0,1,0.. (nC 0)
    assert cavlc_bits([0, 3, 0, 1, -1, -1, 0, 1] + [0] * 8, 0) == "000010001110010111101101"
    # empty blocks: coeff_token 1 (nC 0..1), 11 (2..3), 1111 (4..7), 000011 (>= 8)
    assert cavlc_bits([0] * 16, 0) == "1"
    assert cavlc_bits([0] * 16, 2) == "11"
    assert cavlc_bits([0] * 16, 4) == "1111"
    assert cavlc_bits([0] * 16, 8) == "000011"
    # chroma DC (nC = -1), a single +1 at position 0: coeff_token "1" (Table 9-5,
    # TrailingOnes 1, TotalCoeff 1), sign "0", total_zeros "1" (Table 9-9a, zeros 0)
    assert cavlc_bits([1, 0, 0, 0], 0) == "101"


def test_intra4x4_predictors_known_answers():
    top = np.array([10, 20, 30, 40, 50, 60, 70, 80], np.uint8)
    left = np.array([15, 25, 35, 45], np.uint8)
    out = np.zeros(16, np.uint8)
    L = lib()
    L.ora_predict_intra4x4(0, ptr(top), ptr(left), 5, 1, 1, ptr(out))
    assert (out.reshape(4, 4) == np.tile(top[:4], (4, 1))).all()
    L.ora_predict_intra4x4(1, ptr(top), ptr(left), 5, 1, 1, ptr(out))
    assert (out.reshape(4, 4) == np.repeat(left[:, None], 4, 1)).all()
    L.ora_predict_intra4x4(2, ptr(top), ptr(left), 5, 1, 1, ptr(out))
    assert (out == (100 + 120 + 4) >> 3).all()
    # Diagonal down-left: pred[y][x] = (p[x+y] + 2 p[x+y+1] + p[x+y+2] + 2) >> 2 (8.3.1.2.4)
    L.ora_predict_intra4x4(3, ptr(top), ptr(left), 5, 1, 0, ptr(out))
    t = top.astype(int)
    for y in range(4):
        for x in range(4):
            k = x + y
            want = (t[6] + 3 * t[7] + 2) >> 2 if k == 6 else (t[k] + 2 * t[k + 1] + t[k + 2] + 2) >> 2
            assert out[y * 4 + x] == want


def test_compact_half_sample_planes_give_the_spec_luma_mc():
    # compact.rs:87 RefLuma::src pairs the b / h / j planes (compact.rs:24 HpelPlanes::build)
    # into every quarter-sample position; the result must be the standard's luma MC
    # (8.4.2.2.1), including blocks reaching into the replicated padding.
    W, H, PAD = 48, 32, 48
    rng = np.random.default_rng(7)
    for it in range(3):
        pic = np.ascontiguousarray(rng.integers(0, 256, (H, W)).astype(np.uint8))
        padded = np.pad(pic, PAD, mode="edge")
        ss = padded.shape[1]
        flat = np.ascontiguousarray(padded).reshape(-1)
        for (x0, y0, bw, bh) in ((0, 0, 16, 16), (16, 8, 8, 8), (32, 16, 16, 16), (8, 24, 8, 8)):
            for mx in range(-4 * (x0 + 28), 4 * (W - bw - x0 + 28) + 1, 37):
                for my in (-4 * (y0 + 28), -5, 0, 3, 6, 4 * (H - bh - y0 + 28)):
                    got = np.zeros((bh, bw), np.uint8)
                    assert lib().ora_compact_pred(ptr(pic), W, H, x0, y0, bw, bh, mx, my, ptr(got)) == 0
                    want = np.zeros((16, 16), np.uint8)
                    off = (PAD + y0 + (my >> 2)) * ss + PAD + x0 + (mx >> 2)
                    lib().ora_luma_mc(ptr(want), 16, ptr(flat[off:]), ss, bw, bh, mx & 3, my & 3)
                    assert (got == want[:bh, :bw]).all(), (x0, y0, mx, my)
