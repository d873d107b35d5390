"""Pins of the oracle's High-profile path (adaptive 8x8 transform) by what the mathematics and
the standard fix, independently of the oracle's own code:

* the 8x8 quantizer multipliers follow from the standard's normAdjust8x8 and the squared norms of
  the 8x8 integer transform's basis rows: mf = round(2^24 / (normAdjust8x8 x N_class));
* 8x8 dequantization is ITU-T H.264 8.5.13.1 with flat scaling (LevelScale8x8 = 16 x
  normAdjust8x8) for every QP and position;
* the inverse 8x8 transform is 8.5.13.2 (the row and column equations written out here);
* the transform choice of a coded P macroblock (enc.rs:3170-3179: the 8x8 transform when it
  leaves no more non-zero levels than the 4x4 one, tried only when the 4x4 residual is coded)
  is checked end to end on known-answer macroblocks whose levels, decision and reconstruction
  are computed here from the standard's transforms.
"""
import numpy as np
import pytest

from oracle_lib import OracleEncoder, lib, ptr
from test_oracle_kernels import CF, M8, idct4_spec

RNG = np.random.default_rng(8080)

# ITU-T H.264 Table 8-16 (normAdjust8x8, v_mn) and the 4x4 counterpart of 8.5.9 (v_mn).
NORM8 = [[20, 18, 32, 19, 25, 24], [22, 19, 35, 21, 28, 26], [26, 23, 42, 24, 33, 31],
         [28, 25, 45, 26, 35, 33], [32, 28, 51, 30, 40, 38], [36, 32, 58, 34, 46, 43]]
NORM4 = [[10, 16, 13], [11, 18, 14], [13, 20, 16], [14, 23, 18], [16, 25, 20], [18, 29, 23]]
# The 4x4 forward quantizer multipliers (JM / x264 quant_mf; pinned end to end by the golden clips).
MF4 = [[13107, 5243, 8066], [11916, 4660, 7490], [10082, 4194, 6554], [9362, 3647, 5825], [8192, 3355, 5243],
       [7282, 2893, 4559]]
ZZ4 = [0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15]
ZZ8 = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
       21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
       61, 54, 47, 55, 62, 63]


def class8(i, j):
    """The position classes of normAdjust8x8 (8.5.9, equation 8-317)."""
    if i % 4 == 0 and j % 4 == 0:
        return 0
    if i % 2 == 1 and j % 2 == 1:
        return 1
    if i % 4 == 2 and j % 4 == 2:
        return 2
    if (i % 4 == 0 and j % 2 == 1) or (i % 2 == 1 and j % 4 == 0):
        return 3
    if (i % 4 == 0 and j % 4 == 2) or (i % 4 == 2 and j % 4 == 0):
        return 4
    return 5


def class4(i, j):
    return 0 if (i % 2 == 0 and j % 2 == 0) else (1 if (i % 2 == 1 and j % 2 == 1) else 2)


ROW_NORM2 = (M8 * M8).sum(axis=1)  # 512, 578, 320, 578, 512, 578, 320, 578


def basis_norm(c):
    """N_class: the product of the two basis rows' squared norms / 4096 for a class's positions."""
    vals = {ROW_NORM2[i] * ROW_NORM2[j] / 4096 for i in range(8) for j in range(8) if class8(i, j) == c}
    assert len(vals) == 1
    return vals.pop()


MF8 = [[int(round(2 ** 24 / (NORM8[q][c] * basis_norm(c)))) for c in range(6)] for q in range(6)]


def test_basis_norms_are_the_closed_form():
    # 512 x 512, 578 x 578, 320 x 320, 512 x 578, 512 x 320 and 578 x 320, over 4096.
    assert [basis_norm(c) for c in range(6)] == [64, 578 * 578 / 4096, 25, 72.25, 40, 578 * 320 / 4096]


def test_quant8_multipliers_follow_the_basis_norms():
    for rem in range(6):
        for pos in range(64):
            c = class8(pos >> 3, pos & 7)
            assert lib().ora_quant8_mf(rem, pos) == MF8[rem][c], (rem, pos)
            assert lib().ora_dequant8_norm(rem, pos) == NORM8[rem][c], (rem, pos)


def dequant8_spec(levels, qp):
    """8.5.13.1 with flat weights: LevelScale8x8 = 16 x normAdjust8x8."""
    d = np.zeros(64, np.int64)
    for pos in range(64):
        ls = 16 * NORM8[qp % 6][class8(pos >> 3, pos & 7)]
        c = int(levels[pos])
        if qp >= 36:
            d[pos] = (c * ls) << (qp // 6 - 6)
        else:
            d[pos] = (c * ls + (1 << (5 - qp // 6))) >> (6 - qp // 6)
    return d


def test_dequant8_matches_level_scale_8x8():
    out = np.zeros(64, np.int32)
    for qp in range(52):
        for _ in range(4):
            # Levels small enough that every product stays inside int16 (no clamping).
            lim = max(1, 32767 // (16 * 58 << max(0, qp // 6 - 6)) // 2)
            lv = RNG.integers(-lim, lim + 1, 64).astype(np.int16)
            lib().ora_dequant8(ptr(lv), qp, ptr(out))
            assert (out == dequant8_spec(lv, qp)).all(), qp
    # Out-of-range products are clamped to int16 (h264.rs:1403).
    lv = np.full(64, 2000, np.int16)
    lib().ora_dequant8(ptr(lv), 51, ptr(out))
    assert (out == 32767).all()


def idct8_rows(d):
    """One pass of 8.5.13.2 (equations 8-338 .. 8-353) over each row of d."""
    out = np.zeros_like(d)
    for r in range(8):
        x = [int(v) for v in d[r]]
        e0 = x[0] + x[4]
        e1 = -x[3] + x[5] - x[7] - (x[7] >> 1)
        e2 = x[0] - x[4]
        e3 = x[1] + x[7] - x[3] - (x[3] >> 1)
        e4 = (x[2] >> 1) - x[6]
        e5 = -x[1] + x[7] + x[5] + (x[5] >> 1)
        e6 = x[2] + (x[6] >> 1)
        e7 = x[3] + x[5] + x[1] + (x[1] >> 1)
        f0 = e0 + e6
        f1 = e1 + (e7 >> 2)
        f2 = e2 + e4
        f3 = e3 + (e5 >> 2)
        f4 = e2 - e4
        f5 = (e3 >> 2) - e5
        f6 = e0 - e6
        f7 = e7 - (e1 >> 2)
        out[r] = [f0 + f7, f2 + f5, f4 + f3, f6 + f1, f6 - f1, f4 - f3, f2 - f5, f0 - f7]
    return out


def idct8_spec(d):
    """Rows, then columns, then r = (h + 32) >> 6 (8.5.13.2)."""
    h = idct8_rows(idct8_rows(np.asarray(d, np.int64).reshape(8, 8)).T).T
    return (h + 32) >> 6


def test_inverse_dct8_matches_the_standard():
    out = np.zeros(64, np.int32)
    for _ in range(300):
        d = RNG.integers(-3000, 3001, 64).astype(np.int32)
        lib().ora_idct8(ptr(d), ptr(out))
        assert (out.reshape(8, 8) == idct8_spec(d)).all()
    # DC only: every residual sample is (D + 32) >> 6.
    for dc in (-4096, -97, 0, 31, 32, 100, 4095):
        d = np.zeros(64, np.int32)
        d[0] = dc
        lib().ora_idct8(ptr(d), ptr(out))
        assert (out == (dc + 32) >> 6).all()


def test_8x8_path_inverts_the_forward_transform():
    # Forward transform, quantization at QP 0 and the inverse path return the residual to
    # within the quantizer's rounding (a sign, index or scaling slip would not).
    fo = np.zeros(64, np.int16)
    d = np.zeros(64, np.int32)
    r = np.zeros(64, np.int32)
    for _ in range(50):
        x = RNG.integers(-60, 61, 64).astype(np.int16)
        lib().ora_fdct8(ptr(x), ptr(fo))
        mf = np.array([MF8[0][class8(p >> 3, p & 7)] for p in range(64)])
        lv = (np.sign(fo) * ((np.abs(fo.astype(np.int64)) * mf + (1 << 16) // 2) >> 16)).astype(np.int16)
        lib().ora_dequant8(ptr(lv), 0, ptr(d))
        lib().ora_idct8(ptr(d), ptr(r))
        assert np.abs(r - x).max() <= 2


def quantize(c, mf, shift, bias):
    a = np.minimum((np.abs(c) * mf + bias) >> shift, 32767)
    return np.sign(c) * a


def expected_mb(R, qp):
    """The P_L0_16x16 residual coding of a macroblock with prediction 128 and residual R, from the
    standard's transforms and quantizers: the 4x4 levels (decoding order) and, when the 4x4
    residual is coded, the 8x8 levels and the choice between them (enc.rs:3170-3179). Every
    coded block here holds a level above 1, so no decimation applies (asserted)."""
    s4, s8 = 15 + qp // 6, 16 + qp // 6
    mf4 = np.array([[MF4[qp % 6][class4(i, j)] for j in range(4)] for i in range(4)])
    mf8 = np.array([[MF8[qp % 6][class8(i, j)] for j in range(8)] for i in range(8)])
    q4 = []
    for b in range(16):
        bx, by = (b & 1) + ((b >> 1) & 2), ((b >> 1) & 1) + ((b >> 2) & 2)
        q4.append(quantize(CF @ R[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] @ CF.T, mf4, s4, (1 << s4) // 6))
    q8 = []
    for qd in range(4):
        X = R[(qd >> 1) * 8:(qd >> 1) * 8 + 8, (qd & 1) * 8:(qd & 1) * 8 + 8]
        W = M8 @ X @ M8.T
        assert (W % 64 == 0).all()  # residual samples are multiples of 64: the butterfly is exact
        q8.append(quantize(W // 64, mf8, s8, (1 << s8) // 6))
    cbp4 = cbp8 = 0
    for qd in range(4):
        blks = q4[4 * qd:4 * qd + 4]
        if any((b != 0).any() for b in blks):
            assert any((np.abs(b) > 1).any() for b in blks)
            cbp4 |= 1 << qd
        if (q8[qd] != 0).any():
            assert (np.abs(q8[qd]) > 1).any()
            cbp8 |= 1 << qd
    nnz4 = sum(int((b != 0).sum()) for b in q4)
    nnz8 = sum(int((b != 0).sum()) for b in q8)
    use8 = cbp4 != 0 and nnz8 <= nnz4
    return dict(q4=q4, q8=q8, cbp=cbp8 if use8 else cbp4, use8=use8, nnz4=nnz4, nnz8=nnz8)


def expected_recon(E, qp):
    rec = np.full((16, 16), 128, np.int64)
    if E["use8"]:
        for qd in range(4):
            if (E["cbp"] >> qd) & 1:
                r = idct8_spec(dequant8_spec(E["q8"][qd].reshape(-1), qp))
                rec[(qd >> 1) * 8:(qd >> 1) * 8 + 8, (qd & 1) * 8:(qd & 1) * 8 + 8] += r
    else:
        for b in range(16):
            bx, by = (b & 1) + ((b >> 1) & 2), ((b >> 1) & 1) + ((b >> 2) & 2)
            v = np.array([[NORM4[qp % 6][class4(i, j)] for j in range(4)] for i in range(4)])
            d = E["q4"][b] * v << (qp // 6)  # 8.5.12.1 with flat weights (LevelScale4x4 = 16 v)
            rec[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] += idct4_spec(d)
    return np.clip(rec, 0, 255)


def code_mb(R, qp, w=48, h=48):
    """A flat grey IDR, then a P frame equal to it except macroblock (1, 1) = 128 + R. Every
    vector predicts the same flat block and the mv cost grows away from (0, 0), so the search
    keeps (0, 0); the intra 16x16 prediction is the same flat block and costs more."""
    y0 = np.full((h, w), 128, np.uint8)
    c0 = np.full((h // 2, w // 2), 128, np.uint8)
    ora = OracleEncoder(w, h, qp, profile="high", keyint=1 << 20)
    ora.encode(y0, c0, c0.copy())
    y1 = y0.astype(np.int64)
    y1[16:32, 16:32] += R
    ora.encode(y1.astype(np.uint8), c0, c0.copy())
    m = (w // 16) + 1
    return ora.decisions()[m], ora.levels()[m], ora.recon(deblocked=False)[0][16:32, 16:32]


# (name, qp, 4x4-block values of the 16x16 residual (rows of blocks), the expected comparison)
KATS = [
    ("fewer", 26, [[64, 64, 0, 0], [64, 64, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]], "<"),
    ("equal", 26, [[64, 64, -64, -64], [64, 64, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]], "=="),
    ("more", 28, [[64, -64, 64, 0], [0, 0, 0, 64], [-64, 0, 0, 0], [0, -64, -64, -64]], ">"),
]


@pytest.mark.parametrize("name,qp,blocks,rel", KATS, ids=[k[0] for k in KATS])
def test_transform_choice_known_answers(name, qp, blocks, rel):
    R = np.kron(np.array(blocks), np.ones((4, 4), np.int64))
    E = expected_mb(R, qp)
    assert {"<": E["nnz8"] < E["nnz4"], "==": E["nnz8"] == E["nnz4"], ">": E["nnz8"] > E["nnz4"]}[rel]
    d, lv, rec = code_mb(R, qp)
    assert d["mb_type"] == 1 and tuple(d["mv"]) == (0, 0)
    assert d["transform_8x8"] == int(E["use8"]) and d["cbp_luma"] == E["cbp"]
    if E["use8"]:
        want = np.concatenate([E["q8"][qd].reshape(-1)[ZZ8] for qd in range(4)])
    else:
        want = np.concatenate([E["q4"][b].reshape(-1)[ZZ4] for b in range(16)])
    assert (lv[:256] == want).all()
    assert (rec == expected_recon(E, qp)).all()


def test_no_8x8_when_the_4x4_residual_vanishes():
    # A flat residual of 3 at QP 30: every 4x4 DC level is (48 x 13107 + 2^20/6) >> 20 = 0, while
    # the 8x8 DC level would be (192 x 13107 + 2^21/6) >> 21 = 1. The 8x8 transform is only tried
    # for a coded 4x4 residual, so the macroblock is P_Skip.
    R = np.full((16, 16), 3, np.int64)
    assert (48 * 13107 + (1 << 20) // 6) >> 20 == 0 and (192 * 13107 + (1 << 21) // 6) >> 21 == 1
    d, lv, rec = code_mb(R, 30)
    assert d["mb_type"] == 0 and d["transform_8x8"] == 0 and (lv == 0).all()
    assert (rec == 128).all()
