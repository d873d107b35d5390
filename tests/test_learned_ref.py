"""CPU checks of the proposal network's numpy reference and of the library's learned ABI.

The reference's backward pass is pinned against central finite differences in float64, its
features against a brute-force loop, and its number formats against the formats' definitions
(every E4M3 code, midpoints, saturation; bfloat16 ties). No GPU is used.
"""
import ctypes
import os

import numpy as np
import pytest

import pnet_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_the_learned_abi():
    from ragdoll_amd import learned
    lib = ctypes.CDLL(os.path.join(ROOT, "ragdoll_amd", "libragdoll_hip.so"))
    for name in learned.EXPORTS:
        assert hasattr(lib, name), name
    with open(os.path.join(ROOT, "include", "ragdoll_learned.h")) as f:
        text = f.read()
    declared = {w.split("(")[0] for w in text.split("rgh_status ")[1:]}
    assert declared == set(learned.EXPORTS)


def test_create_checks_arguments_before_the_device():
    from ragdoll_amd import RghError, learned
    L = learned._bind()
    h = ctypes.c_void_p()
    assert L.rgh_pnet_create(0, 0, ctypes.byref(h)) == 1  # RGH_ERR_INVALID_ARGUMENT
    assert not h.value
    with pytest.raises(RghError):
        learned.ProposalNet(max_batch=-5)


def test_layout_constants_agree():
    from ragdoll_amd import learned
    assert learned.PARAMS == R.NPARAM == sum(int(np.prod(s)) for _, s in R.LAYOUT.values())
    assert learned.LAYOUT == R.LAYOUT
    with open(os.path.join(ROOT, "include", "ragdoll_learned.h")) as f:
        text = f.read()
    assert "#define RGH_PNET_PARAMS %d" % R.NPARAM in text


def test_e4m3_every_code_round_trips_and_midpoints_round_to_even():
    codes = np.arange(0x7F)
    vals = R.e4m3_value(codes)
    assert vals[0x7E] == 448.0 and vals[1] == 2.0 ** -9 and vals[8] == 2.0 ** -6
    assert np.all(np.diff(vals) > 0)
    assert np.array_equal(R.e4m3_code(vals), codes)
    assert np.array_equal(R.e4m3_code(-vals), codes | 0x80)
    mid = (vals[:-1] + vals[1:]) / 2
    got = R.e4m3_code(mid)
    want = np.where(codes[:-1] % 2 == 0, codes[:-1], codes[1:])
    assert np.array_equal(got, want)
    assert R.e4m3_code(np.array([1e6, 449.0, 448.0]))[0] == 0x7E
    # Just above / below a midpoint go to the nearer code.
    assert np.array_equal(R.e4m3_code(mid * (1 + 1e-9)), codes[1:])
    assert np.array_equal(R.e4m3_code(mid * (1 - 1e-9)), codes[:-1])


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -9, -2.5], np.float32)
    assert list(R.bf16_bits(x)) == [0x3F80, 0x3F80, 0x3F82, 0x3F80, 0xC020]


def test_features_match_a_brute_force_loop():
    rng = np.random.default_rng(3)
    H, W = 21, 37  # 3 x 2 macroblocks, partial ones, clamped borders everywhere
    cur = rng.integers(0, 256, (H, W))
    prev = rng.integers(0, 256, (H, W))
    got = R.bf16_value(R.features(cur, prev, 37))

    def px(img, x, y):
        return int(img[min(max(y, 0), H - 1), min(max(x, 0), W - 1)])

    def cell(img, x, y):
        return (sum(px(img, x + i, y + j) for j in range(4) for i in range(4)) + 8) >> 4

    for m in range(6):
        x0, y0 = (m % 3) * 16, (m // 3) * 16
        sc = [[cell(cur, x0 + 4 * i, y0 + 4 * j) for i in range(4)] for j in range(4)]
        for dy in range(-6, 7):
            for dx in range(-6, 7):
                s = sum(abs(sc[j][i] - cell(prev, x0 + 4 * (i + dx), y0 + 4 * (j + dy)))
                        for j in range(4) for i in range(4))
                want = R.bf16_round(np.float32(min(s, 2047) / 256.0))
                assert got[m, (dy + 6) * 13 + dx + 6] == want
        for j in range(4):
            for i in range(4):
                assert got[m, 169 + 4 * j + i] == R.bf16_round(np.float32(sc[j][i] / 64.0))
        samples = [px(cur, x0 + i, y0 + j) for j in range(16) for i in range(16)]
        mean = (sum(samples) + 128) >> 8
        mad = (sum(abs(s - mean) for s in samples) + 128) >> 8
        assert got[m, 185] == R.bf16_round(np.float32(mad / 16.0))
        assert got[m, 186] == R.bf16_round(np.float32(37 / 32.0))
        assert not got[m, 187:].any()


def test_labels_follow_the_nearest_cell():
    from ragdoll_amd import DECISION_DTYPE
    d = np.zeros(6, DECISION_DTYPE)
    d["mb_type"] = [0, 1, 1, 2, 3, 1]
    d["mv"] = [[0, 0], [8, -8], [-9, 24], [40, 40], [0, 0], [2000, -2000]]
    disp, mode = R.labels(d)
    # 8 -> cell 1, -8 -> cell 0 (floor((-8 + 8) / 16)), -9 -> -1, 24 -> 2, clamped to +-6.
    assert list(mode) == [0, 1, 1, 2, 2, 1]
    assert list(disp) == [84, 6 * 13 + 7, 8 * 13 + 5, -1, -1, 0 * 13 + 12]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(7)
    from ragdoll_amd.learned import init_params
    P = init_params(1).astype(np.float64)
    P[R.LAYOUT["b1"][0]:R.LAYOUT["b1"][0] + 128] = rng.standard_normal(128) * 0.1
    P[R.LAYOUT["b3"][0]:R.LAYOUT["b3"][0] + 176] = rng.standard_normal(176) * 0.1
    B = 6
    X = rng.random((B, 192)) * 4
    disp = np.array([3, -1, 84, 168, 0, 50])
    mode = np.array([0, 2, 1, 1, 0, 2])
    ld, lm, g = R.loss_and_grads(X, disp, mode, P, emulate=False)

    def loss(p):
        a, b, _ = R.loss_and_grads(X, disp, mode, p, emulate=False)
        return (a + b) / B

    assert np.isclose((ld + lm) / B, loss(P))
    idx = []
    for name, (off, shape) in R.LAYOUT.items():
        idx += list(off + rng.choice(int(np.prod(shape)), 6, replace=False))
    for i in idx:
        e = 1e-6
        p1, p2 = P.copy(), P.copy()
        p1[i] += e
        p2[i] -= e
        fd = (loss(p1) - loss(p2)) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_quantize_reference_reconstructs_the_weights():
    from ragdoll_amd.learned import init_params
    P = init_params(2)
    codes, scales, mul = R.quantize(P, 3.0, 5.0)
    W = R.unpack(P)
    for l, name in enumerate(("W1", "W2", "W3")):
        w = W[name]
        back = R.e4m3_value(codes[l]).T * (scales[l, :w.shape[1]] * mul[l])[None, :]
        # E4M3 keeps 3 mantissa bits: within 1/16 relative, plus the subnormal step.
        assert np.all(np.abs(back - w) <= np.abs(w) / 16 + np.abs(w).max(0) / 448 * 2.0 ** -9 + 1e-12)


@pytest.mark.parametrize("method", ["pchip", "cubic"])
def test_bd_rate_of_a_scaled_curve_is_the_scale(method):
    from ragdoll_amd.rd import bd_rate, psnr
    rate = np.array([1000.0, 1800.0, 3500.0, 7000.0])
    q = np.array([30.0, 33.0, 36.5, 40.0])
    assert abs(bd_rate(rate, q, rate, q, method)) < 1e-9
    assert abs(bd_rate(rate, q, rate * 0.9, q, method) + 10.0) < 1e-6
    assert abs(bd_rate(rate, q, rate * 1.01, q, method) - 1.0) < 1e-6
    # A curve shifted up in PSNR reaches the same quality with fewer bits.
    assert bd_rate(rate, q, rate, q + 0.5, method) < 0
    a = np.full((4, 4), 10, np.uint8)
    assert psnr(a, a) == 100.0
    assert abs(psnr(a, a + 1) - 10 * np.log10(255.0 ** 2)) < 1e-9


def test_pchip_integral_matches_scipy():
    # The piecewise cubic BD-rate's interpolant and its integral against scipy's
    # PchipInterpolator (the same Fritsch-Carlson construction) on random monotone and
    # non-monotone data, over partial intervals.
    from scipy.interpolate import PchipInterpolator
    from ragdoll_amd.rd import pchip_integral
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(3, 8))
        x = np.sort(rng.uniform(25, 45, n))
        if np.any(np.diff(x) < 1e-3):
            continue
        y = np.cumsum(rng.uniform(-0.5 if trial % 3 == 0 else 0.05, 1.0, n))
        lo, hi = np.sort(rng.uniform(x[0], x[-1], 2))
        want = PchipInterpolator(x, y).integrate(lo, hi)
        assert abs(pchip_integral(x, y, lo, hi) - want) < 1e-9 * max(1.0, abs(want)), trial
