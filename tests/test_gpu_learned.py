"""GPU checks of the proposal network's kernels against the numpy reference (tests/pnet_ref.py).

Integer work (features, labels, E4M3 coding, quantized weights) must match exactly; floating
point within stated tolerances: the bfloat16 GEMM within fp32 accumulation error (1e-5 of the
operands' scale), gradients within 1e-2 (relative, per tensor; the device and the reference
round the same bfloat16 intermediates, but a different fp32 summation order can move one
across a rounding boundary), FP8 inference logits within 3e-2 of their range (an activation
rounded to the other side of an E4M3 midpoint moves by 1/16 of itself).
"""
import numpy as np
import pytest

import pnet_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    from ragdoll_amd.learned import ProposalNet
    n = ProposalNet(max_batch=1024)
    yield n
    n.close()


def rand_bf16(rng, shape, scale=1.0):
    return R.bf16_bits((rng.standard_normal(shape) * scale).astype(np.float32))


@pytest.mark.parametrize("M,N,K,trans", [(64, 64, 32, False), (100, 70, 45, False), (37, 176, 130, True),
                                         (192, 128, 300, True)])
def test_gemm_bf16_matches_numpy(net, M, N, K, trans):
    print("gemm", M, N, K, trans, flush=True)
    rng = np.random.default_rng(M * 7 + N)
    hip = net.hip
    A = rand_bf16(rng, (K, M) if trans else (M, K))
    B = rand_bf16(rng, (K, N))
    bias = rng.standard_normal(N).astype(np.float32)
    ldc = N + 3
    C0 = rng.standard_normal((M, ldc)).astype(np.float32)
    mask = R.bf16_bits((rng.random((M, ldc)) > 0.3).astype(np.float32))
    Av = R.bf16_value(A).T if trans else R.bf16_value(A)
    ref = Av @ R.bf16_value(B)
    sam, sak = (1, M) if trans else (K, 1)
    bufs = [hip.upload(A), hip.upload(B), hip.upload(bias), hip.upload(C0), hip.alloc(M * ldc * 2), hip.upload(mask)]
    dA, dB, dbias, dC, dCb, dmask = bufs
    try:
        # Plain product.
        net.gemm_bf16(dA, sam, sak, dB, N, 1, dC, None, ldc, None, None, M, N, K)
        got = hip.download(dC, np.zeros((M, ldc), np.float32))
        tol = 1e-5 * (np.abs(Av) @ np.abs(R.bf16_value(B))) + 1e-6
        assert np.all(np.abs(got[:, :N] - ref) <= tol)
        assert np.array_equal(got[:, N:], C0[:, N:])  # columns beyond N untouched
        # Accumulate + bias + ReLU + mask, fp32 and bfloat16 outputs.
        net.gemm_bf16(dA, sam, sak, dB, N, 1, dC, dCb, ldc, dbias, dmask, M, N, K, relu=1, accumulate=1)
        got2 = hip.download(dC, np.zeros((M, ldc), np.float32))
        want = np.maximum(got[:, :N] + ref + bias, 0) * (R.bf16_value(mask[:, :N]) != 0)
        assert np.all(np.abs(got2[:, :N] - want) <= 2 * tol + 1e-5 * np.abs(want))
        gotb = hip.download(dCb, np.zeros((M, ldc), np.uint16))
        assert np.array_equal(gotb[:, :N], R.bf16_bits(got2[:, :N]))
    finally:
        for b in bufs:
            hip.free(b)


def synth_pair(rng, H, W, shift=(0, 0)):
    """A smooth random texture and a copy displaced by shift (dx, dy) with edge clamping."""
    base = rng.integers(0, 256, (H // 4 + 3, W // 4 + 3)).astype(np.float64)
    big = np.kron(base, np.ones((4, 4)))[:H + 8, :W + 8]
    prev = np.clip(big[4:4 + H, 4:4 + W] + rng.normal(0, 2, (H, W)), 0, 255).astype(np.uint8)
    ys = np.clip(np.arange(H) - shift[1], 0, H - 1)
    xs = np.clip(np.arange(W) - shift[0], 0, W - 1)
    cur = prev[np.ix_(ys, xs)]
    return cur, prev


@pytest.mark.parametrize("H,W", [(48, 64), (70, 100), (90, 160), (16, 16), (1, 1)])
def test_features_match_numpy(net, H, W):
    print("features", H, W, flush=True)
    rng = np.random.default_rng(H * W)
    cur, prev = synth_pair(rng, H, W, (5, -3))
    got = net.features_host(cur, prev, qp=H % 52)
    assert np.array_equal(got, R.features(cur, prev, H % 52))
    noise = rng.integers(0, 256, (H, W)).astype(np.uint8)
    assert np.array_equal(net.features_host(noise, cur, 51), R.features(noise, cur, 51))


def test_labels_match_numpy(net):
    from ragdoll_amd import DECISION_DTYPE
    rng = np.random.default_rng(5)
    d = np.zeros(3000, DECISION_DTYPE)
    d["mb_type"] = rng.integers(0, 6, 3000)
    d["mv"] = rng.integers(-600, 600, (3000, 2))
    got = net.labels_host(d)
    disp, mode = R.labels(d)
    assert np.array_equal(got["disp"], disp) and np.array_equal(got["mode"], mode)


def test_fp8_encoding_matches_the_reference(net):
    rng = np.random.default_rng(11)
    codes = np.arange(0x7F)
    vals = R.e4m3_value(codes)
    mids = (vals[:-1] + vals[1:]) / 2
    x = np.concatenate([vals, -vals[1:], mids, -mids, rng.uniform(-448, 448, 5000),
                        rng.uniform(-0.02, 0.02, 2000), [448.0, -448.0]]).astype(np.float32)
    hip = net.hip
    dx, do = hip.upload(x), hip.alloc(len(x))
    try:
        net.fp8_encode(dx, len(x), do)
        got = hip.download(do, np.zeros(len(x), np.uint8))
    finally:
        hip.free(dx)
        hip.free(do)
    want = R.e4m3_code(x)
    # A negative value that rounds to zero may be coded +0 or -0: both decode to zero.
    same = (got == want) | ((got & 0x7F) == 0) & ((want & 0x7F) == 0)
    bad = np.nonzero(~same)[0]
    assert len(bad) == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:8]]


def make_batch(rng, B):
    X = R.bf16_bits((rng.random((B, 192)) * 3).astype(np.float32))
    X[:, 187:] = 0
    from ragdoll_amd.learned import LABEL_DTYPE
    lab = np.zeros(B, LABEL_DTYPE)
    lab["mode"] = rng.integers(0, 3, B)
    lab["disp"] = np.where(lab["mode"] == 2, -1, rng.integers(0, 169, B))
    return X, lab


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_numpy(net):
    from ragdoll_amd.learned import init_params
    rng = np.random.default_rng(21)
    P = init_params(3)
    P[R.LAYOUT["b1"][0]:R.LAYOUT["b1"][0] + 128] = rng.standard_normal(128).astype(np.float32) * 0.1
    net.set_params(P)
    B = 300
    X, lab = make_batch(rng, B)
    hip = net.hip
    dX, dL = hip.upload(X), hip.upload(lab)
    try:
        g, z = net.gradients(dX, dL, B)
        loss = net.train_step(dX, dL, B, 1e-3, want_loss=True)
    finally:
        hip.free(dX)
        hip.free(dL)
    _, _, zr = R.forward(R.bf16_value(X), P)
    assert np.all(np.abs(z - zr) <= 1e-2 * np.abs(zr).max())
    ld, lm, gr = R.loss_and_grads(R.bf16_value(X), lab["disp"].astype(np.int64), lab["mode"].astype(np.int64), P)
    assert abs(loss[0] - ld) <= 1e-3 * ld and abs(loss[1] - lm) <= 1e-3 * lm
    for name, (off, shape) in R.LAYOUT.items():
        n = int(np.prod(shape))
        e = rel_err(g[off:off + n], gr[off:off + n])
        print(name, "relative error", e, flush=True)
        assert e < 1e-2, name
    # The Adam step from zero moments: p - lr * g / (|g| + eps) elementwise (bias-corrected).
    P1 = net.get_params()
    step = 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(P1, P - step, rtol=1e-5, atol=1e-6)


def test_quantize_and_fp8_inference_match_numpy(net):
    from ragdoll_amd.learned import init_params
    rng = np.random.default_rng(31)
    P = init_params(4)
    P[R.LAYOUT["b3"][0]:R.LAYOUT["b3"][0] + 176] = rng.standard_normal(176).astype(np.float32)
    net.set_params(P)
    n = 700
    X, _ = make_batch(rng, n)
    hip = net.hip
    dX = hip.upload(X)
    try:
        net.quantize(dX, n)
    finally:
        hip.free(dX)
    w8, sc = net.get_quantized()
    mul = [float(v) for v in sc[3 * 176:]]
    h1, h2, _ = R.forward(R.bf16_value(X), P)
    assert np.isclose(mul[0], 1.0)
    assert np.isclose(mul[1], 448.0 / h1.max(), rtol=1e-2) and np.isclose(mul[2], 448.0 / h2.max(), rtol=1e-2)
    codes, scales, _ = R.quantize(P, 448.0 / mul[1], 448.0 / mul[2])
    assert np.array_equal(w8, np.concatenate([c.ravel() for c in codes]))
    assert np.allclose(sc[:3 * 176].reshape(3, 176), scales, rtol=1e-6)
    props, z = net.infer_host(X)
    zr = R.infer(R.bf16_value(X), w8, sc, mul, P)
    span = np.abs(zr).max()
    err = np.abs(z - zr)
    print("fp8 logits: max error %.4g, median %.4g of span %.4g" % (err.max(), np.median(err), span), flush=True)
    assert err.max() <= 3e-2 * span
    # The proposals are the four largest displacement logits of the device's own logits.
    order = np.argsort(-z[:, :169], axis=1, kind="stable")[:, :4]
    assert np.array_equal(props["cell"][:, :, 0], order % 13 - 6)
    assert np.array_equal(props["cell"][:, :, 1], order // 13 - 6)
    zm = z[:, 169:172] - z[:, 169:172].max(1, keepdims=True)
    pm = np.exp(zm) / np.exp(zm).sum(1, keepdims=True)
    assert np.all(np.abs(props["mode_prob"].astype(np.int64) - np.round(pm * 255)) <= 1)
    # The FP8 network stays close to the bfloat16 one it was quantized from.
    zb = net.forward_host(X)
    assert np.abs(z - zb).max() <= 0.1 * np.abs(zb).max()


def test_fused_proposals_equal_features_then_inference():
    # rgh_pnet_propose_batch (features computed into the inference tiles) against
    # rgh_pnet_features_batch + rgh_pnet_infer: the same proposals and logits, bit for bit, for a
    # batch of pictures whose macroblock counts leave partial tiles (and one picture of one MB).
    from ragdoll_amd.learned import PROPOSAL_DTYPE, load_trained
    net = load_trained(max_batch=1024)
    hip = net.hip
    rng = np.random.default_rng(77)
    sizes = [(144, 176), (90, 160), (16, 16), (200, 120), (48, 64)]
    pics, bufs = [], []
    try:
        for k, (H, W) in enumerate(sizes):
            cur, prev = synth_pair(rng, H, W, (k - 2, 3 - k))
            dc, dp = hip.upload(cur), hip.upload(prev)
            bufs += [dc, dp]
            pics.append((dc, W, dp, W, W, H, 20 + 3 * k))
        rows = sum(((W + 15) // 16) * ((H + 15) // 16) for H, W in sizes)
        feat = hip.alloc(rows * 192 * 2)
        p1, p2 = hip.alloc(rows * 16), hip.alloc(rows * 16)
        z1, z2 = hip.alloc(rows * 176 * 4), hip.alloc(rows * 176 * 4)
        bufs += [feat, p1, p2, z1, z2]
        jobs, r = [], 0
        for (c, cs, p, ps, w, h, qp) in pics:
            jobs.append((c, cs, p, ps, w, h, qp, feat + r * 192 * 2))
            r += ((w + 15) // 16) * ((h + 15) // 16)
        net.features_batch(jobs)
        net.infer(feat, rows, p1, z1)
        net.propose_batch(pics, p2, z2)
        a = hip.download(p1, np.zeros(rows, PROPOSAL_DTYPE))
        b = hip.download(p2, np.zeros(rows, PROPOSAL_DTYPE))
        za = hip.download(z1, np.zeros((rows, 176), np.float32))
        zb = hip.download(z2, np.zeros((rows, 176), np.float32))
        assert np.array_equal(za.view(np.uint32), zb.view(np.uint32))
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    finally:
        for q in bufs:
            hip.free(q)
        net.close()


def test_training_learns_a_known_displacement(net):
    """Pictures displaced by a whole number of cells: the network must learn to point there."""
    from ragdoll_amd.learned import LABEL_DTYPE, init_params
    rng = np.random.default_rng(41)
    feats, labs = [], []
    for i in range(40):
        dx, dy = rng.integers(-5, 6, 2)
        # cur(x, y) = prev(x - sx, y - sy): the motion vector (reference - current) is -s.
        cur, prev = synth_pair(rng, 96, 128, (-4 * dx, -4 * dy))
        f = net.features_host(cur, prev).reshape(6, 8, 192)[1:-1, 1:-1].reshape(-1, 192)  # interior macroblocks
        lab = np.zeros(len(f), LABEL_DTYPE)
        lab["disp"] = (dy + 6) * 13 + dx + 6
        lab["mode"] = 1
        feats.append(f)
        labs.append(lab)
    X, lab = np.concatenate(feats), np.concatenate(labs)
    net.set_params(init_params(5))
    hip = net.hip
    dX, dL = hip.upload(X), hip.upload(lab)
    try:
        first = net.train_step(dX, dL, len(X), 3e-3, want_loss=True)
        for _ in range(300):
            net.train_step(dX, dL, len(X), 3e-3)
        last = net.train_step(dX, dL, len(X), 3e-3, want_loss=True)
        net.quantize(dX, len(X))
    finally:
        hip.free(dX)
        hip.free(dL)
    print("loss", first / len(X), "->", last / len(X), flush=True)
    assert last[0] < 0.25 * first[0]
    props, _ = net.infer_host(X)
    cls = (props["cell"][:, 0, 1].astype(int) + 6) * 13 + props["cell"][:, 0, 0] + 6
    acc = float(np.mean(cls == lab["disp"]))
    cells = props["cell"].astype(np.int64)
    top4 = float(np.mean(((cells[:, :, 1] + 6) * 13 + cells[:, :, 0] + 6 == lab["disp"][:, None]).any(axis=1)))
    print("fp8 top-1 %.3f top-4 %.3f" % (acc, top4), flush=True)
    assert acc > 0.9 and top4 > 0.95
