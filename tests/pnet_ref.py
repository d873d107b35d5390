"""numpy reference of the proposal network (include/ragdoll_learned.h), for the tests.

Features are integer arithmetic (exact). The network runs either in float64 (exact, used to
pin the backward pass against finite differences) or emulating the device's arithmetic:
bfloat16 operands and activations (round to nearest even) with fp32-class accumulation for
training, E4M3 operands (round to nearest even, saturating at 448) for inference.
"""
import numpy as np

F_IN, HID, OUT, NCLS, MODE0 = 192, 128, 176, 169, 169
LAYOUT = {"W1": (0, (F_IN, HID)), "b1": (24576, (HID,)), "W2": (24704, (HID, HID)),
          "b2": (41088, (HID,)), "W3": (41216, (HID, OUT)), "b3": (63744, (OUT,))}
NPARAM = 63920


def unpack(p):
    return {k: np.asarray(p[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in LAYOUT.items()}


def pack(d):
    p = np.zeros(NPARAM, np.float64)
    for k, (o, s) in LAYOUT.items():
        p[o:o + int(np.prod(s))] = np.asarray(d[k]).ravel()
    return p


# ---- number formats ----
def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_round(x):
    return bf16_value(bf16_bits(np.asarray(x, np.float32)))


def e4m3_value(code):
    c = np.asarray(code, np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 15
    m = c & 7
    v = np.where(e == 0, m / 8.0 * 2.0 ** -6, (1 + m / 8.0) * 2.0 ** (e - 7))
    return s * v


_E4M3_POS = e4m3_value(np.arange(0x7F))  # codes 0..0x7E, increasing


def e4m3_code(x):
    """Round to nearest even onto E4M3 (|x| saturates at 448)."""
    x = np.asarray(x, np.float64)
    a = np.minimum(np.abs(x), 448.0)
    hi = np.searchsorted(_E4M3_POS, a, side="left")  # first code with value >= a
    hi = np.minimum(hi, 0x7E)
    lo = np.maximum(hi - 1, 0)
    dlo, dhi = a - _E4M3_POS[lo], _E4M3_POS[hi] - a
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (hi % 2 == 0))
    code = np.where(_E4M3_POS[hi] == a, hi, np.where(pick_hi, hi, lo))
    return (code | np.where(np.signbit(x), 0x80, 0)).astype(np.uint8)


# ---- features ----
def features(cur, prev, qp=26):
    """Feature rows ([n][192] bfloat16 bits) of every macroblock."""
    cur = np.asarray(cur, np.int64)
    prev = np.asarray(prev, np.int64)
    H, W = cur.shape
    mb_w, mb_h = (W + 15) // 16, (H + 15) // 16
    out = np.zeros((mb_w * mb_h, F_IN), np.float32)

    def cells(img, x0, y0, n):
        ys = np.clip(y0 + np.arange(4 * n), 0, H - 1)
        xs = np.clip(x0 + np.arange(4 * n), 0, W - 1)
        blk = img[np.ix_(ys, xs)].reshape(n, 4, n, 4).sum(axis=(1, 3))
        return (blk + 8) >> 4

    for m in range(mb_w * mb_h):
        x0, y0 = (m % mb_w) * 16, (m // mb_w) * 16
        rc = cells(prev, x0 - 24, y0 - 24, 16)
        sc = cells(cur, x0, y0, 4)
        win = np.lib.stride_tricks.sliding_window_view(rc, (4, 4))  # [dy][dx][4][4]
        cost = np.abs(win - sc[None, None]).sum(axis=(2, 3)).ravel()
        out[m, :NCLS] = np.minimum(cost, 2047) / 256.0
        out[m, NCLS:NCLS + 16] = sc.ravel() / 64.0
        ys = np.clip(y0 + np.arange(16), 0, H - 1)
        xs = np.clip(x0 + np.arange(16), 0, W - 1)
        blk = cur[np.ix_(ys, xs)]
        mean = (int(blk.sum()) + 128) >> 8
        out[m, NCLS + 16] = ((int(np.abs(blk - mean).sum()) + 128) >> 8) / 16.0
        out[m, NCLS + 17] = qp / 32.0
    return bf16_bits(out)


def labels(decisions):
    t = decisions["mb_type"].astype(np.int64)
    intra = (t == 2) | (t == 3) | (t == 4)
    cx = np.clip((decisions["mv"][:, 0].astype(np.int64) + 8) >> 4, -6, 6)
    cy = np.clip((decisions["mv"][:, 1].astype(np.int64) + 8) >> 4, -6, 6)
    disp = np.where(intra, -1, (cy + 6) * 13 + cx + 6)
    mode = np.where(intra, 2, np.where(t == 0, 0, 1))
    return disp.astype(np.int16), mode.astype(np.int16)


# ---- network ----
def forward(X, P, emulate=True):
    """Activations of the bfloat16 training forward pass. X: feature values (float)."""
    r = bf16_round if emulate else (lambda a: a)
    W = unpack(P)
    X = np.asarray(X, np.float64)
    a1 = X @ r(W["W1"]) + W["b1"]
    h1 = r(np.maximum(a1, 0))
    a2 = h1 @ r(W["W2"]) + W["b2"]
    h2 = r(np.maximum(a2, 0))
    z = h2 @ r(W["W3"]) + W["b3"]
    return h1, h2, z


def _log_softmax(z):
    m = z.max(axis=1, keepdims=True)
    return z - m - np.log(np.exp(z - m).sum(axis=1, keepdims=True))


def loss_and_grads(X, disp, mode, P, emulate=True):
    """(displacement loss sum, mode loss sum, flat gradient of their mean over the batch)."""
    r = bf16_round if emulate else (lambda a: a)
    W = unpack(P)
    X = np.asarray(X, np.float64)
    B = X.shape[0]
    h1, h2, z = forward(X, P, emulate)
    dz = np.zeros_like(z)
    has = disp >= 0
    ls = _log_softmax(z[:, :NCLS])
    ld = -ls[np.arange(B)[has], disp[has]].sum()
    sd = np.exp(ls)
    sd[np.arange(B)[has], disp[has]] -= 1
    dz[has, :NCLS] = sd[has]
    lm_s = _log_softmax(z[:, MODE0:MODE0 + 3])
    lm = -lm_s[np.arange(B), mode].sum()
    sm = np.exp(lm_s)
    sm[np.arange(B), mode] -= 1
    dz[:, MODE0:MODE0 + 3] = sm
    dz /= B
    dzb = r(dz)
    g = {"W3": h2.T @ dzb, "b3": dz.sum(0)}
    dh2 = (dzb @ r(W["W3"]).T) * (h2 > 0)
    dh2b = r(dh2)
    g["W2"], g["b2"] = h1.T @ dh2b, dh2.sum(0)
    dh1 = (dh2b @ r(W["W2"]).T) * (h1 > 0)
    dh1b = r(dh1)
    g["W1"], g["b1"] = X.T @ dh1b, dh1.sum(0)
    return ld, lm, pack(g)


def quantize(P, h1max, h2max):
    """(w8 codes [out][in] per layer, column scales [3][176], input multipliers [3])."""
    W = unpack(P)
    mul = [1.0, 448.0 / h1max if h1max > 0 else 1.0, 448.0 / h2max if h2max > 0 else 1.0]
    mul = [float(np.float32(m)) for m in mul]
    codes, scales = [], np.zeros((3, OUT))
    for l, name in enumerate(("W1", "W2", "W3")):
        w = W[name].astype(np.float32)
        wmax = np.abs(w).max(axis=0)
        ws = np.where(wmax > 0, wmax / np.float32(448.0), 1.0).astype(np.float32)
        inv = (np.float32(1.0) / ws).astype(np.float32)
        codes.append(e4m3_code((w * inv[None, :]).astype(np.float32)).T.copy())
        scales[l, :w.shape[1]] = (ws / np.float32(mul[l])).astype(np.float32)
    return codes, scales, mul


def infer(X, w8, scales, mul, P):
    """FP8 inference logits from feature values X and the device's quantized network."""
    W = unpack(P)
    sizes = [(F_IN, HID), (HID, HID), (HID, OUT)]
    off = 0
    mats = []
    for k, n in sizes:
        mats.append(e4m3_value(w8[off:off + k * n].reshape(n, k)))
        off += k * n
    sc = scales[:3 * OUT].reshape(3, OUT)
    x = e4m3_value(e4m3_code(np.clip(np.asarray(X, np.float32) * np.float32(mul[0]), -448, 448)))
    for l, b in enumerate(("b1", "b2", "b3")):
        n = sizes[l][1]
        acc = x @ mats[l].T
        v = acc * sc[l, :n] + W[b]
        if l == 2:
            return v
        x = e4m3_value(e4m3_code(np.minimum(np.maximum(v, 0) * mul[l + 1], 448.0)))
