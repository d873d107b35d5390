"""ctypes binding of include/ragdoll_learned.h: the learned motion-proposal network.

`ProposalNet` owns one network on a device (fp32 master parameters, Adam state, the FP8
copy) and drives its kernels: per-macroblock features from two source pictures, labels from
analysed decisions, bfloat16 training steps, FP8 quantization and inference into
`rgh_proposal` records that `analyse_frames(..., proposals=...)` consumes. Device buffers come
from the HIP runtime the library links (ragdoll_amd.hip); nothing here computes on the CPU.
"""
import ctypes
import json
import os

import numpy as np

from .abi import DECISION_DTYPE, _check, load_library

FEATURES = 192
HIDDEN = 128
OUTPUTS = 176
CELLS = 13
NCLS = CELLS * CELLS
MODE_OFFSET = 169
PARAMS = 63920
W8_BYTES = 192 * 128 + 128 * 128 + 176 * 128
SCALES = 3 * 176 + 3
# Offsets of the flat parameter vector: W1 [192][128], b1, W2 [128][128], b2, W3 [128][176], b3.
LAYOUT = {"W1": (0, (FEATURES, HIDDEN)), "b1": (24576, (HIDDEN,)), "W2": (24704, (HIDDEN, HIDDEN)),
          "b2": (41088, (HIDDEN,)), "W3": (41216, (HIDDEN, OUTPUTS)), "b3": (63744, (OUTPUTS,))}

LABEL_DTYPE = np.dtype([("disp", "<i2"), ("mode", "<i2")])
PROPOSAL_DTYPE = np.dtype([("cell", "i1", (4, 2)), ("mode_prob", "u1", (3,)), ("reserved", "u1", (5,))])
assert PROPOSAL_DTYPE.itemsize == 16

EXPORTS = ["rgh_pnet_create", "rgh_pnet_destroy", "rgh_pnet_set_launch_limit", "rgh_pnet_set_params", "rgh_pnet_get_params",
           "rgh_pnet_features", "rgh_pnet_features_batch", "rgh_pnet_labels", "rgh_pnet_gather", "rgh_pnet_train_step",
           "rgh_pnet_gradients", "rgh_pnet_forward", "rgh_pnet_quantize", "rgh_pnet_quantize_ranges", "rgh_pnet_infer",
           "rgh_pnet_get_quantized", "rgh_pnet_fp8_encode", "rgh_pnet_gemm_bf16", "rgh_pnet_propose_batch",
           "rgh_analyse_frames_learned"]

_bound = False


class RgPnetPicture(ctypes.Structure):
    _fields_ = [("cur_y", ctypes.c_void_p), ("prev_y", ctypes.c_void_p), ("features", ctypes.c_void_p),
                ("cur_stride", ctypes.c_int32), ("prev_stride", ctypes.c_int32), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("qp", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def _bind():
    global _bound
    L = load_library()
    if _bound:
        return L
    P, I32, I64, F32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    sig = {
        "rgh_pnet_create": [I32, I32, ctypes.POINTER(ctypes.c_void_p)],
        "rgh_pnet_destroy": [P],
        "rgh_pnet_set_launch_limit": [P, I32],
        "rgh_pnet_set_params": [P, P, I64],
        "rgh_pnet_get_params": [P, P, I64],
        "rgh_pnet_features": [P, P, I32, P, I32, I32, I32, I32, P],
        "rgh_pnet_features_batch": [P, P, I32],
        "rgh_pnet_labels": [P, P, I32, P],
        "rgh_pnet_gather": [P, P, P, P, I32, P, P],
        "rgh_pnet_train_step": [P, P, P, I32, F32, P],
        "rgh_pnet_gradients": [P, P, P, I32, P, P],
        "rgh_pnet_forward": [P, P, I32, P],
        "rgh_pnet_quantize": [P, P, I32],
        "rgh_pnet_quantize_ranges": [P, F32, F32],
        "rgh_pnet_infer": [P, P, I32, P, P],
        "rgh_pnet_propose_batch": [P, P, I32, P, P],
        "rgh_pnet_get_quantized": [P, P, P],
        "rgh_pnet_fp8_encode": [P, P, I32, P],
        "rgh_pnet_gemm_bf16": [P, P, I64, I64, P, I64, I64, P, P, I64, P, P, I32, I32, I32, I32, I32],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.restype = I32
        fn.argtypes = args
    _bound = True
    return L


def init_params(seed=0):
    """He-initialised parameters (fp32, flat layout) from a seeded generator; biases zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for name, (off, shape) in LAYOUT.items():
        if name.startswith("W"):
            w = rng.standard_normal(shape).astype(np.float32) * np.float32(np.sqrt(2.0 / shape[0]))
            p[off:off + w.size] = w.ravel()
    return p


class ProposalNet:
    """One proposal network on a device. Pointer arguments named *_dev are device addresses."""

    def __init__(self, max_batch=4096, device=0, params=None):
        from .hip import Hip
        self.L = _bind()
        self.hip = Hip()
        h = ctypes.c_void_p()
        _check(self.L.rgh_pnet_create(device, max_batch, ctypes.byref(h)), "rgh_pnet_create")
        self.h = h
        self.max_batch = max_batch
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "h", None):
            self.L.rgh_pnet_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_launch_limit(self, max_workgroups):
        _check(self.L.rgh_pnet_set_launch_limit(self.h, max_workgroups), "rgh_pnet_set_launch_limit")

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.shape != (PARAMS,):
            raise ValueError("expected %d parameters" % PARAMS)
        _check(self.L.rgh_pnet_set_params(self.h, p.ctypes.data, PARAMS), "rgh_pnet_set_params")

    def get_params(self):
        p = np.zeros(PARAMS, np.float32)
        _check(self.L.rgh_pnet_get_params(self.h, p.ctypes.data, PARAMS), "rgh_pnet_get_params")
        return p

    def features(self, cur_y_dev, cur_stride, prev_y_dev, prev_stride, width, height, qp, out_dev):
        _check(self.L.rgh_pnet_features(self.h, cur_y_dev, cur_stride, prev_y_dev, prev_stride, width, height, qp,
                                        out_dev), "rgh_pnet_features")

    def features_batch(self, pictures):
        """pictures: (cur_y_dev, cur_stride, prev_y_dev, prev_stride, width, height, qp, out_dev) tuples."""
        arr = (RgPnetPicture * max(1, len(pictures)))()
        for i, (c, cs, p, ps, w, h, qp, out) in enumerate(pictures):
            arr[i] = RgPnetPicture(c, p, out, cs, ps, w, h, qp, 0)
        _check(self.L.rgh_pnet_features_batch(self.h, arr, len(pictures)), "rgh_pnet_features_batch")

    def propose_batch(self, pictures, proposals_dev, logits_dev=None):
        """Features and FP8 inference in one pass. pictures: (cur_y_dev, cur_stride, prev_y_dev,
        prev_stride, width, height, qp) tuples; their proposals fill consecutive rows of
        proposals_dev, picture after picture."""
        arr = (RgPnetPicture * max(1, len(pictures)))()
        for i, (c, cs, p, ps, w, h, qp) in enumerate(pictures):
            arr[i] = RgPnetPicture(c, p, None, cs, ps, w, h, qp, 0)
        _check(self.L.rgh_pnet_propose_batch(self.h, arr, len(pictures), proposals_dev, logits_dev),
               "rgh_pnet_propose_batch")

    def labels(self, decisions_dev, n, out_dev):
        _check(self.L.rgh_pnet_labels(self.h, decisions_dev, n, out_dev), "rgh_pnet_labels")

    def gather(self, feat_dev, lab_dev, idx_dev, n, feat_out_dev, lab_out_dev):
        _check(self.L.rgh_pnet_gather(self.h, feat_dev, lab_dev, idx_dev, n, feat_out_dev, lab_out_dev),
               "rgh_pnet_gather")

    def train_step(self, feat_dev, lab_dev, batch, lr, want_loss=False):
        loss = np.zeros(2, np.float32)
        _check(self.L.rgh_pnet_train_step(self.h, feat_dev, lab_dev, batch, lr,
                                          loss.ctypes.data if want_loss else None), "rgh_pnet_train_step")
        return loss if want_loss else None

    def gradients(self, feat_dev, lab_dev, batch):
        g = np.zeros(PARAMS, np.float32)
        z = np.zeros((batch, OUTPUTS), np.float32)
        _check(self.L.rgh_pnet_gradients(self.h, feat_dev, lab_dev, batch, g.ctypes.data, z.ctypes.data),
               "rgh_pnet_gradients")
        return g, z

    def forward(self, feat_dev, n, logits_dev):
        _check(self.L.rgh_pnet_forward(self.h, feat_dev, n, logits_dev), "rgh_pnet_forward")

    def quantize(self, calib_dev, n):
        _check(self.L.rgh_pnet_quantize(self.h, calib_dev, n), "rgh_pnet_quantize")

    def quantize_ranges(self, h1_max, h2_max):
        _check(self.L.rgh_pnet_quantize_ranges(self.h, h1_max, h2_max), "rgh_pnet_quantize_ranges")

    def activation_maxima(self):
        """The hidden-activation maxima the FP8 copy was built with."""
        _, sc = self.get_quantized()
        return [float(448.0 / sc[3 * 176 + 1]), float(448.0 / sc[3 * 176 + 2])]

    def infer(self, feat_dev, n, proposals_dev, logits_dev=None):
        _check(self.L.rgh_pnet_infer(self.h, feat_dev, n, proposals_dev, logits_dev), "rgh_pnet_infer")

    def get_quantized(self):
        w8 = np.zeros(W8_BYTES, np.uint8)
        sc = np.zeros(SCALES, np.float32)
        _check(self.L.rgh_pnet_get_quantized(self.h, w8.ctypes.data, sc.ctypes.data), "rgh_pnet_get_quantized")
        return w8, sc

    def fp8_encode(self, x_dev, n, out_dev):
        _check(self.L.rgh_pnet_fp8_encode(self.h, x_dev, n, out_dev), "rgh_pnet_fp8_encode")

    def gemm_bf16(self, A, sam, sak, B, sbk, sbn, C, Cb, ldc, bias, mask, M, N, K, relu=0, accumulate=0):
        _check(self.L.rgh_pnet_gemm_bf16(self.h, A, sam, sak, B, sbk, sbn, C, Cb, ldc, bias, mask, M, N, K, relu,
                                         accumulate), "rgh_pnet_gemm_bf16")

    # ---- host-array conveniences (upload, run, download) ----
    def features_host(self, cur_y, prev_y, qp=26):
        """Features of every macroblock of two host luma planes (same size)."""
        cur_y = np.ascontiguousarray(cur_y, np.uint8)
        prev_y = np.ascontiguousarray(prev_y, np.uint8)
        if cur_y.shape != prev_y.shape:
            raise ValueError("pictures differ in size")
        H, W = cur_y.shape
        n = ((W + 15) // 16) * ((H + 15) // 16)
        hip = self.hip
        c, p, f = hip.upload(cur_y), hip.upload(prev_y), hip.alloc(n * FEATURES * 2)
        try:
            self.features(c, W, p, W, W, H, qp, f)
            return hip.download(f, np.zeros((n, FEATURES), np.uint16))
        finally:
            for q in (c, p, f):
                hip.free(q)

    def infer_host(self, features):
        """(proposals, logits) of host feature rows through the FP8 network."""
        x = np.ascontiguousarray(features, np.uint16)
        n = x.shape[0]
        hip = self.hip
        xd, pd, zd = hip.upload(x), hip.alloc(n * 16), hip.alloc(n * OUTPUTS * 4)
        try:
            self.infer(xd, n, pd, zd)
            props = hip.download(pd, np.zeros(n, PROPOSAL_DTYPE))
            return props, hip.download(zd, np.zeros((n, OUTPUTS), np.float32))
        finally:
            for q in (xd, pd, zd):
                hip.free(q)

    def forward_host(self, features):
        x = np.ascontiguousarray(features, np.uint16)
        n = x.shape[0]
        hip = self.hip
        xd, zd = hip.upload(x), hip.alloc(n * OUTPUTS * 4)
        try:
            self.forward(xd, n, zd)
            return hip.download(zd, np.zeros((n, OUTPUTS), np.float32))
        finally:
            hip.free(xd)
            hip.free(zd)

    def labels_host(self, decisions):
        d = np.ascontiguousarray(decisions, DECISION_DTYPE)
        n = len(d)
        hip = self.hip
        dd, ld = hip.upload(d), hip.alloc(n * 4)
        try:
            self.labels(dd, n, ld)
            return hip.download(ld, np.zeros(n, LABEL_DTYPE))
        finally:
            hip.free(dd)
            hip.free(ld)


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposal_net.npz")


def trained(path=WEIGHTS):
    """(parameters, [h1_max, h2_max], training record) of the trained network shipped with the
    package (written by ragdoll_amd.train_proposals)."""
    with np.load(path, allow_pickle=False) as d:
        return d["params"].astype(np.float32), [float(v) for v in d["act_max"]], json.loads(str(d["meta"]))


def load_trained(device=0, max_batch=4096, path=WEIGHTS):
    """A ProposalNet holding the shipped parameters, quantized to FP8 as it was evaluated."""
    params, act_max, _ = trained(path)
    net = ProposalNet(max_batch=max_batch, device=device, params=params)
    net.quantize_ranges(*act_max)
    return net


class Proposer:
    """Learned proposals for batches of device pictures: their features and FP8 inference over all
    their macroblocks in one pass (rgh_pnet_propose_batch), into a buffer kept for reuse."""

    def __init__(self, net, max_rows):
        self.net = net
        self.max_rows = max_rows
        self.props = net.hip.alloc(max_rows * 16)

    def close(self):
        if self.props:
            self.net.hip.free(self.props)
            self.props = None

    def run(self, pictures):
        """pictures: (cur_y_dev, cur_stride, prev_y_dev, prev_stride, width, height, qp) tuples.
        Returns the device address of each picture's proposal records (valid until the next run)."""
        rows, out = 0, []
        for (c, cs, p, ps, w, h, qp) in pictures:
            out.append(self.props + rows * 16)
            rows += ((w + 15) // 16) * ((h + 15) // 16)
        if rows > self.max_rows:
            raise ValueError("%d macroblocks exceed the proposer's %d rows" % (rows, self.max_rows))
        self.net.propose_batch(pictures, self.props)
        return out
