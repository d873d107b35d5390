"""ctypes binding of include/ragdoll_hip.h (the library is loaded on first use)."""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libragdoll_hip.so")
LEVELS_PER_MB = 408
PROFILES = {"baseline": 0, "main": 1, "high": 2}
SPEEDS = {"balanced": 0, "compact": 1, "fast": 2, "partitions": 3}
FLAG_NO_UNDEBLOCKED_COPY = 1
FLAG_COMPACT = 2
FLAG_FAST = 4
FLAG_PARTITIONS = 8

DECISION_DTYPE = np.dtype([
    ("mb_type", "u1"), ("i16_mode", "u1"), ("chroma_mode", "u1"), ("cbp_luma", "u1"),
    ("cbp_chroma", "u1"), ("transform_8x8", "u1"), ("qp", "u1"), ("flags", "u1"),
    ("mv", "<i2", (2,)), ("mvp", "<i2", (2,)), ("cost", "<u4"), ("n_sad", "<u2"),
    ("n_subpel", "<u2"), ("n_satd", "<u2"), ("n_hpel", "<u2"), ("nnz", "<u2"), ("reserved", "<u2"),
])
assert DECISION_DTYPE.itemsize == 32

EXPORTS = ["rgh_create", "rgh_destroy", "rgh_set_reference", "rgh_analyse_p_frame", "rgh_analyse_idr",
           "rgh_analyse_batch", "rgh_analyse_frames", "rgh_analyse_frames_noalloc", "rgh_write_p_slice_cavlc",
           "rgh_write_idr_slice_cavlc", "rgh_write_p_slice_cabac", "rgh_write_idr_slice_cabac", "rgh_write_headers",
           "rgh_get_outputs", "rgh_get_mb_dims", "rgh_synchronize", "rgh_get_stream", "rgh_enable_timing",
           "rgh_get_timing", "rgh_status_str", "rgh_last_error"]


class RgConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("qp", ctypes.c_int32), ("me_range", ctypes.c_int32), ("qpel_iters", ctypes.c_int32),
                ("device", ctypes.c_int32), ("flags", ctypes.c_int32), ("ring", ctypes.c_int32)]


class RgSliceConfig(ctypes.Structure):
    _fields_ = [("mb_w", ctypes.c_int32), ("mb_h", ctypes.c_int32), ("qp", ctypes.c_int32),
                ("frame_num", ctypes.c_int32), ("poc_lsb", ctypes.c_int32), ("speed", ctypes.c_int32)]


class RgStreamConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("qp", ctypes.c_int32), ("fps_num", ctypes.c_int32), ("fps_den", ctypes.c_int32),
                ("speed", ctypes.c_int32)]


class RgI420(ctypes.Structure):
    _fields_ = [("y", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("stride_y", ctypes.c_int32), ("stride_u", ctypes.c_int32), ("stride_v", ctypes.c_int32),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class RgI420Out(ctypes.Structure):
    _fields_ = [("y", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("stride_y", ctypes.c_int32), ("stride_u", ctypes.c_int32), ("stride_v", ctypes.c_int32)]


class RghError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        detail = _lib.rgh_last_error().decode()
        super().__init__("%s: %s (status %d)%s" % (what, _lib.rgh_status_str(status).decode(), status,
                                                    " — " + detail if detail else ""))


_lib = None


def load_library(path=LIB_PATH):
    """Loads libragdoll_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libragdoll_hip.so is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    P, I32 = ctypes.c_void_p, ctypes.c_int32
    L.rgh_create.restype = I32
    L.rgh_create.argtypes = [ctypes.POINTER(RgConfig), ctypes.POINTER(ctypes.c_void_p)]
    L.rgh_destroy.restype = I32
    L.rgh_destroy.argtypes = [P]
    L.rgh_set_reference.restype = I32
    L.rgh_set_reference.argtypes = [P, ctypes.POINTER(RgI420)]
    L.rgh_analyse_p_frame.restype = I32
    L.rgh_analyse_p_frame.argtypes = [P, ctypes.POINTER(RgI420), P, P, ctypes.POINTER(RgI420Out)]
    L.rgh_analyse_idr.restype = I32
    L.rgh_analyse_idr.argtypes = [P, ctypes.POINTER(RgI420), P, P, ctypes.POINTER(RgI420Out)]
    L.rgh_analyse_batch.restype = I32
    L.rgh_analyse_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), I32, ctypes.POINTER(RgI420)]
    L.rgh_analyse_frames_learned.restype = I32
    L.rgh_analyse_frames_learned.argtypes = [ctypes.POINTER(ctypes.c_void_p), I32, I32, ctypes.POINTER(RgI420),
                                             ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), I32, I32,
                                             I32]
    for name in ("rgh_analyse_frames", "rgh_analyse_frames_noalloc"):
        getattr(L, name).restype = I32
        getattr(L, name).argtypes = [ctypes.POINTER(ctypes.c_void_p), I32, I32, ctypes.POINTER(RgI420),
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), I32, I32, I32]
    L.rgh_write_p_slice_cavlc.restype = I32
    L.rgh_write_p_slice_cavlc.argtypes = [ctypes.POINTER(RgSliceConfig), P, P, P, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_int64)]
    L.rgh_write_idr_slice_cavlc.restype = I32
    L.rgh_write_idr_slice_cavlc.argtypes = [ctypes.POINTER(RgSliceConfig), I32, P, P, P, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]
    L.rgh_write_p_slice_cabac.restype = I32
    L.rgh_write_p_slice_cabac.argtypes = [ctypes.POINTER(RgSliceConfig), I32, P, P, P, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_int64)]
    L.rgh_write_idr_slice_cabac.restype = I32
    L.rgh_write_idr_slice_cabac.argtypes = [ctypes.POINTER(RgSliceConfig), I32, I32, P, P, P, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]
    L.rgh_write_headers.restype = I32
    L.rgh_write_headers.argtypes = [ctypes.POINTER(RgStreamConfig), P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.rgh_get_outputs.restype = I32
    L.rgh_get_outputs.argtypes = [P, P, P, ctypes.POINTER(RgI420Out), I32]
    L.rgh_get_mb_dims.restype = I32
    L.rgh_get_mb_dims.argtypes = [P, ctypes.POINTER(I32), ctypes.POINTER(I32)]
    L.rgh_synchronize.restype = I32
    L.rgh_synchronize.argtypes = [P]
    L.rgh_get_stream.restype = P
    L.rgh_get_stream.argtypes = [P]
    L.rgh_enable_timing.restype = I32
    L.rgh_enable_timing.argtypes = [P, I32]
    L.rgh_get_timing.restype = I32
    L.rgh_get_timing.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(I32)]
    L.rgh_status_str.restype = ctypes.c_char_p
    L.rgh_status_str.argtypes = [I32]
    L.rgh_last_error.restype = ctypes.c_char_p
    L.rgh_last_error.argtypes = []
    _lib = L
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(status, what):
    if status != 0:
        raise RghError(status, what)


def host_i420(y, u, v, width, height):
    return RgI420(_ptr(y), _ptr(u), _ptr(v), y.strides[0], u.strides[0], v.strides[0], width, height)


class StreamAnalyzer:
    """One stream's P-frame analysis context (mirror of an H264Encoder's P path)."""

    def __init__(self, width, height, qp, profile="baseline", me_range=0, qpel_iters=0, device=0,
                 keep_undeblocked=True, ring=0, speed="balanced"):
        self.L = load_library()
        flags = ((0 if keep_undeblocked else FLAG_NO_UNDEBLOCKED_COPY) | (FLAG_COMPACT if speed == "compact" else 0) |
                 (FLAG_FAST if speed == "fast" else 0) | (FLAG_PARTITIONS if speed == "partitions" else 0))
        cfg = RgConfig(width, height, PROFILES[profile], qp, me_range, qpel_iters, device, flags, ring)
        h = ctypes.c_void_p()
        _check(self.L.rgh_create(ctypes.byref(cfg), ctypes.byref(h)), "rgh_create")
        self.h = h
        self.width, self.height, self.qp = width, height, qp
        mw, mh = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L.rgh_get_mb_dims(self.h, ctypes.byref(mw), ctypes.byref(mh)), "rgh_get_mb_dims")
        self.mb_w, self.mb_h = mw.value, mh.value

    def close(self):
        if getattr(self, "h", None):
            self.L.rgh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_reference(self, y, u, v):
        y, u, v = (np.ascontiguousarray(p, np.uint8) for p in (y, u, v))
        W, H = self.mb_w * 16, self.mb_h * 16
        if y.shape != (H, W):
            raise ValueError("reference must be at coded size %dx%d" % (W, H))
        _check(self.L.rgh_set_reference(self.h, ctypes.byref(host_i420(y, u, v, W, H))), "rgh_set_reference")

    def analyse_idr(self, y, u, v, want_levels=True, want_recon=True):
        """Analyses an IDR picture (it becomes the reference); returns (decisions, levels,
        (y, u, v) deblocked recon)."""
        return self.analyse(y, u, v, want_levels, want_recon, _entry="rgh_analyse_idr")

    def analyse(self, y, u, v, want_levels=True, want_recon=True, _entry="rgh_analyse_p_frame"):
        """Analyses one P frame; returns (decisions, levels, (y, u, v) deblocked recon)."""
        y, u, v = (np.ascontiguousarray(p, np.uint8) for p in (y, u, v))
        n = self.mb_w * self.mb_h
        dec = np.zeros(n, DECISION_DTYPE)
        lv = np.zeros((n, LEVELS_PER_MB), np.int16) if want_levels else None
        rec = None
        out = None
        if want_recon:
            W, H = self.mb_w * 16, self.mb_h * 16
            rec = (np.zeros((H, W), np.uint8), np.zeros((H // 2, W // 2), np.uint8),
                   np.zeros((H // 2, W // 2), np.uint8))
            out = RgI420Out(_ptr(rec[0]), _ptr(rec[1]), _ptr(rec[2]), W, W // 2, W // 2)
        _check(getattr(self.L, _entry)(self.h, ctypes.byref(host_i420(y, u, v, self.width, self.height)),
                                       _ptr(dec), _ptr(lv), ctypes.byref(out) if out else None), _entry)
        return dec, lv, rec

    def outputs(self, which_recon=1, want_levels=True):
        n = self.mb_w * self.mb_h
        W, H = self.mb_w * 16, self.mb_h * 16
        dec = np.zeros(n, DECISION_DTYPE)
        lv = np.zeros((n, LEVELS_PER_MB), np.int16) if want_levels else None
        rec = (np.zeros((H, W), np.uint8), np.zeros((H // 2, W // 2), np.uint8), np.zeros((H // 2, W // 2), np.uint8))
        out = RgI420Out(_ptr(rec[0]), _ptr(rec[1]), _ptr(rec[2]), W, W // 2, W // 2)
        _check(self.L.rgh_get_outputs(self.h, _ptr(dec), _ptr(lv), ctypes.byref(out), which_recon), "rgh_get_outputs")
        return dec, lv, rec

    def enable_timing(self, on=True):
        _check(self.L.rgh_enable_timing(self.h, 1 if on else 0), "rgh_enable_timing")

    def timing(self):
        """(schedule ms, 0.0, frame-step launches) of the last timed call."""
        a, d, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_int32()
        _check(self.L.rgh_get_timing(self.h, ctypes.byref(a), ctypes.byref(d), ctypes.byref(n)), "rgh_get_timing")
        return a.value, d.value, n.value

    def stream(self):
        return self.L.rgh_get_stream(self.h)

    def synchronize(self):
        _check(self.L.rgh_synchronize(self.h), "rgh_synchronize")


def analyse_batch(analyzers, device_frames):
    """One P frame for each analyzer from device-resident I420 planes.

    device_frames: sequence of (y_ptr, u_ptr, v_ptr, stride_y, stride_u, stride_v) of device
    memory holding each stream's source frame."""
    L = load_library()
    n = len(analyzers)
    arr = (ctypes.c_void_p * n)(*[a.h.value for a in analyzers])
    srcs = (RgI420 * n)()
    for i, (a, f) in enumerate(zip(analyzers, device_frames)):
        srcs[i] = RgI420(f[0], f[1], f[2], f[3], f[4], f[5], a.width, a.height)
    _check(L.rgh_analyse_batch(arr, n, srcs), "rgh_analyse_batch")


FRAMES_FIRST_IDR = 1
FRAMES_PERSISTENT = 2
FRAMES_GRAPH = 4


class RgLearnedConfig(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("skip_gate", ctypes.c_int32), ("intra_gate", ctypes.c_int32),
                ("fallback_pct", ctypes.c_int32), ("intra_first", ctypes.c_int32)]


def analyse_frames(analyzers, device_frames, decisions, levels, lag=0, groups=0, first_idr=False, proposals=None,
                   learned=None, noalloc=False, persistent=False, graph=False, keyint=0):
    """n_frames consecutive P frames of every analyzer in one pipelined schedule.

    device_frames[i][k]: (y_ptr, u_ptr, v_ptr, stride_y, stride_u, stride_v) of frame k of
    analyzer i in device memory; decisions[i][k] / levels[i][k]: device addresses receiving its
    mb_w*mb_h decision records / levels. With proposals[i][k] (device addresses of rgh_proposal
    records; None for an IDR frame 0) the search runs from learned proposals
    (rgh_analyse_frames_learned) with learned = dict(iters, skip_gate, intra_gate, fallback_pct).
    noalloc: the plain form through rgh_analyse_frames_noalloc (no allocation by the library).
    persistent: one persistent launch (RGH_FRAMES_PERSISTENT) instead of the per-step launches.
    graph: the planned per-step launches as one captured HIP graph (RGH_FRAMES_GRAPH), replayed by
    later calls with the same launches. keyint (with first_idr): frames 0, keyint, 2 keyint, ... are
    IDR pictures (RGH_FRAMES_KEYINT), several GOPs of each analyzer back to back in one call."""
    L = load_library()
    n = len(analyzers)
    nf = len(device_frames[0])
    arr = (ctypes.c_void_p * n)(*[a.h.value for a in analyzers])
    srcs = (RgI420 * (n * nf))()
    dp = (ctypes.c_void_p * (n * nf))()
    lp = (ctypes.c_void_p * (n * nf))()
    for i, a in enumerate(analyzers):
        if len(device_frames[i]) != nf or len(decisions[i]) != nf or len(levels[i]) != nf:
            raise ValueError("every analyzer needs the same number of frames and outputs")
        for k, f in enumerate(device_frames[i]):
            srcs[i * nf + k] = RgI420(f[0], f[1], f[2], f[3], f[4], f[5], a.width, a.height)
            dp[i * nf + k] = decisions[i][k]
            lp[i * nf + k] = levels[i][k]
    flags = ((FRAMES_FIRST_IDR if first_idr else 0) | (FRAMES_PERSISTENT if persistent else 0) |
             (FRAMES_GRAPH if graph else 0) | (int(keyint) << 16))
    if proposals is None:
        entry = "rgh_analyse_frames_noalloc" if noalloc else "rgh_analyse_frames"
        _check(getattr(L, entry)(arr, n, nf, srcs, dp, lp, lag, groups, flags), entry)
        return
    if noalloc:
        raise ValueError("the no-allocation entry point serves the plain form only")
    pp = (ctypes.c_void_p * (n * nf))()
    for i in range(n):
        for k in range(nf):
            pp[i * nf + k] = proposals[i][k]
    cfg = RgLearnedConfig(**(learned or {}))
    _check(L.rgh_analyse_frames_learned(arr, n, nf, srcs, pp, ctypes.byref(cfg), dp, lp, lag, groups, flags),
           "rgh_analyse_frames_learned")


def write_p_slice_cavlc(decisions, levels, mb_w, mb_h, qp, frame_num, poc_lsb, speed="balanced"):
    """The Annex B NAL unit (bytes) of one analysed P picture, Constrained Baseline CAVLC."""
    L = load_library()
    dec = np.ascontiguousarray(decisions, DECISION_DTYPE)
    lv = np.ascontiguousarray(levels, np.int16)
    if dec.shape != (mb_w * mb_h,) or lv.shape != (mb_w * mb_h, LEVELS_PER_MB):
        raise ValueError("decisions/levels do not match the macroblock grid")
    cfg = RgSliceConfig(mb_w, mb_h, qp, frame_num, poc_lsb, SPEEDS[speed])
    cap = mb_w * mb_h * 1024 + 64
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_int64()
    _check(L.rgh_write_p_slice_cavlc(ctypes.byref(cfg), _ptr(dec), _ptr(lv), _ptr(out), cap, ctypes.byref(n)),
           "rgh_write_p_slice_cavlc")
    return bytes(out[:n.value])


def write_idr_slice_cavlc(decisions, levels, mb_w, mb_h, qp, idr_pic_id, speed="balanced"):
    """The Annex B NAL unit (bytes) of one analysed IDR picture, Constrained Baseline CAVLC."""
    L = load_library()
    dec = np.ascontiguousarray(decisions, DECISION_DTYPE)
    lv = np.ascontiguousarray(levels, np.int16)
    if dec.shape != (mb_w * mb_h,) or lv.shape != (mb_w * mb_h, LEVELS_PER_MB):
        raise ValueError("decisions/levels do not match the macroblock grid")
    cfg = RgSliceConfig(mb_w, mb_h, qp, 0, 0, SPEEDS[speed])
    cap = mb_w * mb_h * 1024 + 64
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_int64()
    _check(L.rgh_write_idr_slice_cavlc(ctypes.byref(cfg), idr_pic_id, _ptr(dec), _ptr(lv), _ptr(out), cap,
                                       ctypes.byref(n)), "rgh_write_idr_slice_cavlc")
    return bytes(out[:n.value])


def _slice_args(decisions, levels, mb_w, mb_h):
    dec = np.ascontiguousarray(decisions, DECISION_DTYPE)
    lv = np.ascontiguousarray(levels, np.int16)
    if dec.shape != (mb_w * mb_h,) or lv.shape != (mb_w * mb_h, LEVELS_PER_MB):
        raise ValueError("decisions/levels do not match the macroblock grid")
    cap = mb_w * mb_h * 1024 + 64
    return dec, lv, cap, np.zeros(cap, np.uint8), ctypes.c_int64()


def write_p_slice_cabac(decisions, levels, mb_w, mb_h, qp, frame_num, poc_lsb, profile="high", speed="balanced"):
    """The Annex B NAL unit (bytes) of one analysed P picture, Main or High CABAC."""
    L = load_library()
    dec, lv, cap, out, n = _slice_args(decisions, levels, mb_w, mb_h)
    cfg = RgSliceConfig(mb_w, mb_h, qp, frame_num, poc_lsb, SPEEDS[speed])
    _check(L.rgh_write_p_slice_cabac(ctypes.byref(cfg), PROFILES[profile], _ptr(dec), _ptr(lv), _ptr(out), cap,
                                     ctypes.byref(n)), "rgh_write_p_slice_cabac")
    return bytes(out[:n.value])


def write_idr_slice_cabac(decisions, levels, mb_w, mb_h, qp, idr_pic_id, profile="high", speed="balanced"):
    """The Annex B NAL unit (bytes) of one analysed IDR picture, Main or High CABAC."""
    L = load_library()
    dec, lv, cap, out, n = _slice_args(decisions, levels, mb_w, mb_h)
    cfg = RgSliceConfig(mb_w, mb_h, qp, 0, 0, SPEEDS[speed])
    _check(L.rgh_write_idr_slice_cabac(ctypes.byref(cfg), PROFILES[profile], idr_pic_id, _ptr(dec), _ptr(lv),
                                       _ptr(out), cap, ctypes.byref(n)), "rgh_write_idr_slice_cabac")
    return bytes(out[:n.value])


def write_headers(width, height, qp, profile="baseline", fps=(30, 1), speed="balanced"):
    """SPS and PPS NAL units (bytes, Annex B) of a stream."""
    L = load_library()
    cfg = RgStreamConfig(width, height, PROFILES[profile], qp, fps[0], fps[1], SPEEDS[speed])
    out = np.zeros(1024, np.uint8)
    n = ctypes.c_int64()
    _check(L.rgh_write_headers(ctypes.byref(cfg), _ptr(out), 1024, ctypes.byref(n)), "rgh_write_headers")
    return bytes(out[:n.value])
