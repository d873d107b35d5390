"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

LEVELS_PER_MB = 408

DECISION_DTYPE = np.dtype([
    ("mb_type", "u1"), ("i16_mode", "u1"), ("chroma_mode", "u1"), ("cbp_luma", "u1"),
    ("cbp_chroma", "u1"), ("transform_8x8", "u1"), ("qp", "u1"), ("flags", "u1"),
    ("mv", "<i2", (2,)), ("mvp", "<i2", (2,)), ("cost", "<u4"), ("n_sad", "<u2"),
    ("n_subpel", "<u2"), ("n_satd", "<u2"), ("n_hpel", "<u2"), ("nnz", "<u2"), ("reserved", "<u2"),
])
assert DECISION_DTYPE.itemsize == 32

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(LIB_PATH)
        P, I, L64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_long
        L.ora_create.restype = P
        L.ora_create.argtypes = [I, I, I, I, I, I, I]
        L.ora_destroy.argtypes = [P]
        L.ora_set_me_range.argtypes = [P, I]
        L.ora_compact_pred.restype = I
        L.ora_compact_pred.argtypes = [P, I, I, I, I, I, I, I, I, P]
        L.ora_set_compact.restype = I
        L.ora_set_compact.argtypes = [P, I]
        L.ora_set_fast.restype = I
        L.ora_set_fast.argtypes = [P, I]
        L.ora_set_partitions.restype = I
        L.ora_set_partitions.argtypes = [P, I]
        L.ora_encode_frame.restype = L64
        L.ora_encode_frame.argtypes = [P, P, P, P, I, I, I, P, L64]
        L.ora_headers.restype = L64
        L.ora_headers.argtypes = [P, P, L64]
        L.ora_mb_w.argtypes = [P]
        L.ora_mb_h.argtypes = [P]
        L.ora_last_was_idr.argtypes = [P]
        L.ora_get_decisions.argtypes = [P, P]
        L.ora_get_levels.argtypes = [P, P]
        L.ora_get_recon.argtypes = [P, I, P, P, P]
        for n in ("ora_sad16", "ora_satd16", "ora_satd4"):
            getattr(L, n).restype = ctypes.c_uint32
            getattr(L, n).argtypes = [P, I, P, I]
        L.ora_fdct4.argtypes = [P, P]
        L.ora_idct4.argtypes = [P, P]
        L.ora_fdct8.argtypes = [P, P]
        L.ora_luma_mc.argtypes = [P, I, P, I, I, I, I, I]
        L.ora_chroma_mc8.argtypes = [P, P, I, I, I]
        L.ora_window_pred.argtypes = [P, P, I, I, I]
        L.ora_decimate_score.restype = ctypes.c_uint32
        L.ora_decimate_score.argtypes = [P, I]
        L.ora_dequant8.argtypes = [P, I, P]
        L.ora_idct8.argtypes = [P, P]
        L.ora_quant8_mf.argtypes = [I, I]
        L.ora_dequant8_norm.argtypes = [I, I]
        L.ora_cavlc_block.restype = I
        L.ora_cavlc_block.argtypes = [P, I, I, P, I]
        L.ora_predict_intra4x4.argtypes = [I, P, P, I, I, I, P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


PROFILES = {"baseline": 0, "main": 1, "high": 2}


class OracleEncoder:
    """Mirror of `H264Encoder::new(w, h, fps, qp).with_profile(p)` with `keyint` (the oracle)."""

    def __init__(self, width, height, qp, fps=(30, 1), profile="baseline", keyint=30, me_range=16, speed="balanced"):
        self.L = lib()
        self.h = self.L.ora_create(width, height, fps[0], fps[1], qp, PROFILES[profile], keyint)
        if not self.h:
            raise ValueError("unsupported dimensions %dx%d" % (width, height))
        self.L.ora_set_me_range(self.h, me_range)
        if speed == "compact" and self.L.ora_set_compact(self.h, 1) != 0:
            raise ValueError("the Compact preset codes B pictures with CABAC")
        if speed == "fast" and self.L.ora_set_fast(self.h, 1) != 0:
            raise ValueError("the Fast preset does not combine with Compact")
        if speed == "partitions" and self.L.ora_set_partitions(self.h, 1) != 0:
            raise ValueError("partitions extend the Balanced preset only")
        self.width, self.height, self.qp = width, height, qp
        self.mb_w = self.L.ora_mb_w(self.h)
        self.mb_h = self.L.ora_mb_h(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_destroy(self.h)
            self.h = None

    def encode(self, y, u, v):
        y, u, v = (np.ascontiguousarray(p, np.uint8) for p in (y, u, v))
        cap = self.width * self.height * 4 + 4096
        out = np.empty(cap, np.uint8)
        n = self.L.ora_encode_frame(self.h, ptr(y), ptr(u), ptr(v), y.shape[1], u.shape[1],
                                    v.shape[1], ptr(out), cap)
        assert n >= 0
        return bytes(out[:n])

    def headers(self):
        """SPS and PPS NAL units (Annex B) as the oracle writes them before an IDR."""
        out = np.empty(1024, np.uint8)
        n = self.L.ora_headers(self.h, ptr(out), 1024)
        assert n > 0
        return bytes(out[:n])

    @property
    def last_was_idr(self):
        return bool(self.L.ora_last_was_idr(self.h))

    def decisions(self):
        d = np.zeros(self.mb_w * self.mb_h, DECISION_DTYPE)
        self.L.ora_get_decisions(self.h, ptr(d))
        return d

    def levels(self):
        lv = np.zeros((self.mb_w * self.mb_h, LEVELS_PER_MB), np.int16)
        self.L.ora_get_levels(self.h, ptr(lv))
        return lv

    def recon(self, deblocked=True):
        W, H = self.mb_w * 16, self.mb_h * 16
        y = np.zeros((H, W), np.uint8)
        u = np.zeros((H // 2, W // 2), np.uint8)
        v = np.zeros((H // 2, W // 2), np.uint8)
        self.L.ora_get_recon(self.h, 1 if deblocked else 0, ptr(y), ptr(u), ptr(v))
        return y, u, v


def read_nals(path):
    import struct
    data = open(path, "rb").read()
    out, off = [], 0
    while off < len(data):
        n = struct.unpack(">I", data[off:off + 4])[0]
        out.append(data[off + 4:off + 4 + n])
        off += 4 + n
    return out


def split_annexb(data):
    """NAL units (without start codes) of an Annex-B byte stream."""
    out, i, n = [], 0, len(data)
    starts = []
    while i + 3 <= n:
        if data[i] == 0 and data[i + 1] == 0 and data[i + 2] == 1:
            starts.append(i + 3)
            i += 3
        else:
            i += 1
    for k, s in enumerate(starts):
        e = starts[k + 1] - 3 if k + 1 < len(starts) else n
        # strip trailing zero of a 4-byte start code
        while e > s and k + 1 < len(starts) and data[e - 1] == 0:
            e -= 1
        out.append(data[s:e])
    return out
