"""Benchmark: 1080p P-frame analysis throughput (macroblocks/s) on one MI355X.

Workload (the plain GPU form at scale): 1920x1080 I420, 120 frames of seeded synthetic content
(20 objects, |v| <= 32 px/frame, 10% scene cuts, noise sigma 3), High profile decisions
(adaptive 8x8 transform), QP 24/30/36, keyint 30: 3 QPs x 4 GOPs = 12 independent GOP streams,
each an IDR then 29 P frames. A step is one P frame of each of the 12 streams, analysed by one
batched set of launches (rgh_analyse_batch): every macroblock's decision, quantized levels and
reconstruction, then the deblock and border extension that make the next reference. Sources
are resident in HBM before the timed region.

The IDR frames are not coded by this path: each GOP's first reference is its IDR source frame
installed with rgh_set_reference (outside the timed region).

Prints one JSON line (see the keys below). --gpus above 1 is refused: this build measures one
GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MB_BYTES = 1184  # algorithmic HBM bytes per P macroblock: 384 source + 384 reference + 384 recon + 32 decision
HBM_PEAK_GBS = 8000.0
GOP = 30
QPS = (24, 30, 36)
W, H = 1920, 1080


class Hip:
    """The ROCm HIP runtime that libragdoll_hip.so itself links (one runtime per process: a
    second copy, such as the one bundled with PyTorch, must not be initialised beside it)."""

    def __init__(self):
        self.L = None
        for name in ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.L is None:
            raise RuntimeError("libamdhip64.so.7 not found")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed with HIP error %d" % (what, rc))

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), arr.nbytes), "hipMalloc")
        self.check(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy")  # host to device
        return p.value

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))

    def synchronize(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json, copy-calibrated), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            k = json.load(f)["kernels"][kernel]
        return round(k["fetch_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(mib, repeats=4):
    """A plain device-to-device copy of known size: reading FETCH_SIZE / WRITE_SIZE for it gives
    the factor between those counters and the bytes a kernel moves."""
    hip = Hip()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.check(hip.L.hipMalloc(ctypes.byref(a), n), "hipMalloc")
    hip.check(hip.L.hipMalloc(ctypes.byref(b), n), "hipMalloc")
    hip.check(hip.L.hipMemset(a, 1, ctypes.c_size_t(n)), "hipMemset")
    for _ in range(repeats):
        hip.check(hip.L.hipMemcpy(b, a, ctypes.c_size_t(n), 3), "hipMemcpy")  # device to device
    hip.synchronize()
    hip.free(a.value)
    hip.free(b.value)
    print(json.dumps({"copy_bytes": n, "repeats": repeats}), flush=True)


def render_frames(n_frames):
    from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
    kw = dict(CONFIGS["1080p"])
    kw.pop("width")
    kw.pop("height")
    g = SceneGenerator(W, H, seed=config_seed(2), **kw)
    return [g.next_frame() for _ in range(n_frames)]


def cpu_baseline(frames, seconds=12.0):
    """The CPU oracle (the reference's Balanced encoder restated in C, built -O3) on a bounded
    sample of the same workload: P frames of the 12 streams on all host cores, one thread per
    stream (like the reference's parallel_gop_map), until `seconds` of wall time pass."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleEncoder
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    streams = [(qp, g) for qp in QPS for g in range(len(frames) // GOP)]
    threads = min(cores, len(streams))
    done = [0] * threads
    deadline = [0.0]

    def work(k):
        for s in range(k, len(streams), threads):
            qp, g = streams[s]
            enc = OracleEncoder(W, H, qp, profile="high", keyint=1 << 20)
            enc.encode(*frames[g * GOP])  # IDR, untimed share is small and counted as setup
            for i in range(1, GOP):
                if time.perf_counter() > deadline[0]:
                    return
                enc.encode(*frames[g * GOP + i])
                done[k] += 1

    t0 = time.perf_counter()
    deadline[0] = t0 + seconds
    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    mbs = sum(done) * (W // 16) * ((H + 15) // 16)
    return {"value": mbs / el, "unit": "macroblocks/s", "cores": threads, "kind": "port",
            "sample": "%d P frames (1080p, High decisions, QP 24/30/36, one thread per GOP stream) in %.1f s "
                      "by the C oracle (Balanced encoder restated, IDR + P + deblock; no entropy coding)" % (
                          sum(done), el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (independent GOP streams would "
              "shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.steps + args.warmup > GOP - 1:
        print("steps + warmup must be <= %d (P frames per GOP)" % (GOP - 1), file=sys.stderr)
        sys.exit(2)

    from ragdoll_amd import StreamAnalyzer, analyse_batch

    n_frames = 120
    t0 = time.time()
    frames = render_frames(n_frames)
    print("rendered %d frames in %.1f s" % (n_frames, time.time() - t0), file=sys.stderr, flush=True)

    hip = Hip()
    cw, ch = W // 2, H // 2
    dy = hip.upload(np.stack([f[0] for f in frames]))
    du = hip.upload(np.stack([f[1] for f in frames]))
    dv = hip.upload(np.stack([f[2] for f in frames]))
    hip.synchronize()

    n_gops = n_frames // GOP
    streams = [(qp, g) for qp in QPS for g in range(n_gops)]
    ans = []
    for qp, g in streams:
        a = StreamAnalyzer(W, H, qp, profile="high", keep_undeblocked=False)
        y, u, v = frames[g * GOP]
        W16, H16 = a.mb_w * 16, a.mb_h * 16
        pad = lambda p, ww, hh: np.pad(p, ((0, hh - p.shape[0]), (0, ww - p.shape[1])), mode="edge")
        a.set_reference(pad(y, W16, H16), pad(u, W16 // 2, H16 // 2), pad(v, W16 // 2, H16 // 2))
        ans.append(a)
    ans[0].enable_timing(True)
    mbs_per_frame = ans[0].mb_w * ans[0].mb_h

    def frame_ptrs(i):
        return (dy + i * W * H, du + i * cw * ch, dv + i * cw * ch, W, cw, cw)

    def step(k):
        analyse_batch(ans, [frame_ptrs(g * GOP + 1 + k) for (_, g) in streams])

    for k in range(args.warmup):
        step(k)
    hip.synchronize()
    analyse_ms = deblock_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        step(k)
        a_ms, d_ms, n_l = ans[0].timing()
        analyse_ms += a_ms
        deblock_ms += d_ms
        launches += n_l
    hip.synchronize()
    el = time.perf_counter() - t0

    total_mbs = args.steps * len(streams) * mbs_per_frame
    value = total_mbs / el
    avg_launch_s = analyse_ms / 1e3 / launches
    mbs_per_launch = total_mbs / launches
    achieved = MB_BYTES * mbs_per_launch / avg_launch_s / 1e9
    out = {
        "metric": "1080p ME macroblocks/sec, 1 GPU: plain vs learned-proposal at <=1% BD-rate",
        "value": round(value, 1),
        "unit": "macroblocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "plain GPU form: 1920x1080 P-frame analysis, High decisions, QP 24/30/36, "
                               "keyint 30, 12 GOP streams batched per step, bit-exact with the CPU oracle",
                   "form": "plain", "streams": len(streams), "mbs_per_frame": mbs_per_frame,
                   "me_range": 16, "analyse_ms": round(analyse_ms, 3), "deblock_ms": round(deblock_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("rgh_k_analyse_step"),
                     "kernel": "rgh_k_analyse_step", "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "mbs_per_launch": round(mbs_per_launch, 2), "bytes_per_mb": MB_BYTES},
    }
    for a in ans:
        a.close()
    for p in (dy, du, dv):
        hip.free(p)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(frames, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
