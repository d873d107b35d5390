"""Benchmark: 1080p P-frame analysis throughput (macroblocks/s) on one MI355X.

Workload (the plain GPU form at scale): 1920x1080 I420, 120 frames of seeded synthetic content
(20 objects, |v| <= 32 px/frame, 10% scene cuts, noise sigma 3), High profile decisions
(adaptive 8x8 transform), QP 24/30/36, search range +-64 (64 diamond iterations, quarter-pel
refinement), keyint 30: 3 QPs x 4 GOPs = 12 independent GOP streams,
each an IDR then 29 P frames. A step is the whole job: all 30 frames of the 12 streams (12 IDR
and 348 P frames) analysed by one pipelined schedule (rgh_analyse_frames with each stream's first
frame its IDR): every macroblock's decision, quantized levels and reconstruction, and the
deblocking and border extension that make each next reference, with up to ~13 frames of a
stream in flight. Sources are resident in HBM before the timed region; decisions and levels
stay in HBM. The value counts the P ("ME") macroblocks, 2,839,680 per step, over the time of
the whole job, IDR pictures included.

The schedule's frame lag (wavefront steps between consecutive frames of a stream) and its number
of concurrent groups (schedules on separate HIP streams) are picked by a small autotuner during
warm-up unless --lag / --groups are given. The persistent schedule (one launch per job) is timed
beside it on the same job (--compare-persistent 0 skips it), and so is the per-step schedule with
each group's planned launches captured as one HIP graph and replayed (--compare-graph 0 skips
it; --graph 1 makes it the timed form).

Prints one JSON line (see the keys below). --gpus above 1 is refused: this build measures one
GPU.
"""
import argparse
import ctypes
import functools
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ragdoll_amd.hip import Hip  # noqa: E402

MB_BYTES = 1184  # algorithmic HBM bytes per P macroblock: 384 source + 384 reference + 384 recon + 32 decision
IDR_MB_BYTES = 800  # per IDR macroblock: no reference read
HBM_PEAK_GBS = 8000.0
# Integer VALU peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes per clock at 2.4 GHz = 39.3 T lane
# operations per second (one wave64 VALU instruction per CU per clock).
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
GOP = 30
QPS = (24, 30, 36)
ME_RANGE = 64  # the headline's search range: +-64 (64 diamond iterations)
W, H = 1920, 1080


def profile_record(kernel):
    """The committed rocprofv3 counter figures of `kernel` on this workload
    (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(mbs_per_round):
    """HBM-side bytes of the analysis kernels from the committed counter profile of this workload
    (raw L2-to-fabric read and write requests by size, per macroblock of the profiled job): per
    round of launches (one launch per group, the unit of avg_launch_us) as the profiled bytes per
    macroblock times this run's macroblocks per round, per macroblock, and their ratio to the
    algorithmic bytes; None without the profile."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        return {"per_round": round(t["bytes_per_macroblock"] * mbs_per_round),
                "per_macroblock": t["bytes_per_macroblock"], "ratio_to_algorithmic": t["traffic_ratio"]}
    except (OSError, KeyError, ValueError):
        return None


def valu_figure(mbs, device_s):
    """Integer VALU throughput of the job: the counted INT32 + INT64 VALU instructions per
    macroblock (SQ_INSTS_VALU_INT32 / _INT64 of the analysis kernels over the macroblocks of the
    profiled run, every deferred attempt included) x 64 lanes x the job's macroblocks, over the
    job's device time, against the integer VALU peak; None without the committed counters.
    mbs: {record: macroblocks per job}, the record "job" covering all analysis kernels."""
    ops = 0.0
    per = {}
    for kernel, n in mbs.items():
        k = profile_record(kernel)
        if not k or "valu_int_per_mb" not in k:
            return None
        ops += k["valu_int_per_mb"] * 64 * n
        per[kernel] = {"valu_insts_per_mb": k.get("valu_per_mb"), "valu_int_insts_per_mb": k["valu_int_per_mb"],
                       "valu_busy": k.get("valu_busy")}
    tops = ops / device_s / 1e12
    return {"achieved": round(tops, 3), "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s (int32 lane ops)",
            "frac": tops / VALU_PEAK_TOPS, "kernels": per}


def calibrate_copy(mib, repeats=4):
    """A plain device-to-device copy of known size, for checking what the HBM-side counters
    (the L2's read/write requests by size, FETCH_SIZE / WRITE_SIZE) read for bytes a kernel
    must move."""
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


def calibrate_gather(mib, repeats=4):
    """A gather of one 4-byte word from every 128-byte line of a buffer of known size (torch's
    index_select, in a process of its own): what the read counters give for it tells their unit
    for scattered reads like the analysis's, beside the streaming copy."""
    import torch
    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.int32, device="cuda")
    idx = torch.arange(0, n // 4, 32, device="cuda")
    for _ in range(repeats):
        out = torch.index_select(src, 0, idx)
    torch.cuda.synchronize()
    print(json.dumps({"gather_bytes_spanned": n, "lines": int(idx.numel()), "repeats": repeats,
                      "checksum": int(out.sum().item())}), flush=True)


@functools.lru_cache(maxsize=1)
def render_frames(n_frames):
    from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
    kw = dict(CONFIGS["1080p"])
    kw.pop("width")
    kw.pop("height")
    g = SceneGenerator(W, H, seed=config_seed(2), **kw)
    return [g.next_frame() for _ in range(n_frames)]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frames, seconds=12.0, single_seconds=5.0, max_threads=16):
    """The CPU oracle (the reference's Balanced encoder restated in C, built -O3 for x86-64-v3, the
    reference's AVX2 level) on a bounded sample of the same workload (1080p, High decisions, QP
    24/30/36, +-64 search), timed on this host's cores:
    * single_thread: one thread coding the first 8 frames of each GOP stream in turn for
      `single_seconds`;
    * all_cores: one thread per usable core (up to `max_threads`, the host share of one GPU), each
      coding its own GOP streams (the 12 streams dealt round-robin, repeated), like the
      reference's parallel_gop_map, for `seconds`.
    Each thread times its IDR and P frames; a figure is the sum of the threads' P-macroblock
    rates over that time (the GPU figure's definition)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleEncoder
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    streams = [(qp, g) for qp in QPS for g in range(len(frames) // GOP)]
    mbf = (W // 16) * ((H + 15) // 16)

    def run(threads, secs, frames_per_gop=GOP):
        done = [0] * threads
        busy = [0.0] * threads
        deadline = time.perf_counter() + secs

        def work(k):
            s = k
            while True:
                qp, g = streams[s % len(streams)]
                s += threads
                enc = OracleEncoder(W, H, qp, profile="high", keyint=1 << 20, me_range=ME_RANGE)
                for i in range(frames_per_gop):
                    t = time.perf_counter()
                    if t > deadline:
                        return
                    enc.encode(*frames[g * GOP + i])
                    busy[k] += time.perf_counter() - t
                    done[k] += i > 0

        t0 = time.perf_counter()
        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        el = time.perf_counter() - t0
        return sum(done[k] * mbf / busy[k] for k in range(threads) if busy[k] > 0), sum(done), el

    # One thread visits every stream (the first 8 frames of each GOP, QP 24/30/36 alike), so the
    # single-thread sample holds the same content mix as the all-cores one.
    single, n1, el1 = run(1, single_seconds, frames_per_gop=8)
    threads = max(1, min(usable, max_threads))
    allc, n, el = run(threads, seconds)
    return {"value": round(allc, 1), "unit": "macroblocks/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "usable_cores": usable,
            "single_thread": round(single, 1), "all_cores": round(allc, 1),
            "sample": "C oracle (Balanced encoder restated: IDR and P analysis, reconstruction, deblocking; "
                      "-O3 -march=x86-64-v3) on the headline workload (1080p, High decisions, QP 24/30/36, "
                      "+-64 search): single_thread = %d P frames (the first 8 frames of each GOP in turn) on 1 thread "
                      "in %.1f s; all_cores = "
                      "%d P frames of the 12 GOP streams on %d threads (one stream at a time each) in %.1f s"
                      % (n1, el1, n, threads, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lag", type=int, default=0, help="frame lag in wavefront steps (0: autotune)")
    ap.add_argument("--groups", type=int, default=0, help="concurrent schedules (0: autotune)")
    ap.add_argument("--workload", choices=["headline", "batched", "compact", "fast", "partitions"], default="headline",
                    help="headline: the 1080p metric (default); batched: the 24-stream uneven-size set; "
                         "compact: the headline streams with the Compact preset (Baseline); "
                         "fast: the headline streams with the Fast preset (High)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--end-to-end", type=int, default=1, help="also time the step with its host transfers (1: yes)")
    ap.add_argument("--compare-persistent", type=int, default=1,
                    help="also time the persistent schedule (one launch per job) on the same job (1: yes)")
    ap.add_argument("--continuous", type=int, default=1,
                    help="run the timed steps back to back in one call, every GOP stream starting its next GOP "
                         "behind its last frames (1), or one call per step (0)")
    ap.add_argument("--isolated", type=int, default=1,
                    help="also time the steps one call each, beside the back-to-back value (1: yes)")
    ap.add_argument("--graph", type=int, default=0,
                    help="issue the per-step schedule's planned launches as one replayed HIP graph (1: yes)")
    ap.add_argument("--compare-graph", type=int, default=1,
                    help="also time the graph form of the schedule on the same job (1: yes)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal form")
    ap.add_argument("--learned-iters", type=int, default=2)
    ap.add_argument("--skip-gate", type=int, default=26)
    ap.add_argument("--intra-gate", type=int, default=0)
    ap.add_argument("--fallback-pct", type=int, default=105)
    ap.add_argument("--intra-first", type=int, default=153)
    ap.add_argument("--proposal-groups", type=int, default=0,
                    help="workgroup cap of the FP8 inference launch beside the analysis (0: none)")
    ap.add_argument("--bd-frames", type=int, default=120, help="frames per QP of the BD-rate sample (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration)")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="MIB",
                    help="only run a one-word-per-line gather over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.calibrate_gather:
        calibrate_gather(args.calibrate_gather)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (independent GOP streams would "
              "shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        gops, label = batched_gops(), "batched"
    elif args.workload == "compact":
        gops, label = [dict(g, profile="baseline", speed="compact") for g in headline_gops()], "compact"
    elif args.workload == "fast":
        gops, label = [dict(g, speed="fast") for g in headline_gops()], "fast"
    elif args.workload == "partitions":
        gops, label = partition_gops(), "partitions"
        args.compare_persistent = 0  # no persistent kernel for partitions: the per-step schedule runs them
    else:
        gops, label = headline_gops(), "headline"
    out = run_workload(gops, args)
    if label == "partitions":
        # The same job with 16x16 partitions only (the reference's Balanced decisions), and what the
        # partitions gain in rate for the same quality.
        args.end_to_end = args.compare_graph = 0
        plain = run_workload([dict(g, speed="balanced") for g in gops], args)
        out["config"]["without_partitions"] = {"value": plain["value"], "ms_per_step": plain["ms_per_step"],
                                               "vs_partitions": round(plain["value"] / out["value"], 3)}
        if args.bd_frames:
            out["config"]["bd_rate"] = partition_bd_rate(gops[0]["frames"])
    if label == "headline" and not args.no_learned and args.bd_frames:
        out["learned"]["bd_rate"] = bd_rate_sample(args)
    if label == "headline" and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(render_frames(120), args.cpu_seconds)
    print(json.dumps(out), flush=True)


def learned_config(args):
    return dict(iters=args.learned_iters, skip_gate=args.skip_gate, intra_gate=args.intra_gate,
                fallback_pct=args.fallback_pct, intra_first=args.intra_first)


def learned_ceiling(gops, ans, srcs, decs, lvls, sched, args, total_mbs, hip, analyse_frames):
    """The learned form's speed with perfect proposals, proposal cost excluded: every P
    macroblock's proposal is the plain analysis's own result (its vector as all four cells, its
    mode as a certain mode), and the search takes it with no diamond iteration and no fallback
    (iters 0, fallback 100000%, skip probe / intra check only for macroblocks proposed as such).
    That is the most the learned form could gain from any proposals on these analysis kernels."""
    from ragdoll_amd import DECISION_DTYPE
    from ragdoll_amd.learned import PROPOSAL_DTYPE
    analyse_frames(ans, srcs, decs, lvls, lag=sched[0], groups=sched[1], first_idr=True)
    bufs, P = [], []
    for i, gp in enumerate(gops):
        n = ans[i].mb_w * ans[i].mb_h
        row = [None]
        for k in range(1, len(gp["frames"])):
            d = hip.download(decs[i][k], np.zeros(n, DECISION_DTYPE))
            p = np.zeros(n, PROPOSAL_DTYPE)
            cell = np.stack([np.round(d["mv"][:, 0] / 16.0), np.round(d["mv"][:, 1] / 16.0)], 1)
            cell = np.clip(cell, -127, 127).astype(np.int8)
            intra = (d["mb_type"] == 2) | (d["mb_type"] == 3)
            cell[intra] = 0
            p["cell"] = cell[:, None, :]
            p["mode_prob"][:, 0] = np.where(d["mb_type"] == 0, 255, 0)
            p["mode_prob"][:, 1] = np.where(d["mb_type"] == 1, 255, 0)
            p["mode_prob"][:, 2] = np.where(intra, 255, 0)
            buf = hip.upload(p.view(np.uint8))
            bufs.append(buf)
            row.append(buf)
        P.append(row)
    cfg = dict(iters=0, skip_gate=128, intra_gate=128, fallback_pct=100000, intra_first=128)

    def step(k=1):
        t = time.perf_counter()
        if k == 1:
            analyse_frames(ans, srcs, decs, lvls, lag=sched[0], groups=sched[1], first_idr=True, proposals=P,
                           learned=cfg)
        else:  # k steps back to back, as the plain form's timed steps run
            analyse_frames(ans, [x * k for x in srcs], [x * k for x in decs], [x * k for x in lvls], lag=sched[0],
                           groups=sched[1], first_idr=True, proposals=[r * k for r in P], learned=cfg,
                           keyint=len(gops[0]["frames"]))
        return time.perf_counter() - t

    if args.continuous:
        step(2)
        el = step(args.steps)
    else:
        for _ in range(max(1, args.warmup)):
            step()
        el = sum(step() for _ in range(args.steps))
    for b in bufs:
        hip.free(b)
    return {"value": round(total_mbs * args.steps / el, 1), "ms_per_step": round(el / args.steps * 1e3, 3),
            "config": cfg, "steps_back_to_back": bool(args.continuous),
            "note": "proposals = the plain analysis's own vectors and modes (not timed)"}


def bd_rate_sample(args):
    """BD-rate of the learned form against the plain one on the headline content: the first
    --bd-frames frames (GOPs of 30, each an IDR and P frames) coded at QP 22/27/32/37 by both,
    High profile, +-64 search, CABAC bits (rgh_write_*_slice_cabac) against mean luma PSNR,
    piecewise cubic (pchip) interpolation of log rate in PSNR, with the least-squares cubic beside
    it."""
    from ragdoll_amd.learned import Proposer, load_trained
    from ragdoll_amd.rd import bd_rate, code_gop
    frames = render_frames(120)[:args.bd_frames]
    gops = [frames[k:k + GOP] for k in range(0, len(frames), GOP)]
    net = load_trained()
    proposer = Proposer(net, ((W + 15) // 16) * ((H + 15) // 16))
    t0 = time.time()
    curves = {"plain": [], "learned": []}
    for qp in (22, 27, 32, 37):
        for form in curves:
            parts = [code_gop(g, W, H, qp, profile="high", me_range=ME_RANGE,
                              proposer=proposer if form == "learned" else None,
                              learned=learned_config(args) if form == "learned" else None) for g in gops]
            curves[form].append({"bytes": sum(p["bytes"] for p in parts),
                                 "psnr_y": float(np.mean([v for p in parts for v in p["frame_psnr_y"]])),
                                 "plain_search": sum(p["plain_search"] for p in parts),
                                 "p_macroblocks": sum(p["p_macroblocks"] for p in parts)})
        print("bd-rate sample: QP %d done at %.1f s" % (qp, time.time() - t0), file=sys.stderr, flush=True)
    proposer.close()
    net.close()
    rate = {k: [c["bytes"] for c in v] for k, v in curves.items()}
    q = {k: [c["psnr_y"] for c in v] for k, v in curves.items()}
    out = {"bd_rate_pct": round(bd_rate(rate["plain"], q["plain"], rate["learned"], q["learned"]), 3),
           "bd_rate_pct_cubic": round(bd_rate(rate["plain"], q["plain"], rate["learned"], q["learned"], "cubic"), 3),
           "method": "piecewise cubic (pchip) interpolation of log rate in luma PSNR",
           "sample": "first %d frames of the 1080p content (%d GOPs of an IDR and P frames), QP 22/27/32/37, "
                     "High profile, +-64 search, CABAC bits vs mean luma PSNR" % (len(frames), len(gops)),
           "qps": [22, 27, 32, 37], "plain_bytes": rate["plain"], "learned_bytes": rate["learned"],
           "plain_psnr_y": [round(v, 4) for v in q["plain"]], "learned_psnr_y": [round(v, 4) for v in q["learned"]],
           "learned_plain_search": [c["plain_search"] for c in curves["learned"]],
           "p_macroblocks_per_qp": curves["learned"][0]["p_macroblocks"], "seconds": round(time.time() - t0, 1)}
    print("bd-rate sample: %s" % json.dumps(out), file=sys.stderr, flush=True)
    return out


def headline_gops():
    """The 12 GOP streams of the 1080p workload: 120 frames at QP 24/30/36, keyint 30."""
    t0 = time.time()
    frames = render_frames(120)
    print("rendered 120 frames in %.1f s" % (time.time() - t0), file=sys.stderr, flush=True)
    return [dict(w=W, h=H, qp=qp, me=ME_RANGE, frames=frames[g * GOP:(g + 1) * GOP]) for qp in QPS for g in range(4)]


def partition_gops():
    """The 720p configuration with partitions: 1280x720, 60 frames of the 720p content (12 objects,
    |v| <= 16 px/frame, a 0.75 px/frame pan) as two GOPs of 30, High, QP 26, +-32 search,
    16x16/16x8/8x16/8x8 P partitions with the 8x8 transform."""
    from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
    kw = dict(CONFIGS["720p"])
    w, h = kw.pop("width"), kw.pop("height")
    t0 = time.time()
    g = SceneGenerator(w, h, seed=config_seed(1), **kw)
    frames = [g.next_frame() for _ in range(2 * GOP)]
    print("rendered 60 frames of 720p in %.1f s" % (time.time() - t0), file=sys.stderr, flush=True)
    return [dict(w=w, h=h, qp=26, me=32, speed="partitions", frames=frames[k * GOP:(k + 1) * GOP]) for k in range(2)]


def partition_bd_rate(frames):
    """BD-rate of the partition form against 16x16 partitions only on one GOP of the 720p content:
    QP 22/26/30/34, High, +-32, CABAC bits of the library's writer against mean luma PSNR (pchip)."""
    from ragdoll_amd.rd import bd_rate, code_gop
    t0 = time.time()
    qps = (22, 26, 30, 34)
    cur = {"balanced": [], "partitions": []}
    for qp in qps:
        for form in cur:
            cur[form].append(code_gop(frames, frames[0][0].shape[1], frames[0][0].shape[0], qp, profile="high",
                                      me_range=32, speed=form))
    rate = {k: [c["bytes"] for c in v] for k, v in cur.items()}
    q = {k: [c["psnr_y"] for c in v] for k, v in cur.items()}
    out = {"bd_rate_pct": round(bd_rate(rate["balanced"], q["balanced"], rate["partitions"], q["partitions"]), 3),
           "method": "piecewise cubic (pchip) interpolation of log rate in luma PSNR",
           "sample": "first GOP (an IDR and 29 P frames) of the 720p content, QP 22/26/30/34, High, +-32, CABAC "
                     "bits vs mean luma PSNR; negative: partitions need fewer bits for the same quality",
           "qps": list(qps), "bytes_16x16": rate["balanced"], "bytes_partitions": rate["partitions"],
           "psnr_y_16x16": [round(v, 4) for v in q["balanced"]],
           "psnr_y_partitions": [round(v, 4) for v in q["partitions"]],
           "mb_types_partitions": cur["partitions"][1]["mb_types"], "seconds": round(time.time() - t0, 1)}
    print("partition bd-rate: %s" % json.dumps(out), file=sys.stderr, flush=True)
    return out


def batched_gops():
    """The batched configuration: 24 streams of uneven size, 60 frames each (two GOPs)."""
    from ragdoll_amd.synth import SceneGenerator, batched_stream
    t0 = time.time()
    gops = []
    for sid in range(24):
        kw, seed, qp, me = batched_stream(sid)
        kw = dict(kw)
        w, h = kw.pop("width"), kw.pop("height")
        g = SceneGenerator(w, h, seed=seed, **kw)
        frames = [g.next_frame() for _ in range(2 * GOP)]
        gops += [dict(w=w, h=h, qp=qp, me=me, frames=frames[k * GOP:(k + 1) * GOP]) for k in range(2)]
    print("rendered the batched streams in %.1f s" % (time.time() - t0), file=sys.stderr, flush=True)
    return gops


def run_workload(gops, args):
    """Times every GOP (its IDR, then GOP - 1 P frames) analysed by one rgh_analyse_frames call
    per step; returns the JSON record."""
    from ragdoll_amd import LEVELS_PER_MB, StreamAnalyzer, analyse_frames

    hip = Hip()
    uploaded = {}

    def dev(p):
        if id(p) not in uploaded:
            uploaded[id(p)] = (hip.upload(p), p)
        return uploaded[id(p)][0]

    ans, srcs, decs, lvls = [], [], [], []
    for gp in gops:
        a = StreamAnalyzer(gp["w"], gp["h"], gp["qp"], profile=gp.get("profile", "high"), me_range=gp["me"],
                           keep_undeblocked=False, ring=len(gp["frames"]) + 1, speed=gp.get("speed", "balanced"))
        ans.append(a)
        srcs.append([(dev(y), dev(u), dev(v), y.shape[1], u.shape[1], v.shape[1]) for y, u, v in gp["frames"]])
        n = a.mb_w * a.mb_h
        decs.append([hip.alloc(n * 32) for _ in gp["frames"]])
        lvls.append([hip.alloc(n * LEVELS_PER_MB * 2) for _ in gp["frames"]])
    ans[0].enable_timing(True)
    hip.synchronize()
    total_mbs = sum(a.mb_w * a.mb_h * (len(gp["frames"]) - 1) for a, gp in zip(ans, gops))
    idr_mbs = sum(a.mb_w * a.mb_h for a in ans)
    step_bytes = total_mbs * MB_BYTES + idr_mbs * IDR_MB_BYTES

    cpu = [0.0]  # host CPU time of the process over the steps since the last reset

    gop_len = len(gops[0]["frames"])
    assert all(len(gp["frames"]) == gop_len for gp in gops)

    def step(sched, persistent=False, graph=None, k=1):
        """k steps: one call per step (k = 1), or k steps in one call, every stream running its GOP
        k times back to back (an IDR every gop_len frames, RGH_FRAMES_KEYINT): a stream's next GOP
        starts at the call's lag behind the last frame of the GOP before, as a stream of GOPs runs.
        The output buffers of frame f serve frame f + gop_len too (no two are in flight at once)."""
        lag, groups = sched
        graph = bool(args.graph) if graph is None else graph
        c = time.process_time()
        t = time.perf_counter()
        if k == 1:
            analyse_frames(ans, srcs, decs, lvls, lag=lag, groups=groups, first_idr=True, persistent=persistent,
                           graph=graph and not persistent)
        else:
            analyse_frames(ans, [x * k for x in srcs], [x * k for x in decs], [x * k for x in lvls], lag=lag,
                           groups=groups, first_idr=True, keyint=gop_len, graph=graph)
        el = time.perf_counter() - t
        cpu[0] += time.process_time() - c
        ms, _, launches = ans[0].timing()
        return el, ms, launches

    # Warm-up, with the per-step schedule's frame lag and concurrent groups autotuned unless given.
    lags = [args.lag] if args.lag else [16, 18, 20, 22, 24]
    groups = [args.groups] if args.groups else [2, 3]
    tune = {}
    tk = 2 if args.continuous else 1
    for lag in lags:
        for g in groups:
            tune[(lag, g)] = step((lag, g), k=tk)[0] / tk
    sched = min(tune, key=tune.get)
    # The persistent schedule (one launch per job) on the same job, for comparison: its lag tuned,
    # then timed over the same number of steps.
    persist = None
    if args.compare_persistent:
        pt = {lag: step((lag, 1), True)[0] for lag in (16, 20, 24, 32)}
        plag = min(pt, key=pt.get)
        el_p = 0.0
        launches_p = 0
        for _ in range(args.steps):
            e, _, nl = step((plag, 1), True)
            el_p += e
            launches_p += nl
        persist = {"value": round(total_mbs * args.steps / el_p, 1), "ms_per_step": round(el_p / args.steps * 1e3, 3),
                   "lag": plag, "launches_per_step": launches_p / args.steps,
                   "tuning_s": {"lag %d" % k: round(v, 4) for k, v in pt.items()}}
    # The planned launches of each schedule as one captured HIP graph, replayed: the first call
    # captures, the timed ones replay.
    graphed = None
    if args.compare_graph and not args.graph:
        step(sched, graph=True)
        cpu[0] = 0.0
        el_g = 0.0
        for _ in range(args.steps):
            el_g += step(sched, graph=True)[0]
        graphed = {"value": round(total_mbs * args.steps / el_g, 1), "ms_per_step": round(el_g / args.steps * 1e3, 3),
                   "host_cpu_ms_per_step": round(cpu[0] / args.steps * 1e3, 3)}
    # The steps one call each (every job ramps up from its IDR frames and down to its last frames).
    isolated = None
    if args.isolated and args.continuous:
        for _ in range(args.warmup):
            step(sched)
        el_i = 0.0
        for _ in range(args.steps):
            el_i += step(sched)[0]
        isolated = {"value": round(total_mbs * args.steps / el_i, 1),
                    "ms_per_step": round(el_i / args.steps * 1e3, 3), "note": "each step its own rgh_analyse_frames call"}
    # The timed steps: back to back in one call (continuous), or one call each.
    if args.warmup:
        if args.continuous:
            step(sched, k=max(2, args.warmup))
        else:
            for _ in range(args.warmup):
                step(sched)
    el = dev_ms = 0.0
    launches = 0
    cpu[0] = 0.0
    if args.continuous:
        el, dev_ms, launches = step(sched, k=args.steps)
    else:
        for _ in range(args.steps):
            e, ms, nl = step(sched)
            el += e
            dev_ms += ms
            launches += nl
    host_cpu_ms = cpu[0] / args.steps * 1e3

    value = total_mbs * args.steps / el
    e2e = end_to_end(hip, gops, ans, srcs, decs, lvls, uploaded, sched, total_mbs, args) if args.end_to_end else None
    avg_launch_s = dev_ms / 1e3 / launches
    mbs_per_launch = (total_mbs + idr_mbs) * args.steps / launches
    achieved = step_bytes * args.steps / launches / avg_launch_s / 1e9
    headline = args.workload == "headline"
    traffic = measured_traffic(mbs_per_launch) if headline else None
    metric = {"headline": "1080p ME macroblocks/sec, 1 GPU: plain vs learned-proposal at <=1% BD-rate",
              "batched": "batched 24-stream ME macroblocks/sec, 1 GPU, plain form",
              "compact": "1080p ME macroblocks/sec, 1 GPU, Compact preset",
              "fast": "1080p ME macroblocks/sec, 1 GPU, Fast preset",
              "partitions": "720p ME macroblocks/sec, 1 GPU, High with 16x16/16x8/8x16/8x8 P partitions"}[args.workload]
    workload = {
        "headline": "plain GPU form: 1920x1080 analysis, High decisions, QP 24/30/36, search range +-64 (64 "
                    "diamond iterations) with quarter-pel refinement, keyint 30, 12 GOP streams x (1 IDR + 29 P "
                    "frames) per step in one pipelined schedule, bit-exact with the CPU oracle; value counts P "
                    "macroblocks" + ("; the timed steps run back to back in one schedule, each stream starting "
                                     "its next GOP behind its last frames (IDR every 30 frames)"
                                     if args.continuous else ""),
        "batched": "plain GPU form: 24 streams (8 x 640x360, 8 x 1280x720, 8 x 1920x1080), 60 frames each (2 GOPs, "
                   "keyint 30), seeded QP 22-38, search range scaled by width, High decisions, all 48 GOPs (IDR + P "
                   "frames) in one pipelined schedule; value counts P macroblocks",
        "compact": "Compact preset (hexagon search, range 16, 8x8 partitions, I_4x4 in P; Constrained Baseline "
                   "decisions): 1920x1080, QP 24/30/36, keyint 30, 12 GOP streams x (1 IDR + 29 P frames) per step in one "
                   "pipelined schedule, bit-exact with the CPU oracle; value counts P macroblocks",
        "fast": "Fast preset (diamond search +-64 with a half-pel step, Intra 16x16 only, no loop filter; High "
                "decisions): 1920x1080, QP 24/30/36, keyint 30, 12 GOP streams x (1 IDR + 29 P frames) per step in "
                "one pipelined schedule, bit-exact with the CPU oracle; value counts P macroblocks",
        "partitions": "plain GPU form with P partitions 16x16/16x8/8x16/8x8 (the partition extension of the "
                      "Balanced search) and the 8x8 transform: 1280x720, 60 frames as 2 GOP streams x (1 IDR + 29 P "
                      "frames), High decisions, QP 26, search range +-32 (32 diamond iterations), one pipelined "
                      "schedule per step, bit-exact with the CPU oracle; value counts P macroblocks"}[args.workload]
    kernel = {"compact": "rgh_k_compact_step", "partitions": "rgh_k_parts_step"}.get(args.workload, "rgh_k_frame_step")
    out = {
        "metric": metric,
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
        "config": {"workload": workload,
                   "form": "plain", "gops": len(gops), "p_macroblocks_per_step": total_mbs,
                   "idr_macroblocks_per_step": idr_mbs,
                   "all_macroblocks_per_s": round((total_mbs + idr_mbs) * args.steps / el, 1),
                   "schedule": "per-step launches" + (" (one captured HIP graph)" if args.graph else ""), "lag": sched[0], "groups": sched[1],
                   "steps_back_to_back": bool(args.continuous),
                   "isolated_steps": isolated,
                   "launches_per_step": launches / args.steps,
                   "schedule_tuning_s": {"%d/%d" % k: round(v, 4) for k, v in tune.items()},
                   "persistent_schedule": persist,
                   "graph_schedule": graphed,
                   "host_cpu_ms_per_step": round(host_cpu_ms, 3),
                   "device_ms_per_step": round(dev_ms / args.steps, 3),
                   "end_to_end": e2e},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["per_round"] if traffic else None,
                     "traffic_per_macroblock": traffic["per_macroblock"] if traffic else None,
                     "traffic_ratio": traffic["ratio_to_algorithmic"] if traffic else None,
                     "kernel": kernel + (" (+ rgh_k_intra_step in launches with IDR frames)" if args.workload == "compact"
                                         else " (rgh_k_mixed_parts_step in launches with IDR frames)"
                                         if args.workload == "partitions"
                                         else " (rgh_k_mixed_step in launches with IDR frames)"),
                     "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "mbs_per_launch": round(mbs_per_launch, 2), "bytes_per_p_mb": MB_BYTES,
                     "bytes_per_idr_mb": IDR_MB_BYTES},
    }
    if headline:
        out["valu"] = valu_figure({"job": total_mbs + idr_mbs}, dev_ms / 1e3 / args.steps)
        if not args.no_learned:
            out["learned"] = run_learned(gops, ans, srcs, decs, lvls, sched, args, total_mbs, value, hip)
    for a in ans:
        a.close()
    for p in [q for q, _ in uploaded.values()] + [q for d in decs for q in d] + [q for l in lvls for q in l]:
        hip.free(p)
    return out


def end_to_end(hip, gops, ans, srcs, decs, lvls, uploaded, sched, total_mbs, args):
    """The step with its host transfers: every distinct source frame uploaded from pinned host
    memory (the 12 streams share the 120 frames: 4 GOPs x 3 QPs), the analysis, and every frame's
    decision records and levels downloaded into pinned host memory, each copy synchronous."""
    from ragdoll_amd import LEVELS_PER_MB
    planes = [(dp, arr) for dp, arr in uploaded.values()]
    src_bytes = sum(a.nbytes for _, a in planes)
    outs = []
    for a, d, l in zip(ans, decs, lvls):
        n = a.mb_w * a.mb_h
        outs += [(p, n * 32) for p in d] + [(p, n * LEVELS_PER_MB * 2) for p in l]
    out_bytes = sum(b for _, b in outs)
    hs, hsrc = hip.host_alloc(src_bytes)
    ho, hout = hip.host_alloc(out_bytes)
    off = 0
    offs = []
    for _, arr in planes:
        hsrc[off:off + arr.nbytes] = np.ascontiguousarray(arr).reshape(-1).view(np.uint8)
        offs.append(off)
        off += arr.nbytes

    def one():
        t = time.perf_counter()
        for (dp, arr), o in zip(planes, offs):
            hip.memcpy(dp, hs + o, arr.nbytes, 1)
        t1 = time.perf_counter()
        analyse_frames_e2e(sched)
        t2 = time.perf_counter()
        o = 0
        for p, b in outs:
            hip.memcpy(ho + o, p, b, 2)
            o += b
        t3 = time.perf_counter()
        return t3 - t, t1 - t, t2 - t1, t3 - t2

    def analyse_frames_e2e(s):
        from ragdoll_amd import analyse_frames
        analyse_frames(ans, srcs, decs, lvls, lag=s[0], groups=s[1], first_idr=True)

    one()
    tot = h2d = ana = d2h = 0.0
    for _ in range(args.steps):
        a, b, c, d = one()
        tot += a
        h2d += b
        ana += c
        d2h += d
    hip.host_free(hs)
    hip.host_free(ho)
    k = args.steps
    return {"value": round(total_mbs * k / tot, 1), "unit": "macroblocks/s", "ms_per_step": round(tot / k * 1e3, 3),
            "h2d_ms": round(h2d / k * 1e3, 3), "analysis_ms": round(ana / k * 1e3, 3), "d2h_ms": round(d2h / k * 1e3, 3),
            "h2d_bytes": src_bytes, "d2h_bytes": out_bytes,
            "note": "distinct source frames up from pinned host memory, then every frame's decisions and levels "
                    "(408 int16 per macroblock, stored densely) down to pinned host memory; synchronous copies"}


def run_learned(gops, ans, srcs, decs, lvls, sched, args, total_mbs, plain_value, hip):
    """The learned-proposal form on the same streams and schedule. A step computes the proposals
    of every P frame (features from the frame and its previous source frame and FP8 inference over
    all their macroblocks, in one pass) and analyses every GOP from them (rgh_analyse_frames_learned):
    that serial step is the value. The pipelined form, where a step analyses from proposals
    computed beside the previous step's analysis on the network's own HIP stream (as a stream of
    GOP waves could run them), is timed too: on MI355X the concurrent kernels slow the
    latency-bound analysis launches about as much as the overlap saves."""
    from ragdoll_amd import DECISION_DTYPE, analyse_frames
    from ragdoll_amd.learned import Proposer, load_trained
    net = load_trained()
    pics = []
    for i, gp in enumerate(gops):
        for k in range(1, len(gp["frames"])):
            pics.append((srcs[i][k][0], srcs[i][k][3], srcs[i][k - 1][0], srcs[i][k - 1][3], gp["w"], gp["h"], gp["qp"]))
    proposers = [Proposer(net, total_mbs), Proposer(net, total_mbs)]
    cfg = learned_config(args)

    def analyse(props):
        it = iter(props)
        P = [[None] + [next(it) for _ in gp["frames"][1:]] for gp in gops]
        analyse_frames(ans, srcs, decs, lvls, lag=sched[0], groups=sched[1], first_idr=True, proposals=P,
                       learned=cfg)

    def serial_step():
        t = time.perf_counter()
        props = proposers[0].run(pics)
        t1 = time.perf_counter()
        analyse(props)
        return time.perf_counter() - t, t1 - t

    gop_len = len(gops[0]["frames"])
    per_gop = gop_len - 1

    def continuous_step(provs, k):
        """k steps back to back as the plain form runs them: every step's proposals (one pass
        each, into its own buffer), then one call in which each stream runs its GOP k times."""
        t = time.perf_counter()
        props = [provs[s].run(pics) for s in range(k)]
        t1 = time.perf_counter()
        P = [sum(([None] + props[s][i * per_gop:(i + 1) * per_gop] for s in range(k)), []) for i in range(len(gops))]
        analyse_frames(ans, [x * k for x in srcs], [x * k for x in decs], [x * k for x in lvls], lag=sched[0],
                       groups=sched[1], first_idr=True, proposals=P, learned=cfg, keyint=gop_len)
        return time.perf_counter() - t, t1 - t

    def pipelined_step(b, props):
        box = {}
        th = threading.Thread(target=lambda: box.setdefault("p", proposers[b].run(pics)))
        t = time.perf_counter()
        th.start()
        analyse(props)
        th.join()
        return time.perf_counter() - t, box["p"]

    for _ in range(max(1, args.warmup)):
        serial_step()
    el_serial = prop = 0.0
    for _ in range(args.steps):
        e, p = serial_step()
        el_serial += e
        prop += p
    isolated = {"value": round(total_mbs * args.steps / el_serial, 1),
                "ms_per_step": round(el_serial / args.steps * 1e3, 3), "note": "each step its own call"}
    if args.continuous:
        provs = [Proposer(net, total_mbs) for _ in range(max(args.steps, 2))]
        continuous_step(provs, 2)
        el_serial, prop = continuous_step(provs, args.steps)
        for q in provs:
            q.close()
    net.set_launch_limit(args.proposal_groups)
    props = proposers[0].run(pics)
    b = 1
    for _ in range(max(1, args.warmup)):
        _, props = pipelined_step(b, props)
        b ^= 1
    el = 0.0
    for _ in range(args.steps):
        e, props = pipelined_step(b, props)
        el += e
        b ^= 1
    ceiling = learned_ceiling(gops, ans, srcs, decs, lvls, sched, args, total_mbs, hip, analyse_frames)
    # The proposal pass alone (mean images, then features and the FP8 network over every P
    # macroblock of the step in one k_propose launch on the whole device), for its roofline: timed
    # per call around the synchronous pass, so the figure includes the mean images and the
    # descriptor copies (the kernel trace has k_propose's own duration).
    net.set_launch_limit(0)
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        proposers[0].run(pics)
    infer_s = (time.perf_counter() - t) / reps
    flops_mb = 2 * (192 * 128 + 128 * 128 + 128 * 176)
    flops = total_mbs * flops_mb
    # Decision statistics of the last step's P frames.
    counts = np.zeros(5, np.int64)
    types = np.zeros(8, np.int64)
    for i, gp in enumerate(gops):
        n = ans[i].mb_w * ans[i].mb_h
        for k in range(1, len(gp["frames"])):
            d = hip.download(decs[i][k], np.zeros(n, DECISION_DTYPE))
            types += np.bincount(d["mb_type"], minlength=8)[:8]
            for bit in range(5):
                counts[bit] += int(np.count_nonzero(d["flags"] & (1 << bit)))
    for q in proposers:
        q.close()
    net.close()
    value = total_mbs * args.steps / el_serial
    ceiling["vs_plain"] = round(ceiling["value"] / plain_value, 4)
    return {"value": round(value, 1), "unit": "macroblocks/s", "vs_plain": round(value / plain_value, 4),
            "steps_back_to_back": bool(args.continuous), "isolated_steps": isolated,
            "ceiling": ceiling,
            "ms_per_step": round(el_serial / args.steps * 1e3, 3),
            "pipelined_value": round(total_mbs * args.steps / el, 1),
            "pipelined_ms_per_step": round(el / args.steps * 1e3, 3),
            "proposal_ms_per_step": round(prop / args.steps * 1e3, 3),
            "config": cfg, "lag": sched[0], "groups": sched[1], "proposal_groups": args.proposal_groups,
            "plain_search_pct": round(100.0 * counts[1] / total_mbs, 3),
            "skip_probe_gated_pct": round(100.0 * counts[2] / total_mbs, 3),
            "intra_gated_pct": round(100.0 * counts[3] / total_mbs, 3),
            "full_pel_only_pct": round(100.0 * counts[4] / total_mbs, 3),
            "mb_types": types.tolist(),
            "mfma_roofline": {"bound": "mfma", "kernel": "k_propose (features + FP8 proposal network, one pass; "
                                                          "timed with k_means and the descriptor copies)",
                              "achieved": round(flops / infer_s / 1e12, 3), "peak": 5000.0, "unit": "TFLOP/s",
                              "frac": flops / infer_s / 1e12 / 5000.0, "flops_per_mb": flops_mb,
                              "launch_ms": round(infer_s * 1e3, 3)}}


if __name__ == "__main__":
    main()
