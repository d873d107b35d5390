"""Trains the learned motion-proposal network on the GPU and writes its parameters.

Data: synthetic 1920x1080 sequences from the benchmark's generator (20 objects, |v| <= 32
px/frame, 10% scene cuts, noise sigma 3) under training seeds distinct from the benchmark's,
coded in GOPs of 30 (an IDR then 29 P frames) at QPs 22-37 by the plain GPU analysis (High
decisions, the form that is bit-exact with the reference encoder). Every P macroblock gives one
example: its features (rgh_pnet_features from the frame and the previous source frame, on the
GPU) and its label (rgh_pnet_labels: the cell nearest the plain search's vector, and
skip/inter/intra).

Training: Adam on the bfloat16 MFMA kernels (rgh_pnet_train_step), shuffled mini-batches
gathered on the device, warm-up then cosine decay. The FP8 copy is calibrated on a sample of
the training features and scored on held-out GOPs (top-1 / top-4 cell accuracy, mode
accuracy).

    python -m ragdoll_amd.train_proposals --frames 4000 --out proposal_net.npz
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

GOP = 30
W, H = 1920, 1080
TRAIN_QPS = (22, 24, 27, 30, 32, 36, 37)
EVAL_QPS = (24, 30, 36)
TRAIN_CONFIG_ID, EVAL_CONFIG_ID = 6, 7


def _render(args):
    """Frames (y, u, v) of one synthetic 1080p GOP."""
    config_id, stream_id = args
    from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
    kw = dict(CONFIGS["1080p"])
    kw.pop("width")
    kw.pop("height")
    g = SceneGenerator(W, H, seed=config_seed(config_id, stream_id), **kw)
    return [g.next_frame() for _ in range(GOP)]


def collect(net, gops, qps, pool, chunk=8, log=print):
    """Features and labels (device buffers) of every P macroblock of the given GOPs
    ((config_id, stream_id) pairs), coded at qps[i % len(qps)]."""
    from ragdoll_amd import LEVELS_PER_MB, StreamAnalyzer, analyse_frames
    hip = net.hip
    mb = ((W + 15) // 16) * ((H + 15) // 16)
    rows = len(gops) * (GOP - 1) * mb
    feat, lab = hip.alloc(rows * 192 * 2), hip.alloc(rows * 4)
    row = 0
    t0 = time.time()
    pending = pool.imap(_render, gops, chunksize=1)
    for c0 in range(0, len(gops), chunk):
        part = [next(pending) for _ in gops[c0:c0 + chunk]]
        ans, srcs, decs, lvls, bufs = [], [], [], [], []
        for i, frames in enumerate(part):
            qp = qps[(c0 + i) % len(qps)]
            a = StreamAnalyzer(W, H, qp, profile="high", me_range=16, keep_undeblocked=False, ring=GOP + 1)
            ans.append(a)
            fr = []
            for y, u, v in frames:
                p = (hip.upload(y), hip.upload(u), hip.upload(v))
                bufs += list(p)
                fr.append((p[0], p[1], p[2], y.shape[1], u.shape[1], v.shape[1]))
            srcs.append(fr)
            decs.append([hip.alloc(mb * 32) for _ in frames])
            lvls.append([hip.alloc(mb * LEVELS_PER_MB * 2) for _ in frames])
        analyse_frames(ans, srcs, decs, lvls, first_idr=True)
        for i, a in enumerate(ans):
            for k in range(1, GOP):
                net.features(srcs[i][k][0], W, srcs[i][k - 1][0], W, W, H, a.qp, feat + row * 384)
                net.labels(decs[i][k], mb, lab + row * 4)
                row += mb
        for a in ans:
            a.close()
        for q in bufs + [q for d in decs for q in d] + [q for l in lvls for q in l]:
            hip.free(q)
        log("collected %d / %d GOPs (%d examples, %.0f s)" % (min(c0 + chunk, len(gops)), len(gops), row,
                                                              time.time() - t0))
    return feat, lab, row


def evaluate(net, feat, lab, n, batch=1 << 16):
    """FP8 top-1 / top-4 cell accuracy over motion-labelled rows, mode accuracy over all rows."""
    from ragdoll_amd.learned import LABEL_DTYPE, PROPOSAL_DTYPE
    hip = net.hip
    labels = hip.download(lab, np.zeros(n, LABEL_DTYPE))
    props = np.zeros(n, PROPOSAL_DTYPE)
    out = hip.alloc(batch * 16)
    for r in range(0, n, batch):
        k = min(batch, n - r)
        net.infer(feat + r * 384, k, out)
        props[r:r + k] = hip.download(out, np.zeros(k, PROPOSAL_DTYPE))
    hip.free(out)
    cells = props["cell"].astype(np.int64)
    cls = (cells[:, :, 1] + 6) * 13 + cells[:, :, 0] + 6
    has = labels["disp"] >= 0
    top1 = float(np.mean(cls[has, 0] == labels["disp"][has]))
    top4 = float(np.mean((cls[has] == labels["disp"][has, None]).any(axis=1)))
    mode = float(np.mean(np.argmax(props["mode_prob"], axis=1) == labels["mode"]))
    counts = np.bincount(labels["mode"], minlength=3).tolist()
    return {"examples": int(n), "top1": top1, "top4": top4, "mode_accuracy": mode, "mode_counts": counts}


def train(net, feat, lab, n, epochs, batch, lr, log=print):
    hip = net.hip
    xb, lb, ib = hip.alloc(batch * 384), hip.alloc(batch * 4), hip.alloc(n * 4)
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    warm = min(300, total // 10)
    rng = np.random.default_rng(0)
    history = []
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        hip.upload_to(ib, rng.permutation(n).astype(np.int32))
        acc = np.zeros(2)
        seen = 0
        for s in range(steps_per_epoch):
            rate = lr * (step + 1) / warm if step < warm else \
                lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * (step - warm) / max(1, total - warm))))
            net.gather(feat, lab, ib + s * batch * 4, batch, xb, lb)
            want = s % 100 == 0 or s == steps_per_epoch - 1
            loss = net.train_step(xb, lb, batch, rate, want_loss=want)
            if want:
                acc += loss / batch
                seen += 1
            if s % 1000 == 0 or s == steps_per_epoch - 1:
                log("epoch %d step %d/%d lr %.2e loss %.4f %.4f (%.0f s)" % (ep, s, steps_per_epoch, rate,
                                                                           loss[0] / batch, loss[1] / batch,
                                                                           time.time() - t0))
            step += 1
        history.append((acc / max(seen, 1)).tolist())
    for q in (xb, lb, ib):
        hip.free(q)
    return history


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=4000, help="training frames (GOPs of 30)")
    ap.add_argument("--eval-gops", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    ctx = mp.get_context("spawn")
    with ctx.Pool(args.workers) as pool:
        from ragdoll_amd.learned import ProposalNet, init_params
        net = ProposalNet(max_batch=args.batch)
        net.set_params(init_params(args.seed))
        n_gops = max(1, round(args.frames / GOP))
        t0 = time.time()
        feat, lab, n = collect(net, [(TRAIN_CONFIG_ID, s) for s in range(n_gops)], TRAIN_QPS, pool, log=log)
        efeat, elab, en = collect(net, [(EVAL_CONFIG_ID, s) for s in range(args.eval_gops)], EVAL_QPS, pool, log=log)
    t_data = time.time() - t0
    t0 = time.time()
    history = train(net, feat, lab, n, args.epochs, args.batch, args.lr, log=log)
    t_train = time.time() - t0
    # FP8 calibration on rows spread over the whole training set.
    hip = net.hip
    k = min(n, args.batch)
    idx, xb, lb = hip.upload(np.linspace(0, n - 1, k).astype(np.int32)), hip.alloc(k * 384), hip.alloc(k * 4)
    net.gather(feat, lab, idx, k, xb, lb)
    net.quantize(xb, k)
    for q in (idx, xb, lb):
        hip.free(q)
    result = {"train": evaluate(net, feat, lab, min(n, 1 << 20)), "held_out": evaluate(net, efeat, elab, en),
              "frames": n_gops * GOP, "gops": n_gops, "qps": list(TRAIN_QPS), "epochs": args.epochs,
              "batch": args.batch, "lr": args.lr, "loss_history": history, "data_s": round(t_data, 1),
              "train_s": round(t_train, 1)}
    log(json.dumps(result))
    np.savez(args.out, params=net.get_params(), act_max=np.array(net.activation_maxima(), np.float32),
             meta=np.array(json.dumps(result)))
    for q in (feat, lab, efeat, elab):
        net.hip.free(q)
    net.close()


if __name__ == "__main__":
    sys.exit(main())
