"""Rate-distortion measurement of the GPU analysis: PSNR, CAVLC bits and Bjontegaard delta rate.

Used to score the learned-proposal search against the plain one: each form codes the same GOPs
at several QPs (Constrained Baseline, CAVLC, through rgh_write_*_slice_cavlc), and the BD-rate
of the learned form's (bits, luma PSNR) curve against the plain form's says how many more bits
it spends for the same quality.
"""
import numpy as np


def psnr(a, b, peak=255.0):
    """PSNR in dB of two equally shaped sample arrays (100 dB when identical)."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    mse = float(np.mean(d * d))
    return 100.0 if mse == 0.0 else 10.0 * np.log10(peak * peak / mse)


def bd_rate(rate_a, psnr_a, rate_b, psnr_b):
    """Bjontegaard delta rate of curve b against curve a, in percent (positive: b needs more
    bits for the same PSNR). Each curve's log rate is fitted by a cubic in PSNR; the mean
    log-rate difference over the overlapping PSNR interval is returned as a rate ratio."""
    ra, pa = np.log(np.asarray(rate_a, np.float64)), np.asarray(psnr_a, np.float64)
    rb, pb = np.log(np.asarray(rate_b, np.float64)), np.asarray(psnr_b, np.float64)
    if len(ra) < 4 or len(rb) < 4:
        raise ValueError("BD-rate needs four rate points per curve")
    fa, fb = np.polyfit(pa, ra, 3), np.polyfit(pb, rb, 3)
    lo, hi = max(pa.min(), pb.min()), min(pa.max(), pb.max())
    if not hi > lo:
        raise ValueError("the curves' PSNR ranges do not overlap")
    ia, ib = np.polyint(fa), np.polyint(fb)
    mean_a = (np.polyval(ia, hi) - np.polyval(ia, lo)) / (hi - lo)
    mean_b = (np.polyval(ib, hi) - np.polyval(ib, lo)) / (hi - lo)
    return float((np.exp(mean_b - mean_a) - 1.0) * 100.0)


def code_gop(frames, width, height, qp, proposer=None, learned=None):
    """Codes one GOP (frame 0 IDR, the rest P) with the GPU analysis, one frame per call, and
    the CAVLC writer. With `proposer` (a learned.Proposer) every P frame is searched from
    learned proposals (rgh_analyse_frames_learned with `learned` as its config).

    Returns a dict: bytes (headers + slices), psnr_y (mean luma PSNR of the deblocked
    reconstruction at display size), per-frame lists, and the P frames' decision statistics."""
    from .abi import (LEVELS_PER_MB, StreamAnalyzer, analyse_frames, write_headers, write_idr_slice_cavlc,
                      write_p_slice_cavlc)
    from .hip import Hip
    hip = Hip()
    a = StreamAnalyzer(width, height, qp, profile="baseline", keep_undeblocked=False, ring=2)
    n = a.mb_w * a.mb_h
    dec, lv = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
    total = len(write_headers(width, height, qp))
    sizes, psnrs = [], []
    flags = np.zeros(16, np.int64)
    types = np.zeros(8, np.int64)
    inter_cost = 0
    prev = None
    bufs = []
    try:
        for k, (y, u, v) in enumerate(frames):
            cur = (hip.upload(y), hip.upload(u), hip.upload(v))
            bufs += list(cur)
            src = [[(cur[0], cur[1], cur[2], y.shape[1], u.shape[1], v.shape[1])]]
            if k == 0:
                analyse_frames([a], src, [[dec]], [[lv]], first_idr=True)
            elif proposer is not None:
                props = proposer.run([(cur[0], y.shape[1], prev[0], y.shape[1], width, height, qp)])
                analyse_frames([a], src, [[dec]], [[lv]], proposals=[[props[0]]], learned=learned)
            else:
                analyse_frames([a], src, [[dec]], [[lv]])
            d, l, rec = a.outputs(which_recon=1)
            if k == 0:
                nal = write_idr_slice_cavlc(d, l, a.mb_w, a.mb_h, qp, 0)
            else:
                nal = write_p_slice_cavlc(d, l, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63)
                types += np.bincount(d["mb_type"], minlength=8)[:8]
                inter_cost += int(d["cost"][d["mb_type"] == 1].astype(np.int64).sum())
                for b in range(4):
                    flags[b] += int(np.count_nonzero(d["flags"] & (1 << b)))
            total += len(nal)
            sizes.append(len(nal))
            psnrs.append(psnr(rec[0][:height, :width], y))
            if prev is not None:
                for q in prev:
                    hip.free(q)
                    bufs.remove(q)
            prev = cur
    finally:
        for q in bufs:
            hip.free(q)
        hip.free(dec)
        hip.free(lv)
        a.close()
    p_mbs = n * max(0, len(frames) - 1)
    return {"bytes": total, "psnr_y": float(np.mean(psnrs)), "frame_bytes": sizes, "frame_psnr_y": psnrs,
            "p_macroblocks": p_mbs, "mb_types": types.tolist(),
            "inter_cost_per_mb": inter_cost / max(1, int(types[1])),
            "plain_search": int(flags[1]), "skip_probe_gated": int(flags[2]), "intra_gated": int(flags[3])}
