"""Rate-distortion measurement of the GPU analysis: PSNR, coded bits and Bjontegaard delta rate.

Used to score the learned-proposal search against the plain one: each form codes the same GOPs
at several QPs (High profile CABAC through rgh_write_*_slice_cabac, or Constrained Baseline
CAVLC), and the BD-rate of the learned form's (bits, luma PSNR) curve against the plain form's
says how many more bits it spends for the same quality.
"""
import numpy as np


def psnr(a, b, peak=255.0):
    """PSNR in dB of two equally shaped sample arrays (100 dB when identical)."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    mse = float(np.mean(d * d))
    return 100.0 if mse == 0.0 else 10.0 * np.log10(peak * peak / mse)


def pchip_slopes(x, y):
    """Node derivatives of the monotone piecewise cubic Hermite interpolant (Fritsch and
    Carlson 1980, with the three-point end formula of Moler's pchip): zero where the secant
    slopes change sign, else their weighted harmonic mean."""
    h = np.diff(x)
    delta = np.diff(y) / h
    n = len(x)
    d = np.zeros(n)
    for k in range(1, n - 1):
        if delta[k - 1] * delta[k] > 0:
            w1, w2 = 2 * h[k] + h[k - 1], h[k] + 2 * h[k - 1]
            d[k] = (w1 + w2) / (w1 / delta[k - 1] + w2 / delta[k])

    def end(h0, h1, d0, d1):
        v = ((2 * h0 + h1) * d0 - h0 * d1) / (h0 + h1)
        if np.sign(v) != np.sign(d0):
            return 0.0
        if np.sign(d0) != np.sign(d1) and abs(v) > 3 * abs(d0):
            return 3 * d0
        return v

    if n == 2:
        d[:] = delta[0]
    else:
        d[0] = end(h[0], h[1], delta[0], delta[1])
        d[-1] = end(h[-1], h[-2], delta[-1], delta[-2])
    return d


def pchip_integral(x, y, lo, hi):
    """Integral over [lo, hi] (inside [x[0], x[-1]]) of the pchip interpolant of (x, y)."""
    d = pchip_slopes(x, y)
    total = 0.0
    for k in range(len(x) - 1):
        a, b = max(lo, x[k]), min(hi, x[k + 1])
        if b <= a:
            continue
        h = x[k + 1] - x[k]
        delta = (y[k + 1] - y[k]) / h
        c0, c1 = y[k], d[k]
        c2 = (3 * delta - 2 * d[k] - d[k + 1]) / h
        c3 = (d[k] + d[k + 1] - 2 * delta) / (h * h)

        def anti(s):
            return c0 * s + c1 * s * s / 2 + c2 * s ** 3 / 3 + c3 * s ** 4 / 4

        total += anti(b - x[k]) - anti(a - x[k])
    return total


def bd_rate(rate_a, psnr_a, rate_b, psnr_b, method="pchip"):
    """Bjontegaard delta rate of curve b against curve a, in percent (positive: b needs more
    bits for the same PSNR): the mean difference of the curves' log rates over their common
    PSNR interval, as a rate ratio. method "pchip": each curve's log rate is the piecewise cubic
    (monotone Hermite) interpolant of its points in PSNR; "cubic": one least-squares cubic in
    PSNR per curve (the original Bjontegaard fit)."""
    ra, pa = np.log(np.asarray(rate_a, np.float64)), np.asarray(psnr_a, np.float64)
    rb, pb = np.log(np.asarray(rate_b, np.float64)), np.asarray(psnr_b, np.float64)
    if len(ra) < 4 or len(rb) < 4:
        raise ValueError("BD-rate needs four rate points per curve")
    lo, hi = max(pa.min(), pb.min()), min(pa.max(), pb.max())
    if not hi > lo:
        raise ValueError("the curves' PSNR ranges do not overlap")
    if method == "pchip":
        oa, ob = np.argsort(pa), np.argsort(pb)
        if np.any(np.diff(pa[oa]) <= 0) or np.any(np.diff(pb[ob]) <= 0):
            raise ValueError("a curve repeats a PSNR value")
        mean_a = pchip_integral(pa[oa], ra[oa], lo, hi) / (hi - lo)
        mean_b = pchip_integral(pb[ob], rb[ob], lo, hi) / (hi - lo)
    elif method == "cubic":
        ia, ib = np.polyint(np.polyfit(pa, ra, 3)), np.polyint(np.polyfit(pb, rb, 3))
        mean_a = (np.polyval(ia, hi) - np.polyval(ia, lo)) / (hi - lo)
        mean_b = (np.polyval(ib, hi) - np.polyval(ib, lo)) / (hi - lo)
    else:
        raise ValueError("unknown BD-rate method %r" % method)
    return float((np.exp(mean_b - mean_a) - 1.0) * 100.0)


def code_gop(frames, width, height, qp, proposer=None, learned=None, profile="baseline", me_range=0, speed="balanced"):
    """Codes one GOP (frame 0 IDR, the rest P) with the GPU analysis, one frame per call, and
    the library's slice writer: CAVLC for "baseline", CABAC for "main" / "high". With `proposer`
    (a learned.Proposer) every P frame is searched from learned proposals
    (rgh_analyse_frames_learned with `learned` as its config).

    Returns a dict: bytes (headers + slices), psnr_y (mean luma PSNR of the deblocked
    reconstruction at display size), per-frame lists, and the P frames' decision statistics."""
    from .abi import (LEVELS_PER_MB, StreamAnalyzer, analyse_frames, write_headers, write_idr_slice_cabac,
                      write_idr_slice_cavlc, write_p_slice_cabac, write_p_slice_cavlc)
    from .hip import Hip
    hip = Hip()
    a = StreamAnalyzer(width, height, qp, profile=profile, me_range=me_range, keep_undeblocked=False, ring=2,
                       speed=speed)
    n = a.mb_w * a.mb_h
    dec, lv = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
    total = len(write_headers(width, height, qp, profile=profile, speed=speed))
    cabac = profile != "baseline"
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
                nal = (write_idr_slice_cabac(d, l, a.mb_w, a.mb_h, qp, 0, profile=profile, speed=speed) if cabac
                       else write_idr_slice_cavlc(d, l, a.mb_w, a.mb_h, qp, 0, speed=speed))
            else:
                nal = (write_p_slice_cabac(d, l, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, profile=profile,
                                           speed=speed) if cabac
                       else write_p_slice_cavlc(d, l, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, speed=speed))
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
