"""Shared driver of the GPU-vs-oracle parity tests.

The oracle codes the IDR frame; its deblocked reconstruction becomes the GPU context's
reference (rgh_set_reference), or the GPU analyses the IDR itself (rgh_analyse_idr). Every
following P frame is analysed by both, and the GPU's decision records, levels, undeblocked and
deblocked reconstructions must equal the oracle's bit for bit.
"""
import numpy as np

from oracle_lib import OracleEncoder


def first_difference(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    bad = np.nonzero(a.reshape(a.shape[0], -1) != b.reshape(b.shape[0], -1))[0] if a.ndim > 1 else np.nonzero(a != b)[0]
    return int(bad[0]) if len(bad) else None


def check_frame(gpu, ora, dec, lv, rec, label, i):
    """The GPU's decisions, levels and both reconstructions of frame i against the oracle's."""
    odec, olv = ora.decisions(), ora.levels()
    bad = np.nonzero(dec.view(np.uint8).reshape(-1, 32) != odec.view(np.uint8).reshape(-1, 32))
    if len(bad[0]):
        m = int(bad[0][0])
        raise AssertionError("%s frame %d: decision of MB %d (x %d, y %d) differs\n gpu %s\n ora %s" % (
            label, i, m, m % ora.mb_w, m // ora.mb_w, dec[m], odec[m]))
    m = first_difference(lv, olv)
    assert m is None, "%s frame %d: levels of MB %d differ\n gpu %s\n ora %s" % (label, i, m, lv[m], olv[m])
    _, _, pre = gpu.outputs(which_recon=0, want_levels=False)
    for p, (g, o) in enumerate(zip(pre, ora.recon(deblocked=False))):
        if not np.array_equal(g, o):
            yy, xx = np.nonzero(g != o)
            raise AssertionError("%s frame %d: undeblocked plane %d differs first at (%d, %d)" % (
                label, i, p, xx[0], yy[0]))
    for p, (g, o) in enumerate(zip(rec, ora.recon(deblocked=True))):
        if not np.array_equal(g, o):
            yy, xx = np.nonzero(g != o)
            raise AssertionError("%s frame %d: deblocked plane %d differs first at (%d, %d)" % (
                label, i, p, xx[0], yy[0]))


def run_parity(frames, width, height, qp, profile="baseline", me_range=16, label="", keyint=1 << 20,
               gpu_idr=False, speed="balanced"):
    """Codes frames with the oracle and the GPU context side by side. IDR pictures are set as
    the GPU's reference from the oracle's reconstruction, or, with gpu_idr, analysed on the GPU
    (rgh_analyse_idr) and checked like P pictures."""
    from ragdoll_amd import StreamAnalyzer
    ora = OracleEncoder(width, height, qp, profile=profile, keyint=keyint, me_range=me_range, speed=speed)
    gpu = StreamAnalyzer(width, height, qp, profile=profile, me_range=me_range, speed=speed)
    stats = {"frames": 0, "mbs": 0, "skip": 0, "inter": 0, "intra": 0, "idr": 0, "i4": 0, "p8x8": 0, "p16x8": 0,
             "p8x16": 0}
    try:
        for i, (y, u, v) in enumerate(frames):
            ora.encode(y, u, v)
            if ora.last_was_idr:
                if not gpu_idr:
                    gpu.set_reference(*ora.recon(deblocked=True))
                    continue
                print("parity %s frame %d (IDR)" % (label, i), flush=True)
                dec, lv, rec = gpu.analyse_idr(y, u, v)
                check_frame(gpu, ora, dec, lv, rec, label, i)
                stats["idr"] += 1
                stats["i4"] += int((dec["mb_type"] == 3).sum())
                continue
            print("parity %s frame %d" % (label, i), flush=True)
            dec, lv, rec = gpu.analyse(y, u, v)
            check_frame(gpu, ora, dec, lv, rec, label, i)
            stats["frames"] += 1
            stats["mbs"] += len(dec)
            stats["skip"] += int((dec["mb_type"] == 0).sum())
            stats["inter"] += int((dec["mb_type"] == 1).sum())
            stats["intra"] += int((dec["mb_type"] == 2).sum())
            stats["i4"] += int((dec["mb_type"] == 3).sum())
            stats["p8x8"] += int((dec["mb_type"] == 5).sum())
            stats["p16x8"] += int((dec["mb_type"] == 6).sum())
            stats["p8x16"] += int((dec["mb_type"] == 7).sum())
    finally:
        gpu.close()
    return stats
