"""End to end on the GPU: the reference's golden bitstreams from GPU decisions, and CABAC GOPs.

Each GOP's IDR is coded by the oracle (the IDR is not on this path); every P picture is
analysed on the GPU (single-frame and pipelined entry points) and entropy-coded by the
library's CAVLC writer. The slice NAL units must equal the reference's committed golden clips
(qa/corpus/synth/*, tests/golden/*.nals) byte for byte.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import OracleEncoder, read_nals, split_annexb
from ragdoll_amd import DECISION_DTYPE, LEVELS_PER_MB, StreamAnalyzer, analyse_frames, write_p_slice_cavlc
from ragdoll_amd.hip import Hip
from ragdoll_amd.synth import fixture_frame

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
META = {k: m for k, m in json.load(open(os.path.join(HERE, "golden", "golden.json"))).items()
        if "qp" in m}  # the encoded clips (the y4m entry is raw generator content)


def golden(name):
    m = META[name]
    return m, read_nals(os.path.join(HERE, "golden", name + ".nals"))[m["n_param_sets"]:]


def nal_of(dec, lv, a, qp, k):
    units = split_annexb(write_p_slice_cavlc(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63))
    assert len(units) == 1
    return units[0]


@pytest.mark.parametrize("name", sorted(META))
def test_golden_clip_from_gpu_single_frames(name):
    m, slices = golden(name)
    W, H = m["width"], m["height"]
    print("bitstream single %s" % name, flush=True)
    for g0 in range(0, m["frames"], m["gop"]):
        ora = OracleEncoder(W, H, m["qp"], fps=(m["fps"], 1), keyint=m["gop"] + 1)
        ora.encode(*fixture_frame(W, H, g0))
        a = StreamAnalyzer(W, H, m["qp"])
        try:
            a.set_reference(*ora.recon(deblocked=True))
            for i in range(g0 + 1, min(g0 + m["gop"], m["frames"])):
                dec, lv, _ = a.analyse(*fixture_frame(W, H, i), want_recon=False)
                got = nal_of(dec, lv, a, m["qp"], i - g0)
                assert got == slices[i], "%s frame %d: slice differs (%d vs %d bytes)" % (
                    name, i, len(got), len(slices[i]))
        finally:
            a.close()


@pytest.mark.parametrize("name", sorted(META))
def test_golden_clip_from_gpu_pipelined_gops(name):
    # Every GOP of the clip in one rgh_analyse_frames call (all GOPs' P frames in flight).
    m, slices = golden(name)
    W, H = m["width"], m["height"]
    print("bitstream pipelined %s" % name, flush=True)
    hip = Hip()
    gops = list(range(0, m["frames"], m["gop"]))
    ans, bufs, plan = [], [], []
    try:
        for g0 in gops:
            ora = OracleEncoder(W, H, m["qp"], fps=(m["fps"], 1), keyint=m["gop"] + 1)
            ora.encode(*fixture_frame(W, H, g0))
            a = StreamAnalyzer(W, H, m["qp"], keep_undeblocked=False, ring=m["gop"] + 1)
            a.set_reference(*ora.recon(deblocked=True))
            ans.append(a)
            n = a.mb_w * a.mb_h
            srcs, decs, lvs, idx = [], [], [], []
            for i in range(g0 + 1, min(g0 + m["gop"], m["frames"])):
                y, u, v = fixture_frame(W, H, i)
                dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
                d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
                bufs += [dy, du, dv, d, l]
                srcs.append((dy, du, dv, y.shape[1], u.shape[1], v.shape[1]))
                decs.append(d)
                lvs.append(l)
                idx.append(i)
            plan.append((srcs, decs, lvs, idx, n))
        nf = min(len(p[0]) for p in plan)
        analyse_frames(ans, [p[0][:nf] for p in plan], [p[1][:nf] for p in plan], [p[2][:nf] for p in plan],
                       lag=4, groups=2)
        for a, g0, (srcs, decs, lvs, idx, n) in zip(ans, gops, plan):
            for k in range(nf):
                dec = hip.download(decs[k], np.zeros(n, DECISION_DTYPE))
                lv = hip.download(lvs[k], np.zeros((n, LEVELS_PER_MB), np.int16))
                i = idx[k]
                got = nal_of(dec, lv, a, m["qp"], i - g0)
                assert got == slices[i], "%s frame %d: slice differs" % (name, i)
    finally:
        for a in ans:
            a.close()
        for b in bufs:
            hip.free(b)


# Main / High CABAC: a whole GOP analysed on the GPU (IDR, then P frames) and written by the
# library's CABAC writer must equal the oracle's stream NAL unit for NAL unit (the oracle's
# CABAC slices are pinned by the independent parser of tests/cabac_decoder.py, test_cabac.py).
@pytest.mark.parametrize("cfg,profile,qp,me", [("cif", "high", 26, 16), ("cif", "main", 34, 16),
                                                ("cif", "high", 18, 24), ("720p", "high", 26, 32)])
def test_cabac_gop_from_gpu(cfg, profile, qp, me):
    from ragdoll_amd import write_idr_slice_cabac, write_p_slice_cabac
    from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
    kw = dict(CONFIGS[cfg])
    W, H = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(W, H, seed=config_seed({"cif": 0, "720p": 1}[cfg], 40 + qp), **kw)
    frames = [g.next_frame() for _ in range(5)]
    print("cabac %s %s qp %d me %d" % (cfg, profile, qp, me), flush=True)
    ora = OracleEncoder(W, H, qp, profile=profile, keyint=1 << 20, me_range=me)
    want = [split_annexb(ora.encode(*f))[-1] for f in frames]
    a = StreamAnalyzer(W, H, qp, profile=profile, me_range=me)
    try:
        for k, f in enumerate(frames):
            if k == 0:
                dec, lv, _ = a.analyse_idr(*f, want_recon=False)
                nal = write_idr_slice_cabac(dec, lv, a.mb_w, a.mb_h, qp, 0, profile=profile)
            else:
                dec, lv, _ = a.analyse(*f, want_recon=False)
                nal = write_p_slice_cabac(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, profile=profile)
            got = split_annexb(nal)
            assert len(got) == 1
            assert got[0] == want[k], "%s %s qp %d frame %d: slice differs (%d vs %d bytes)" % (
                cfg, profile, qp, k, len(got[0]), len(want[k]))
    finally:
        a.close()
