"""GPU IDR analysis (rgh_analyse_idr: I_16x16 / I_4x4 decisions, levels, reconstructions) vs
the oracle, bit for bit, and whole GOPs analysed on the GPU from their IDR on."""
import json
import os

import numpy as np
import pytest

from oracle_lib import OracleEncoder, read_nals, split_annexb
from parity import run_parity
from ragdoll_amd import (DECISION_DTYPE, LEVELS_PER_MB, StreamAnalyzer, analyse_frames, write_headers,
                         write_idr_slice_cavlc, write_p_slice_cavlc)
from ragdoll_amd.hip import Hip
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
META = {k: m for k, m in json.load(open(os.path.join(HERE, "golden", "golden.json"))).items()
        if "qp" in m}  # the encoded clips (the y4m entry is raw generator content)


def scene_frames(w, h, n, seed, **kw):
    g = SceneGenerator(w, h, seed=seed, **kw)
    return [g.next_frame() for _ in range(n)]


def cif(n, seed):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    return scene_frames(w, h, n, seed, **kw), w, h


@pytest.mark.parametrize("w,h", [(160, 90), (320, 180)])
def test_golden_clip_content_idr(w, h):
    frames = [fixture_frame(w, h, i) for i in range(6)]
    s = run_parity(frames, w, h, 20, keyint=3, gpu_idr=True, label="idr fixture %dx%d" % (w, h))
    assert s["idr"] == 2 and s["i4"] > 0


@pytest.mark.parametrize("qp", [1, 6, 18, 28, 40, 51])
def test_cif_idr_qp_sweep(qp):
    frames, w, h = cif(3, config_seed(2, qp))
    s = run_parity(frames, w, h, qp, keyint=2, gpu_idr=True, label="idr cif qp%d" % qp)
    assert s["idr"] == 2


@pytest.mark.parametrize("profile", ["main", "high"])
def test_cif_idr_cabac_profiles(profile):
    frames, w, h = cif(3, config_seed(2, 77))
    run_parity(frames, w, h, 26, profile=profile, keyint=2, gpu_idr=True, label="idr cif " + profile)


@pytest.mark.parametrize("w,h", [(18, 18), (50, 30), (130, 66)])
def test_idr_odd_sizes(w, h):
    frames = scene_frames(w, h, 3, 98, n_objects=2, max_speed=6, noise_sigma=2)
    run_parity(frames, w, h, 24, keyint=2, gpu_idr=True, label="idr %dx%d" % (w, h))


def test_720p_idr_high():
    kw = dict(CONFIGS["720p"])
    w, h = kw.pop("width"), kw.pop("height")
    run_parity(scene_frames(w, h, 2, config_seed(3), **kw), w, h, 30, profile="high", keyint=4, gpu_idr=True,
               label="idr 720p")


@pytest.mark.parametrize("name", sorted(META))
def test_golden_clip_entirely_from_gpu(name):
    # Every picture of the reference's golden clips (IDRs included) analysed on the GPU and
    # entropy-coded by the library: the clip's parameter sets and slices byte for byte.
    m = META[name]
    gold = read_nals(os.path.join(HERE, "golden", name + ".nals"))
    W, H = m["width"], m["height"]
    assert split_annexb(write_headers(W, H, m["qp"], m["profile"], fps=(m["fps"], 1))) == gold[:m["n_param_sets"]]
    slices = gold[m["n_param_sets"]:]
    print("bitstream from GPU %s" % name, flush=True)
    a = StreamAnalyzer(W, H, m["qp"])
    try:
        for i in range(m["frames"]):
            k = i % m["gop"]
            if k == 0:
                dec, lv, _ = a.analyse_idr(*fixture_frame(W, H, i), want_recon=False)
                nal = write_idr_slice_cavlc(dec, lv, a.mb_w, a.mb_h, m["qp"], 0)
            else:
                dec, lv, _ = a.analyse(*fixture_frame(W, H, i), want_recon=False)
                nal = write_p_slice_cavlc(dec, lv, a.mb_w, a.mb_h, m["qp"], k & 15, (2 * k) & 63)
            assert split_annexb(nal) == [slices[i]], "%s frame %d" % (name, i)
    finally:
        a.close()


@pytest.mark.parametrize("lag,groups", [(0, 0), (3, 2)])
def test_pipelined_gops_from_idr(lag, groups):
    # rgh_analyse_frames with its first frame an IDR: two streams' whole GOPs in one call,
    # against the oracle's decisions and levels frame by frame.
    hip = Hip()
    streams = []
    for s in range(2):
        frames, w, h = cif(4, config_seed(4, s))
        streams.append(frames)
    qp = 27
    ans, bufs, srcs, decs, lvs = [], [], [], [], []
    try:
        for frames in streams:
            a = StreamAnalyzer(w, h, qp, ring=6)
            ans.append(a)
            n = a.mb_w * a.mb_h
            s_, d_, l_ = [], [], []
            for y, u, v in frames:
                dy, du, dv = hip.upload(y), hip.upload(u), hip.upload(v)
                d, l = hip.alloc(n * 32), hip.alloc(n * LEVELS_PER_MB * 2)
                bufs += [dy, du, dv, d, l]
                s_.append((dy, du, dv, y.shape[1], u.shape[1], v.shape[1]))
                d_.append(d)
                l_.append(l)
            srcs.append(s_)
            decs.append(d_)
            lvs.append(l_)
        print("pipelined from IDR lag %d groups %d" % (lag, groups), flush=True)
        analyse_frames(ans, srcs, decs, lvs, lag=lag, groups=groups, first_idr=True)
        for si, frames in enumerate(streams):
            ora = OracleEncoder(w, h, qp, keyint=1 << 20)
            n = ans[si].mb_w * ans[si].mb_h
            for i, (y, u, v) in enumerate(frames):
                ora.encode(y, u, v)
                dec = hip.download(decs[si][i], np.zeros(n, DECISION_DTYPE))
                lv = hip.download(lvs[si][i], np.zeros((n, LEVELS_PER_MB), np.int16))
                odec = ora.decisions()
                bad = np.nonzero(dec.view(np.uint8).reshape(-1, 32) != odec.view(np.uint8).reshape(-1, 32))[0]
                assert len(bad) == 0, "stream %d frame %d: MB %d differs\n gpu %s\n ora %s" % (
                    si, i, bad[0], dec[bad[0]], odec[bad[0]])
                assert np.array_equal(lv, ora.levels()), "stream %d frame %d: levels differ" % (si, i)
            # The last picture's deblocked reconstruction.
            _, _, rec = ans[si].outputs(which_recon=1, want_levels=False)
            for g, o in zip(rec, ora.recon(deblocked=True)):
                assert np.array_equal(g, o), "stream %d: final reconstruction differs" % si
    finally:
        for a in ans:
            a.close()
        for b in bufs:
            hip.free(b)


def test_misaligned_output_buffers_are_rejected():
    # rgh_analyse_frames stores records and clears levels as 64-bit words: buffers that are not
    # 8-byte aligned are refused before any work starts; aligned ones are analysed.
    from ragdoll_amd import RghError
    hip = Hip()
    frames, w, h = cif(1, config_seed(4, 0))
    y, u, v = frames[0]
    a = StreamAnalyzer(w, h, 27, ring=3)
    n = a.mb_w * a.mb_h
    src = (hip.upload(y), hip.upload(u), hip.upload(v))
    d, l = hip.alloc(n * 32 + 16), hip.alloc(n * LEVELS_PER_MB * 2 + 16)
    frame = [[(src[0], src[1], src[2], y.shape[1], u.shape[1], v.shape[1])]]
    try:
        print("misaligned output buffers", flush=True)
        with pytest.raises(RghError):
            analyse_frames([a], frame, [[d]], [[l + 2]], first_idr=True)
        with pytest.raises(RghError):
            analyse_frames([a], frame, [[d + 4]], [[l]], first_idr=True)
        analyse_frames([a], frame, [[d]], [[l]], first_idr=True)
        dec = hip.download(d, np.zeros(n, DECISION_DTYPE))
        assert (dec["mb_type"] >= 2).all()  # an IDR picture: I_16x16 or I_4x4 throughout
    finally:
        a.close()
        for q in list(src) + [d, l]:
            hip.free(q)
