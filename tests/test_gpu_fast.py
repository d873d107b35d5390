"""GPU Fast-preset analysis (enc.rs:503 H264Speed::Fast) vs the oracle, bit for bit, and the
bitstreams the library writes from the GPU's records.

Fast differs from Balanced in: the P_Skip probe only beside a skipped left or top neighbour, a
half-pel step toward the cheaper side of the full-pel minimum (no quarter-pel refinement), the
Intra 16x16 trial only when its SAD undercuts the inter SAD by the 7:4 gate, no plane modes and
no Intra 4x4, an intra bias of 16 lambda, the P-picture chroma mode from the luma mode, and no
loop filter.
"""
import numpy as np
import pytest

from oracle_lib import OracleEncoder, split_annexb
from parity import run_parity
from ragdoll_amd import (RghError, StreamAnalyzer, analyse_batch, write_headers, write_idr_slice_cabac, write_idr_slice_cavlc,
                         write_p_slice_cabac, write_p_slice_cavlc)
from ragdoll_amd.hip import Hip
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame
from test_gpu_pipeline import run_pipeline

pytestmark = pytest.mark.gpu


def scene(name, n, seed_extra=0):
    kw = dict(CONFIGS[name])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(8, seed_extra), **kw)
    return w, h, [g.next_frame() for _ in range(n)]


@pytest.mark.parametrize("profile", ["baseline", "high"])
@pytest.mark.parametrize("qp", [10, 24, 38, 51])
def test_fast_cif(profile, qp):
    w, h, frames = scene("cif", 5, qp)
    s = run_parity(frames, w, h, qp, profile=profile, speed="fast", label="fast cif %s qp%d" % (profile, qp))
    assert s["frames"] == 4 and s["i4"] == 0
    assert s["inter"] > 0 or qp == 51


@pytest.mark.parametrize("me_range", [4, 32, 64])
def test_fast_search_ranges(me_range):
    w, h, frames = scene("cif", 4, 100 + me_range)
    s = run_parity(frames, w, h, 26, profile="high", me_range=me_range, speed="fast",
                   label="fast cif me%d" % me_range)
    assert s["frames"] == 3


def test_fast_gpu_idr_and_fixture_content():
    frames = [fixture_frame(320, 180, i) for i in range(5)]
    s = run_parity(frames, 320, 180, 22, profile="main", speed="fast", gpu_idr=True, keyint=3,
                   label="fast fixture idr")
    assert s["idr"] == 2 and s["i4"] == 0


@pytest.mark.parametrize("w,h", [(18, 18), (50, 30)])
def test_fast_odd_sizes(w, h):
    g = SceneGenerator(w, h, seed=91, n_objects=2, max_speed=6, noise_sigma=2)
    frames = [g.next_frame() for _ in range(4)]
    run_parity(frames, w, h, 28, speed="fast", label="fast %dx%d" % (w, h))


def test_fast_720p():
    w, h, frames = scene("720p", 3)
    s = run_parity(frames, w, h, 26, profile="high", me_range=32, speed="fast", label="fast 720p")
    assert s["frames"] == 2


def test_fast_pipelined():
    # Fast streams in the pipelined schedule (two concurrent groups), then the persistent schedule.
    w, h, cif = scene("cif", 5, 7)
    streams = [(cif, w, h, 26, "high", 16, "fast"), (cif, w, h, 20, "main", 16, "fast"),
               (cif, w, h, 34, "baseline", 32, "fast")]
    run_pipeline(streams, lag=8, ring=4, groups=2, label="fast pipelined")
    run_pipeline(streams, lag=8, ring=4, label="fast persistent", persistent=True)


def test_one_preset_per_call():
    # Each preset's P pictures run their own kernels: a call that mixes presets is refused.
    w, h, frames = scene("cif", 2, 9)
    hip = Hip()
    a = StreamAnalyzer(w, h, 26, speed="fast")
    b = StreamAnalyzer(w, h, 26)
    planes = [hip.upload(p) for p in frames[1]]
    try:
        for s in (a, b):
            s.set_reference(*frames[0])
        src = tuple(planes) + tuple(p.shape[1] for p in frames[1])
        with pytest.raises(RghError, match="one preset"):
            analyse_batch([a, b], [src, src])
    finally:
        a.close()
        b.close()
        for p in planes:
            hip.free(p)


@pytest.mark.parametrize("profile", ["baseline", "high"])
def test_fast_bitstream_from_gpu(profile):
    # The oracle's Fast stream from GPU records (IDR analysed on the GPU) and the library's writers.
    w, h, frames = scene("cif", 5, 55)
    qp = 28
    ora = OracleEncoder(w, h, qp, profile=profile, keyint=1 << 20, speed="fast")
    a = StreamAnalyzer(w, h, qp, profile=profile, speed="fast")
    try:
        for k, (y, u, v) in enumerate(frames):
            want = ora.encode(y, u, v)
            if k == 0:
                dec, lv, _ = a.analyse_idr(y, u, v, want_recon=False)
                if profile == "baseline":
                    got = write_headers(w, h, qp, speed="fast") + write_idr_slice_cavlc(
                        dec, lv, a.mb_w, a.mb_h, qp, 0, speed="fast")
                else:
                    got = write_headers(w, h, qp, profile=profile, speed="fast") + write_idr_slice_cabac(
                        dec, lv, a.mb_w, a.mb_h, qp, 0, profile=profile, speed="fast")
            else:
                dec, lv, _ = a.analyse(y, u, v, want_recon=False)
                if profile == "baseline":
                    got = write_p_slice_cavlc(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, speed="fast")
                else:
                    got = write_p_slice_cabac(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, profile=profile,
                                              speed="fast")
            assert split_annexb(got) == split_annexb(want), "%s frame %d" % (profile, k)
    finally:
        a.close()
