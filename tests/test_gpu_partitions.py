"""GPU analysis with P partitions (RGH_FLAG_PARTITIONS: 16x8, 8x16 and 8x8 partitions on the
Balanced search) vs the oracle, bit for bit, and the bitstreams the library writes from the GPU's
records.

The extension is defined by the oracle (parts_analyse in oracle/h264_oracle.c), whose streams the
independent CABAC parser reads back to the oracle's records, motion vectors derived by the
standard's 8.4.1 included (tests/test_cabac.py). Here the GPU must reproduce the oracle's
decisions (the quadrant record layout), levels and both reconstructions, including the partition
predictors (directional for 16x8 / 8x16), the per-quadrant chroma prediction and the deblocking of
motion edges inside macroblocks. The 720p case is the configuration 1280x720 High, QP 26, +-32
search, partitions 16x16/16x8/8x16/8x8 with the 8x8 transform.
"""
import pytest

from oracle_lib import OracleEncoder, split_annexb
from parity import run_parity
from ragdoll_amd import (StreamAnalyzer, write_headers, write_idr_slice_cabac, write_idr_slice_cavlc,
                         write_p_slice_cabac, write_p_slice_cavlc)
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed
from test_gpu_pipeline import run_pipeline

pytestmark = pytest.mark.gpu


def scene(name, n, seed_extra=0):
    kw = dict(CONFIGS[name])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(9, seed_extra), **kw)
    return w, h, [g.next_frame() for _ in range(n)]


@pytest.mark.parametrize("profile", ["baseline", "high"])
@pytest.mark.parametrize("qp", [12, 26, 40])
def test_partitions_cif(profile, qp):
    w, h, frames = scene("cif", 5, qp)
    s = run_parity(frames, w, h, qp, profile=profile, me_range=32, speed="partitions",
                   label="parts cif %s qp%d" % (profile, qp))
    assert s["frames"] == 4
    if qp <= 26:
        assert s["p8x8"] > 0 and s["p16x8"] > 0 and s["p8x16"] > 0, s


def test_partitions_720p_high_qp26_me32():
    w, h, frames = scene("720p", 5, 26)
    s = run_parity(frames, w, h, 26, profile="high", me_range=32, speed="partitions", label="parts 720p qp26")
    assert s["frames"] == 4 and s["p8x8"] + s["p16x8"] + s["p8x16"] > 0, s


def test_partitions_gpu_idr_and_ranges():
    w, h, frames = scene("cif", 5, 77)
    s = run_parity(frames, w, h, 22, profile="main", me_range=4, speed="partitions", gpu_idr=True, keyint=3,
                   label="parts gpu idr me4")
    assert s["idr"] == 2
    s = run_parity(frames, w, h, 30, profile="high", me_range=64, speed="partitions", label="parts me64")
    assert s["frames"] == 4


@pytest.mark.parametrize("w,h", [(18, 18), (50, 30), (130, 66)])
def test_partitions_odd_sizes(w, h):
    g = SceneGenerator(w, h, seed=93, n_objects=3, max_speed=6, noise_sigma=2)
    frames = [g.next_frame() for _ in range(4)]
    run_parity(frames, w, h, 24, profile="high", speed="partitions", label="parts %dx%d" % (w, h))


def test_partitions_pipelined():
    # Partition streams in the pipelined schedule: lag 1 (most macroblocks defer, including at the
    # partition region's readiness check) and lag 12 in two concurrent groups.
    w, h, cif = scene("cif", 6, 11)
    streams = [(cif, w, h, 24, "high", 32, "partitions"), (cif, w, h, 18, "main", 16, "partitions"),
               (cif, w, h, 30, "baseline", 24, "partitions")]
    run_pipeline(streams, lag=1, ring=4, label="parts lag 1")
    run_pipeline(streams, lag=12, ring=4, groups=2, label="parts lag 12")


@pytest.mark.parametrize("profile", ["baseline", "high"])
def test_partition_bitstream_from_gpu(profile):
    # The oracle's stream from GPU records (IDR analysed on the GPU) and the library's writers.
    w, h, frames = scene("cif", 5, 55)
    qp = 24
    ora = OracleEncoder(w, h, qp, profile=profile, keyint=1 << 20, me_range=32, speed="partitions")
    a = StreamAnalyzer(w, h, qp, profile=profile, me_range=32, speed="partitions")
    kinds = set()
    try:
        for k, (y, u, v) in enumerate(frames):
            want = ora.encode(y, u, v)
            if k == 0:
                dec, lv, _ = a.analyse_idr(y, u, v, want_recon=False)
                if profile == "baseline":
                    got = write_headers(w, h, qp, speed="partitions") + write_idr_slice_cavlc(
                        dec, lv, a.mb_w, a.mb_h, qp, 0, speed="partitions")
                else:
                    got = write_headers(w, h, qp, profile=profile, speed="partitions") + write_idr_slice_cabac(
                        dec, lv, a.mb_w, a.mb_h, qp, 0, profile=profile, speed="partitions")
            else:
                dec, lv, _ = a.analyse(y, u, v, want_recon=False)
                kinds |= set(int(t) for t in dec["mb_type"])
                if profile == "baseline":
                    got = write_p_slice_cavlc(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, speed="partitions")
                else:
                    got = write_p_slice_cabac(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63, profile=profile,
                                              speed="partitions")
            assert split_annexb(got) == split_annexb(want), "%s frame %d" % (profile, k)
    finally:
        a.close()
    assert {5, 6, 7} <= kinds, kinds
