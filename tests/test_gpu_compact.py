"""GPU Compact-preset P analysis (hexagon search, 8x8 partitions, I_4x4 in P, partition-aware
deblocking, chroma QP offset 1) vs the oracle, bit for bit, and the bitstream the library writes
from the GPU's records."""
import numpy as np
import pytest

from oracle_lib import OracleEncoder, split_annexb
from parity import run_parity
from ragdoll_amd import StreamAnalyzer, write_headers, write_idr_slice_cavlc, write_p_slice_cavlc
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame

pytestmark = pytest.mark.gpu


def scene_frames(w, h, n, seed, **kw):
    g = SceneGenerator(w, h, seed=seed, **kw)
    return [g.next_frame() for _ in range(n)]


def cif(n, seed):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    return scene_frames(w, h, n, seed, **kw), w, h


@pytest.mark.parametrize("w,h", [(160, 90), (320, 180)])
def test_compact_golden_content(w, h):
    frames = [fixture_frame(w, h, i) for i in range(6)]
    s = run_parity(frames, w, h, 20, speed="compact", label="compact fixture %dx%d" % (w, h))
    assert s["frames"] == 5


@pytest.mark.parametrize("qp", [12, 22, 28, 36, 46])
def test_compact_cif_qp_sweep(qp):
    frames, w, h = cif(5, config_seed(5, qp))
    s = run_parity(frames, w, h, qp, speed="compact", label="compact cif qp%d" % qp)
    assert s["frames"] == 4


def test_compact_cif_gpu_idr_and_partitions():
    # The IDR analysed on the GPU too. Streams start at their IDR (one encoder per GOP, as the
    # reference's GopRingEncoder runs them): after a mid-stream IDR the reference would take the
    # 16x16 vectors of the picture two before it as candidates, the GPU context the IDR's zeros.
    frames, w, h = cif(6, config_seed(5, 99))
    s = run_parity(frames, w, h, 24, speed="compact", gpu_idr=True, label="compact cif idr")
    assert s["idr"] == 1 and s["p8x8"] > 0 and s["inter"] > 0


@pytest.mark.parametrize("w,h", [(18, 18), (50, 30), (130, 66)])
def test_compact_odd_sizes(w, h):
    run_parity(scene_frames(w, h, 4, 97, n_objects=2, max_speed=6, noise_sigma=2), w, h, 24, speed="compact",
               label="compact %dx%d" % (w, h))


def test_compact_720p():
    kw = dict(CONFIGS["720p"])
    w, h = kw.pop("width"), kw.pop("height")
    run_parity(scene_frames(w, h, 3, config_seed(6), **kw), w, h, 28, speed="compact", label="compact 720p")


def test_compact_bitstream_from_gpu():
    # The oracle's Compact stream (SPS, PPS, IDR, P slices) from GPU records and the writers.
    frames, w, h = cif(5, config_seed(7))
    qp = 26
    ora = OracleEncoder(w, h, qp, keyint=1 << 20, speed="compact")
    a = StreamAnalyzer(w, h, qp, speed="compact")
    try:
        for k, (y, u, v) in enumerate(frames):
            nals = split_annexb(ora.encode(y, u, v))
            if k == 0:
                dec, lv, _ = a.analyse_idr(y, u, v, want_recon=False)
                got = split_annexb(write_headers(w, h, qp, speed="compact") +
                                   write_idr_slice_cavlc(dec, lv, a.mb_w, a.mb_h, qp, 0))
            else:
                dec, lv, _ = a.analyse(y, u, v, want_recon=False)
                got = split_annexb(write_p_slice_cavlc(dec, lv, a.mb_w, a.mb_h, qp, k & 15, (2 * k) & 63,
                                                       speed="compact"))
            assert got == nals, "frame %d" % k
    finally:
        a.close()
