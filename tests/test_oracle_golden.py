"""The oracle against the reference's committed golden bitstreams.

qa/corpus/synth/clip_160x90_30fps.mp4 and clip_320x180_25fps.mkv (pinned by sha256 in
qa/corpus/manifest.toml:9-22) were produced by ragdoll's `synth` verb: `synth_video_frame`
content encoded by `transcode_file(.., "h264", qp 20)` — one H264Encoder per 15-frame GOP with
keyint 16 (crates/ragdoll-pipeline/src/lib.rs:575-590), Balanced, Constrained Baseline. Their
SPS, PPS and slice NAL units were extracted into tests/golden/*.nals by
tests/golden/extract_golden.py.
"""
import json
import os

import pytest

from oracle_lib import OracleEncoder, read_nals, split_annexb
from ragdoll_amd.synth import fixture_frame

HERE = os.path.dirname(os.path.abspath(__file__))
META = json.load(open(os.path.join(HERE, "golden", "golden.json")))


@pytest.mark.parametrize("name", sorted(META))
def test_oracle_reproduces_golden_clip(name):
    m = META[name]
    gold = read_nals(os.path.join(HERE, "golden", name + ".nals"))
    params, slices = gold[:m["n_param_sets"]], gold[m["n_param_sets"]:]
    assert len(slices) == m["frames"]
    for g0 in range(0, m["frames"], m["gop"]):
        enc = OracleEncoder(m["width"], m["height"], m["qp"], fps=(m["fps"], 1), keyint=m["gop"] + 1)
        for i in range(g0, min(g0 + m["gop"], m["frames"])):
            nals = split_annexb(enc.encode(*fixture_frame(m["width"], m["height"], i)))
            if i == g0:
                assert [n[0] & 0x1F for n in nals] == [7, 8, 5]
                assert nals[0] == params[0], "SPS differs"
                assert nals[1] == params[1], "PPS differs"
            else:
                assert [n[0] & 0x1F for n in nals] == [1]
            assert nals[-1] == slices[i], "%s frame %d: slice NAL differs (%d vs %d bytes)" % (
                name, i, len(nals[-1]), len(slices[i]))


def test_golden_sample_sizes_match_probe_snapshot():
    # crates/ragdoll-cli/tests/golden/probe_gop_map.txt lists the MP4 sample sizes.
    sizes = [277, 44, 52, 139, 66, 75, 76, 102, 89, 77, 100, 101, 90, 61, 74, 271, 43, 50, 58, 130, 71, 80, 68,
             114, 72, 81, 78, 110, 86, 85]
    assert META["clip_160x90_30fps"]["sample_sizes"] == sizes


def test_fixture_generator_matches_reference_y4m_sizes():
    # qa/corpus/synth/clip_64x36_30fps.y4m holds synth_video_frame(64, 36, 0..5); its frame
    # size pins the plane geometry (64*36 + 2*32*18 bytes).
    y, u, v = fixture_frame(64, 36, 3)
    assert y.shape == (36, 64) and u.shape == (18, 32) and v.shape == (18, 32)
