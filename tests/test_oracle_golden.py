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

import numpy as np
import pytest

from oracle_lib import OracleEncoder, read_nals, split_annexb
from ragdoll_amd.synth import fixture_frame

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))
CLIPS = sorted(k for k, m in META.items() if "qp" in m)  # the encoded clips (the y4m is raw content)


@pytest.mark.parametrize("name", CLIPS)
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


def read_y4m(path):
    """Frames (y, u, v) of a 4:2:0 YUV4MPEG2 file."""
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"\n")
    assert head.startswith(b"YUV4MPEG2 ")
    tags = {t[:1]: t[1:] for t in head.split()[1:]}
    w, h = int(tags[b"W"]), int(tags[b"H"])
    cw, ch = (w + 1) // 2, (h + 1) // 2
    frames = []
    while rest:
        fh, _, rest = rest.partition(b"\n")
        assert fh.startswith(b"FRAME")
        buf = np.frombuffer(rest[:w * h + 2 * cw * ch], np.uint8)
        rest = rest[w * h + 2 * cw * ch:]
        frames.append((buf[:w * h].reshape(h, w), buf[w * h:w * h + cw * ch].reshape(ch, cw),
                       buf[w * h + cw * ch:].reshape(ch, cw)))
    return frames


def test_fixture_generator_matches_reference_y4m_bytes():
    # qa/corpus/synth/clip_64x36_30fps.y4m holds synth_video_frame(64, 36, 0..5), the content of
    # the golden clips; the restated generator must give the same bytes.
    frames = read_y4m(os.path.join(GOLDEN, "clip_64x36_30fps.y4m"))
    assert len(frames) == META["clip_64x36_30fps"]["frames"] == 6
    for i, planes in enumerate(frames):
        for p, (want, got) in enumerate(zip(planes, fixture_frame(64, 36, i))):
            assert np.array_equal(want, got), "frame %d plane %d differs" % (i, p)
