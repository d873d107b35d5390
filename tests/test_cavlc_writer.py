"""The library's CAVLC writers (rgh_write_p_slice_cavlc, rgh_write_idr_slice_cavlc,
rgh_write_headers) against the reference's bytes.

Decisions and levels from the oracle are entropy-coded by the product writer; the slice NAL
units must equal the reference's committed golden clips (qa/corpus/synth/*, extracted into
tests/golden/*.nals) byte for byte, and the oracle's own slices on other content and QPs.
Runs on the CPU: the writer is host code of libragdoll_hip.so.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import OracleEncoder, read_nals, split_annexb
from ragdoll_amd import write_headers, write_idr_slice_cavlc, write_p_slice_cavlc
from ragdoll_amd.synth import CONFIGS, SceneGenerator, config_seed, fixture_frame

HERE = os.path.dirname(os.path.abspath(__file__))
META = {k: m for k, m in json.load(open(os.path.join(HERE, "golden", "golden.json"))).items()
        if "qp" in m}  # the encoded clips (the y4m entry is raw generator content)


def p_slice(ora, k):
    """The product writer's NAL for the oracle's last encoded picture, the k-th after its IDR."""
    nal = write_p_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, ora.qp, k & 15, (2 * k) & 63)
    units = split_annexb(nal)
    assert len(units) == 1
    return units[0]


@pytest.mark.parametrize("name", sorted(META))
def test_writer_reproduces_golden_p_slices(name):
    m = META[name]
    gold = read_nals(os.path.join(HERE, "golden", name + ".nals"))
    slices = gold[m["n_param_sets"]:]
    checked = 0
    for g0 in range(0, m["frames"], m["gop"]):
        ora = OracleEncoder(m["width"], m["height"], m["qp"], fps=(m["fps"], 1), keyint=m["gop"] + 1)
        for i in range(g0, min(g0 + m["gop"], m["frames"])):
            ora.encode(*fixture_frame(m["width"], m["height"], i))
            if i == g0:
                continue
            got = p_slice(ora, i - g0)
            assert got == slices[i], "%s frame %d: slice differs (%d vs %d bytes)" % (name, i, len(got), len(slices[i]))
            checked += 1
    assert checked == m["frames"] - (m["frames"] + m["gop"] - 1) // m["gop"]


@pytest.mark.parametrize("qp", [1, 10, 20, 28, 36, 51])
def test_writer_matches_oracle_slices(qp):
    # Synthetic CIF content with motion, skips, intra and large levels at low QP.
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(0, 100 + qp), **kw)
    ora = OracleEncoder(w, h, qp, keyint=1 << 20)
    types = set()
    for k in range(5):
        nals = split_annexb(ora.encode(*g.next_frame()))
        if k == 0:
            continue
        assert p_slice(ora, k) == nals[-1], "qp %d frame %d" % (qp, k)
        types |= set(np.unique(ora.decisions()["mb_type"]).tolist())
    assert 2 in types or qp >= 36


def test_writer_rejects_cabac_records():
    ora = OracleEncoder(64, 48, 26, profile="high", keyint=1 << 20)
    ora.encode(*fixture_frame(64, 48, 0))
    ora.encode(*fixture_frame(64, 48, 1))
    dec = ora.decisions().copy()
    dec["transform_8x8"][:] = 0
    dec["mb_type"][0] = 1
    dec["transform_8x8"][0] = 1
    from ragdoll_amd import RghError
    with pytest.raises(RghError):
        write_p_slice_cavlc(dec, ora.levels(), ora.mb_w, ora.mb_h, 26, 1, 2)


@pytest.mark.parametrize("name", sorted(META))
def test_writer_reproduces_golden_parameter_sets(name):
    m = META[name]
    gold = read_nals(os.path.join(HERE, "golden", name + ".nals"))
    units = split_annexb(write_headers(m["width"], m["height"], m["qp"], m["profile"], fps=(m["fps"], 1)))
    assert units == gold[:m["n_param_sets"]]


@pytest.mark.parametrize("name", sorted(META))
def test_writer_reproduces_golden_idr_slices(name):
    m = META[name]
    gold = read_nals(os.path.join(HERE, "golden", name + ".nals"))
    slices = gold[m["n_param_sets"]:]
    for g0 in range(0, m["frames"], m["gop"]):
        ora = OracleEncoder(m["width"], m["height"], m["qp"], fps=(m["fps"], 1), keyint=m["gop"] + 1)
        ora.encode(*fixture_frame(m["width"], m["height"], g0))
        assert ora.last_was_idr
        dec = ora.decisions()
        assert (dec["mb_type"] == 3).any(), "the clip's IDR codes intra 4x4 macroblocks"
        units = split_annexb(write_idr_slice_cavlc(dec, ora.levels(), ora.mb_w, ora.mb_h, ora.qp, 0))
        assert units == [slices[g0]], "%s frame %d" % (name, g0)


@pytest.mark.parametrize("qp", [1, 2, 12, 26, 40, 51])
@pytest.mark.parametrize("profile", ["baseline", "main", "high"])
def test_writer_matches_oracle_idr_slices_and_headers(qp, profile):
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    w, h = w - 6, h - 10  # cropping in the SPS
    g = SceneGenerator(w, h, seed=config_seed(1, 200 + qp), **kw)
    ora = OracleEncoder(w, h, qp, profile=profile, keyint=2, fps=(24000, 1001))
    for k in range(3):
        nals = split_annexb(ora.encode(*g.next_frame()))
        if k == 1:
            continue
        assert write_headers(w, h, qp, profile, fps=(24000, 1001)) == ora.headers()
        if profile != "baseline":
            continue  # CABAC slices: not written by this library
        assert split_annexb(ora.headers()) == nals[:2]
        got = split_annexb(write_idr_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, k // 2))
        assert got == nals[2:], "qp %d frame %d" % (qp, k)


@pytest.mark.parametrize("qp", [4, 14, 22, 28, 36, 46])
def test_writer_matches_oracle_compact_slices(qp):
    # The Compact preset's P slices (P_8x8 partitions with their own predictors, I_4x4 in P) and
    # parameter sets (chroma_qp_index_offset 1), from the oracle's records.
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(3, qp), **kw)
    ora = OracleEncoder(w, h, qp, keyint=1 << 20, speed="compact")
    types = set()
    for k in range(5):
        nals = split_annexb(ora.encode(*g.next_frame()))
        if k == 0:
            assert split_annexb(write_headers(w, h, qp, speed="compact")) == nals[:2]
            got = split_annexb(write_idr_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, 0))
            assert got == nals[2:]
            continue
        nal = write_p_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, k & 15, (2 * k) & 63,
                                  speed="compact")
        assert split_annexb(nal) == nals[-1:], "qp %d frame %d" % (qp, k)
        types |= set(np.unique(ora.decisions()["mb_type"]).tolist())
    assert 5 in types and 1 in types


@pytest.mark.parametrize("qp", [0, 8, 20, 30, 42])
def test_writer_matches_oracle_fast_slices(qp):
    # The Fast preset (enc.rs:503): half-pel vectors, Intra 16x16 only, the loop filter off in
    # every slice header (disable_deblocking_filter_idc 1, also at QP > 0).
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(4, qp), **kw)
    ora = OracleEncoder(w, h, qp, keyint=4, speed="fast")
    types = set()
    for k in range(6):
        nals = split_annexb(ora.encode(*g.next_frame()))
        if ora.last_was_idr:
            assert split_annexb(write_headers(w, h, qp, speed="fast")) == nals[:2]
            got = write_idr_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, k // 4, speed="fast")
            assert split_annexb(got) == nals[2:], "qp %d frame %d" % (qp, k)
            continue
        gop = k % 4
        nal = write_p_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, gop, (2 * gop) & 63,
                                  speed="fast")
        assert split_annexb(nal) == nals[-1:], "qp %d frame %d" % (qp, k)
        types |= set(np.unique(ora.decisions()["mb_type"]).tolist())
    assert 3 not in types  # no Intra 4x4 in Fast
    assert 1 in types


@pytest.mark.parametrize("qp", [10, 24, 34, 46])
def test_writer_matches_oracle_partition_slices(qp):
    # The partition extension (16x8 / 8x16 / 8x8 P partitions on the Balanced search): mb_type
    # ue(0..3), sub_mb_type ue(0) x 4, each partition's mvd against its directional / median
    # predictor, from the oracle's records in the quadrant layout.
    kw = dict(CONFIGS["cif"])
    w, h = kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=config_seed(5, qp), **kw)
    ora = OracleEncoder(w, h, qp, keyint=5, speed="partitions", me_range=32)
    types = set()
    for k in range(6):
        nals = split_annexb(ora.encode(*g.next_frame()))
        if ora.last_was_idr:
            assert split_annexb(write_headers(w, h, qp, speed="partitions")) == nals[:2]
            got = write_idr_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, k // 5,
                                        speed="partitions")
            assert split_annexb(got) == nals[2:], "qp %d frame %d" % (qp, k)
            continue
        gop = k % 5
        nal = write_p_slice_cavlc(ora.decisions(), ora.levels(), ora.mb_w, ora.mb_h, qp, gop, (2 * gop) & 63,
                                  speed="partitions")
        assert split_annexb(nal) == nals[-1:], "qp %d frame %d" % (qp, k)
        types |= set(np.unique(ora.decisions()["mb_type"]).tolist())
    if qp <= 34:
        assert {5, 6, 7} <= types, types
