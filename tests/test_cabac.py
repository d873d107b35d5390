"""CABAC (Main / High) entropy coding, pinned by an independent parser (tests/cabac_decoder.py,
written from ITU-T H.264 7.3 and 9.3).

* The parser reads every I and P slice of two libx264 CABAC streams that the reference's
  decoder tests hold (crates/ragdoll-codec/tests/data/h264) to their exact ends: all P
  partition shapes, up to 4 reference pictures, the 8x8 transform in I and P slices. That pins
  the arithmetic decoding engine, the context tables (m, n) of I slices and of
  cabac_init_idc 0, the binarizations and the ctxIdxInc derivations against an encoder written
  elsewhere; a wrong table entry makes the parse fail (shown below by planting one).
* The oracle's Main and High streams (IDR and P pictures of moving content over a range of
  QPs) parse to their ends with the standard's stop-bit position, and the parsed syntax equals
  the oracle's decision records and levels.
* The engine's tables satisfy the probability model of 9.3.1.1 / Table 9-44.
"""
import math

import numpy as np
import pytest

import cabac_decoder as cd
from oracle_lib import OracleEncoder
from ragdoll_amd.synth import CONFIGS, SceneGenerator

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")
X264 = ["cabac_temporal_direct_8x8_cqm.264", "cabac_bpyramid_weighted.264"]


def x264_stream(name):
    return open(GOLDEN + "/" + name, "rb").read()


@pytest.mark.parametrize("name", X264)
def test_parses_libx264_slices_to_their_ends(name):
    pics = cd.parse_stream(x264_stream(name), strict_end=False)
    assert len(pics) == 12
    coded = [p for p in pics if not p.get("skipped")]
    assert [p["type"] for p in coded] == [2, 0, 0, 0, 0]
    kinds = set()
    for p in coded:
        assert len(p["mbs"]) == 48
        kinds |= {m["type"] for m in p["mbs"]}
        for m in p["mbs"]:
            if "ref_idx" in m:
                assert all(0 <= r < p["headers"][0]["num_ref_l0"] for r in m["ref_idx"])
            for mode in m.get("intra_modes", []):
                assert 0 <= mode <= 8
    assert {"skip", "P16x16", "P16x8", "P8x16", "P8x8", "I_NxN", "I16"} <= kinds
    if "8x8" in name:
        assert any(m["t8"] for p in coded for m in p["mbs"] if m["type"] == "I_NxN")
        assert any(m["t8"] for p in coded for m in p["mbs"] if m["type"].startswith("P"))
    else:
        assert max(p["headers"][0]["num_ref_l0"] for p in coded) == 4


@pytest.mark.parametrize("table,ctx", [(0, 3), (0, 68), (0, 105), (0, 227), (1, 11), (1, 14), (1, 40), (1, 54),
                                       (1, 73), (1, 399), (1, 402), (1, 426)])
def test_a_wrong_context_initialisation_breaks_the_parse(table, ctx, monkeypatch):
    # One (m, n) pair moved by a few steps of probability: the parse of the libx264 streams no
    # longer ends where the slices end (or reads impossible syntax).
    mn = [list(t) for t in cd.INIT_MN]
    m, n = mn[table][ctx]
    mn[table][ctx] = (m, n + 12 if n < 100 else n - 12)
    monkeypatch.setattr(cd, "INIT_MN", mn)
    with pytest.raises((cd.ParseError, IndexError, KeyError)):
        for name in X264:
            cd.parse_stream(x264_stream(name), strict_end=False)


def test_a_wrong_8x8_significance_offset_breaks_the_parse(monkeypatch):
    sig = list(cd.SIG8)
    sig[5], sig[9] = sig[9], sig[5]
    monkeypatch.setattr(cd, "SIG8", sig)
    with pytest.raises((cd.ParseError, IndexError, KeyError)):
        cd.parse_stream(x264_stream(X264[0]), strict_end=False)


def test_range_table_follows_the_probability_model():
    # 9.3.1.1: pLPS(s) = 0.5 a^s with a = (0.01875 / 0.5)^(1/63); rangeTabLPS[s][q] is pLPS(s)
    # times the interval's representative range 288 + 64 q (Table 9-44), at most 128 for q = 0
    # (an LPS range stays below half of the smallest range, 256). transIdxMPS is s + 1 up to 62;
    # transIdxLPS is the state nearest to a pLPS + (1 - a).
    a = (0.01875 / 0.5) ** (1 / 63)
    for s in range(63):
        p = 0.5 * a ** s
        for q in range(4):
            want = min(p * (288 + 64 * q), 128 if q == 0 else 1e9)
            assert abs(cd.RANGE_LPS[s][q] - want) <= 1.5, (s, q, cd.RANGE_LPS[s][q], want)
        if s > 0:
            assert all(cd.RANGE_LPS[s][q] <= cd.RANGE_LPS[s - 1][q] for q in range(4))
        target = a * p + (1 - a)
        near = min(range(63), key=lambda k: abs(math.log(0.5 * a ** k) - math.log(target)))
        assert abs(cd.NEXT_LPS[s] - near) <= 1, (s, cd.NEXT_LPS[s], near)
    assert cd.RANGE_LPS[63] == [2, 2, 2, 2] and cd.NEXT_LPS[63] == 63  # the terminate state


# ---------------------------------------------------------------------------------------------
# The oracle's CABAC streams
# ---------------------------------------------------------------------------------------------
def level_record(rec):
    """The parsed levels of one macroblock in the records' layout (408 int16)."""
    lv = np.zeros(408, np.int16)
    if rec["type"] in ("skip",):
        return lv
    L = rec.get("levels")
    if not L:
        return lv
    if "luma8" in L:
        for q, blk in enumerate(L["luma8"]):
            lv[q * 64:(q + 1) * 64] = blk
    if "luma4" in L:
        for b, blk in enumerate(L["luma4"]):
            if len(blk) == 15:
                lv[b * 16 + 1:b * 16 + 16] = blk
            else:
                lv[b * 16:b * 16 + 16] = blk
    if "i16dc" in L:
        lv[256:272] = L["i16dc"]
    for p in range(2):
        lv[272 + 4 * p:276 + 4 * p] = L["cdc"][p]
        for k in range(4):
            o = 280 + (p * 4 + k) * 16
            lv[o + 1:o + 16] = L["cac"][p][k]
    return lv


def check_picture(pic, dec, lev, label):
    mbs = pic["mbs"]
    assert len(mbs) == len(dec), label
    for i, (m, d) in enumerate(zip(mbs, dec)):
        where = "%s MB %d" % (label, i)
        kind = {"skip": 0, "P16x16": 1, "I16": 2, "I_NxN": 3}[m["type"]]
        assert kind == d["mb_type"], where
        if kind == 0:
            continue
        assert m["dqp"] == 0, where
        assert m["cbp_luma"] == d["cbp_luma"] and m["cbp_chroma"] == d["cbp_chroma"], where
        assert m["t8"] == (1 if (d["transform_8x8"] and d["cbp_luma"]) else 0), where
        if kind == 1:
            assert m["mvd"] == [(int(d["mv"][0]) - int(d["mvp"][0]), int(d["mv"][1]) - int(d["mvp"][1]))], where
            assert m["ref_idx"] == [0], where
        else:
            assert m["chroma_mode"] == d["chroma_mode"], where
        if kind == 2:
            assert m["i16_mode"] == d["i16_mode"], where
        if kind == 3:
            packed = int(np.frombuffer(d.tobytes()[8:16], np.uint64)[0])
            assert m["intra_modes"] == [(packed >> (4 * b)) & 15 for b in range(16)], where
        assert np.array_equal(level_record(m), lev[i]), where + ": levels differ"


def scene(w, h, n, seed):
    kw = dict(CONFIGS["cif"])
    kw.pop("width"), kw.pop("height")
    g = SceneGenerator(w, h, seed=seed, **kw)
    return [g.next_frame() for _ in range(n)]


@pytest.mark.parametrize("profile", ["main", "high"])
@pytest.mark.parametrize("qp", [12, 24, 36, 48])
def test_oracle_cabac_stream_parses_to_its_records(profile, qp):
    w, h = 176, 144
    frames = scene(w, h, 4, 700 + qp)
    enc = OracleEncoder(w, h, qp, profile=profile, keyint=3)
    stream = b""
    records = []
    for f in frames:
        stream += enc.encode(*f)
        records.append((enc.decisions(), enc.levels()))
    pics = cd.parse_stream(stream, strict_end=True)
    assert [p["type"] for p in pics] == [2, 0, 0, 2]
    for k, (pic, (dec, lev)) in enumerate(zip(pics, records)):
        check_picture(pic, dec, lev, "%s qp %d frame %d" % (profile, qp, k))
    if profile == "high" and qp <= 24:
        assert any(d["transform_8x8"] and d["cbp_luma"] for dec, _ in records[1:3] for d in dec)


# ---------------------------------------------------------------------------------------------
# The partition extension (16x8 / 8x16 / 8x8 P partitions; records in the quadrant layout)
# ---------------------------------------------------------------------------------------------
PART_KIND = {"skip": 0, "P16x16": 1, "I16": 2, "I_NxN": 3, "P8x8": 5, "P16x8": 6, "P8x16": 7}


def quad_vectors(d):
    """The four quadrant vectors (raster order) of a partition-mode record (words 2-5)."""
    return np.frombuffer(d.tobytes()[8:24], np.int16).reshape(4, 2)


def check_parts_picture(pic, dec, lev, label):
    """Types, coded block patterns, transform flags and levels of every macroblock as parsed,
    and every P-picture vector as the parser derives it (8.4.1) from the parsed mvds."""
    mbs = pic["mbs"]
    assert len(mbs) == len(dec), label
    mvs = cd.derive_motion(pic) if pic["type"] == 0 else None
    w4 = 4 * pic["pic"].mb_w
    for i, (m, d) in enumerate(zip(mbs, dec)):
        where = "%s MB %d" % (label, i)
        kind = PART_KIND[m["type"]]
        assert kind == d["mb_type"], where
        if kind == 5:
            assert m["sub"] == [0, 0, 0, 0], where
        if mvs is not None and kind in (0, 1, 5, 6, 7):
            gx, gy = 4 * (i % pic["pic"].mb_w), 4 * (i // pic["pic"].mb_w)
            got = [mvs[(gy + 2 * (k >> 1)) * w4 + gx + 2 * (k & 1)] for k in range(4)]
            assert [tuple(int(c) for c in v) for v in quad_vectors(d)] == got, where
        if kind == 0:
            continue
        assert m["dqp"] == 0, where
        assert m["cbp_luma"] == d["cbp_luma"] and m["cbp_chroma"] == d["cbp_chroma"], where
        assert m["t8"] == (1 if (d["transform_8x8"] and d["cbp_luma"]) else 0), where
        if kind in (2, 3):
            assert m["chroma_mode"] == d["chroma_mode"], where
        if kind == 2:
            assert m["i16_mode"] == d["i16_mode"], where
        assert np.array_equal(level_record(m), lev[i]), where + ": levels differ"


def test_derive_motion_follows_the_standard_on_known_cases():
    # A hand-built picture of 2x2 macroblocks: the directional rules and the P_Skip zero rule.
    pic = {"type": 0, "pic": cd.Picture(2, 2), "mbs": [
        {"addr": 0, "type": "P16x8", "intra": False, "mvd": [(8, 4), (-3, 2)]},
        {"addr": 1, "type": "P8x16", "intra": False, "mvd": [(1, 1), (2, -2)]},
        {"addr": 2, "type": "skip", "intra": False},
        {"addr": 3, "type": "P8x8", "intra": False, "sub": [0, 0, 0, 0], "mvd": [(0, 0), (4, 0), (0, 4), (1, 1)]},
    ]}
    mv = cd.derive_motion(pic)
    at = lambda x, y: mv[y * 8 + x]
    # MB 0 upper half: no neighbours at all -> median of zeros.
    assert at(0, 0) == (8, 4)
    # MB 0 lower half: A unavailable, B = the upper half (C unavailable, D unavailable) -> only B
    # matches -> B = (8, 4).
    assert at(0, 2) == (5, 6)
    # MB 1 left half (8x16 part 0): A = MB 0's upper right block (8, 4), directional -> A.
    assert at(4, 0) == (9, 5)
    # MB 1 right half (8x16 part 1): C = (16, -1) outside, D = (3, -1) outside (picture top);
    # B outside; A = MB 1's left half (9, 5) -> B and C from A -> directional C = (9, 5).
    assert at(6, 0) == (11, 3)
    # MB 2 (skip): A unavailable (picture edge) -> zero vector.
    assert at(0, 4) == (0, 0)
    # MB 3 quadrant 0: A = MB 2 (0, 0), B = MB 1 lower-left (9, 5), C = MB 1 lower-right
    # block 2 (11, 3)? C = (16 + 8, 15) -> x 6 in 4x4 units at row 3: MB 1's right half (11, 3)
    # -> median((0,0), (9,5), (11,3)) = (9, 3).
    assert at(4, 4) == (9, 3)
    # quadrant 1: A = quadrant 0 (9, 3), B = MB 1 (11, 3), C outside the picture -> D = MB 1's
    # block at x 5 = (9, 5) -> median = (9, 3) -> + (4, 0).
    assert at(6, 4) == (13, 3)
    # quadrant 2: A = MB 2 (0, 0), B = quadrant 0 (9, 3), C = quadrant 1 (13, 3) -> (9, 3).
    assert at(4, 6) == (9, 7)
    # quadrant 3: A = quadrant 2 (9, 7), B = quadrant 1 (13, 3), C unavailable -> D = quadrant 0
    # (9, 3) -> median = (9, 3) -> + (1, 1).
    assert at(6, 6) == (10, 4)


@pytest.mark.parametrize("profile", ["main", "high"])
@pytest.mark.parametrize("qp", [14, 26, 38])
def test_oracle_partition_stream_parses_to_its_records(profile, qp):
    w, h = 208, 144
    frames = scene(w, h, 5, 1200 + qp)
    enc = OracleEncoder(w, h, qp, profile=profile, keyint=4, speed="partitions", me_range=32)
    stream = b""
    records = []
    for f in frames:
        stream += enc.encode(*f)
        records.append((enc.decisions(), enc.levels()))
    pics = cd.parse_stream(stream, strict_end=True)
    assert [p["type"] for p in pics] == [2, 0, 0, 0, 2]
    for k, (pic, (dec, lev)) in enumerate(zip(pics, records)):
        check_parts_picture(pic, dec, lev, "%s qp %d frame %d" % (profile, qp, k))
    kinds = set(int(t) for dec, _ in records[1:4] for t in dec["mb_type"])
    assert {5, 6, 7} <= kinds, kinds


# ---------------------------------------------------------------------------------------------
# The product's writer (rgh_write_p_slice_cabac / rgh_write_idr_slice_cabac) against the oracle
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("profile", ["main", "high"])
@pytest.mark.parametrize("qp", [10, 22, 34, 46])
def test_writer_matches_the_oracle(profile, qp):
    from ragdoll_amd import write_headers, write_idr_slice_cabac, write_p_slice_cabac
    w, h, keyint = 208, 112, 4
    frames = scene(w, h, 6, 800 + qp)
    enc = OracleEncoder(w, h, qp, profile=profile, keyint=keyint)
    for k, f in enumerate(frames):
        want = enc.encode(*f)
        dec, lev = enc.decisions(), enc.levels()
        gop = k % keyint
        if enc.last_was_idr:
            got = write_headers(w, h, qp, profile=profile) + write_idr_slice_cabac(
                dec, lev, enc.mb_w, enc.mb_h, qp, k // keyint, profile=profile)
        else:
            got = write_p_slice_cabac(dec, lev, enc.mb_w, enc.mb_h, qp, gop, (2 * gop) & 63, profile=profile)
        assert got == want, "%s qp %d frame %d: %d vs %d bytes" % (profile, qp, k, len(got), len(want))


@pytest.mark.parametrize("profile", ["main", "high"])
@pytest.mark.parametrize("qp", [6, 26, 40])
def test_fast_writer_matches_the_oracle(profile, qp):
    # Fast records (half-pel vectors, Intra 16x16 only) and slice headers with the loop filter off.
    from ragdoll_amd import write_headers, write_idr_slice_cabac, write_p_slice_cabac
    w, h, keyint = 176, 144, 3
    frames = scene(w, h, 5, 900 + qp)
    enc = OracleEncoder(w, h, qp, profile=profile, keyint=keyint, speed="fast")
    stream = b""
    for k, f in enumerate(frames):
        want = enc.encode(*f)
        stream += want
        dec, lev = enc.decisions(), enc.levels()
        assert not (dec["mb_type"] == 3).any()
        gop = k % keyint
        if enc.last_was_idr:
            got = write_headers(w, h, qp, profile=profile, speed="fast") + write_idr_slice_cabac(
                dec, lev, enc.mb_w, enc.mb_h, qp, k // keyint, profile=profile, speed="fast")
        else:
            got = write_p_slice_cabac(dec, lev, enc.mb_w, enc.mb_h, qp, gop, (2 * gop) & 63, profile=profile,
                                      speed="fast")
        assert got == want, "%s qp %d frame %d: %d vs %d bytes" % (profile, qp, k, len(got), len(want))
    pics = cd.parse_stream(stream, strict_end=True)
    assert len(pics) == len(frames)
    assert all(p["headers"][0]["disable_deblock"] == 1 for p in pics)


@pytest.mark.parametrize("profile", ["main", "high"])
@pytest.mark.parametrize("qp", [12, 26, 40])
def test_partition_writer_matches_the_oracle(profile, qp):
    from ragdoll_amd import write_headers, write_idr_slice_cabac, write_p_slice_cabac
    w, h, keyint = 208, 144, 4
    frames = scene(w, h, 6, 1300 + qp)
    enc = OracleEncoder(w, h, qp, profile=profile, keyint=keyint, speed="partitions", me_range=32)
    kinds = set()
    for k, f in enumerate(frames):
        want = enc.encode(*f)
        dec, lev = enc.decisions(), enc.levels()
        gop = k % keyint
        if enc.last_was_idr:
            got = write_headers(w, h, qp, profile=profile, speed="partitions") + write_idr_slice_cabac(
                dec, lev, enc.mb_w, enc.mb_h, qp, k // keyint, profile=profile, speed="partitions")
        else:
            got = write_p_slice_cabac(dec, lev, enc.mb_w, enc.mb_h, qp, gop, (2 * gop) & 63, profile=profile,
                                      speed="partitions")
            kinds |= set(int(t) for t in dec["mb_type"])
        assert got == want, "%s qp %d frame %d: %d vs %d bytes" % (profile, qp, k, len(got), len(want))
    if qp <= 26:
        assert {5, 6, 7} <= kinds, kinds


def test_writer_refuses_what_it_cannot_write():
    from ragdoll_amd import RghError, write_idr_slice_cabac, write_p_slice_cabac
    from ragdoll_amd.abi import DECISION_DTYPE, LEVELS_PER_MB
    dec = np.zeros(4, DECISION_DTYPE)
    lev = np.zeros((4, LEVELS_PER_MB), np.int16)
    dec["qp"] = 26
    assert write_p_slice_cabac(dec, lev, 2, 2, 26, 1, 2)  # all P_Skip
    with pytest.raises(RghError):
        write_p_slice_cabac(dec, lev, 2, 2, 26, 1, 2, profile="baseline")
    bad = dec.copy()
    bad["mb_type"][1] = 1
    bad["transform_8x8"][1] = 1
    bad["cbp_luma"][1] = 1
    with pytest.raises(RghError):  # the 8x8 transform needs High
        write_p_slice_cabac(bad, lev, 2, 2, 26, 1, 2, profile="main")
    bad = dec.copy()
    bad["mb_type"][2] = 4  # I_PCM
    with pytest.raises(RghError):
        write_p_slice_cabac(bad, lev, 2, 2, 26, 1, 2)
    with pytest.raises(RghError):  # an IDR picture holds intra records only
        write_idr_slice_cabac(dec, lev, 2, 2, 26, 0)
