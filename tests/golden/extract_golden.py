"""Extracts the reference's committed golden H.264 clips into small fixtures.

The ragdoll repository keeps two encoder outputs under qa/corpus/synth/ (pinned by sha256 in
qa/corpus/manifest.toml:9-22): clip_160x90_30fps.mp4 (30 frames, two 15-frame GOPs) and
clip_320x180_25fps.mkv (15 frames). Both are `synth_video_frame` content
(crates/ragdoll-pipeline/src/lib.rs:2919-2952) encoded by `transcode_file(.., "h264", 20, 0)`:
Balanced, Constrained Baseline CAVLC, QP 20.

This script reads those files (run it where the reference checkout is present) and writes, per
clip, `<name>.nals`: the SPS, the PPS and every slice NAL unit in decoding order, each as a
4-byte big-endian length followed by the escaped NAL unit, plus `golden.json` with the clip
parameters. The tests compare the oracle's output against these bytes.

It also copies qa/corpus/synth/clip_64x36_30fps.y4m (manifest.toml:23-29; raw
synth_video_frame(64, 36, 0..5), the reference's content generator) as `clip_64x36_30fps.y4m`,
which pins the test-content generator byte for byte, and two libx264 CABAC streams of the
reference's decoder fixtures (crates/ragdoll-codec/tests/data/h264) for the CABAC parser.
"""
import hashlib
import json
import os
import struct
import sys

REF = os.environ.get("RAGDOLL_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def mp4_boxes(b, off, end):
    while off < end:
        size, typ = struct.unpack(">I4s", b[off:off + 8])
        hdr = 8
        if size == 1:
            size = struct.unpack(">Q", b[off + 8:off + 16])[0]
            hdr = 16
        elif size == 0:
            size = end - off
        yield typ.decode("latin1"), off + hdr, off + size
        off += size


def mp4_find(b, off, end, path):
    containers = {"moov", "trak", "mdia", "minf", "stbl"}
    for typ, s, e in mp4_boxes(b, off, end):
        if typ == path[0]:
            if len(path) == 1:
                return s, e
            if typ in containers:
                return mp4_find(b, s, e, path[1:])
            if typ == "stsd":
                return mp4_find(b, s + 8, e, path[1:])
            if typ == "avc1":
                return mp4_find(b, s + 78, e, path[1:])
    raise KeyError(path)


def parse_avcc(rec):
    """Returns (sps_list, pps_list, nal_length_size) of an AVCDecoderConfigurationRecord."""
    length_size = (rec[4] & 3) + 1
    n_sps = rec[5] & 0x1F
    off = 6
    sps, pps = [], []
    for _ in range(n_sps):
        n = struct.unpack(">H", rec[off:off + 2])[0]
        sps.append(rec[off + 2:off + 2 + n])
        off += 2 + n
    n_pps = rec[off]
    off += 1
    for _ in range(n_pps):
        n = struct.unpack(">H", rec[off:off + 2])[0]
        pps.append(rec[off + 2:off + 2 + n])
        off += 2 + n
    return sps, pps, length_size


def split_avcc_sample(sample, length_size):
    out, off = [], 0
    while off < len(sample):
        n = int.from_bytes(sample[off:off + length_size], "big")
        out.append(sample[off + length_size:off + length_size + n])
        off += length_size + n
    return out


def mp4_extract(path):
    b = open(path, "rb").read()
    trak = ["moov", "trak", "mdia", "minf", "stbl"]
    s, e = mp4_find(b, 0, len(b), trak + ["stsd", "avc1", "avcC"])
    sps, pps, ls = parse_avcc(b[s:e])
    s, e = mp4_find(b, 0, len(b), trak + ["stsz"])
    sample_size, count = struct.unpack(">II", b[s + 4:s + 12])
    sizes = [sample_size] * count if sample_size else list(
        struct.unpack(">%dI" % count, b[s + 12:s + 12 + 4 * count]))
    s, e = mp4_find(b, 0, len(b), trak + ["stco"])
    n_chunks = struct.unpack(">I", b[s + 4:s + 8])[0]
    offsets = struct.unpack(">%dI" % n_chunks, b[s + 8:s + 8 + 4 * n_chunks])
    s, e = mp4_find(b, 0, len(b), trak + ["stsc"])
    n_ent = struct.unpack(">I", b[s + 4:s + 8])[0]
    stsc = [struct.unpack(">III", b[s + 8 + 12 * i:s + 20 + 12 * i]) for i in range(n_ent)]
    samples, k = [], 0
    for ci, coff in enumerate(offsets):
        per = [spc for (first, spc, _) in stsc if first <= ci + 1][-1]
        for _ in range(per):
            samples.append(b[coff:coff + sizes[k]])
            coff += sizes[k]
            k += 1
    nals = []
    for smp in samples:
        nals += split_avcc_sample(smp, ls)
    return sps, pps, nals, sizes


def ebml_vint(b, off, strip=True):
    first = b[off]
    length = 1
    mask = 0x80
    while not (first & mask):
        mask >>= 1
        length += 1
    val = first & (mask - 1) if strip else first
    for i in range(1, length):
        val = (val << 8) | b[off + i]
    return val, length


def ebml_elements(b, off, end):
    while off < end:
        eid, n1 = ebml_vint(b, off, strip=False)
        size, n2 = ebml_vint(b, off + n1)
        s = off + n1 + n2
        unknown = size == (1 << (7 * n2)) - 1
        e = end if unknown else s + size
        yield eid, s, e
        off = e


def mkv_extract(path):
    b = open(path, "rb").read()
    SEGMENT, TRACKS, TRACKENTRY, CODECPRIV, CLUSTER = 0x18538067, 0x1654AE6B, 0xAE, 0x63A2, 0x1F43B675
    SIMPLEBLOCK, BLOCKGROUP, BLOCK = 0xA3, 0xA0, 0xA1
    codec_private, blocks = None, []

    def walk(off, end):
        nonlocal codec_private
        for eid, s, e in ebml_elements(b, off, end):
            if eid in (SEGMENT, TRACKS, TRACKENTRY, CLUSTER, BLOCKGROUP):
                walk(s, e)
            elif eid == CODECPRIV:
                codec_private = b[s:e]
            elif eid in (SIMPLEBLOCK, BLOCK):
                _, n = ebml_vint(b, s)
                blocks.append(b[s + n + 3:e])
    walk(0, len(b))
    sps, pps, ls = parse_avcc(codec_private)
    nals = []
    for blk in blocks:
        nals += split_avcc_sample(blk, ls)
    return sps, pps, nals, [len(x) for x in blocks]


CLIPS = [
    # name, file, width, height, fps, frames, gop (frames per independent encoder)
    ("clip_160x90_30fps", "clip_160x90_30fps.mp4", 160, 90, 30, 30, 15),
    ("clip_320x180_25fps", "clip_320x180_25fps.mkv", 320, 180, 25, 15, 15),
]


def main():
    meta = {}
    for name, fname, w, h, fps, frames, gop in CLIPS:
        path = os.path.join(REF, "qa", "corpus", "synth", fname)
        data = open(path, "rb").read()
        sps, pps, nals, sizes = (mp4_extract if fname.endswith(".mp4") else mkv_extract)(path)
        assert len(nals) == frames, (name, len(nals))
        with open(os.path.join(HERE, name + ".nals"), "wb") as f:
            for nal in sps + pps + nals:
                f.write(struct.pack(">I", len(nal)))
                f.write(nal)
        meta[name] = {
            "source": "qa/corpus/synth/" + fname,
            "source_sha256": hashlib.sha256(data).hexdigest(),
            "width": w, "height": h, "fps": fps, "frames": frames, "gop": gop,
            "qp": 20, "profile": "baseline",
            "n_param_sets": len(sps) + len(pps),
            "sample_sizes": sizes,
        }
        print(name, "frames", len(nals), "bytes", sum(len(n) for n in nals))
    y4m = open(os.path.join(REF, "qa", "corpus", "synth", "clip_64x36_30fps.y4m"), "rb").read()
    with open(os.path.join(HERE, "clip_64x36_30fps.y4m"), "wb") as f:
        f.write(y4m)
    meta["clip_64x36_30fps"] = {"source": "qa/corpus/synth/clip_64x36_30fps.y4m",
                                "source_sha256": hashlib.sha256(y4m).hexdigest(), "width": 64, "height": 36,
                                "frames": 6}
    # Two progressive CABAC streams written by libx264 that the reference's decoder tests hold
    # (crates/ragdoll-codec/tests/h264_reference.rs:11-30, data in tests/data/h264): an
    # independent encoder's I and P slices for the CABAC parser of the tests.
    for fname, note in (("cabac_bpyramid_weighted.264",
                         "Main, CABAC, 3 references, B pyramid, explicit weighted prediction, 128x96"),
                        ("cabac_temporal_direct_8x8_cqm.264",
                         "High, CABAC, 8x8 transform, default scaling matrices, temporal direct, 128x96")):
        data = open(os.path.join(REF, "crates", "ragdoll-codec", "tests", "data", "h264", fname), "rb").read()
        with open(os.path.join(HERE, fname), "wb") as f:
            f.write(data)
        meta[fname[:-4]] = {"source": "crates/ragdoll-codec/tests/data/h264/" + fname,
                            "source_sha256": hashlib.sha256(data).hexdigest(), "encoder": "libx264",
                            "note": note, "width": 128, "height": 96, "frames": 12}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
