"""A parser of H.264 CABAC slices written from ITU-T H.264 clauses 7.3 (syntax) and 9.3
(CABAC parsing: context initialisation 9.3.1.1, binarizations 9.3.2, ctxIdx derivation
9.3.3.1, the arithmetic decoding engine 9.3.3.2), independent of the encoders it checks.

Test infrastructure. It parses the I and P slices of progressive 4:2:0 streams (frame
macroblocks only, one slice group, no data partitioning) into syntax element values, and
checks that every slice ends exactly where its end_of_slice_flag says: the last bit the
arithmetic decoder reads for end_of_slice_flag = 1 is the rbsp_stop_one_bit, followed only by
alignment zero bits (and cabac_zero_words). B slices are recognised and skipped. The context
initialisation values (m, n) and the engine's rangeTabLPS / transIdxLPS tables are read from
oracle/h264_cabac_tables.h, the only thing shared with the encoders: parsing streams written by
another encoder (libx264, the reference's decoder fixtures) to their exact ends pins them.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_tables(path=os.path.join(ROOT, "oracle", "h264_cabac_tables.h")):
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)

    def body(name):
        i = text.index(name)
        j = text.index("};", i)
        return [int(v) for v in re.findall(r"-?\d+", text[text.index("=", i):j])]

    rng = body("CABAC_RANGE_LPS")
    nxt = body("CABAC_NEXT_LPS")
    mn = body("CABAC_INIT_MN")
    assert len(rng) == 256 and len(nxt) == 64 and len(mn) == 4 * 460 * 2
    range_lps = [rng[4 * s:4 * s + 4] for s in range(64)]
    init = [[(mn[(t * 460 + i) * 2], mn[(t * 460 + i) * 2 + 1]) for i in range(460)] for t in range(4)]
    return range_lps, nxt, init


RANGE_LPS, NEXT_LPS, INIT_MN = load_tables()

# Table 9-43, frame-coded 8x8 blocks: ctxIdxInc of significant_coeff_flag and
# last_significant_coeff_flag by scan position.
SIG8 = [0, 1, 2, 3, 4, 5, 5, 4, 4, 3, 3, 4, 4, 4, 5, 5, 4, 4, 4, 4, 3, 3, 6, 7, 7, 7, 8, 9, 10, 9, 8, 7,
        7, 6, 11, 12, 13, 11, 6, 7, 8, 9, 14, 10, 9, 8, 6, 11, 12, 13, 11, 6, 9, 14, 10, 9, 11, 12, 13, 11,
        14, 10, 12]
LAST8 = [0] + [1] * 15 + [2] * 16 + [3] * 8 + [4] * 8 + [5] * 4 + [6] * 4 + [7] * 4 + [8] * 3
assert len(LAST8) == 63

# ctxIdxOffset + ctxBlockCatOffset per ctxBlockCat (Tables 9-34 and 9-40), frame coding.
CBF_BASE = [85, 89, 93, 97, 101, None]
SIG_BASE = [105, 120, 134, 149, 152, 402]
LAST_BASE = [166, 181, 195, 210, 213, 417]
ABS_BASE = [227, 237, 247, 257, 266, 426]


class ParseError(Exception):
    pass


# ---------------------------------------------------------------------------------------------
# NAL units and RBSP bits
# ---------------------------------------------------------------------------------------------
def nal_units(data):
    """Annex B byte stream -> list of NAL unit payloads (header byte included)."""
    out, i, n = [], 0, len(data)
    starts = []
    while i + 3 <= n:
        if data[i] == 0 and data[i + 1] == 0 and data[i + 2] == 1:
            starts.append(i + 3)
            i += 3
        else:
            i += 1
    for k, s in enumerate(starts):
        e = starts[k + 1] - 3 if k + 1 < len(starts) else n
        while e > s and data[e - 1] == 0:  # trailing zero bytes belong to the next start code
            e -= 1
        out.append(bytes(data[s:e]))
    return out


def rbsp_of(nal):
    """The RBSP of a NAL unit (emulation prevention bytes removed), header byte dropped."""
    out = bytearray()
    zeros = 0
    for b in nal[1:]:
        if zeros >= 2 and b == 3:
            zeros = 0
            continue
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
    return bytes(out)


class Bits:
    def __init__(self, data):
        self.data = data
        self.pos = 0  # bit position
        self.n = len(data) * 8

    def u(self, k):
        v = 0
        for _ in range(k):
            if self.pos >= self.n:
                raise ParseError("read past the end of the RBSP")
            v = (v << 1) | ((self.data[self.pos >> 3] >> (7 - (self.pos & 7))) & 1)
            self.pos += 1
        return v

    def ue(self):
        z = 0
        while self.u(1) == 0:
            z += 1
            if z > 31:
                raise ParseError("ue(v) too long")
        return (1 << z) - 1 + self.u(z)

    def se(self):
        k = self.ue()
        return (k + 1) // 2 if k & 1 else -(k // 2)

    def more_rbsp_data(self):
        # Data remain when a 1 bit exists after the current position other than the last 1 bit
        # of the RBSP (the rbsp_stop_one_bit).
        last = self.n - 1
        while last >= 0 and not (self.data[last >> 3] >> (7 - (last & 7))) & 1:
            last -= 1
        return self.pos < last


# ---------------------------------------------------------------------------------------------
# Parameter sets and slice headers (7.3.2.1, 7.3.2.2, 7.3.3)
# ---------------------------------------------------------------------------------------------
def skip_scaling_list(b, size):
    last, nxt = 8, 8
    for _ in range(size):
        if nxt != 0:
            nxt = (last + b.se() + 256) % 256
        last = last if nxt == 0 else nxt


def parse_sps(r):
    b = Bits(r)
    s = {"profile_idc": b.u(8)}
    b.u(16)  # constraint flags, reserved, level_idc
    s["id"] = b.ue()
    s["chroma_format_idc"] = 1
    if s["profile_idc"] in (100, 110, 122, 244, 44, 83, 86, 118, 128, 138, 139, 134, 135):
        s["chroma_format_idc"] = b.ue()
        if s["chroma_format_idc"] == 3:
            b.u(1)
        b.ue()
        b.ue()
        b.u(1)
        if b.u(1):
            for i in range(8 if s["chroma_format_idc"] != 3 else 12):
                if b.u(1):
                    skip_scaling_list(b, 16 if i < 6 else 64)
    s["log2_max_frame_num"] = b.ue() + 4
    s["poc_type"] = b.ue()
    if s["poc_type"] == 0:
        s["log2_max_poc_lsb"] = b.ue() + 4
    elif s["poc_type"] == 1:
        s["delta_pic_order_always_zero"] = b.u(1)
        b.se()
        b.se()
        for _ in range(b.ue()):
            b.se()
    b.ue()  # max_num_ref_frames
    b.u(1)
    s["mb_w"] = b.ue() + 1
    map_h = b.ue() + 1
    s["frame_mbs_only"] = b.u(1)
    if not s["frame_mbs_only"]:
        s["mbaff"] = b.u(1)
    s["mb_h"] = map_h * (2 - s["frame_mbs_only"])
    s["direct_8x8_inference"] = b.u(1)
    return s


def parse_pps(r, spss):
    b = Bits(r)
    p = {"id": b.ue(), "sps_id": b.ue()}
    sps = spss[p["sps_id"]]
    p["cabac"] = b.u(1)
    p["bottom_field_pic_order"] = b.u(1)
    if b.ue() != 0:
        raise ParseError("slice groups are not parsed")
    p["num_ref_l0"] = b.ue() + 1
    p["num_ref_l1"] = b.ue() + 1
    p["weighted_pred"] = b.u(1)
    p["weighted_bipred"] = b.u(2)
    p["init_qp"] = 26 + b.se()
    b.se()
    p["cqo"] = b.se()
    p["deblock_ctrl"] = b.u(1)
    p["constrained_intra"] = b.u(1)
    p["redundant_pic_cnt"] = b.u(1)
    p["t8x8"] = 0
    if b.more_rbsp_data():
        p["t8x8"] = b.u(1)
        if b.u(1):
            for i in range(6 + (2 if sps["chroma_format_idc"] != 3 else 6) * p["t8x8"]):
                if b.u(1):
                    skip_scaling_list(b, 16 if i < 6 else 64)
        b.se()
    return p


def parse_slice_header(b, nal_type, ref_idc, spss, ppss):
    h = {"first_mb": b.ue(), "slice_type": b.ue() % 5}
    pps = ppss[b.ue()]
    sps = spss[pps["sps_id"]]
    h["pps"], h["sps"] = pps, sps
    h["frame_num"] = b.u(sps["log2_max_frame_num"])
    if not sps["frame_mbs_only"]:
        if b.u(1):
            raise ParseError("field pictures are not parsed")
    if sps.get("mbaff"):
        raise ParseError("MBAFF is not parsed")
    if nal_type == 5:
        h["idr_pic_id"] = b.ue()
    if sps["poc_type"] == 0:
        h["poc_lsb"] = b.u(sps["log2_max_poc_lsb"])
        if pps["bottom_field_pic_order"]:
            b.se()
    elif sps["poc_type"] == 1 and not sps["delta_pic_order_always_zero"]:
        b.se()
        if pps["bottom_field_pic_order"]:
            b.se()
    if pps["redundant_pic_cnt"]:
        b.ue()
    st = h["slice_type"]
    if st == 1:
        b.u(1)
    h["num_ref_l0"] = pps["num_ref_l0"]
    if st in (0, 1, 3):
        if b.u(1):
            h["num_ref_l0"] = b.ue() + 1
            if st == 1:
                b.ue()
    if st not in (2, 4):
        for _ in range(2 if st == 1 else 1):
            if b.u(1):
                while True:
                    idc = b.ue()
                    if idc == 3:
                        break
                    b.ue()
    if (pps["weighted_pred"] and st in (0, 3)) or (pps["weighted_bipred"] == 1 and st == 1):
        b.ue()
        if sps["chroma_format_idc"]:
            b.ue()
        for _ in range(2 if st == 1 else 1):
            for _ in range(h["num_ref_l0"]):
                if b.u(1):
                    b.se()
                    b.se()
                if sps["chroma_format_idc"] and b.u(1):
                    for _ in range(4):
                        b.se()
    if ref_idc:
        if nal_type == 5:
            b.u(2)
        elif b.u(1):
            while True:
                op = b.ue()
                if op == 0:
                    break
                if op in (1, 3):
                    b.ue()
                if op == 2:
                    b.ue()
                if op in (3, 6):
                    b.ue()
                if op == 4:
                    b.ue()
    h["cabac_init_idc"] = b.ue() if pps["cabac"] and st not in (2, 4) else 0
    h["qp"] = pps["init_qp"] + b.se()
    if st in (3, 4):
        raise ParseError("SP/SI slices are not parsed")
    h["disable_deblock"] = 0
    if pps["deblock_ctrl"]:
        h["disable_deblock"] = b.ue()
        if h["disable_deblock"] != 1:
            b.se()
            b.se()
    return h


# ---------------------------------------------------------------------------------------------
# The arithmetic decoding engine (9.3.1.2, 9.3.3.2)
# ---------------------------------------------------------------------------------------------
class Engine:
    def __init__(self, bits, slice_qp, table):
        self.b = bits
        q = min(max(slice_qp, 0), 51)
        self.state = []
        for m, n in INIT_MN[table]:
            pre = min(max(((m * q) >> 4) + n, 1), 126)
            self.state.append([63 - pre, 0] if pre <= 63 else [pre - 64, 1])
        self.start()

    def start(self):
        self.range = 510
        self.offset = self.b.u(9)
        if self.offset >= 510:
            raise ParseError("codIOffset 510 or 511 at initialisation")

    def decision(self, ctx):
        st = self.state[ctx]
        s, mps = st
        lps = RANGE_LPS[s][(self.range >> 6) & 3]
        self.range -= lps
        if self.offset >= self.range:
            bin_ = 1 - mps
            self.offset -= self.range
            self.range = lps
            if s == 0:
                st[1] = 1 - mps
            st[0] = NEXT_LPS[s]
        else:
            bin_ = mps
            st[0] = min(s + 1, 62) if s < 63 else 63
        while self.range < 256:
            self.range <<= 1
            self.offset = (self.offset << 1) | self.b.u(1)
        return bin_

    def bypass(self):
        self.offset = (self.offset << 1) | self.b.u(1)
        if self.offset >= self.range:
            self.offset -= self.range
            return 1
        return 0

    def terminate(self):
        self.range -= 2
        if self.offset >= self.range:
            return 1
        while self.range < 256:
            self.range <<= 1
            self.offset = (self.offset << 1) | self.b.u(1)
        return 0

    # Binarizations (9.3.2)
    def tu(self, ctxs, cmax):
        v = 0
        while v < cmax and self.decision(ctxs(v)):
            v += 1
        return v

    def eg(self, k):
        v = 0
        while self.bypass():
            v += 1 << k
            k += 1
            if k > 31:
                raise ParseError("Exp-Golomb suffix too long")
        while k:
            k -= 1
            v += self.bypass() << k
        return v


# ---------------------------------------------------------------------------------------------
# Slice data
# ---------------------------------------------------------------------------------------------
# Luma 4x4 block index -> (x, y) in 4x4 units inside the macroblock (6.4.3).
BLK_XY = [((b & 1) + ((b >> 1) & 2), ((b >> 1) & 1) + ((b >> 2) & 2)) for b in range(16)]

# P macroblock partitions (Table 7-13): (x, y, w, h) in samples.
P_PARTS = {0: [(0, 0, 16, 16)], 1: [(0, 0, 16, 8), (0, 8, 16, 8)], 2: [(0, 0, 8, 16), (8, 0, 8, 16)]}
# P sub-macroblock partitions (Table 7-17) relative to the sub-macroblock.
SUB_PARTS = {0: [(0, 0, 8, 8)], 1: [(0, 0, 8, 4), (0, 4, 8, 4)], 2: [(0, 0, 4, 8), (4, 0, 4, 8)],
             3: [(0, 0, 4, 4), (4, 0, 4, 4), (0, 4, 4, 4), (4, 4, 4, 4)]}


class Picture:
    """Per-picture state that neighbouring macroblocks read (one entry per macroblock, or per
    4x4 luma / chroma block on picture-wide grids)."""

    def __init__(self, mb_w, mb_h):
        n = mb_w * mb_h
        self.mb_w, self.mb_h = mb_w, mb_h
        self.slice = [-1] * n
        self.mb = [None] * n
        self.cbf4 = [0] * (n * 16)          # luma 4x4 coded_block_flag (8x8 blocks: inferred 1)
        self.ccbf = [[0] * (n * 4), [0] * (n * 4)]
        self.ref = [-1] * (n * 16)          # ref_idx_l0 per 4x4 block (-1 intra / none)
        self.mvd = [(0, 0)] * (n * 16)      # mvd_l0 per 4x4 block
        self.i4 = [-1] * (n * 16)           # Intra4x4 / Intra8x8 prediction modes (-1: not I_NxN)

    def avail(self, m, sl):
        return 0 <= m < len(self.slice) and self.slice[m] == sl


class SliceParser:
    def __init__(self, pic, hdr, bits, sl_id, strict_end=True):
        self.p, self.h, self.sl = pic, hdr, sl_id
        self.strict_end = strict_end
        self.pps = hdr["pps"]
        st = hdr["slice_type"]
        self.is_i = st == 2
        self.e = Engine(bits, hdr["qp"], 0 if self.is_i else 1 + hdr["cabac_init_idc"])
        self.prev_dqp_nz = False
        self.mbs = []

    # Neighbouring macroblocks A (left) and B (above) of macroblock m, or None.
    def nb(self, m):
        w = self.p.mb_w
        a = m - 1 if m % w and self.p.avail(m - 1, self.sl) else None
        b = m - w if m >= w and self.p.avail(m - w, self.sl) else None
        return a, b

    # The macroblock and the grid index of the 4x4 luma block covering luma sample (x, y) of the
    # picture, if it is available to the current slice.
    def blk_at(self, x, y):
        if x < 0 or y < 0 or x >= 16 * self.p.mb_w or y >= 16 * self.p.mb_h:
            return None, None
        m = (y >> 4) * self.p.mb_w + (x >> 4)
        if not self.p.avail(m, self.sl):
            return None, None
        return m, (y >> 2) * 4 * self.p.mb_w + (x >> 2)

    def parse(self):
        w = self.p.mb_w
        m = self.h["first_mb"]
        while True:
            if m >= w * self.p.mb_h:
                raise ParseError("slice runs past the picture")
            self.p.slice[m] = self.sl
            skip = False
            if not self.is_i:
                a, b = self.nb(m)
                inc = sum(1 for n in (a, b) if n is not None and self.p.mb[n]["type"] != "skip")
                skip = self.e.decision(11 + inc)
            if skip:
                rec = {"type": "skip", "t8": 0, "cbp_luma": 0, "cbp_chroma": 0, "intra": False, "i16": False}
                self.p.mb[m] = rec
                for i in self.blocks(m):
                    self.p.ref[i] = 0
                    self.p.mvd[i] = (0, 0)
                self.prev_dqp_nz = False
            else:
                rec = self.macroblock(m)
            rec["addr"] = m
            self.mbs.append(rec)
            if self.e.terminate():
                break
            m += 1
        # The encoding process of 9.3.4.5 (EncodeFlush) ends the arithmetic codeword so that the
        # last bit the decoder reads for end_of_slice_flag = 1 is a 1, the rbsp_stop_one_bit,
        # followed by alignment zeros (and cabac_zero_words). libx264 writes that 1 and then
        # its own stop bit and alignment: with strict_end False a single further 1 bit within
        # the next byte is accepted.
        bits = self.e.b
        if bits.pos < 1 or not (bits.data[(bits.pos - 1) >> 3] >> (7 - ((bits.pos - 1) & 7))) & 1:
            raise ParseError("the arithmetic codeword does not end in a 1 bit")
        end = bits.pos
        tail = []
        while bits.pos < bits.n and (bits.pos & 7 or not tail or bits.pos < end + 8):
            tail.append(bits.u(1))
            if bits.pos & 7 == 0 and (self.strict_end or 1 in tail):
                break
        if self.strict_end and any(tail):
            raise ParseError("non-zero alignment bits after rbsp_stop_one_bit")
        if not self.strict_end and tail.count(1) > 1:
            raise ParseError("more than one 1 bit after the arithmetic codeword")
        rest = bits.data[bits.pos >> 3:]
        if any(rest):  # only cabac_zero_words may follow
            raise ParseError("%d non-zero bytes after the slice data" % len(rest))
        return self.mbs

    def blocks(self, m):
        w4 = 4 * self.p.mb_w
        x0, y0 = 4 * (m % self.p.mb_w), 4 * (m // self.p.mb_w)
        return [(y0 + dy) * w4 + x0 + dx for dy in range(4) for dx in range(4)]

    # -- mb_type ---------------------------------------------------------------------------
    def intra_type(self, in_p, m):
        d = self.e.decision
        if in_p:
            b0 = d(17)
        else:
            a, b = self.nb(m)
            inc = sum(1 for n in (a, b) if n is not None and self.p.mb[n]["type"] != "I_NxN")
            b0 = d(3 + inc)
        if not b0:
            return 0
        if self.e.terminate():
            return 25
        o = [18, 19, 19, 20, 20] if in_p else [6, 7, 8, 9, 10]
        luma = d(o[0])
        cc = d(o[1])
        if cc:
            cc += d(o[2])
        mode = (d(o[3]) << 1) | d(o[4])
        return 1 + mode + 4 * cc + 12 * luma

    def p_type(self):
        d = self.e.decision
        if d(14):
            return 5 + self.intra_type(True, None)
        if d(15):
            return 1 if d(17) else 2
        return 3 if d(16) else 0

    # -- macroblock_layer (7.3.5) ----------------------------------------------------------------
    def macroblock(self, m):
        p, d = self.p, self.e.decision
        t = self.intra_type(False, m) if self.is_i else self.p_type()
        rec = {"t8": 0, "intra": False, "i16": False}
        p.mb[m] = rec
        blocks = self.blocks(m)
        it = t if self.is_i else t - 5
        if not self.is_i and t < 5:
            rec["type"] = ["P16x16", "P16x8", "P8x16", "P8x8"][t]
        elif it == 0:
            rec["type"] = "I_NxN"
        elif it == 25:
            rec["type"] = "PCM"
        else:
            rec["type"] = "I16"
            k = it - 1
            rec["i16_mode"], rec["cbp_chroma"], rec["cbp_luma"] = k & 3, (k >> 2) % 3, 15 if k >= 12 else 0
        rec["intra"] = rec["type"] in ("I_NxN", "I16", "PCM")
        rec["i16"] = rec["type"] == "I16"
        if rec["intra"]:
            for i in blocks:
                p.ref[i] = -1
                p.mvd[i] = (0, 0)
        if rec["type"] == "PCM":
            b = self.e.b
            while b.pos & 7:
                if b.u(1):
                    raise ParseError("pcm_alignment_zero_bit is 1")
            rec["pcm"] = [b.u(8) for _ in range(384)]
            self.e.start()
            rec.update(cbp_luma=15, cbp_chroma=2)
            for i in blocks:
                p.cbf4[i] = 1
            self.fill_chroma_cbf(m, 1)
            rec["dc_cbf"] = 1
            rec["cdc_cbf"] = [1, 1]
            rec["chroma_mode"] = 0
            self.prev_dqp_nz = False
            return rec
        no_sub_lt8 = True
        if rec["type"] == "P8x8":
            subs = [self.sub_type() for _ in range(4)]
            rec["sub"] = subs
            no_sub_lt8 = all(s == 0 for s in subs)
            self.inter_pred(m, rec, [(8 * (k & 1), 8 * (k >> 1), s) for k, s in enumerate(subs)])
        elif rec["type"].startswith("P"):
            self.inter_pred(m, rec, None)
        else:
            if rec["type"] == "I_NxN":
                if self.pps["t8x8"]:
                    rec["t8"] = self.t8_flag(m)
                n = 4 if rec["t8"] else 16
                rec["modes"] = [(1, 0) if d(68) else (0, d(69) | (d(69) << 1) | (d(69) << 2)) for _ in range(n)]
                self.derive_modes(m, rec)
            rec["chroma_mode"] = self.chroma_mode(m)
        if rec["type"] != "I16":
            rec["cbp_luma"], rec["cbp_chroma"] = self.cbp(m)
            if rec["cbp_luma"] and self.pps["t8x8"] and rec["type"] != "I_NxN" and no_sub_lt8:
                rec["t8"] = self.t8_flag(m)
        if rec["cbp_luma"] or rec["cbp_chroma"] or rec["i16"]:
            a = self.prev_dqp_nz
            v = self.e.tu(lambda i: 60 + a if i == 0 else (62 if i == 1 else 63), 1 << 30)
            rec["dqp"] = (v + 1) // 2 if v & 1 else -(v // 2)
            self.prev_dqp_nz = v != 0
            self.residual(m, rec)
        else:
            rec["dqp"] = 0
            self.prev_dqp_nz = False
            for i in blocks:
                p.cbf4[i] = 0
        return rec

    def t8_flag(self, m):
        a, b = self.nb(m)
        inc = sum(1 for n in (a, b) if n is not None and self.p.mb[n]["t8"])
        return self.e.decision(399 + inc)

    def sub_type(self):  # Table 9-38, P sub_mb_type, ctxIdx 21-23
        d = self.e.decision
        if d(21):
            return 0
        if not d(22):
            return 1
        return 2 if d(23) else 3

    def chroma_mode(self, m):
        a, b = self.nb(m)
        inc = sum(1 for n in (a, b) if n is not None and self.p.mb[n]["intra"] and self.p.mb[n]["type"] != "PCM"
                  and self.p.mb[n].get("chroma_mode", 0) != 0)
        d = self.e.decision
        if not d(64 + inc):
            return 0
        if not d(67):
            return 1
        return 3 if d(67) else 2

    def derive_modes(self, m, rec):
        # 8.3.1.1 / 8.3.2.1: the predicted mode is the smaller of the neighbours' (DC when a
        # neighbour is missing, 2 for a neighbour that is not I_NxN).
        p, w4 = self.p, 4 * self.p.mb_w
        x0, y0 = 16 * (m % p.mb_w), 16 * (m // p.mb_w)
        eight = rec["t8"]
        out = []
        for k, (flag, rem) in enumerate(rec["modes"]):
            if eight:
                bx, by = 8 * (k & 1), 8 * (k >> 1)
            else:
                bx, by = 4 * BLK_XY[k][0], 4 * BLK_XY[k][1]
            x, y = x0 + bx, y0 + by
            ma, ia = self.blk_at(x - 1, y)
            mb_, ib = self.blk_at(x, y - 1)
            if ma is None or mb_ is None:
                pred = 2
            else:
                def mode_of(mm, ii, is_a):
                    if p.i4[ii] < 0:
                        return 2
                    if eight and not p.mb[mm]["t8"] and mm != m:
                        # An I_4x4 neighbour of an 8x8 block: its 4x4 block 1 (left) or 2 (above)
                        # of the neighbouring 8x8 block.
                        xx, yy = (x - 1, y) if is_a else (x, y - 1)
                        xx, yy = (xx & ~7) + (4 if is_a else 0), (yy & ~7) + (0 if is_a else 4)
                        return p.i4[(yy >> 2) * w4 + (xx >> 2)]
                    return p.i4[ii]
                pred = min(mode_of(ma, ia, True), mode_of(mb_, ib, False))
            mode = pred if flag else (rem if rem < pred else rem + 1)
            out.append(mode)
            n = 2 if eight else 1
            for dy in range(n):
                for dx in range(n):
                    p.i4[((y >> 2) + dy) * w4 + (x >> 2) + dx] = mode
        rec["intra_modes"] = out

    # -- inter prediction syntax (mb_pred / sub_mb_pred) -----------------------------------------
    def inter_pred(self, m, rec, subs):
        p = self.p
        x0, y0 = 16 * (m % p.mb_w), 16 * (m // p.mb_w)
        w4 = 4 * p.mb_w
        parts = []
        if subs is None:
            parts = [[part] for part in P_PARTS[["P16x16", "P16x8", "P8x16"].index(rec["type"])]]
        else:
            for sx, sy, st in subs:
                parts.append([(sx + x, sy + y, ww, hh) for x, y, ww, hh in SUB_PARTS[st]])
        # ref_idx: one per partition (per sub-macroblock for P_8x8)
        refs = []
        for group in parts:
            gx, gy, gw, gh = (group[0][0], group[0][1], 8, 8) if subs is not None else group[0]
            r = 0
            if self.h["num_ref_l0"] > 1:
                inc = 0
                for k, (xx, yy) in enumerate(((x0 + gx - 1, y0 + gy), (x0 + gx, y0 + gy - 1))):
                    mm, ii = self.blk_at(xx, yy)
                    if mm is not None and p.mb[mm]["type"] != "skip" and not p.mb[mm]["intra"] and p.ref[ii] > 0:
                        inc += 1 << k
                d = self.e.decision
                if d(54 + inc):
                    r = 1
                    if d(58):
                        r = 2
                        while d(59):
                            r += 1
                            if r > 32:
                                raise ParseError("ref_idx out of range")
                if r >= self.h["num_ref_l0"]:
                    raise ParseError("ref_idx %d >= %d" % (r, self.h["num_ref_l0"]))
            refs.append(r)
            for yy in range(gy >> 2, (gy + gh) >> 2):
                for xx in range(gx >> 2, (gx + gw) >> 2):
                    p.ref[((y0 >> 2) + yy) * w4 + (x0 >> 2) + xx] = r
        rec["ref_idx"] = refs
        mvds = []
        for group in parts:
            for (px, py, pw, ph) in group:
                v = []
                for comp in range(2):
                    s = 0
                    for xx, yy in ((x0 + px - 1, y0 + py), (x0 + px, y0 + py - 1)):
                        mm, ii = self.blk_at(xx, yy)
                        if mm is not None and p.mb[mm]["type"] != "skip" and not p.mb[mm]["intra"]:
                            s += abs(p.mvd[ii][comp])
                    v.append(self.mvd(comp, s))
                mvds.append(tuple(v))
                for yy in range(py >> 2, (py + ph) >> 2):
                    for xx in range(px >> 2, (px + pw) >> 2):
                        p.mvd[((y0 >> 2) + yy) * w4 + (x0 >> 2) + xx] = tuple(v)
        rec["mvd"] = mvds

    def mvd(self, comp, s):
        base = 47 if comp else 40
        e = self.e
        if not e.decision(base + (0 if s < 3 else (1 if s <= 32 else 2))):
            return 0
        a = 1
        while a < 9 and e.decision(base + 3 + min(a - 1, 3)):
            a += 1
        if a >= 9:
            a = 9 + e.eg(3)
        return -a if e.bypass() else a

    # -- coded_block_pattern (9.3.2.6, 9.3.3.1.1.4) ---------------------------------------------
    def cbp(self, m):
        p, d = self.p, self.e.decision
        a, b = self.nb(m)
        luma = 0
        for q in range(4):
            bx, by = q & 1, q >> 1
            conds = []
            for is_a in (True, False):
                if (is_a and bx) or (not is_a and by):
                    bit = (luma >> (q - 1 if is_a else q - 2)) & 1
                    conds.append(0 if bit else 1)
                    continue
                n = a if is_a else b
                if n is None or p.mb[n]["type"] == "PCM":
                    conds.append(0)
                    continue
                if p.mb[n]["type"] == "skip":
                    conds.append(1)
                    continue
                qn = (by * 2 + 1) if is_a else (2 + bx)
                conds.append(0 if (p.mb[n]["cbp_luma"] >> qn) & 1 else 1)
            luma |= d(73 + conds[0] + 2 * conds[1]) << q

        def cc(n, want):
            if n is None or p.mb[n]["type"] == "skip":
                return 0
            if p.mb[n]["type"] == "PCM":
                return 1
            return 1 if (p.mb[n]["cbp_chroma"] != 0 if want == 0 else p.mb[n]["cbp_chroma"] == 2) else 0

        chroma = 0
        if d(77 + cc(a, 0) + 2 * cc(b, 0)):
            chroma = 1 + d(81 + cc(a, 1) + 2 * cc(b, 1))
        return luma, chroma

    # -- residual (7.3.5.3) ----------------------------------------------------------------------
    def cbf_cond(self, nm, trans_ok, flag, cur_intra):
        # 9.3.3.1.1.9: the neighbour's coded_block_flag, 0 / 1 when it has no such block.
        if nm is None:
            return 1 if cur_intra else 0
        t = self.p.mb[nm]["type"]
        if t == "PCM":
            return 1
        if not trans_ok:
            return 0
        return flag

    def luma_cbf_inc(self, m, x, y, intra):
        p = self.p
        inc = 0
        for k, (xx, yy) in enumerate(((x - 1, y), (x, y - 1))):
            mm, ii = self.blk_at(xx, yy)
            ok = False
            if mm is not None and p.mb[mm]["type"] not in ("skip", "PCM"):
                q = ((yy >> 3) & 1) * 2 + ((xx >> 3) & 1)
                ok = bool((p.mb[mm]["cbp_luma"] >> q) & 1)
            inc += self.cbf_cond(mm, ok, p.cbf4[ii] if ii is not None else 0, intra) << k
        return inc

    def block(self, cat, cbf_inc, n):
        """residual_block_cabac: the n levels of one block (scan order), or None if not coded."""
        d, e = self.e.decision, self.e
        if cat != 5:
            if not d(CBF_BASE[cat] + cbf_inc):
                return None
        sig = [0] * n
        last = n - 1
        i = 0
        while i < n - 1:
            si = SIG8[i] if cat == 5 else (min(i, 2) if cat == 3 else i)
            li = LAST8[i] if cat == 5 else (min(i, 2) if cat == 3 else i)
            if d(SIG_BASE[cat] + si):
                sig[i] = 1
                if d(LAST_BASE[cat] + li):
                    last = i
                    break
            i += 1
        sig[last] = 1
        lv = [0] * n
        eq1 = gt1 = 0
        gmax = 3 if cat == 3 else 4
        base = ABS_BASE[cat]
        for i in range(last, -1, -1):
            if not sig[i]:
                continue
            if d(base + (0 if gt1 else min(4, 1 + eq1))):
                a = 1
                c1 = base + 5 + min(gt1, gmax)
                while a < 14 and d(c1):
                    a += 1
                if a >= 14:
                    a = 14 + e.eg(0)
                gt1 += 1
            else:
                a = 0
                eq1 += 1
            lv[i] = -(a + 1) if e.bypass() else a + 1
        return lv

    def fill_chroma_cbf(self, m, v):
        w2 = 2 * self.p.mb_w
        x0, y0 = 2 * (m % self.p.mb_w), 2 * (m // self.p.mb_w)
        for pl in range(2):
            for k in range(4):
                self.p.ccbf[pl][(y0 + (k >> 1)) * w2 + x0 + (k & 1)] = v

    def residual(self, m, rec):
        p = self.p
        intra = rec["intra"]
        x0, y0 = 16 * (m % p.mb_w), 16 * (m // p.mb_w)
        w4 = 4 * p.mb_w
        a, b = self.nb(m)
        lev = {}
        if rec["i16"]:
            inc = 0
            for k, n in enumerate((a, b)):
                ok = n is not None and p.mb[n]["i16"]
                inc += self.cbf_cond(n, ok, p.mb[n].get("dc_cbf", 0) if ok else 0, True) << k
            dc = self.block(0, inc, 16)
            rec["dc_cbf"] = int(dc is not None)
            lev["i16dc"] = dc or [0] * 16
        for q in range(4):
            coded = (rec["cbp_luma"] >> q) & 1
            if rec["t8"]:
                blk = self.block(5, 0, 64) if coded else None
                lev.setdefault("luma8", [None] * 4)[q] = blk or [0] * 64
                for k in range(4):
                    bx, by = BLK_XY[4 * q + k]
                    p.cbf4[((y0 >> 2) + by) * w4 + (x0 >> 2) + bx] = coded
                continue
            for k in range(4):
                bi = 4 * q + k
                bx, by = BLK_XY[bi]
                gi = ((y0 >> 2) + by) * w4 + (x0 >> 2) + bx
                blk = None
                if coded:
                    inc = self.luma_cbf_inc(m, x0 + 4 * bx, y0 + 4 * by, intra)
                    blk = self.block(1 if rec["i16"] else 2, inc, 15 if rec["i16"] else 16)
                p.cbf4[gi] = int(blk is not None)
                lev.setdefault("luma4", [None] * 16)[bi] = blk or [0] * (15 if rec["i16"] else 16)
        rec["cdc_cbf"] = [0, 0]
        cdc = [[0] * 4, [0] * 4]
        cac = [[[0] * 15 for _ in range(4)] for _ in range(2)]
        if rec["cbp_chroma"]:
            for pl in range(2):
                inc = 0
                for k, n in enumerate((a, b)):
                    ok = n is not None and p.mb[n]["type"] not in ("skip", "PCM") and p.mb[n]["cbp_chroma"] != 0
                    inc += self.cbf_cond(n, ok, p.mb[n]["cdc_cbf"][pl] if ok else 0, intra) << k
                blk = self.block(3, inc, 4)
                rec["cdc_cbf"][pl] = int(blk is not None)
                cdc[pl] = blk or [0] * 4
        w2 = 2 * p.mb_w
        cx0, cy0 = 2 * (m % p.mb_w), 2 * (m // p.mb_w)
        for pl in range(2):
            for k in range(4):
                cx, cy = cx0 + (k & 1), cy0 + (k >> 1)
                f = 0
                if rec["cbp_chroma"] == 2:
                    inc = 0
                    for j, (xx, yy) in enumerate(((cx - 1, cy), (cx, cy - 1))):
                        nm = None
                        if xx >= 0 and yy >= 0:
                            mm = (yy >> 1) * p.mb_w + (xx >> 1)
                            nm = mm if p.avail(mm, self.sl) else None
                        ok = nm is not None and p.mb[nm]["type"] not in ("skip", "PCM") and p.mb[nm]["cbp_chroma"] == 2
                        inc += self.cbf_cond(nm, ok, p.ccbf[pl][yy * w2 + xx] if ok else 0, intra) << j
                    blk = self.block(4, inc, 15)
                    f = int(blk is not None)
                    cac[pl][k] = blk or [0] * 15
                p.ccbf[pl][cy * w2 + cx] = f
        lev["cdc"], lev["cac"] = cdc, cac
        rec["levels"] = lev


def parse_stream(data, stop_after=None, strict_end=True):
    """Parses every I and P slice of an Annex B stream. Returns a list of pictures, each a dict
    with the slice headers and the macroblock records (B pictures: 'skipped')."""
    spss, ppss = {}, {}
    pics = []
    cur = None
    for nal in nal_units(data):
        nt, ref_idc = nal[0] & 31, (nal[0] >> 5) & 3
        r = rbsp_of(nal)
        if nt == 7:
            s = parse_sps(r)
            spss[s["id"]] = s
        elif nt == 8:
            p = parse_pps(r, spss)
            ppss[p["id"]] = p
        elif nt in (1, 5):
            b = Bits(r)
            h = parse_slice_header(b, nt, ref_idc, spss, ppss)
            if h["first_mb"] == 0:
                cur = {"headers": [], "mbs": [], "type": h["slice_type"],
                       "pic": Picture(h["sps"]["mb_w"], h["sps"]["mb_h"])}
                pics.append(cur)
            cur["headers"].append(h)
            if h["slice_type"] == 1:
                cur["skipped"] = True
                continue
            if not h["pps"]["cabac"]:
                raise ParseError("a CAVLC slice")
            while b.pos & 7:  # cabac_alignment_one_bit
                if not b.u(1):
                    raise ParseError("cabac_alignment_one_bit is 0")
            sp = SliceParser(cur["pic"], h, b, len(cur["headers"]) - 1, strict_end)
            cur["mbs"] += sp.parse()
            if stop_after and len(pics) >= stop_after:
                break
    for pic in pics:
        if not pic.get("skipped") and len(pic["mbs"]) != pic["pic"].mb_w * pic["pic"].mb_h:
            raise ParseError("a picture's slices cover %d of %d macroblocks"
                             % (len(pic["mbs"]), pic["pic"].mb_w * pic["pic"].mb_h))
    return pics


def derive_motion(pic):
    """Luma motion vectors of every 4x4 block of a parsed P picture that uses one reference
    picture, from the decoded mvds (8.4.1): P_Skip by 8.4.1.1, every partition by 8.4.1.3 (the
    median rule, the directional rules of 16x8 and 8x16 partitions, B and C taken from A when only
    A is available), with the neighbouring partitions A, B, C (else D) of 6.4.11.7, a neighbour
    being available when it lies in the picture and is already decoded. Returns a list over the 4x4
    blocks in raster order: (mvx, mvy) in quarter samples, or None for intra blocks. Sub-macroblock
    partitions smaller than 8x8 are not derived (NotImplementedError)."""
    P = pic["pic"]
    w4, h4 = 4 * P.mb_w, 4 * P.mb_h
    mv = [None] * (w4 * h4)
    done = [False] * (w4 * h4)

    def nbr(x, y):  # (available, refIdx, mv)
        if x < 0 or y < 0 or x >= w4 or y >= h4 or not done[y * w4 + x]:
            return (False, -1, (0, 0))
        v = mv[y * w4 + x]
        return (True, -1, (0, 0)) if v is None else (True, 0, v)

    def med(a, b, c):
        return max(min(a, b), min(max(a, b), c))

    def mvp(x, y, pw, shape, part):
        A, B, C = nbr(x - 1, y), nbr(x, y - 1), nbr(x + pw, y - 1)
        if not C[0]:
            C = nbr(x - 1, y - 1)
        if not B[0] and not C[0] and A[0]:
            B = C = A
        if shape == "P16x8" and part == 0 and B[1] == 0:
            return B[2]
        if shape == "P16x8" and part == 1 and A[1] == 0:
            return A[2]
        if shape == "P8x16" and part == 0 and A[1] == 0:
            return A[2]
        if shape == "P8x16" and part == 1 and C[1] == 0:
            return C[2]
        hits = [n for n in (A, B, C) if n[1] == 0]
        if len(hits) == 1:
            return hits[0][2]
        return (med(A[2][0], B[2][0], C[2][0]), med(A[2][1], B[2][1], C[2][1]))

    def fill(x, y, bw, bh, v):
        for yy in range(y, y + bh):
            for xx in range(x, x + bw):
                mv[yy * w4 + xx] = v
                done[yy * w4 + xx] = True

    for rec in pic["mbs"]:
        m = rec["addr"]
        gx, gy = 4 * (m % P.mb_w), 4 * (m // P.mb_w)
        t = rec["type"]
        if rec["intra"]:
            fill(gx, gy, 4, 4, None)
        elif t == "skip":
            A, B = nbr(gx - 1, gy), nbr(gx, gy - 1)
            if not A[0] or not B[0] or (A[1] == 0 and A[2] == (0, 0)) or (B[1] == 0 and B[2] == (0, 0)):
                v = (0, 0)
            else:
                v = mvp(gx, gy, 4, "P16x16", 0)
            fill(gx, gy, 4, 4, v)
        else:
            if t == "P8x8":
                if any(s != 0 for s in rec["sub"]):
                    raise NotImplementedError("sub-macroblock partitions below 8x8")
                parts = [(8 * (k & 1), 8 * (k >> 1), 8, 8) for k in range(4)]
            else:
                parts = P_PARTS[["P16x16", "P16x8", "P8x16"].index(t)]
            for k, ((px, py, pw, ph), d) in enumerate(zip(parts, rec["mvd"])):
                p = mvp(gx + px // 4, gy + py // 4, pw // 4, t, k)
                fill(gx + px // 4, gy + py // 4, pw // 4, ph // 4, (p[0] + d[0], p[1] + d[1]))
    return mv
