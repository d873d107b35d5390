// Entropy coding of analysed pictures with CABAC: the Main / High slices a ragdoll encoder writes
// for the decisions and levels the GPU analysis returns. Host code.
//
// Restates the slice header of SliceHeader::write (crates/ragdoll-codec/src/h264.rs:1071-1127,
// cabac_init_idc 0), the cabac_alignment_one_bit and slice data of encode_frame_body
// (enc.rs:2346-2440: end_of_slice_flag after every macroblock), the per-macroblock syntax of
// begin_coded_mb / code_skip / code_inter / code_i4 / code_i16 (enc.rs:2779-3306, :3441-3720),
// encode_cabac_cbp_and_residuals and encode_cabac_chroma_residuals (h264.rs:3327-3552), and the
// CabacEncoder engine with its binarizations (cabac.rs:504-1130), as ITU-T H.264 9.3 defines
// them.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/ragdoll_hip.h"
#include "rgh_bits.h"
#include "rgh_cabac_tables.h"
#include "rgh_motion.h"

namespace {

using namespace rgh_bits;
using namespace rgh_cabac;

int absi(int v) { return v < 0 ? -v : v; }
int mini(int a, int b) { return a < b ? a : b; }

// The arithmetic coder of 9.3.4.2, byte-wise. `low` holds the 10-bit codILow window and, above
// it, the bits shifted out of the window that are not written yet; `pend` + 8 counts them, the
// top one being the slot a carry out of them lands in. A byte leaves once its 8 bits and the
// carry slot above them are there: the carry adds one to the byte written before it. A byte of
// all ones stays back (a later carry would turn it to 0x00 and reach the byte before it) until
// the next byte settles the run. The first bit of the stream (suppressed by 9.3.4.2's
// firstBitFlag) is the first byte's carry slot.
class Arith {
  public:
    Arith(std::vector<uint8_t> &out, int slice_qp, int table) : out_(out), base_(out.size()) {
        const int q = slice_qp < 0 ? 0 : (slice_qp > 51 ? 51 : slice_qp);
        for (int i = 0; i < 460; i++) {
            int pre = ((kInitMn[table][i][0] * q) >> 4) + kInitMn[table][i][1];
            pre = pre < 1 ? 1 : (pre > 126 ? 126 : pre);
            ctx_[i] = (uint8_t)(pre <= 63 ? (63 - pre) << 1 : ((pre - 64) << 1) | 1);
        }
    }

    void bin(int c, int b) {  // EncodeDecision
        const int s = ctx_[c] >> 1, mps = ctx_[c] & 1;
        const uint32_t r_lps = kRangeLps[s][(range_ >> 6) & 3];
        range_ -= r_lps;
        if (b != mps) {
            low_ += range_;
            range_ = r_lps;
            ctx_[c] = (uint8_t)((kNextLps[s] << 1) | (s == 0 ? 1 - mps : mps));
        } else {
            ctx_[c] = (uint8_t)(((s < 62 ? s + 1 : s) << 1) | mps);
        }
        renorm();
    }
    void bypass(int b) {  // EncodeBypass
        low_ = (low_ << 1) + (b ? range_ : 0);
        pend_++;
        drain();
    }
    void terminate(int b) {  // EncodeTerminate; b = 1 also flushes (EncodeFlush)
        range_ -= 2;
        if (!b) {
            renorm();
            return;
        }
        low_ += range_;
        // EncodeFlush: bits 9..1 of codILow, then a 1 in place of bit 0 (rbsp_stop_one_bit),
        // then zero bits to the byte boundary.
        low_ = (low_ | 1u) << 10;
        pend_ += 10;
        drain();
        if (pend_ > -8) {
            low_ <<= -pend_;
            pend_ = 0;
            drain();
        }
        for (; held_; held_--) out_.push_back(0xFF);
    }

    // Binarization helpers (9.3.2).
    void eg_suffix(uint32_t v, int k) {  // UEGk suffix
        while (v >= (1u << k)) {
            bypass(1);
            v -= 1u << k;
            k++;
        }
        bypass(0);
        while (k--) bypass((int)((v >> k) & 1));
    }

  private:
    void renorm() {
        const int s = __builtin_clz(range_) - 23;
        range_ <<= s;
        low_ <<= s;
        pend_ += s;
        drain();
    }
    void drain() {
        while (pend_ >= 0) {
            const int sh = pend_ + 10;
            const uint32_t v = low_ >> sh;  // carry slot and one byte
            low_ &= (1u << sh) - 1;
            pend_ -= 8;
            const uint32_t byte = v & 0xFF, carry = v >> 8;
            if (byte == 0xFF) {  // no carry can come with it: the bits below a carry are zero
                held_++;
                continue;
            }
            if (carry && out_.size() > base_) out_.back()++;
            for (; held_; held_--) out_.push_back(carry ? 0x00 : 0xFF);
            out_.push_back((uint8_t)byte);
        }
    }

    std::vector<uint8_t> &out_;
    const size_t base_;  // the first byte of the arithmetic codeword
    uint32_t low_ = 0, range_ = 510;
    int pend_ = -9;
    uint32_t held_ = 0;
    uint8_t ctx_[460];  // pStateIdx << 1 | valMPS
};

// Level record offsets (RGH_LEVELS_PER_MB layout).
constexpr int kLuma = 0, kLumaDc = 256, kChromaDc = 272, kChromaAc = 280;

// ctxIdxOffset + ctxBlockCatOffset per ctxBlockCat, frame coding (Tables 9-34, 9-40).
constexpr int kCbf[5] = {85, 89, 93, 97, 101};
constexpr int kSig[6] = {105, 120, 134, 149, 152, 402};
constexpr int kLast[6] = {166, 181, 195, 210, 213, 417};
constexpr int kAbs[6] = {227, 237, 247, 257, 266, 426};
// Table 9-43, frame-coded 8x8 blocks: significant_coeff_flag and last_significant_coeff_flag
// ctxIdxInc by scan position.
constexpr uint8_t kSig8[63] = {0, 1,  2,  3,  4,  5,  5,  4,  4,  3,  3, 4,  4,  4,  5,  5,  4,  4,  4,  4,  3,
                               3, 6,  7,  7,  7,  8,  9,  10, 9,  8,  7, 7,  6,  11, 12, 13, 11, 6,  7,  8,  9,
                               14, 10, 9, 8, 6,  11, 12, 13, 11, 6, 9, 14, 10, 9,  11, 12, 13, 11, 14, 10, 12};
constexpr uint8_t kLast8[63] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2,
                                2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 4, 4,
                                4, 4, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 8, 8, 8};

// One residual block (9.3.2.3 / 9.3.3.1.3): coded_block_flag (none for 8x8 blocks), the
// significance map, then coeff_abs_level_minus1 / sign from the last level down. Returns the
// coded_block_flag.
int residual_block(Arith &a, int cat, int cbf_inc, const int16_t *sc, int n) {
    int last = n - 1;
    while (last >= 0 && sc[last] == 0) last--;
    if (cat < 5) {
        a.bin(kCbf[cat] + cbf_inc, last >= 0);
        if (last < 0) return 0;
    }
    for (int i = 0; i <= last && i < n - 1; i++) {
        const int inc_s = cat == 5 ? kSig8[i] : (cat == 3 ? mini(i, 2) : i);
        const int inc_l = cat == 5 ? kLast8[i] : (cat == 3 ? mini(i, 2) : i);
        a.bin(kSig[cat] + inc_s, sc[i] != 0);
        if (sc[i]) a.bin(kLast[cat] + inc_l, i == last);
    }
    int ones = 0, bigger = 0;
    const int cap = cat == 3 ? 3 : 4, base = kAbs[cat];
    for (int i = last; i >= 0; i--) {
        const int v = sc[i];
        if (!v) continue;
        const int m1 = absi(v) - 1;
        a.bin(base + (bigger ? 0 : mini(4, 1 + ones)), m1 > 0);
        if (m1 > 0) {
            const int c = base + 5 + mini(bigger, cap);
            const int prefix = mini(m1, 14);
            for (int k = 1; k < prefix; k++) a.bin(c, 1);
            if (m1 < 14)
                a.bin(c, 0);
            else
                a.eg_suffix((uint32_t)(m1 - 14), 0);
            bigger++;
        } else {
            ones++;
        }
        a.bypass(v < 0);
    }
    return 1;
}

// The picture's coded state that later macroblocks' contexts read (9.3.3.1.1).
struct Neighbours {
    int mw, mh;
    std::vector<uint8_t> skip, intra, i16, cbpl, cbpc, cmode, t8, dc_cbf, cdc_cbf[2];
    std::vector<uint8_t> cbf4;     // luma 4x4 coded_block_flag, (4 mw) x (4 mh); 8x8 blocks fill four
    std::vector<uint8_t> ccbf[2];  // chroma AC coded_block_flag, (2 mw) x (2 mh)
    std::vector<uint8_t> modes;    // intra 4x4 prediction modes, DC outside I_4x4 macroblocks
    std::vector<int16_t> mvd;      // mvd (x, y) per 4x4 block
    Neighbours(int w, int h)
        : mw(w), mh(h), skip((size_t)w * h), intra((size_t)w * h), i16((size_t)w * h), cbpl((size_t)w * h),
          cbpc((size_t)w * h), cmode((size_t)w * h), t8((size_t)w * h), dc_cbf((size_t)w * h),
          cdc_cbf{std::vector<uint8_t>((size_t)w * h), std::vector<uint8_t>((size_t)w * h)},
          cbf4((size_t)w * h * 16), ccbf{std::vector<uint8_t>((size_t)w * h * 4), std::vector<uint8_t>((size_t)w * h * 4)},
          modes((size_t)w * h * 16, 2), mvd((size_t)w * h * 32, 0) {}

    // coded_block_flag condition of the luma 4x4 block at grid (gx, gy) for a block of an intra
    // (outside the picture: 1) or inter (0) macroblock: a skipped macroblock or an uncoded 8x8
    // quadrant gives 0, otherwise the block's flag (an 8x8-transform quadrant counts 1).
    int luma(int gx, int gy, int cur_intra) const {
        if (gx < 0 || gy < 0) return cur_intra;
        const int m = (gy >> 2) * mw + (gx >> 2);
        if (skip[(size_t)m]) return 0;
        const int q = ((gy >> 1) & 1) * 2 + ((gx >> 1) & 1);
        if (!((cbpl[(size_t)m] >> q) & 1)) return 0;
        return cbf4[(size_t)gy * 4 * mw + gx];
    }
    int chroma_ac(int p, int cx, int cy, int cur_intra) const {
        if (cx < 0 || cy < 0) return cur_intra;
        const int m = (cy >> 1) * mw + (cx >> 1);
        if (skip[(size_t)m] || cbpc[(size_t)m] != 2) return 0;
        return ccbf[p][(size_t)cy * 2 * mw + cx];
    }
};

int blk_x(int b) { return (b & 1) + ((b >> 1) & 2); }
int blk_y(int b) { return ((b >> 1) & 1) + ((b >> 2) & 2); }

// Chroma DC and AC blocks of a macroblock (h264.rs:3478 encode_cabac_chroma_residuals).
void chroma_residual(Arith &a, Neighbours &nb, int mb_x, int mb_y, int intra, int cbpc, const int16_t *lv) {
    const int m = mb_y * nb.mw + mb_x;
    if (cbpc >= 1)
        for (int p = 0; p < 2; p++) {
            const int ca = mb_x > 0 ? (!nb.skip[(size_t)m - 1] && nb.cdc_cbf[p][(size_t)m - 1]) : intra;
            const int cb = mb_y > 0 ? (!nb.skip[(size_t)(m - nb.mw)] && nb.cdc_cbf[p][(size_t)(m - nb.mw)]) : intra;
            nb.cdc_cbf[p][(size_t)m] = (uint8_t)residual_block(a, 3, ca + 2 * cb, lv + kChromaDc + 4 * p, 4);
        }
    if (cbpc >= 2)
        for (int p = 0; p < 2; p++)
            for (int k = 0; k < 4; k++) {
                const int cx = 2 * mb_x + (k & 1), cy = 2 * mb_y + (k >> 1);
                const int inc = nb.chroma_ac(p, cx - 1, cy, intra) + 2 * nb.chroma_ac(p, cx, cy - 1, intra);
                nb.ccbf[p][(size_t)cy * 2 * nb.mw + cx] =
                    (uint8_t)residual_block(a, 4, inc, lv + kChromaAc + (p * 4 + k) * 16 + 1, 15);
            }
}

// coded_block_pattern, transform_size_8x8_flag (inter, High), mb_qp_delta and the residual of a
// non-I_16x16 macroblock (h264.rs:3327 encode_cabac_cbp_and_residuals). dqp_nz: whether the
// previous macroblock's mb_qp_delta was non-zero (always 0 here: one QP per slice).
void cbp_and_residual(Arith &a, Neighbours &nb, int mb_x, int mb_y, int intra, bool t8_syntax, int cbpl, int cbpc,
                      bool use8, const int16_t *lv, int &dqp_nz) {
    const int mw = nb.mw, m = mb_y * mw + mb_x;
    for (int q = 0; q < 4; q++) {
        const int qx = q & 1, qy = q >> 1;
        int ca, cb;
        if (qx)
            ca = !((cbpl >> (q - 1)) & 1);
        else
            ca = mb_x > 0 && (nb.skip[(size_t)m - 1] || !((nb.cbpl[(size_t)m - 1] >> (2 * qy + 1)) & 1));
        if (qy)
            cb = !((cbpl >> (q - 2)) & 1);
        else
            cb = mb_y > 0 && (nb.skip[(size_t)(m - mw)] || !((nb.cbpl[(size_t)(m - mw)] >> (2 + qx)) & 1));
        a.bin(73 + ca + 2 * cb, (cbpl >> q) & 1);
    }
    const int la = mb_x > 0 && !nb.skip[(size_t)m - 1] ? nb.cbpc[(size_t)m - 1] : 0;
    const int ta = mb_y > 0 && !nb.skip[(size_t)(m - mw)] ? nb.cbpc[(size_t)(m - mw)] : 0;
    a.bin(77 + (la != 0) + 2 * (ta != 0), cbpc != 0);
    if (cbpc) a.bin(81 + (la == 2) + 2 * (ta == 2), cbpc == 2);
    nb.cbpl[(size_t)m] = (uint8_t)cbpl;
    nb.cbpc[(size_t)m] = (uint8_t)cbpc;
    if (!cbpl && !cbpc) {
        dqp_nz = 0;
        return;
    }
    if (t8_syntax && cbpl)
        a.bin(399 + (mb_x > 0 && nb.t8[(size_t)m - 1]) + (mb_y > 0 && nb.t8[(size_t)(m - mw)]), use8);
    a.bin(60 + (dqp_nz != 0), 0);  // mb_qp_delta = 0
    dqp_nz = 0;
    const size_t gs = (size_t)4 * mw;
    if (use8) {
        for (int q = 0; q < 4; q++) {
            const int on = (cbpl >> q) & 1;
            if (on) residual_block(a, 5, 0, lv + kLuma + 64 * q, 64);
            const int gx = 4 * mb_x + 2 * (q & 1), gy = 4 * mb_y + 2 * (q >> 1);
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) nb.cbf4[(size_t)(gy + dy) * gs + gx + dx] = (uint8_t)on;
        }
    } else {
        for (int b = 0; b < 16; b++) {
            const int gx = 4 * mb_x + blk_x(b), gy = 4 * mb_y + blk_y(b);
            int f = 0;
            if ((cbpl >> (b >> 2)) & 1)
                f = residual_block(a, 2, nb.luma(gx - 1, gy, intra) + 2 * nb.luma(gx, gy - 1, intra), lv + kLuma + 16 * b, 16);
            nb.cbf4[(size_t)gy * gs + gx] = (uint8_t)f;
        }
    }
    chroma_residual(a, nb, mb_x, mb_y, intra, cbpc, lv);
}

// mb_type of an intra macroblock (cabac.rs:738, Table 9-36): in an I slice with the neighbours'
// context (ctxIdx 3-10), after the P-slice prefix with the fixed contexts 17-20. type: 0 I_NxN,
// 1..24 I_16x16.
void intra_mb_type(Arith &a, bool in_p, int inc, int type) {
    a.bin(in_p ? 17 : 3 + inc, type != 0);
    if (!type) return;
    a.terminate(0);  // not I_PCM
    const int k = type - 1, mode = k & 3, cc = (k >> 2) % 3;
    a.bin(in_p ? 18 : 6, k >= 12);
    a.bin(in_p ? 19 : 7, cc != 0);
    if (cc) a.bin(in_p ? 19 : 8, cc == 2);
    a.bin(in_p ? 20 : 9, mode >> 1);
    a.bin(in_p ? 20 : 10, mode & 1);
}

void chroma_pred_mode(Arith &a, const Neighbours &nb, int mb_x, int mb_y, int mode) {
    const int m = mb_y * nb.mw + mb_x;
    const int ca = mb_x > 0 && nb.intra[(size_t)m - 1] && nb.cmode[(size_t)m - 1];
    const int cb = mb_y > 0 && nb.intra[(size_t)(m - nb.mw)] && nb.cmode[(size_t)(m - nb.mw)];
    a.bin(64 + ca + cb, mode != 0);
    if (mode) {
        a.bin(67, mode != 1);
        if (mode != 1) a.bin(67, mode == 3);
    }
}

// mvd_l0 component (cabac.rs:940): UEG3 with uCoff 9, signed; ctxIdxInc of the first bin from
// the neighbours' absolute mvd sum.
void mvd_component(Arith &a, int comp, int sum, int v) {
    const int base = comp ? 47 : 40, m = absi(v);
    a.bin(base + (sum < 3 ? 0 : (sum <= 32 ? 1 : 2)), m != 0);
    if (!m) return;
    const int prefix = mini(m, 9);
    for (int i = 1; i < prefix; i++) a.bin(base + 3 + mini(i - 1, 3), 1);
    if (m < 9)
        a.bin(base + 3 + mini(m - 1, 3), 0);
    else
        a.eg_suffix((uint32_t)(m - 9), 3);
    a.bypass(v < 0);
}

// An I_16x16 macroblock after its mb_type (enc.rs:3587 code_i16): chroma mode, mb_qp_delta, the
// DC block, AC blocks and chroma.
void i16_body(Arith &a, Neighbours &nb, int mb_x, int mb_y, const rgh_mb_decision &d, const int16_t *lv, int &dqp_nz) {
    const int mw = nb.mw, m = mb_y * mw + mb_x;
    nb.intra[(size_t)m] = nb.i16[(size_t)m] = 1;
    chroma_pred_mode(a, nb, mb_x, mb_y, d.chroma_mode);
    nb.cmode[(size_t)m] = d.chroma_mode;
    a.bin(60 + (dqp_nz != 0), 0);  // mb_qp_delta = 0
    dqp_nz = 0;
    const int ca = mb_x > 0 ? nb.dc_cbf[(size_t)m - 1] : 1, cb = mb_y > 0 ? nb.dc_cbf[(size_t)(m - mw)] : 1;
    nb.dc_cbf[(size_t)m] = (uint8_t)residual_block(a, 0, ca + 2 * cb, lv + kLumaDc, 16);
    nb.cbpl[(size_t)m] = d.cbp_luma;
    nb.cbpc[(size_t)m] = d.cbp_chroma;
    if (d.cbp_luma)
        for (int b = 0; b < 16; b++) {
            const int gx = 4 * mb_x + blk_x(b), gy = 4 * mb_y + blk_y(b);
            const int inc = nb.luma(gx - 1, gy, 1) + 2 * nb.luma(gx, gy - 1, 1);
            nb.cbf4[(size_t)gy * 4 * mw + gx] = (uint8_t)residual_block(a, 1, inc, lv + kLuma + 16 * b + 1, 15);
        }
    chroma_residual(a, nb, mb_x, mb_y, 1, d.cbp_chroma, lv);
}

rgh_status check_common(const rgh_slice_config *cfg, int32_t profile, const void *decisions, const void *levels,
                        const void *out, const int64_t *out_len, int64_t cap) {
    if (!cfg || !decisions || !levels || !out || !out_len || cap < 0) return RGH_ERR_INVALID_ARGUMENT;
    if (cfg->mb_w <= 0 || cfg->mb_h <= 0 || cfg->mb_w > 512 || cfg->mb_h > 512) return RGH_ERR_INVALID_DIMENSIONS;
    if (cfg->qp < 1 || cfg->qp > 51) return RGH_ERR_INVALID_ARGUMENT;
    if (profile != RGH_PROFILE_MAIN && profile != RGH_PROFILE_HIGH) return RGH_ERR_INVALID_ARGUMENT;
    // Compact with CABAC codes B pictures (h264.rs:3075), not written here.
    if (cfg->speed != RGH_SPEED_BALANCED && cfg->speed != RGH_SPEED_FAST && cfg->speed != RGH_SPEED_PARTITIONS)
        return RGH_ERR_UNSUPPORTED;
    return RGH_OK;
}

// Slice header tail from cabac_init_idc on, and the alignment before the slice data. The Fast
// preset's slices disable the loop filter (enc.rs:2334).
void header_tail(BitWriter &w, bool p_slice, int qp, bool fast) {
    if (p_slice) w.ue(0);  // cabac_init_idc
    w.se(0);               // slice_qp_delta: the PPS carries the slice QP
    const int disable = qp == 0 || fast ? 1 : 0;
    w.ue((uint32_t)disable);
    if (disable != 1) {
        w.se(0);
        w.se(0);
    }
    w.align_ones();
}

}  // namespace

extern "C" rgh_status rgh_write_p_slice_cabac(const rgh_slice_config *cfg, int32_t profile,
                                              const rgh_mb_decision *decisions, const int16_t *levels, uint8_t *out,
                                              int64_t cap, int64_t *out_len) {
    rgh_status st = check_common(cfg, profile, decisions, levels, out, out_len, cap);
    if (st != RGH_OK) return st;
    if (cfg->frame_num < 0 || cfg->poc_lsb < 0) return RGH_ERR_INVALID_ARGUMENT;
    const int mw = cfg->mb_w, mh = cfg->mb_h;
    const bool high = profile == RGH_PROFILE_HIGH, parts = cfg->speed == RGH_SPEED_PARTITIONS;
    // Validate every record before writing: the writer has no way back.
    for (int i = 0; i < mw * mh; i++) {
        const rgh_mb_decision &d = decisions[i];
        if (d.mb_type == RGH_MB_P_SKIP) continue;
        if (d.qp != cfg->qp) return RGH_ERR_UNSUPPORTED;
        const bool inter = d.mb_type == RGH_MB_P_L0_16X16 || (parts && rgh_motion::shape_of(d.mb_type) > 0);
        if (!inter && d.mb_type != RGH_MB_I16X16) return RGH_ERR_UNSUPPORTED;
        if (d.cbp_chroma > 2 || (d.cbp_luma & ~15) || d.chroma_mode > 3) return RGH_ERR_INVALID_ARGUMENT;
        if (d.mb_type == RGH_MB_I16X16 && (d.i16_mode > 3 || (d.cbp_luma != 0 && d.cbp_luma != 15)))
            return RGH_ERR_INVALID_ARGUMENT;
        if (d.transform_8x8 && (!high || !inter)) return RGH_ERR_INVALID_ARGUMENT;
    }
    BitWriter w;
    // first_mb_in_slice, slice_type P, pps id, frame_num (4 bits), pic_order_cnt_lsb (6 bits),
    // num_ref_idx_active_override_flag, ref_pic_list_modification_flag_l0,
    // adaptive_ref_pic_marking_mode_flag.
    w.ue(0);
    w.ue(5);
    w.ue(0);
    w.put((uint32_t)cfg->frame_num & 15u, 4);
    w.put((uint32_t)cfg->poc_lsb & 63u, 6);
    w.put(0, 3);
    header_tail(w, true, cfg->qp, cfg->speed == RGH_SPEED_FAST);
    std::vector<uint8_t> &rbsp = w.bytes();
    Arith a(rbsp, cfg->qp, 1);
    Neighbours nb(mw, mh);
    rgh_motion::MotionGrid mg(parts ? 4 * mw : 0, parts ? 4 * mh : 0);
    int dqp_nz = 0;
    const size_t gs = (size_t)4 * mw;
    for (int mb_y = 0; mb_y < mh; mb_y++)
        for (int mb_x = 0; mb_x < mw; mb_x++) {
            const int m = mb_y * mw + mb_x;
            const rgh_mb_decision &d = decisions[m];
            const int16_t *lv = levels + (size_t)m * RGH_LEVELS_PER_MB;
            const int ca = mb_x > 0 && !nb.skip[(size_t)m - 1], cb = mb_y > 0 && !nb.skip[(size_t)(m - mw)];
            a.bin(11 + ca + cb, d.mb_type == RGH_MB_P_SKIP);  // mb_skip_flag
            const int shape = rgh_motion::shape_of(d.mb_type);
            int q[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            if (parts) {
                // Records in the quadrant layout; partitions fill their blocks as they are coded.
                if (d.mb_type != RGH_MB_I16X16) rgh_motion::quads_of(d, q);
                if (shape <= 0) mg.fill(4 * mb_x, 4 * mb_y, 4, 4, d.mb_type == RGH_MB_I16X16 ? -1 : 0, q[0][0], q[0][1]);
            }
            if (d.mb_type == RGH_MB_P_SKIP) {
                nb.skip[(size_t)m] = 1;
                dqp_nz = 0;
            } else if (parts && shape >= 0) {
                // mb_type (Table 9-37: P_L0_16x16 0 0 0, P_L0_L0_16x8 0 1 1, P_L0_L0_8x16 0 1 0,
                // P_8x8 0 0 1; ctxIdx 14, 15, then 16 or 17 after a 1), sub_mb_type P_L0_8x8 "1"
                // (ctxIdx 21) four times, then each partition's mvd against its 8.4.1.3 predictor.
                a.bin(14, 0);
                const bool two = shape == rgh_motion::S16x8 || shape == rgh_motion::S8x16;
                a.bin(15, two);
                if (two)
                    a.bin(17, shape == rgh_motion::S16x8);
                else
                    a.bin(16, shape == rgh_motion::S8x8);
                if (shape == rgh_motion::S8x8)
                    for (int k = 0; k < 4; k++) a.bin(21, 1);
                for (int p = 0; p < rgh_motion::part_count(shape); p++) {
                    int bx, by, pw, ph;
                    rgh_motion::part_rect(shape, p, bx, by, pw, ph);
                    const int gx = 4 * mb_x + bx / 4, gy = 4 * mb_y + by / 4, k = (bx >> 3) + (by >> 3) * 2;
                    int mvp[2];
                    mg.pred(gx, gy, pw / 4, shape, p, mvp);
                    const int mvd[2] = {q[k][0] - mvp[0], q[k][1] - mvp[1]};
                    for (int c = 0; c < 2; c++) {
                        int sum = 0;
                        if (gx > 0) sum += absi(nb.mvd[((size_t)gy * gs + gx - 1) * 2 + c]);
                        if (gy > 0) sum += absi(nb.mvd[((size_t)(gy - 1) * gs + gx) * 2 + c]);
                        mvd_component(a, c, sum, mvd[c]);
                    }
                    for (int y = 0; y < ph / 4; y++)
                        for (int x = 0; x < pw / 4; x++) {
                            nb.mvd[((size_t)(gy + y) * gs + gx + x) * 2] = (int16_t)mvd[0];
                            nb.mvd[((size_t)(gy + y) * gs + gx + x) * 2 + 1] = (int16_t)mvd[1];
                        }
                    if (shape != rgh_motion::S16x16) mg.fill(gx, gy, pw / 4, ph / 4, 0, q[k][0], q[k][1]);
                }
                const bool use8 = d.transform_8x8 && d.cbp_luma;
                nb.t8[(size_t)m] = use8;
                cbp_and_residual(a, nb, mb_x, mb_y, 0, high, d.cbp_luma, d.cbp_chroma, use8, lv, dqp_nz);
            } else if (d.mb_type == RGH_MB_P_L0_16X16) {
                a.bin(14, 0);  // P_L0_16x16: 0 0 0
                a.bin(15, 0);
                a.bin(16, 0);
                const int gx = 4 * mb_x, gy = 4 * mb_y;
                const int mvd[2] = {d.mv[0] - d.mvp[0], d.mv[1] - d.mvp[1]};
                for (int c = 0; c < 2; c++) {
                    int sum = 0;
                    if (gx > 0) sum += absi(nb.mvd[((size_t)gy * gs + gx - 1) * 2 + c]);
                    if (gy > 0) sum += absi(nb.mvd[((size_t)(gy - 1) * gs + gx) * 2 + c]);
                    mvd_component(a, c, sum, mvd[c]);
                }
                const bool use8 = d.transform_8x8 && d.cbp_luma;
                nb.t8[(size_t)m] = use8;
                cbp_and_residual(a, nb, mb_x, mb_y, 0, high, d.cbp_luma, d.cbp_chroma, use8, lv, dqp_nz);
                for (int y = 0; y < 4; y++)
                    for (int x = 0; x < 4; x++) {
                        nb.mvd[((size_t)(gy + y) * gs + gx + x) * 2] = (int16_t)mvd[0];
                        nb.mvd[((size_t)(gy + y) * gs + gx + x) * 2 + 1] = (int16_t)mvd[1];
                    }
            } else {  // I_16x16 in a P slice: the intra prefix, then the I-slice type
                a.bin(14, 1);
                intra_mb_type(a, true, 0, 1 + d.i16_mode + 4 * d.cbp_chroma + (d.cbp_luma ? 12 : 0));
                i16_body(a, nb, mb_x, mb_y, d, lv, dqp_nz);
            }
            a.terminate(m + 1 == mw * mh);  // end_of_slice_flag
        }
    std::vector<uint8_t> nal;
    emit_nal(nal, 3, 1, rbsp);
    return deliver(nal, out, cap, out_len);
}

extern "C" rgh_status rgh_write_idr_slice_cabac(const rgh_slice_config *cfg, int32_t profile, int32_t idr_pic_id,
                                                const rgh_mb_decision *decisions, const int16_t *levels, uint8_t *out,
                                                int64_t cap, int64_t *out_len) {
    rgh_status st = check_common(cfg, profile, decisions, levels, out, out_len, cap);
    if (st != RGH_OK) return st;
    if (idr_pic_id < 0) return RGH_ERR_INVALID_ARGUMENT;
    const int mw = cfg->mb_w, mh = cfg->mb_h;
    for (int i = 0; i < mw * mh; i++) {
        const rgh_mb_decision &d = decisions[i];
        if (d.qp != cfg->qp || d.transform_8x8) return RGH_ERR_UNSUPPORTED;
        if (d.mb_type != RGH_MB_I4X4 && d.mb_type != RGH_MB_I16X16) return RGH_ERR_UNSUPPORTED;
        if (d.cbp_chroma > 2 || (d.cbp_luma & ~15) || d.chroma_mode > 3) return RGH_ERR_INVALID_ARGUMENT;
        if (d.mb_type == RGH_MB_I16X16 && (d.i16_mode > 3 || (d.cbp_luma != 0 && d.cbp_luma != 15)))
            return RGH_ERR_INVALID_ARGUMENT;
        if (d.mb_type == RGH_MB_I4X4) {
            uint64_t packed;
            memcpy(&packed, d.mv, sizeof(packed));
            for (int b = 0; b < 16; b++)
                if (((packed >> (4 * b)) & 15) > 8) return RGH_ERR_INVALID_ARGUMENT;
        }
    }
    BitWriter w;
    // first_mb_in_slice, slice_type I, pps id, frame_num 0, idr_pic_id, pic_order_cnt_lsb 0,
    // no_output_of_prior_pics_flag, long_term_reference_flag.
    w.ue(0);
    w.ue(7);
    w.ue(0);
    w.put(0, 4);
    w.ue((uint32_t)idr_pic_id);
    w.put(0, 6);
    w.put(0, 2);
    header_tail(w, false, cfg->qp, cfg->speed == RGH_SPEED_FAST);
    std::vector<uint8_t> &rbsp = w.bytes();
    Arith a(rbsp, cfg->qp, 0);
    Neighbours nb(mw, mh);
    int dqp_nz = 0;
    const size_t gs = (size_t)4 * mw;
    const bool high = profile == RGH_PROFILE_HIGH;
    for (int mb_y = 0; mb_y < mh; mb_y++)
        for (int mb_x = 0; mb_x < mw; mb_x++) {
            const int m = mb_y * mw + mb_x;
            const rgh_mb_decision &d = decisions[m];
            const int16_t *lv = levels + (size_t)m * RGH_LEVELS_PER_MB;
            // I-slice mb_type context: the neighbours that are not I_NxN.
            const int inc = (mb_x > 0 && nb.i16[(size_t)m - 1]) + (mb_y > 0 && nb.i16[(size_t)(m - mw)]);
            if (d.mb_type == RGH_MB_I16X16) {
                intra_mb_type(a, false, inc, 1 + d.i16_mode + 4 * d.cbp_chroma + (d.cbp_luma ? 12 : 0));
                i16_body(a, nb, mb_x, mb_y, d, lv, dqp_nz);
            } else {  // I_4x4 (enc.rs:3441 code_i4)
                intra_mb_type(a, false, inc, 0);
                if (high) a.bin(399 + (mb_x > 0 && nb.t8[(size_t)m - 1]) + (mb_y > 0 && nb.t8[(size_t)(m - mw)]), 0);
                uint64_t packed;
                memcpy(&packed, d.mv, sizeof(packed));
                for (int b = 0; b < 16; b++) {
                    const int gx = 4 * mb_x + blk_x(b), gy = 4 * mb_y + blk_y(b);
                    const int mode = (int)((packed >> (4 * b)) & 15);
                    int pm = 2;  // 8.3.1.1: DC at the picture's edges, else the smaller neighbour mode
                    if (gx > 0 && gy > 0) pm = mini(nb.modes[(size_t)gy * gs + gx - 1], nb.modes[(size_t)(gy - 1) * gs + gx]);
                    a.bin(68, mode == pm);
                    if (mode != pm) {
                        const int rem = mode < pm ? mode : mode - 1;
                        for (int i = 0; i < 3; i++) a.bin(69, (rem >> i) & 1);
                    }
                    nb.modes[(size_t)gy * gs + gx] = (uint8_t)mode;
                }
                nb.intra[(size_t)m] = 1;
                chroma_pred_mode(a, nb, mb_x, mb_y, d.chroma_mode);
                nb.cmode[(size_t)m] = d.chroma_mode;
                cbp_and_residual(a, nb, mb_x, mb_y, 1, false, d.cbp_luma, d.cbp_chroma, false, lv, dqp_nz);
            }
            a.terminate(m + 1 == mw * mh);
        }
    std::vector<uint8_t> nal;
    emit_nal(nal, 3, 5, rbsp);
    return deliver(nal, out, cap, out_len);
}
