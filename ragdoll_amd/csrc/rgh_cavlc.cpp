// Entropy coding of analysed pictures: the CAVLC (Constrained Baseline) slices a ragdoll encoder
// writes for the decisions and levels the GPU analysis returns, and the stream's SPS / PPS.
// Host code.
//
// Restates the syntax writers of crates/ragdoll-codec/src/h264/enc.rs: the slice header of
// SliceHeader::write (crates/ragdoll-codec/src/h264.rs:1071-1127) as encode_frame_body sets it
// for a P picture (enc.rs:2234-2330), mb_skip_run and the per-macroblock syntax of
// begin_coded_mb / code_inter / code_i16 (enc.rs:2779-3306), write_cavlc_residuals and
// write_cavlc_chroma with their nC prediction (enc.rs:779-1054), the CAVLC block coder
// (h264.rs:2215 encode_cavlc_block) and the NAL escaping of enc.rs:2643-2672.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/ragdoll_hip.h"
#include "rgh_bits.h"
#include "rgh_cavlc_tables.h"
#include "rgh_motion.h"

namespace {

using namespace rgh_cavlc;
using namespace rgh_bits;
using namespace rgh_motion;

int absi(int v) { return v < 0 ? -v : v; }

// One residual block in CAVLC (9.2): coeff_token, trailing-one signs, levels with the adaptive
// suffix length, total_zeros, run_before. sc holds maxc coefficients in scan order; nc < 0
// selects the chroma DC tables.
void cavlc_block(BitWriter &w, const int16_t *sc, int maxc, int nc) {
    const bool cdc = nc < 0;
    const int tbl = nc < 2 ? 0 : (nc < 4 ? 1 : (nc < 8 ? 2 : 3));
    int levels[16], runs[16], n = 0, total_zeros = 0;
    int last = maxc - 1;
    while (last >= 0 && sc[last] == 0) last--;
    if (last < 0) {
        if (cdc)
            w.put(kCdcTokenBits[0], kCdcTokenLen[0]);
        else
            w.put(kCoeffTokenBits[tbl][0], kCoeffTokenLen[tbl][0]);
        return;
    }
    // Levels from the highest frequency down, each with the zeros below it up to the next one.
    int run = 0;
    for (int i = last; i >= 0; i--) {
        if (sc[i] == 0) {
            run++;
            continue;
        }
        if (n > 0) {
            runs[n - 1] = run;
            total_zeros += run;
        }
        levels[n++] = sc[i];
        run = 0;
    }
    runs[n - 1] = run;
    total_zeros += run;
    int t1 = 0;
    while (t1 < n && t1 < 3 && absi(levels[t1]) == 1) t1++;
    const int tok = 4 * n + t1;
    if (cdc)
        w.put(kCdcTokenBits[tok], kCdcTokenLen[tok]);
    else
        w.put(kCoeffTokenBits[tbl][tok], kCoeffTokenLen[tbl][tok]);
    for (int i = 0; i < t1; i++) w.put(levels[i] < 0 ? 1u : 0u, 1);
    uint32_t suffix_len = (n > 10 && t1 < 3) ? 1u : 0u;
    for (int i = t1; i < n; i++) {
        const int lv = levels[i];
        int code = lv > 0 ? 2 * lv - 2 : -2 * lv - 1;
        if (i == t1 && t1 < 3) code -= 2;  // the first level after fewer than 3 trailing ones
        const uint32_t c = (uint32_t)code;
        if (c < (14u << suffix_len)) {
            w.zeros((int)(c >> suffix_len));
            w.put(1, 1);
            if (suffix_len) w.put(c & ((1u << suffix_len) - 1u), (int)suffix_len);
        } else if (suffix_len == 0 && c < 30) {  // level_prefix 14, 4-bit suffix
            w.zeros(14);
            w.put(1, 1);
            w.put(c - 14, 4);
        } else if (suffix_len > 0 && c < (15u << suffix_len)) {
            w.zeros(14);
            w.put(1, 1);
            w.put(c - (14u << suffix_len), (int)suffix_len);
        } else {  // level_prefix 15, 12-bit suffix
            uint32_t esc = c - (15u << suffix_len);
            if (suffix_len == 0) esc -= 15;
            w.zeros(15);
            w.put(1, 1);
            w.put(esc & 0xFFFu, 12);
        }
        if (suffix_len == 0) suffix_len = 1;
        if ((uint32_t)absi(lv) > (3u << (suffix_len - 1)) && suffix_len < 6) suffix_len++;
    }
    if (n < maxc) {
        if (cdc)
            w.put(kCdcTzBits[n - 1][total_zeros], kCdcTzLen[n - 1][total_zeros]);
        else
            w.put(kTotalZerosBits[n - 1][total_zeros], kTotalZerosLen[n - 1][total_zeros]);
    }
    int zeros_left = total_zeros;
    for (int i = 0; i < n - 1 && zeros_left > 0; i++) {
        const int t = (zeros_left < 7 ? zeros_left : 7) - 1;
        w.put(kRunBits[t][runs[i]], kRunLen[t][runs[i]]);
        zeros_left -= runs[i];
    }
}

int count_nz(const int16_t *v, int n) {
    int c = 0;
    for (int i = 0; i < n; i++) c += v[i] != 0;
    return c;
}

// Non-zero coefficient counts of the picture's 4x4 blocks, for nC (9.2.1): luma on a
// (4 mb_w) x (4 mb_h) grid, each chroma plane on (2 mb_w) x (2 mb_h); zero until coded.
struct NnzGrid {
    int gw, gh;
    std::vector<uint8_t> v;
    NnzGrid(int w, int h) : gw(w), gh(h), v((size_t)w * h, 0) {}
    uint8_t &at(int x, int y) { return v[(size_t)y * gw + x]; }
    int nc(int x, int y) {
        const bool l = x > 0, t = y > 0;
        if (l && t) return (at(x - 1, y) + at(x, y - 1) + 1) >> 1;
        if (l) return at(x - 1, y);
        if (t) return at(x, y - 1);
        return 0;
    }
};

int blk_x(int b) { return (b & 1) + ((b >> 1) & 2); }
int blk_y(int b) { return ((b >> 1) & 1) + ((b >> 2) & 2); }

// Level record offsets (RGH_LEVELS_PER_MB layout).
constexpr int kLuma = 0, kLumaDc = 256, kChromaDc = 272, kChromaAc = 280;

void write_chroma(BitWriter &w, NnzGrid *cg, int mb_x, int mb_y, int cbpc, const int16_t *lv) {
    if (cbpc >= 1)
        for (int p = 0; p < 2; p++) cavlc_block(w, lv + kChromaDc + 4 * p, 4, -1);
    if (cbpc >= 2)
        for (int p = 0; p < 2; p++)
            for (int b = 0; b < 4; b++) {
                const int x = 2 * mb_x + (b & 1), y = 2 * mb_y + (b >> 1);
                const int16_t *ac = lv + kChromaAc + (p * 4 + b) * 16 + 1;
                cavlc_block(w, ac, 15, cg[p].nc(x, y));
                cg[p].at(x, y) = (uint8_t)count_nz(ac, 15);
            }
}

uint32_t inter_cbp_code(int cbp) {
    for (uint32_t i = 0; i < 48; i++)
        if (kGolombToInterCbp[i] == cbp) return i;
    return 0;
}

uint32_t intra4_cbp_code(int cbp) {
    for (uint32_t i = 0; i < 48; i++)
        if (kGolombToIntra4x4Cbp[i] == cbp) return i;
    return 3;
}

// Luma 4x4 residual blocks of a coded 8x8 pattern (enc.rs:976 write_cavlc_residuals).
void write_luma4x4(BitWriter &w, NnzGrid &lg, int mb_x, int mb_y, int cbpl, const int16_t *lv) {
    for (int i8 = 0; i8 < 4; i8++)
        for (int k = 0; k < 4; k++) {
            const int b = i8 * 4 + k, x = 4 * mb_x + blk_x(b), y = 4 * mb_y + blk_y(b);
            if (!(cbpl & (1 << i8))) {
                lg.at(x, y) = 0;
                continue;
            }
            const int16_t *zz = lv + kLuma + 16 * b;
            cavlc_block(w, zz, 16, lg.nc(x, y));
            lg.at(x, y) = (uint8_t)count_nz(zz, 16);
        }
}

// I_16x16 macroblock body after mb_type: chroma mode, mb_qp_delta, DC, AC, chroma
// (enc.rs:3587 code_i16).
void write_i16_body(BitWriter &w, NnzGrid &lg, NnzGrid *cg, int mb_x, int mb_y, const rgh_mb_decision &d,
                    const int16_t *lv) {
    w.ue(d.chroma_mode);
    w.se(0);  // mb_qp_delta: fixed QP
    cavlc_block(w, lv + kLumaDc, 16, lg.nc(4 * mb_x, 4 * mb_y));
    if (d.cbp_luma)
        for (int b = 0; b < 16; b++) {
            const int x = 4 * mb_x + blk_x(b), y = 4 * mb_y + blk_y(b);
            const int16_t *ac = lv + kLuma + 16 * b + 1;
            cavlc_block(w, ac, 15, lg.nc(x, y));
            lg.at(x, y) = (uint8_t)count_nz(ac, 15);
        }
    write_chroma(w, cg, mb_x, mb_y, d.cbp_chroma, lv);
}

// An I_4x4 macroblock (enc.rs:3441 code_i4) after its mb_type: the 16 prediction modes as
// predicted-mode flags and 3-bit remainders (8.3.1.1: the smaller of the left and top modes, DC
// at the picture's edges; `modes` holds the coded modes, DC elsewhere), the chroma mode, the
// intra CBP and the residual.
rgh_status write_i4_body(BitWriter &w, NnzGrid &lg, NnzGrid *cg, std::vector<uint8_t> &modes, int mb_w, int mb_x,
                         int mb_y, const rgh_mb_decision &d, const int16_t *lv) {
    const int cbpl = d.cbp_luma, cbp = cbpl | (d.cbp_chroma << 4), gs = 4 * mb_w;
    uint64_t packed;
    memcpy(&packed, d.mv, sizeof(packed));
    for (int b = 0; b < 16; b++) {
        const int gx = 4 * mb_x + blk_x(b), gy = 4 * mb_y + blk_y(b);
        const int m = (int)((packed >> (4 * b)) & 15);
        if (m > 8) return RGH_ERR_INVALID_ARGUMENT;
        int pm = 2;
        if (gx > 0 && gy > 0) {
            const int a = modes[(size_t)gy * gs + gx - 1], t = modes[(size_t)(gy - 1) * gs + gx];
            pm = a < t ? a : t;
        }
        if (m == pm) {
            w.put(1, 1);
        } else {
            w.put(0, 1);
            w.put((uint32_t)(m < pm ? m : m - 1), 3);
        }
        modes[(size_t)gy * gs + gx] = (uint8_t)m;
    }
    w.ue(d.chroma_mode);
    w.ue(intra4_cbp_code(cbp));
    if (cbp) {
        w.se(0);  // mb_qp_delta: fixed QP
        write_luma4x4(w, lg, mb_x, mb_y, cbpl, lv);
        write_chroma(w, cg, mb_x, mb_y, d.cbp_chroma, lv);
    }
    return RGH_OK;
}

}  // namespace

extern "C" rgh_status rgh_write_p_slice_cavlc(const rgh_slice_config *cfg, const rgh_mb_decision *decisions,
                                              const int16_t *levels, uint8_t *out, int64_t cap, int64_t *out_len) {
    if (!cfg || !decisions || !levels || !out || !out_len || cap < 0) return RGH_ERR_INVALID_ARGUMENT;
    const int mb_w = cfg->mb_w, mb_h = cfg->mb_h;
    if (mb_w <= 0 || mb_h <= 0 || mb_w > 512 || mb_h > 512) return RGH_ERR_INVALID_DIMENSIONS;
    if (cfg->qp < 0 || cfg->qp > 51 || cfg->frame_num < 0 || cfg->poc_lsb < 0 ||
        (cfg->speed < RGH_SPEED_BALANCED || cfg->speed > RGH_SPEED_PARTITIONS))
        return RGH_ERR_INVALID_ARGUMENT;
    const bool compact = cfg->speed == RGH_SPEED_COMPACT, parts = cfg->speed == RGH_SPEED_PARTITIONS;
    const bool quad = compact || parts;  // records in the quadrant layout
    BitWriter w;
    // Slice header (h264.rs:1071 SliceHeader::write): first_mb_in_slice, slice_type P (5),
    // pic_parameter_set_id, frame_num (4 bits), pic_order_cnt_lsb (6 bits),
    // num_ref_idx_active_override_flag, ref_pic_list_modification_flag_l0,
    // adaptive_ref_pic_marking_mode_flag, slice_qp_delta, disable_deblocking_filter_idc and,
    // when the filter is on, its two offsets.
    w.ue(0);
    w.ue(5);
    w.ue(0);
    w.put((uint32_t)cfg->frame_num & 15u, 4);
    w.put((uint32_t)cfg->poc_lsb & 63u, 6);
    w.put(0, 1);
    w.put(0, 1);
    w.put(0, 1);
    w.se(0);
    // The Fast preset's slices disable the loop filter (enc.rs:2334).
    const int disable = cfg->qp == 0 || cfg->speed == RGH_SPEED_FAST ? 1 : 0;
    w.ue((uint32_t)disable);
    if (disable != 1) {
        w.se(0);
        w.se(0);
    }
    NnzGrid lg(4 * mb_w, 4 * mb_h);
    NnzGrid cg[2] = {NnzGrid(2 * mb_w, 2 * mb_h), NnzGrid(2 * mb_w, 2 * mb_h)};
    MotionGrid mg(quad ? 4 * mb_w : 0, quad ? 4 * mb_h : 0);
    std::vector<uint8_t> modes((size_t)16 * mb_w * mb_h, 2);
    uint32_t skip_run = 0;
    for (int mb_y = 0; mb_y < mb_h; mb_y++)
        for (int mb_x = 0; mb_x < mb_w; mb_x++) {
            const int idx = mb_y * mb_w + mb_x;
            const rgh_mb_decision &d = decisions[idx];
            const int16_t *lv = levels + (size_t)idx * RGH_LEVELS_PER_MB;
            int q[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            const int shape = shape_of(d.mb_type);
            if (quad) {
                const bool intra = d.mb_type == RGH_MB_I16X16 || d.mb_type == RGH_MB_I4X4;
                if (!intra) quads_of(d, q);
                if (shape <= S16x16)  // partitions fill their blocks as they are coded
                    mg.fill(4 * mb_x, 4 * mb_y, 4, 4, intra ? -1 : 0, q[0][0], q[0][1]);
            }
            if (d.mb_type == RGH_MB_P_SKIP) {
                skip_run++;
                continue;
            }
            if (d.qp != cfg->qp || d.transform_8x8) return RGH_ERR_UNSUPPORTED;
            const int cbpl = d.cbp_luma, cbpc = d.cbp_chroma;
            if (cbpc > 2 || (cbpl & ~15)) return RGH_ERR_INVALID_ARGUMENT;
            w.ue(skip_run);
            skip_run = 0;
            if (d.mb_type == RGH_MB_P_L0_16X16 || (compact && d.mb_type == RGH_MB_P_8X8) || (parts && shape >= 0)) {
                const int cbp = cbpl | (cbpc << 4);
                if (!quad) {
                    w.ue(0);
                    w.se(d.mv[0] - d.mvp[0]);
                    w.se(d.mv[1] - d.mvp[1]);
                } else {
                    // mb_type (P_L0_16x16 0, P_L0_L0_16x8 1, P_L0_L0_8x16 2, P_8x8 3 with four
                    // P_L0_8x8 sub-macroblocks: enc/inter.rs:331-358), then each partition's mvd
                    // against its predictor, which sees the partitions coded before it.
                    w.ue((uint32_t)shape);
                    if (shape == S8x8)
                        for (int k = 0; k < 4; k++) w.ue(0);
                    for (int p = 0; p < part_count(shape); p++) {
                        int bx, by, pw, ph;
                        part_rect(shape, p, bx, by, pw, ph);
                        const int gx = 4 * mb_x + bx / 4, gy = 4 * mb_y + by / 4, k = (bx >> 3) + (by >> 3) * 2;
                        int mvp[2];
                        mg.pred(gx, gy, pw / 4, shape, p, mvp);
                        w.se(q[k][0] - mvp[0]);
                        w.se(q[k][1] - mvp[1]);
                        if (shape != S16x16) mg.fill(gx, gy, pw / 4, ph / 4, 0, q[k][0], q[k][1]);
                    }
                }
                w.ue(inter_cbp_code(cbp));
                if (cbp) {
                    w.se(0);  // mb_qp_delta: fixed QP
                    write_luma4x4(w, lg, mb_x, mb_y, cbpl, lv);
                    write_chroma(w, cg, mb_x, mb_y, cbpc, lv);
                }
            } else if (d.mb_type == RGH_MB_I16X16) {
                if (d.i16_mode > 3 || d.chroma_mode > 3 || (cbpl != 0 && cbpl != 15)) return RGH_ERR_INVALID_ARGUMENT;
                w.ue(5u + 1u + d.i16_mode + 4u * (uint32_t)cbpc + (cbpl ? 12u : 0u));
                write_i16_body(w, lg, cg, mb_x, mb_y, d, lv);
            } else if (compact && d.mb_type == RGH_MB_I4X4) {
                if (d.chroma_mode > 3) return RGH_ERR_INVALID_ARGUMENT;
                w.ue(5);  // I_NxN in a P slice
                rgh_status st = write_i4_body(w, lg, cg, modes, mb_w, mb_x, mb_y, d, lv);
                if (st != RGH_OK) return st;
            } else {
                return RGH_ERR_UNSUPPORTED;  // I_PCM; I_4x4 is produced by the Compact analysis only
            }
        }
    if (skip_run) w.ue(skip_run);
    w.trailing();
    std::vector<uint8_t> nal;
    emit_nal(nal, 3, 1, w.bytes());
    return deliver(nal, out, cap, out_len);
}

extern "C" rgh_status rgh_write_idr_slice_cavlc(const rgh_slice_config *cfg, int32_t idr_pic_id,
                                                const rgh_mb_decision *decisions, const int16_t *levels, uint8_t *out,
                                                int64_t cap, int64_t *out_len) {
    if (!cfg || !decisions || !levels || !out || !out_len || cap < 0 || idr_pic_id < 0) return RGH_ERR_INVALID_ARGUMENT;
    const int mb_w = cfg->mb_w, mb_h = cfg->mb_h;
    if (mb_w <= 0 || mb_h <= 0 || mb_w > 512 || mb_h > 512) return RGH_ERR_INVALID_DIMENSIONS;
    if (cfg->qp < 0 || cfg->qp > 51 ||
        (cfg->speed < RGH_SPEED_BALANCED || cfg->speed > RGH_SPEED_PARTITIONS))
        return RGH_ERR_INVALID_ARGUMENT;
    BitWriter w;
    // Slice header of an IDR picture: first_mb_in_slice, slice_type I (7), pps id, frame_num 0,
    // idr_pic_id, pic_order_cnt_lsb 0, no_output_of_prior_pics / long_term_reference flags,
    // slice_qp_delta, disable_deblocking_filter_idc (+ offsets).
    w.ue(0);
    w.ue(7);
    w.ue(0);
    w.put(0, 4);
    w.ue((uint32_t)idr_pic_id);
    w.put(0, 6);
    w.put(0, 1);
    w.put(0, 1);
    w.se(0);
    const int disable = cfg->qp == 0 || cfg->speed == RGH_SPEED_FAST ? 1 : 0;
    w.ue((uint32_t)disable);
    if (disable != 1) {
        w.se(0);
        w.se(0);
    }
    NnzGrid lg(4 * mb_w, 4 * mb_h);
    NnzGrid cg[2] = {NnzGrid(2 * mb_w, 2 * mb_h), NnzGrid(2 * mb_w, 2 * mb_h)};
    // Intra 4x4 prediction modes per 4x4 block (DC for I_16x16) for the predicted mode.
    std::vector<uint8_t> modes((size_t)16 * mb_w * mb_h, 2);
    for (int mb_y = 0; mb_y < mb_h; mb_y++)
        for (int mb_x = 0; mb_x < mb_w; mb_x++) {
            const int idx = mb_y * mb_w + mb_x;
            const rgh_mb_decision &d = decisions[idx];
            const int16_t *lv = levels + (size_t)idx * RGH_LEVELS_PER_MB;
            if (d.qp != cfg->qp || d.transform_8x8 || d.cbp_chroma > 2 || d.chroma_mode > 3) return RGH_ERR_UNSUPPORTED;
            if (d.mb_type == RGH_MB_I4X4) {
                w.ue(0);
                rgh_status st = write_i4_body(w, lg, cg, modes, mb_w, mb_x, mb_y, d, lv);
                if (st != RGH_OK) return st;
            } else if (d.mb_type == RGH_MB_I16X16) {
                const int cbpl = d.cbp_luma;
                if (d.i16_mode > 3 || (cbpl != 0 && cbpl != 15)) return RGH_ERR_INVALID_ARGUMENT;
                w.ue(1u + d.i16_mode + 4u * (uint32_t)d.cbp_chroma + (cbpl ? 12u : 0u));
                write_i16_body(w, lg, cg, mb_x, mb_y, d, lv);
            } else {
                return RGH_ERR_UNSUPPORTED;
            }
        }
    w.trailing();
    std::vector<uint8_t> nal;
    emit_nal(nal, 3, 5, w.bytes());
    return deliver(nal, out, cap, out_len);
}

extern "C" rgh_status rgh_write_headers(const rgh_stream_config *cfg, uint8_t *out, int64_t cap, int64_t *out_len) {
    if (!cfg || !out || !out_len || cap < 0) return RGH_ERR_INVALID_ARGUMENT;
    if (cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 1) || (cfg->height & 1) || cfg->width > 8192 ||
        cfg->height > 8192)
        return RGH_ERR_INVALID_DIMENSIONS;
    if (cfg->profile < RGH_PROFILE_BASELINE || cfg->profile > RGH_PROFILE_HIGH || cfg->qp < 0 || cfg->qp > 51 ||
        cfg->fps_num <= 0 || cfg->fps_den <= 0 ||
        (cfg->speed < RGH_SPEED_BALANCED || cfg->speed > RGH_SPEED_PARTITIONS))
        return RGH_ERR_INVALID_ARGUMENT;
    const int cqo = cfg->speed == RGH_SPEED_COMPACT ? 1 : 0;  // h264.rs:3087 sync_reorder
    const int mb_w = (cfg->width + 15) / 16, mb_h = (cfg->height + 15) / 16;
    const int profile_idc = cfg->profile == RGH_PROFILE_BASELINE ? 66 : (cfg->profile == RGH_PROFILE_MAIN ? 77 : 100);
    std::vector<uint8_t> nal;
    {
        // h264.rs:556 Sps::to_rbsp: level 4.0, poc type 0, 2 reference frames, cropping to the
        // display size, VUI with square pixels, limited-range BT.709 and the frame rate.
        BitWriter w;
        w.put((uint32_t)profile_idc, 8);
        w.put(cfg->profile == RGH_PROFILE_BASELINE ? 0xC0u : (cfg->profile == RGH_PROFILE_MAIN ? 0x40u : 0x00u), 8);
        w.put(40, 8);
        w.ue(0);
        if (profile_idc == 100) {
            w.ue(1);  // chroma_format_idc 4:2:0
            w.ue(0);
            w.ue(0);
            w.put(0, 1);
            w.put(0, 1);
        }
        w.ue(0);  // log2_max_frame_num - 4
        w.ue(0);  // pic_order_cnt_type
        w.ue(2);  // log2_max_pic_order_cnt_lsb - 4
        w.ue(2);  // max_num_ref_frames
        w.put(0, 1);
        w.ue((uint32_t)(mb_w - 1));
        w.ue((uint32_t)(mb_h - 1));
        w.put(1, 1);  // frame_mbs_only
        w.put(1, 1);  // direct_8x8_inference
        const int crop_r = (mb_w * 16 - cfg->width) / 2, crop_b = (mb_h * 16 - cfg->height) / 2;
        w.put(crop_r || crop_b ? 1u : 0u, 1);
        if (crop_r || crop_b) {
            w.ue(0);
            w.ue((uint32_t)crop_r);
            w.ue(0);
            w.ue((uint32_t)crop_b);
        }
        w.put(1, 1);  // vui_parameters_present
        w.put(1, 1);  // aspect_ratio_info_present: 1:1
        w.put(1, 8);
        w.put(0, 1);  // overscan_info_present
        w.put(1, 1);  // video_signal_type_present: unspecified format, limited range, BT.709
        w.put(5, 3);
        w.put(0, 1);
        w.put(1, 1);
        w.put(1, 8);
        w.put(1, 8);
        w.put(1, 8);
        w.put(0, 1);  // chroma_loc_info_present
        w.put(1, 1);  // timing_info_present
        w.put((uint32_t)cfg->fps_den, 32);
        w.put((uint32_t)cfg->fps_num * 2u, 32);
        w.put(1, 1);  // fixed_frame_rate
        w.put(0, 1);  // nal_hrd
        w.put(0, 1);  // vcl_hrd
        w.put(0, 1);  // pic_struct_present
        w.put(0, 1);  // bitstream_restriction
        w.trailing();
        emit_nal(nal, 3, 7, w.bytes());
    }
    {
        // h264.rs:806 Pps::to_rbsp: CABAC for Main/High, initial QP the stream's, deblocking
        // control present, 8x8 transform for High.
        BitWriter w;
        w.ue(0);
        w.ue(0);
        w.put(cfg->profile != RGH_PROFILE_BASELINE ? 1u : 0u, 1);
        w.put(0, 1);
        w.ue(0);
        w.ue(0);
        w.ue(0);
        w.put(0, 1);
        w.put(0, 2);
        w.se(cfg->qp - 26);
        w.se(0);
        w.se(cqo);  // chroma_qp_index_offset
        w.put(1, 1);
        w.put(0, 1);
        w.put(0, 1);
        if (cfg->profile == RGH_PROFILE_HIGH) {
            w.put(1, 1);
            w.put(0, 1);
            w.se(cqo);
        }
        w.trailing();
        emit_nal(nal, 3, 8, w.bytes());
    }
    return deliver(nal, out, cap, out_len);
}
