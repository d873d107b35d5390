/*
 * CPU oracle: a plain-C restatement of ragdoll's Balanced H.264 encoder (IDR + P frames,
 * CAVLC and CABAC, in-loop deblocking), written from the reference's source text.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (ragdoll_amd/, include/) links, imports or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker or the timed CPU baseline.
 *
 * Reference files restated (paths relative to the ragdoll repository root):
 *   crates/ragdoll-codec/src/h264/enc.rs      frame coder, P analysis, CAVLC/CABAC syntax, deblock
 *   crates/ragdoll-codec/src/cabac.rs         CABAC engine, context tables, binarizations
 *   crates/ragdoll-codec/src/h264.rs          tables, SPS/PPS/slice header, intra predictors,
 *                                             MVP (pred_motion_partition), H264Encoder
 *   crates/ragdoll-simd/src/h264_enc.rs       SAD/SSD/SATD, DCT/quant, hpel window, chroma MC
 *   crates/ragdoll-simd/src/h264_mc.rs        spec luma MC
 *   crates/ragdoll-simd/src/fixed.rs          4x4/8x8 transforms, luma DC Hadamard
 *   crates/ragdoll-simd/src/h264_deblock.rs   edge filters; h264_dsp.rs tables and I4x4 preds
 *   crates/ragdoll-simd/src/h264_dct8.rs, h264_idct.rs   8x8 transform path (High profile)
 *
 * Pinning (see tests/): the Baseline path reproduces the reference's committed golden clips
 * qa/corpus/synth/clip_160x90_30fps.mp4 and clip_320x180_25fps.mkv byte for byte; kernels are
 * further checked against closed forms and brute force. The CABAC (Main/High) configuration
 * runs the same decisions with the 8x8-transform choice and writes CABAC slices; those are
 * pinned by an independent parser of ITU-T H.264 9.3 (tests/cabac_decoder.py) that also reads
 * libx264's CABAC slices to their exact ends.
 */
#ifndef RAGDOLL_H264_ORACLE_H
#define RAGDOLL_H264_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-macroblock decision record, 32 bytes (layout shared with the product's rgh_mb_decision). */
typedef struct ora_mb_decision {
    uint8_t mb_type;      /* 0 P_Skip, 1 P_L0_16x16, 2 I_16x16, 3 I_4x4, 4 I_PCM, 5 P_8x8,
                             6 P_L0_L0_16x8, 7 P_L0_L0_8x16 */
    uint8_t i16_mode;     /* I_16x16 luma mode (0 V, 1 H, 2 DC, 3 plane) */
    uint8_t chroma_mode;  /* intra chroma mode (0 DC, 1 H, 2 V, 3 plane) */
    uint8_t cbp_luma;
    uint8_t cbp_chroma;
    uint8_t transform_8x8;
    uint8_t qp;
    uint8_t flags;        /* bit0: skip decided by the probe (before the search) */
    int16_t mv[2];        /* quarter-pel motion vector (0 for intra) */
    int16_t mvp[2];       /* median predictor used for the mvd */
    uint32_t cost;        /* final motion-estimation cost (SAD-domain) of inter MBs */
    uint16_t n_sad;       /* full-pel 16x16 SAD evaluations */
    uint16_t n_subpel;    /* sub-pel SAD (or SAD-avg) evaluations */
    uint16_t n_satd;      /* 16x16 SATD evaluations (inter + intra plane) */
    uint16_t n_hpel;      /* half-pel windows interpolated */
    uint16_t nnz;         /* non-zero 4x4 luma blocks after coding, bit 4*by+bx (raster) */
    uint16_t reserved;
} ora_mb_decision;

/* Levels per macroblock, int16 elements (layout shared with the product). */
#define ORA_LEVELS_PER_MB 408
#define ORA_LV_LUMA 0        /* 16 blocks x 16 scanned levels, or 4 x 64 (8x8 transform) */
#define ORA_LV_LUMA_DC 256   /* 16 scanned I_16x16 DC levels */
#define ORA_LV_CHROMA_DC 272 /* 2 planes x 4 */
#define ORA_LV_CHROMA_AC 280 /* 2 planes x 4 blocks x 16 scanned levels (slot 0 unused) */

typedef struct ora_enc ora_enc;

/* profile: 0 Constrained Baseline (CAVLC), 1 Main (CABAC), 2 High (CABAC + 8x8 transform). */
ora_enc *ora_create(int width, int height, int fps_num, int fps_den, int qp, int profile,
                    int keyint);
void ora_destroy(ora_enc *e);
/* Diamond iterations of the integer search (16 = the reference's value). */
void ora_set_me_range(ora_enc *e, int iters);
/* H264Encoder::with_speed(H264Speed::Compact) for Baseline (CAVLC) streams; -1 for CABAC. */
int ora_set_compact(ora_enc *e, int on);
/* The Fast preset (H264Speed::Fast): half-pel motion toward the cheaper neighbour, intra 16x16
 * without the plane mode, the Fast skip-probe and intra gates, no loop filter. Returns -1 with
 * Compact. */
int ora_set_fast(ora_enc *e, int on);
/* P partitions on top of the Balanced preset: when a macroblock's 16x16 inter cost exceeds 40
 * lambda, 16x8, 8x16 and 8x8 partitions are searched around the 16x16 vector and the cheapest
 * shape by SATD + lambda x bits is coded (an extension the reference does not define; see the
 * oracle source). P-picture records then hold the four quadrant vectors in words 2-5 (the Compact
 * layout; word 6 zero). Returns -1 with Compact or Fast. */
int ora_set_partitions(ora_enc *e, int on);
/* Encodes one I420 frame. Writes the Annex-B access unit (SPS+PPS before an IDR slice) to out
 * and returns its length; returns -1 on a too small buffer. For CABAC profiles no bitstream is
 * written (returns 0) but every decision, level and reconstruction is produced. */
long ora_encode_frame(ora_enc *e, const uint8_t *y, const uint8_t *u, const uint8_t *v,
                      int stride_y, int stride_u, int stride_v, uint8_t *out, long cap);
long ora_headers(ora_enc *e, uint8_t *out, long cap);
int ora_mb_w(const ora_enc *e);
int ora_mb_h(const ora_enc *e);
int ora_last_was_idr(const ora_enc *e);
/* Decisions / levels of the last frame (mb_w*mb_h records / mb_w*mb_h*ORA_LEVELS_PER_MB). */
void ora_get_decisions(const ora_enc *e, ora_mb_decision *out);
void ora_get_levels(const ora_enc *e, int16_t *out);
/* Reconstruction of the last frame at coded size (mb_w*16 x mb_h*16 luma, half chroma):
 * which = 0 before deblocking, 1 after deblocking (the next frame's reference). */
void ora_get_recon(const ora_enc *e, int which, uint8_t *y, uint8_t *u, uint8_t *v);

/* Kernel-level entry points, for known-answer and property tests. */
uint32_t ora_sad16(const uint8_t *a, int sa, const uint8_t *b, int sb);
uint32_t ora_satd16(const uint8_t *a, int sa, const uint8_t *b, int sb);
uint32_t ora_satd4(const uint8_t *a, int sa, const uint8_t *b, int sb);
void ora_fdct4(const int16_t *in, int16_t *out);
void ora_idct4(const int16_t *in, int16_t *out);
void ora_fdct8(const int16_t *in, int16_t *out);
void ora_luma_mc(uint8_t *dst, int ds, const uint8_t *src, int ss, int w, int h, int mx, int my);
void ora_chroma_mc8(uint8_t *dst, const uint8_t *src, int ss, int fx, int fy);
/* Window-table quarter-pel prediction of the 16x16 block whose integer origin is src (padded
 * plane, stride ss) at full-pel search centre (0,0) and quarter-pel offset (dx,dy), |d| <= 3. */
void ora_window_pred(uint8_t *dst, const uint8_t *src, int ss, int dx, int dy);
uint32_t ora_decimate_score(const int16_t *scanned, int n);
/* High-profile 8x8 pieces: dequantization of 64 raster-order levels at qp (h264.rs:1403), the
 * inverse 8x8 transform's residual ((x + 32) >> 6 of the 2-D transform, simd/h264_idct.rs:268),
 * and the quantizer multiplier / normAdjust8x8 entry of raster position pos for qp % 6 = rem. */
void ora_dequant8(const int16_t *levels, int qp, int32_t *out);
void ora_idct8(const int32_t *in, int32_t *out);
int ora_quant8_mf(int rem, int pos);
int ora_dequant8_norm(int rem, int pos);
/* CAVLC residual_block of n (16, 15 or 4) scanned levels at nC; returns bit count and writes
 * the bits MSB-first into out. */
int ora_cavlc_block(const int16_t *scanned, int n, int nc, uint8_t *out, int cap);
void ora_predict_intra4x4(int mode, const uint8_t *top8, const uint8_t *left4, int tl,
                          int has_top, int has_left, uint8_t *dst);

#ifdef __cplusplus
}
#endif
#endif
