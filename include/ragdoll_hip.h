/*
 * ragdoll_hip — MI355X (gfx950) H.264 P-frame analysis behind a plain C ABI.
 *
 * The library replaces the per-macroblock P analysis of ragdoll's Balanced encoder: the seam
 * `EncState::code_p_mb` inside the raster loop of `encode_frame_body`
 * (crates/ragdoll-codec/src/h264/enc.rs:3011-3191, called at :2423-2429), together with the
 * post-frame reference preparation of that loop (deblock :3776-3881, extend_borders :164-178,
 * cur/refp swap and prev_mb_mv record :2470-2503). A whole P frame is analysed on the GPU in
 * anti-diagonal macroblock wavefront order; decisions, quantized levels and the reconstruction
 * come back bit-exact with the CPU encoder, for the entropy coder to serialise in raster order.
 *
 * Conventions (mirroring the reference):
 *  - Errors: every entry point returns an rgh_status; the codes map one-to-one onto the
 *    reference's RagdollError variants (crates/ragdoll-core/src/error.rs:4-31) used by the
 *    encoder (Unsupported for non-I420 input, enc.rs:2235-2240; InvalidDimensions), plus the
 *    device and allocation failures a GPU adds. Nothing aborts across the boundary.
 *  - Ownership: the caller owns its frames (read-only here); a context owns every persistent
 *    device buffer, allocated once at rgh_create ("kept across frames so steady-state encoding
 *    allocates nothing per macroblock", enc.rs:215-217).
 *  - Threading: a context is used by one thread at a time, like one H264Encoder per GOP worker
 *    (crates/ragdoll-pipeline/src/lib.rs:110-155, 575-590). Concurrency comes from many
 *    contexts; rgh_analyse_batch runs one P frame of each of several contexts in one set of
 *    launches.
 */
#ifndef RAGDOLL_HIP_H
#define RAGDOLL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rgh_status {
    RGH_OK = 0,
    RGH_ERR_INVALID_ARGUMENT = 1,   /* RagdollError::InvalidArgument */
    RGH_ERR_UNSUPPORTED = 2,        /* RagdollError::Unsupported (e.g. non-I420, enc.rs:2235) */
    RGH_ERR_INVALID_DIMENSIONS = 3, /* RagdollError::InvalidDimensions */
    RGH_ERR_NO_REFERENCE = 4,       /* a P frame without a reference picture */
    RGH_ERR_DEVICE = 5,             /* HIP runtime failure */
    RGH_ERR_OOM = 6                 /* device allocation failure */
} rgh_status;

/* Encoder profile (crates/ragdoll-codec/src/h264.rs:30-38 H264Profile). */
typedef enum rgh_profile {
    RGH_PROFILE_BASELINE = 0, /* Constrained Baseline, CAVLC, 4x4 transform */
    RGH_PROFILE_MAIN = 1,     /* CABAC, 4x4 transform */
    RGH_PROFILE_HIGH = 2      /* CABAC, adaptive 8x8 transform (enc.rs:3170-3179) */
} rgh_profile;

/* Context configuration: the fields of H264Encoder::new(w, h, fps, qp).with_profile(p)
 * (h264.rs:3032-3065) that the P analysis reads, plus the search parameters the reference
 * hard-codes (enc.rs:1207 diamond iterations = 16, enc.rs:3106 qpel iterations = 2). */
typedef struct rgh_config {
    int32_t width;      /* display width, even, > 0 */
    int32_t height;     /* display height, even, > 0 */
    int32_t profile;    /* rgh_profile */
    int32_t qp;         /* fixed QP 1..51 (RateControlMode::FixedQp) */
    int32_t me_range;   /* diamond iterations; 0 selects the reference's 16 */
    int32_t qpel_iters; /* quarter-pel iterations; 0 selects the reference's 2 */
    int32_t device;     /* HIP device ordinal */
    int32_t flags;      /* RGH_FLAG_* */
    int32_t ring;       /* reconstructed pictures kept (2..64; 0 selects 2): how many frames of
                           one stream rgh_analyse_frames can have in flight, plus one */
} rgh_config;

/* rgh_config.flags: skip the device copy of each frame's reconstruction before deblocking
 * (rgh_get_outputs with which_recon = 0 then fails with RGH_ERR_INVALID_ARGUMENT). */
#define RGH_FLAG_NO_UNDEBLOCKED_COPY 1
/* rgh_config.flags: the Compact preset (rgh_speed RGH_SPEED_COMPACT); Baseline profile only.
 * A context's stream starts at its IDR (GopRingEncoder's encoder per GOP): the P picture after
 * an IDR takes zero 16x16 search vectors of the previous picture as candidates, where the
 * reference, continuing across a mid-stream IDR, takes those of the picture two before it. */
#define RGH_FLAG_COMPACT 2
/* rgh_config.flags: the Fast preset (rgh_speed RGH_SPEED_FAST): half-pel motion toward the
 * cheaper full-pel neighbour, the skip probe only next to a skipped macroblock, intra 16x16 only
 * when its SAD comes within 1.75x of the inter SAD (and must win by 16 lambda), no plane modes,
 * no intra 4x4, chroma from the luma mode in P pictures, and no loop filter (enc.rs:503, :2270,
 * :3050-3152, :3322, :3605). Not combinable with RGH_FLAG_COMPACT or learned proposals. */
#define RGH_FLAG_FAST 4
/* rgh_config.flags: P partitions on top of the Balanced preset (rgh_speed RGH_SPEED_PARTITIONS),
 * an extension the reference does not define (its Balanced preset codes 16x16 partitions only,
 * enc.rs:3008-3191): when a P macroblock's 16x16 inter cost (SATD + lambda x bits) exceeds 40
 * lambda, the 16x8, 8x16 and 8x8 shapes are searched partition by partition (the 8.4.1.3
 * predictor, directional for 16x8 / 8x16; candidates [predictor, 16x16 vector]; a diamond of at
 * most 8 steps within +-8 samples of the 16x16 full-sample vector; half- and quarter-sample
 * diamonds) and the cheapest shape by the 16x16 SATD of its prediction + lambda x (mvd and
 * mb_type bits) is coded. Every P-picture record then holds its four 8x8 quadrant vectors in
 * words 2-5 (mv, mvp, cost, n_sad/n_subpel; zero for intra) and zero in word 6, and mb_type may be
 * RGH_MB_P_8X8, RGH_MB_P_L0_L0_16X8 or RGH_MB_P_L0_L0_8X16. Any profile; not combinable with
 * RGH_FLAG_COMPACT, RGH_FLAG_FAST or learned proposals; rgh_analyse_frames runs it on the
 * per-step schedule (RGH_FRAMES_PERSISTENT is ignored for such contexts). */
#define RGH_FLAG_PARTITIONS 8

/* An I420 picture: three planes with their strides; width/height are display sizes.
 * Chroma planes are (width+1)/2 x (height+1)/2. */
typedef struct rgh_i420 {
    const uint8_t *y;
    const uint8_t *u;
    const uint8_t *v;
    int32_t stride_y;
    int32_t stride_u;
    int32_t stride_v;
    int32_t width;
    int32_t height;
} rgh_i420;

/* A writable I420 picture at coded size (mb_w*16 x mb_h*16, chroma half). */
typedef struct rgh_i420_out {
    uint8_t *y;
    uint8_t *u;
    uint8_t *v;
    int32_t stride_y;
    int32_t stride_u;
    int32_t stride_v;
} rgh_i420_out;

/* Macroblock types of rgh_mb_decision.mb_type. */
enum {
    RGH_MB_P_SKIP = 0,
    RGH_MB_P_L0_16X16 = 1,
    RGH_MB_I16X16 = 2,
    RGH_MB_I4X4 = 3,
    RGH_MB_I_PCM = 4,
    RGH_MB_P_8X8 = 5,        /* four P_L0_8x8 partitions (Compact preset, RGH_FLAG_PARTITIONS) */
    RGH_MB_P_L0_L0_16X8 = 6, /* two 16x8 partitions (RGH_FLAG_PARTITIONS) */
    RGH_MB_P_L0_L0_8X16 = 7  /* two 8x16 partitions (RGH_FLAG_PARTITIONS) */
};

/* Encoder presets (H264Speed, h264.rs). */
typedef enum rgh_speed {
    RGH_SPEED_BALANCED = 0, /* diamond search of 16x16 vectors (enc.rs:1182-1339) */
    RGH_SPEED_COMPACT = 1,  /* hexagon search over half-sample planes, 8x8 partitions, I_4x4 in P,
                               fine inter rounding, chroma QP offset 1 (enc/inter.rs, enc/compact.rs);
                               Constrained Baseline only (with CABAC the preset codes B pictures) */
    RGH_SPEED_FAST = 2,     /* Balanced records with the Fast decisions; slices without the loop
                               filter (disable_deblocking_filter_idc 1, enc.rs:2334) */
    RGH_SPEED_PARTITIONS = 3 /* Balanced with P partitions (RGH_FLAG_PARTITIONS): records in the
                                quadrant layout */
} rgh_speed;

/* Fixed 32-byte decision record per macroblock, raster order. */
typedef struct rgh_mb_decision {
    uint8_t mb_type;       /* RGH_MB_* */
    uint8_t i16_mode;      /* I_16x16 prediction mode (0 V, 1 H, 2 DC, 3 plane) */
    uint8_t chroma_mode;   /* intra chroma prediction mode (0 DC, 1 H, 2 V, 3 plane) */
    uint8_t cbp_luma;      /* coded_block_pattern luma bits (8x8 quadrants) */
    uint8_t cbp_chroma;    /* 0, 1 (DC only) or 2 (DC + AC) */
    uint8_t transform_8x8; /* transform_size_8x8_flag of a coded P_L0_16x16 */
    uint8_t qp;            /* QP_Y of the macroblock */
    uint8_t flags;         /* bit 0: P_Skip found by the probe before the search */
    int16_t mv[2];         /* quarter-pel vector (P_Skip: the skip vector; I_16x16: 0) */
    int16_t mvp[2];        /* motion vector predictor; mvd = mv - mvp. For I_4x4 (IDR
                              pictures) mv and mvp together hold the 16 intra 4x4 prediction
                              modes, 4 bits each: block b (decoding order) in bits 4b..4b+3 of
                              the little-endian 64-bit word at mv[0] */
    uint32_t cost;         /* motion-estimation cost (SAD + lambda * mv bits) of the result */
    uint16_t n_sad;        /* full-pel 16x16 SAD evaluations */
    uint16_t n_subpel;     /* sub-pel SAD evaluations */
    uint16_t n_satd;       /* 16x16 SATD evaluations */
    uint16_t n_hpel;       /* half-pel windows interpolated */
    uint16_t nnz;          /* 4x4 luma blocks holding levels, bit 4*by+bx */
    uint16_t reserved;
} rgh_mb_decision;
/* Compact-preset P pictures use the words from mv on differently: the 32-bit words at byte
 * offsets 8, 12, 16 and 20 hold the quarter-pel vectors (x, y as int16) of the four 8x8
 * quadrants in raster order (all four equal for P_Skip and P_L0_16x16, zero for intra; I_4x4
 * modes as above in the first two), the word at offset 24 the macroblock's 16x16 search vector
 * (the next picture's candidate, enc/inter.rs:549/:610 mvr). No predictor or counters are
 * recorded: the syntax writer derives the predictors (8.4.1.3) from the quadrant vectors. */

/* Quantized levels per macroblock (int16 elements), the reference's scan layout:
 *   [0, 256)   luma: 16 blocks (decoding order) x 16 zig-zag levels; I_16x16 AC blocks keep
 *              slot 0 empty; with the 8x8 transform, 4 blocks x 64 levels (8x8 zig-zag)
 *   [256, 272) I_16x16 luma DC, zig-zag
 *   [272, 280) chroma DC, U then V (2x2 raster)
 *   [280, 408) chroma AC, U then V, 4 blocks x 16 zig-zag levels (slot 0 empty) */
#define RGH_LEVELS_PER_MB 408

/* Creates a context for one stream. Allocates every device buffer it will use.
 * Replaces: H264Encoder::new(..).with_profile(..) (h264.rs:3032-3065) and
 * EncState::prepare's allocation (enc.rs:223-300). */
rgh_status rgh_create(const rgh_config *cfg, void **ctx_out);

/* Frees the context. */
rgh_status rgh_destroy(void *ctx);

/* Installs the reconstruction of an IDR picture coded elsewhere as the reference: the picture
 * (coded size, already deblocked) is padded by edge replication and every previous-frame
 * motion vector is reset to zero (an IDR frame is all intra).
 * Replaces: the post-IDR part of encode_frame_body (enc.rs:2470-2503). */
rgh_status rgh_set_reference(void *ctx, const rgh_i420 *recon);

/* Analyses one P frame whose I420 source is in host memory. Runs every macroblock's
 * decision, residual quantization and reconstruction on the GPU in wavefront order, then
 * deblocks the reconstruction, extends its borders and makes it the next reference.
 * Any output pointer may be NULL. decisions: mb_w*mb_h records; levels: mb_w*mb_h *
 * RGH_LEVELS_PER_MB int16; recon: the deblocked reconstruction at coded size.
 * Replaces: the P-frame body of encode_frame_body (enc.rs:2382-2434 with code_p_mb
 * :3011-3191) and its post-frame reference preparation (:2470-2503). */
rgh_status rgh_analyse_p_frame(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions,
                               int16_t *levels, const rgh_i420_out *recon);

/* Analyses an IDR picture (host source): every macroblock's intra decision (I_16x16 analysis,
 * then I_4x4 when the 16x16 SATD exceeds 40 lambda and it is cheaper), levels and
 * reconstruction, then deblocking and border extension; the result becomes the reference and
 * the co-located vectors of the next P frame are zero. Outputs as rgh_analyse_p_frame.
 * Replaces: the IDR body of encode_frame_body (enc.rs:2382-2434 with code_intra :3311,
 * analyse_i4 :3333, code_i4 :3441, code_i16 :3587) and its post-frame reference preparation
 * (:2470-2503). */
rgh_status rgh_analyse_idr(void *ctx, const rgh_i420 *src, rgh_mb_decision *decisions,
                           int16_t *levels, const rgh_i420_out *recon);

/* Batched form: one P frame of each of n contexts, whose sources are already in device
 * memory, analysed by one set of kernel launches on one HIP stream (all contexts must have
 * been created on the same device and use one preset: Balanced, Compact or Fast,
 * RGH_ERR_INVALID_ARGUMENT otherwise). Results stay on the device; read them with
 * rgh_get_outputs. No counterpart in the reference, whose GOP workers run one encoder per
 * thread (crates/ragdoll-pipeline/src/lib.rs:543-603); this is that wave of GOPs on one GPU. */
rgh_status rgh_analyse_batch(void *const *ctxs, int32_t n, const rgh_i420 *src_device);

/* Pipelined form: n_frames consecutive P frames of each of n contexts (same device and one
 * preset, as rgh_analyse_batch), sources
 * and outputs in device memory, analysed by one schedule of kernel launches. Frame k + 1 of a
 * stream starts `lag` wavefront steps after frame k (0 selects 20) and each of its macroblocks
 * proceeds as soon as the reference samples it can read are final, so up to ring - 1 frames
 * of a stream are in flight at once; the results are identical to analysing the frames one by
 * one. src_device[c * n_frames + k] is frame k of context c; decisions[c * n_frames + k]
 * receives its mb_w*mb_h records and levels[c * n_frames + k] its mb_w*mb_h*RGH_LEVELS_PER_MB
 * levels (device pointers, kept until the call returns; both 8-byte aligned, as hipMalloc
 * returns them). groups > 1 splits the contexts into that
 * many schedules (context i in group i % groups) whose launches run concurrently on their own
 * streams, so one schedule's slowest macroblocks do not leave the device idle (0 selects 1).
 * With RGH_FRAMES_FIRST_IDR in flags, frame 0 of every context is analysed as an IDR picture
 * (as rgh_analyse_idr), so a whole GOP runs in one call; otherwise every context needs a
 * reference.
 * The last frame's reconstruction becomes
 * each context's reference. Returns after the work completes. No counterpart in the reference,
 * which encodes a GOP's frames one after another on one thread
 * (crates/ragdoll-pipeline/src/lib.rs:575-590); x264's frame threads are the analogue. */
#define RGH_FRAMES_FIRST_IDR 1 /* rgh_analyse_frames: frame 0 of every context is an IDR picture */
/* rgh_analyse_frames: run the call as one persistent launch instead of the per-step launch
 * schedule (one kernel launch per wavefront step, in `groups` concurrent schedules). Each stream
 * is homed on one XCD whose waves take its items (macroblock analyses, deblock/pad items) in
 * planned order and wait for the items they read; if an item cannot wait (it needs one planned
 * after it) or the launch gives up, the per-step schedule finishes the call. Results are
 * identical. It serves Balanced, Fast and IDR pictures: calls with Compact contexts, and
 * rgh_analyse_frames_noalloc, use the per-step schedule. On MI355X it is slower than the per-step
 * schedule on the measured workloads (the waits and their cache invalidations cost more than the
 * launches they save), so it is not the default. */
#define RGH_FRAMES_PERSISTENT 2
/* rgh_analyse_frames: issue each schedule's planned launches as one HIP graph on its stream,
 * captured on the first call and replayed by later calls whose launches are the same (same
 * contexts and groups, frame counts, sizes and lag; the job descriptors are uploaded before each
 * replay). The launches that finish deferred frames after the planned ones are issued one by one
 * as always. Results are identical. Not with rgh_analyse_frames_noalloc
 * (RGH_ERR_INVALID_ARGUMENT): the graph is allocated. A schedule's graph is kept by its first
 * context until that context is destroyed or captures another. */
#define RGH_FRAMES_GRAPH 4
/* rgh_analyse_frames: RGH_FRAMES_KEYINT(k) with RGH_FRAMES_FIRST_IDR, 1 <= k <= 32767: frames 0, k,
 * 2k, ... of every context are IDR pictures (H264Encoder's keyint, h264.rs:3032-3065: the frame
 * loop of encode_frame_body, enc.rs:2234, codes frame_count % keyint == 0 as an IDR), so a context
 * runs several GOPs back to back in one call, each GOP's frames following the last ones of the
 * GOP before at the call's lag. Every GOP's results equal those of the GOP analysed alone from its
 * IDR. Such calls run on the per-step schedule (RGH_FRAMES_PERSISTENT is ignored). */
#define RGH_FRAMES_KEYINT(k) ((int32_t)(k) << 16)
rgh_status rgh_analyse_frames(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src_device,
                              rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag,
                              int32_t groups, int32_t flags);

/* rgh_analyse_frames without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device or pinned memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors as rgh_analyse_frames;
 * the call's job descriptors and progress records live in a fixed scratch array of the library
 * on each device (about 0.5 MB, held until the process ends) instead of buffers it allocates,
 * so one call takes at most RGH_NOALLOC_MAX_FRAMES frames in total (n * n_frames; more returns
 * RGH_ERR_INVALID_ARGUMENT). All calls on one device share that array: calls that overlap on
 * one device race and must be serialised by the caller. The contexts must exist (rgh_create
 * allocates their buffers once). */
#define RGH_NOALLOC_MAX_FRAMES 2048
rgh_status rgh_analyse_frames_noalloc(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src_device,
                                      rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag,
                                      int32_t groups, int32_t flags);

/* Slice parameters of rgh_write_p_slice_cavlc: the picture's macroblock grid, the slice QP
 * (equal to the PPS's initial QP and to every macroblock's), frame_num and pic_order_cnt_lsb as
 * the encoder numbers them inside a GOP (k & 15 and 2k & 63 for the k-th picture after the
 * IDR; log2_max_frame_num 4, log2_max_pic_order_cnt_lsb 6). */
typedef struct rgh_slice_config {
    int32_t mb_w;
    int32_t mb_h;
    int32_t qp;
    int32_t frame_num;
    int32_t poc_lsb;
    int32_t speed; /* rgh_speed of the records */
} rgh_slice_config;

/* Entropy-codes one analysed P picture (decisions and levels in host memory, as
 * rgh_get_outputs returns them) as the Constrained Baseline CAVLC slice the reference writes:
 * one Annex B NAL unit (start code, nal_ref_idc 3, type 1, emulation prevention) into out.
 * *out_len receives the size; RGH_ERR_INVALID_ARGUMENT if it exceeds cap, RGH_ERR_UNSUPPORTED
 * for records CAVLC cannot carry here (8x8 transform, a QP other than the slice's).
 * Replaces: the P slice of encode_frame_body (enc.rs:2234-2330 header, mb_skip_run and
 * code_inter / code_i16 syntax :2779-3306, write_cavlc_residuals :976, write_cavlc_chroma
 * :1009, h264.rs:2215 encode_cavlc_block, NAL escaping :2643-2672). */
rgh_status rgh_write_p_slice_cavlc(const rgh_slice_config *cfg, const rgh_mb_decision *decisions,
                                   const int16_t *levels, uint8_t *out, int64_t cap, int64_t *out_len);

/* The IDR counterpart of rgh_write_p_slice_cavlc (frame_num / poc_lsb are 0 for an IDR and
 * are ignored): one Annex B NAL unit of type 5 with idr_pic_id (the GOP's index, as
 * encode_frame_body numbers IDRs: frame_count / keyint). Intra 4x4 modes come from the I_4x4
 * records' mv words. Replaces: the IDR slice of encode_frame_body (enc.rs:2234-2330 header,
 * code_i4 :3441 / code_i16 :3587 syntax, the predicted intra 4x4 mode of 8.3.1.1). */
rgh_status rgh_write_idr_slice_cavlc(const rgh_slice_config *cfg, int32_t idr_pic_id,
                                     const rgh_mb_decision *decisions, const int16_t *levels,
                                     uint8_t *out, int64_t cap, int64_t *out_len);

/* Entropy-codes one analysed P picture with CABAC as the Main (RGH_PROFILE_MAIN) or High
 * (RGH_PROFILE_HIGH: with transform_size_8x8_flag) slice the reference writes: one Annex B NAL
 * unit (nal_ref_idc 3, type 1) holding the slice header with cabac_init_idc 0, the
 * cabac_alignment_one_bits and the arithmetic-coded macroblocks, end_of_slice_flag after each.
 * Records as rgh_write_p_slice_cavlc (P_Skip, P_L0_16x16 with or without the 8x8 transform,
 * I_16x16; one QP); the Compact preset is RGH_ERR_UNSUPPORTED (with CABAC it codes B pictures).
 * *out_len receives the size; RGH_ERR_INVALID_ARGUMENT if it exceeds cap.
 * Replaces: the P slice of encode_frame_body with CABAC (enc.rs:2346-2440; begin_coded_mb /
 * code_skip / code_inter / code_i16 syntax :2779-3306, :3587-3720; h264.rs:3327-3552
 * encode_cabac_cbp_and_residuals; cabac.rs:504-1130 CabacEncoder). */
rgh_status rgh_write_p_slice_cabac(const rgh_slice_config *cfg, int32_t profile,
                                   const rgh_mb_decision *decisions, const int16_t *levels,
                                   uint8_t *out, int64_t cap, int64_t *out_len);

/* The IDR counterpart of rgh_write_p_slice_cabac: one NAL unit of type 5 with idr_pic_id, the
 * I_4x4 (transform_size_8x8_flag 0 under High) and I_16x16 records of rgh_analyse_idr.
 * Replaces: the IDR slice of encode_frame_body with CABAC (code_i4 :3441, code_i16 :3587). */
rgh_status rgh_write_idr_slice_cabac(const rgh_slice_config *cfg, int32_t profile, int32_t idr_pic_id,
                                     const rgh_mb_decision *decisions, const int16_t *levels,
                                     uint8_t *out, int64_t cap, int64_t *out_len);

/* Stream parameters of rgh_write_headers. */
typedef struct rgh_stream_config {
    int32_t width, height; /* display size */
    int32_t profile;       /* rgh_profile */
    int32_t qp;            /* the PPS's initial QP (the fixed QP) */
    int32_t fps_num, fps_den;
    int32_t speed;         /* rgh_speed: chroma_qp_index_offset 1 for Compact */
} rgh_stream_config;

/* The stream's SPS and PPS as two Annex B NAL units (types 7 and 8), as the encoder emits them
 * before each IDR. Replaces: h264.rs:556 Sps::to_rbsp with the encoder's profile settings and
 * h264.rs:806 Pps::to_rbsp. */
rgh_status rgh_write_headers(const rgh_stream_config *cfg, uint8_t *out, int64_t cap, int64_t *out_len);

/* Copies the last analysed frame's results to host memory (any pointer may be NULL). After
 * rgh_analyse_frames the levels are read from that call's last levels buffer, which must still
 * be allocated.
 * which_recon: 0 = before deblocking, 1 = deblocked reference. The undeblocked copy exists only
 * after the single-frame entry points (rgh_analyse_p_frame, rgh_analyse_idr, rgh_analyse_batch)
 * of a context created without RGH_FLAG_NO_UNDEBLOCKED_COPY; otherwise which_recon = 0 returns
 * RGH_ERR_INVALID_ARGUMENT.
 * An analysis call that fails after its launches may have started leaves its contexts without a
 * reference (the next P frame returns RGH_ERR_NO_REFERENCE until rgh_set_reference or an IDR) and
 * without outputs to read (RGH_ERR_NO_REFERENCE here). */
rgh_status rgh_get_outputs(void *ctx, rgh_mb_decision *decisions, int16_t *levels,
                           const rgh_i420_out *recon, int32_t which_recon);

/* Macroblock grid of a context. */
rgh_status rgh_get_mb_dims(void *ctx, int32_t *mb_w, int32_t *mb_h);

/* Waits for the context's queued GPU work. */
rgh_status rgh_synchronize(void *ctx);

/* HIP stream (hipStream_t) the context's work runs on, for callers that time it. */
void *rgh_get_stream(void *ctx);

/* Device-time measurement (HIP events on the context's stream). When enabled on the first
 * context of a call, each analysis call records events around its schedule of frame-step
 * launches (analysis, deblocking and border extension fused); rgh_get_timing waits for the last
 * such call and returns its span in milliseconds, 0 for deblock_ms (no separate deblock pass),
 * and the number of frame-step launches. No counterpart in the reference (its telemetry is
 * wall-clock PipelineTelemetry, crates/ragdoll-pipeline/src/lib.rs:62-76). */
rgh_status rgh_enable_timing(void *ctx, int32_t enable);
rgh_status rgh_get_timing(void *ctx, float *analyse_ms, float *deblock_ms, int32_t *launches);

/* A human-readable description of a status code. */
const char *rgh_status_str(int32_t status);

/* The calling thread's last failure in detail (the failing runtime call and its message), or
 * "" — the text RagdollError::Internal(String) would carry (crates/ragdoll-core/src/error.rs). */
const char *rgh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
