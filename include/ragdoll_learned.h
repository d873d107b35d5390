/* ragdoll_learned.h — learned motion proposals for the P-frame analysis of ragdoll_hip.h.
 *
 * No counterpart in the reference, whose motion search is hand-written
 * (crates/ragdoll-codec/src/enc.rs:1182 full_pel, :3081-3095 candidates). This stage adds a
 * small network that looks at each macroblock and proposes where its motion is, so the search
 * can start there and stop early:
 *
 *   features (per macroblock, from the source picture and the previous source picture)
 *     -> proposal network (correlation features -> 192 -> 128 -> 128 -> 176, ReLU)
 *     -> four candidate displacements + skip/inter/intra probabilities (rgh_proposal)
 *     -> rgh_analyse_frames_learned: integer SAD/SATD/RD arithmetic unchanged, the candidate
 *        list and the search effort taken from the proposal, the plain search as fallback.
 *
 * Training runs on the GPU in bfloat16 (MFMA, fp32 accumulation) through rgh_pnet_train_step;
 * inference runs in FP8 (OCP E4M3, MFMA, fp32 accumulation) through rgh_pnet_infer after
 * rgh_pnet_quantize. All pointers named *_device are device memory; calls return after their
 * work completes unless stated otherwise. A network object is used by one thread at a time.
 */
#ifndef RAGDOLL_LEARNED_H
#define RAGDOLL_LEARNED_H

#include <stdint.h>

#include "ragdoll_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Feature vector of one macroblock (bfloat16 bit patterns, row-major [n][192]):
 *   [0, 169)    cost volume: for dy, dx in [-6, 6] (4-sample cells), the sum over the 4x4
 *               cells of the macroblock's 4x4-mean image of |cell - previous picture's 4x4-mean
 *               at the same cell displaced by (dx, dy) cells|, clamped to 2047, divided by 256;
 *               index (dy + 6) * 13 + (dx + 6). The previous picture's means cover the 64x64
 *               window centred on the macroblock (samples clamped to the picture).
 *   [169, 185)  the macroblock's 4x4-mean cells / 64, raster order.
 *   [185]       mean absolute deviation of the macroblock's luma from its mean / 16.
 *   [186]       the picture's QP / 32.
 *   [187, 192)  zero. */
#define RGH_PNET_FEATURES 192
#define RGH_PNET_HIDDEN 128
#define RGH_PNET_OUTPUTS 176 /* 169 displacement logits, 3 mode logits, 4 unused */
#define RGH_PNET_CELLS 13    /* the displacement grid is 13 x 13 cells of 4 samples: +-24 samples */
#define RGH_PNET_MODE_OFFSET 169 /* index of the first mode logit */
/* Parameters (fp32, flat): W1 [192][128], b1 [128], W2 [128][128], b2 [128], W3 [128][176], b3 [176]. */
#define RGH_PNET_PARAMS 63920

/* Label of one training macroblock: the displacement class (cell (dx, dy) -> (dy+6)*13+dx+6,
 * -1 = no motion label, for intra macroblocks) and the mode (0 P_Skip, 1 inter, 2 intra). */
typedef struct rgh_pnet_label {
    int16_t disp;
    int16_t mode;
} rgh_pnet_label;

/* One macroblock's proposal (16 B). */
typedef struct rgh_proposal {
    int8_t cell[4][2];     /* candidate displacements (x, y) in 4-sample cells, most likely first */
    uint8_t mode_prob[3];  /* softmax of the mode logits x 255, rounded: skip, inter, intra */
    uint8_t reserved[5];
} rgh_proposal;

/* Creates a network on a device with zero parameters; max_batch bounds the batch of
 * rgh_pnet_train_step / rgh_pnet_gradients. */
rgh_status rgh_pnet_create(int32_t device, int32_t max_batch, void **net);
rgh_status rgh_pnet_destroy(void *net);
/* Caps the workgroups of the inference launch (0: one per CU). Each holds 151 KB of LDS, so a
 * cap leaves the other CUs to an analysis running beside the proposals on another stream. */
rgh_status rgh_pnet_set_launch_limit(void *net, int32_t max_workgroups);
/* Parameters in the flat fp32 layout above (host memory); setting them resets the Adam state
 * and the quantized copy. */
rgh_status rgh_pnet_set_params(void *net, const float *params, int64_t n);
rgh_status rgh_pnet_get_params(void *net, float *params, int64_t n);

/* Feature vectors of every macroblock of a width x height luma picture (macroblock grid
 * ceil(width/16) x ceil(height/16), raster order) from the picture and the previous picture
 * (device luma planes). */
rgh_status rgh_pnet_features(void *net, const uint8_t *cur_y_device, int32_t cur_stride,
                             const uint8_t *prev_y_device, int32_t prev_stride, int32_t width,
                             int32_t height, int32_t qp, uint16_t *features_device);

/* One picture of rgh_pnet_features_batch. */
typedef struct rgh_pnet_picture {
    const uint8_t *cur_y;   /* device luma plane of the picture */
    const uint8_t *prev_y;  /* device luma plane of the previous picture */
    uint16_t *features;     /* device rows, ceil(width/16) * ceil(height/16) of them */
    int32_t cur_stride, prev_stride, width, height, qp, reserved;
} rgh_pnet_picture;

/* rgh_pnet_features for n pictures in one launch (pics: host array). */
rgh_status rgh_pnet_features_batch(void *net, const rgh_pnet_picture *pics, int32_t n);

/* Labels from analysed macroblocks (rgh_mb_decision records of ragdoll_hip.h): the nearest
 * cell to the chosen vector (clamped to the grid), and the mode. */
rgh_status rgh_pnet_labels(void *net, const rgh_mb_decision *decisions_device, int32_t n,
                           rgh_pnet_label *labels_device);

/* Copies rows idx[i] of features/labels into rows i of the outputs (batch assembly). */
rgh_status rgh_pnet_gather(void *net, const uint16_t *features_device, const rgh_pnet_label *labels_device,
                           const int32_t *idx_device, int32_t n, uint16_t *features_out_device,
                           rgh_pnet_label *labels_out_device);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the mean cross-entropy of the displacement
 * head (rows with disp >= 0) and the mode head over a batch. Asynchronous unless loss (host,
 * [2]: summed displacement and mode cross-entropy of the batch) is not NULL. */
rgh_status rgh_pnet_train_step(void *net, const uint16_t *features_device, const rgh_pnet_label *labels_device,
                               int32_t batch, float lr, float *loss);

/* Gradients of the same loss at the current parameters without updating them (host outputs:
 * grads [RGH_PNET_PARAMS], logits [batch][176]; either may be NULL). */
rgh_status rgh_pnet_gradients(void *net, const uint16_t *features_device, const rgh_pnet_label *labels_device,
                              int32_t batch, float *grads, float *logits);

/* bfloat16 forward pass (the training arithmetic): logits_device [n][176] fp32. */
rgh_status rgh_pnet_forward(void *net, const uint16_t *features_device, int32_t n, float *logits_device);

/* Builds the FP8 copy: per-output-column weight scales (max |w| / 448) and per-layer activation
 * scales from the largest activation over the calibration features (n rows). */
rgh_status rgh_pnet_quantize(void *net, const uint16_t *calib_features_device, int32_t n);

/* rgh_pnet_quantize with the two hidden layers' activation maxima given (as a previous
 * calibration found them: rgh_pnet_get_quantized's multipliers m give max = 448 / m). */
rgh_status rgh_pnet_quantize_ranges(void *net, float h1_max, float h2_max);

/* FP8 inference of n macroblocks: proposals_device [n]; logits_device [n][176] fp32 may be NULL. */
rgh_status rgh_pnet_infer(void *net, const uint16_t *features_device, int32_t n, rgh_proposal *proposals_device,
                          float *logits_device);

/* Features and FP8 inference of n pictures in one pass (pics: host array; their `features`
 * fields are not used): the proposals of picture i's macroblocks go to consecutive rows of
 * proposals_device, picture after picture (ceil(width/16) * ceil(height/16) rows each), and
 * logits_device [rows][176] fp32 may be NULL. Equals rgh_pnet_features_batch followed by
 * rgh_pnet_infer over the same rows, bit for bit, without the feature rows in device memory.
 * No counterpart in the reference. */
rgh_status rgh_pnet_propose_batch(void *net, const rgh_pnet_picture *pics, int32_t n,
                                  rgh_proposal *proposals_device, float *logits_device);

/* The quantized network (host outputs, either may be NULL): w8 [RGH_PNET_W8_BYTES] bytes (E4M3
 * codes, layers in order, each [out][in]); scales [RGH_PNET_SCALES]: per-column dequantization
 * scales of each layer (176 slots each), then the three layers' input multipliers (an input x
 * is coded as E4M3(min(x * multiplier, 448)); layer l's output column n is
 * accumulator * scale[l][n] + bias). */
#define RGH_PNET_W8_BYTES (192 * 128 + 128 * 128 + 176 * 128)
#define RGH_PNET_SCALES (3 * 176 + 3)
rgh_status rgh_pnet_get_quantized(void *net, uint8_t *w8, float *scales);

/* The device's float -> E4M3 conversion (round to nearest even), as inference codes
 * activations, exported for its tests: |x| <= 448 (device pointers). */
rgh_status rgh_pnet_fp8_encode(void *net, const float *x_device, int32_t n, uint8_t *out_device);

/* Search effort of rgh_analyse_frames_learned (zero fields select the defaults). */
typedef struct rgh_learned_config {
    int32_t iters;        /* diamond iterations after the proposals (default 4) */
    int32_t skip_gate;    /* the skip probe runs when mode_prob[0] >= skip_gate (default 0: always) */
    int32_t intra_gate;   /* the intra 16x16 check runs when mode_prob[2] >= intra_gate (default 0: always) */
    int32_t fallback_pct; /* the plain search runs when the cheapest proposal costs more than
                             fallback_pct percent of the predictor's cost (default 105) */
    int32_t intra_first;  /* when mode_prob[2] >= intra_first, the macroblock keeps its cheapest
                             full-pel candidate (no diamond, no sub-pel refinement) before the
                             intra 16x16 check (default 0: never) */
} rgh_learned_config;

/* rgh_analyse_frames (ragdoll_hip.h) with learned proposals: proposals[c * n_frames + k] holds
 * mb_w * mb_h records for frame k of context c (device memory; NULL only for an IDR frame 0
 * under RGH_FRAMES_FIRST_IDR). Every P macroblock's integer search starts from the predictor
 * and the four proposed cells; when the cheapest proposal costs more than fallback_pct percent
 * of the predictor's cost, the reference's candidates join and the diamond runs its full count
 * (the plain search), otherwise at most `iters` iterations; the skip probe and the intra check
 * run as gated above. SAD, SATD and rate-distortion arithmetic, sub-pel refinement, residual
 * coding, reconstruction and deblocking are the plain path's. Decision flags: bit 1 the plain
 * search ran, bit 2 the skip probe was gated off, bit 3 the intra check was gated off, bit 4
 * the search stopped at full-pel (intra_first).
 * Balanced contexts only (RGH_ERR_UNSUPPORTED for RGH_FLAG_COMPACT). */
rgh_status rgh_analyse_frames_learned(void *const *ctxs, int32_t n, int32_t n_frames, const rgh_i420 *src_device,
                                      const rgh_proposal *const *proposals, const rgh_learned_config *config,
                                      rgh_mb_decision *const *decisions, int16_t *const *levels, int32_t lag,
                                      int32_t groups, int32_t flags);

/* The bfloat16 GEMM behind training, exported for its tests: C[m][n] = sum_k A(m,k) * B(k,n)
 * with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bfloat16 bit patterns), then
 * (+ C if accumulate) (+ bias[n]) (ReLU if relu) (0 where mask[m*ldc+n] == 0 if mask), written
 * to C (fp32) and/or Cb (bfloat16), leading dimension ldc. Device pointers; on the net's stream. */
rgh_status rgh_pnet_gemm_bf16(void *net, const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                              int64_t sbk, int64_t sbn, float *C, uint16_t *Cb, int64_t ldc, const float *bias,
                              const uint16_t *mask, int32_t M, int32_t N, int32_t K, int32_t relu,
                              int32_t accumulate);

#ifdef __cplusplus
}
#endif

#endif
