// Constant tables of ITU-T H.264 and of the Balanced encoder, for the device code.
// Sources (ragdoll repository): crates/ragdoll-codec/src/h264.rs:1174-1420 (zig-zag, quantizer
// and dequantizer tables, chroma QP), h264/enc.rs:64-67 (x264 lambda) and :88-92 (decimation
// weights), crates/ragdoll-simd/src/h264_dsp.rs:16-40 (deblocking thresholds). These are the
// standard's tables; the values must be identical to the reference's.
#pragma once
#include <stdint.h>

#define RGH_TABLE_QUALIFIER __constant__

RGH_TABLE_QUALIFIER uint16_t c_lambda[52] = {
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6, 6, 7, 8, 9,
    10, 11, 13, 14, 16, 18, 20, 23, 25, 29, 32, 36, 40, 45, 51, 57, 64, 72, 81, 91};

RGH_TABLE_QUALIFIER uint8_t c_decimate4[16] = {3, 2, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
RGH_TABLE_QUALIFIER uint8_t c_decimate8[64] = {3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1,
                                               1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0};

RGH_TABLE_QUALIFIER uint8_t c_zigzag4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
RGH_TABLE_QUALIFIER uint8_t c_zigzag8[64] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// LevelScale4x4 normAdjust (v0, v1, v2) and forward multipliers per QP % 6.
RGH_TABLE_QUALIFIER int32_t c_dequant4_init[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                                                     {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
RGH_TABLE_QUALIFIER int32_t c_quant4_mf[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490},
                                                 {10082, 4194, 6554}, {9362, 3647, 5825},
                                                 {8192, 3355, 5243},  {7282, 2893, 4559}};
// 8x8 (High profile): normAdjust8x8 and x264's quant8 multipliers, by position class.
RGH_TABLE_QUALIFIER int32_t c_dequant8_init[6][6] = {
    {20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
    {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};
RGH_TABLE_QUALIFIER int32_t c_quant8_mf[6][6] = {
    {13107, 11428, 20972, 12222, 16777, 15481}, {11916, 10826, 19174, 11058, 14980, 14290},
    {10082, 8943, 15978, 9675, 12710, 11985},   {9362, 8228, 14913, 8931, 11984, 11259},
    {8192, 7346, 13159, 7740, 10486, 9777},     {7282, 6428, 11570, 6830, 9118, 8640}};

RGH_TABLE_QUALIFIER uint8_t c_chroma_qp[52] = {
    0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17,
    18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 29, 30, 31, 32, 32, 33,
    34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

// QP'c of luma QP qp under the picture parameter set's chroma_qp_index_offset (enc.rs:540).
__device__ __forceinline__ int chroma_qp_of(int qp, int cqo) {
    const int q = qp + cqo;
    return c_chroma_qp[q < 0 ? 0 : (q > 51 ? 51 : q)];
}

RGH_TABLE_QUALIFIER uint8_t c_alpha[52] = {
    0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   4,   4,
    5,  6,  7,  8,  9,  10, 12, 13, 15, 17, 20, 22, 25, 28, 32,  36,  40,  45,
    50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
RGH_TABLE_QUALIFIER uint8_t c_beta[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                                          0,  0,  0,  2,  2,  2,  3,  3,  3,  3,  4,  4,  4,
                                          6,  6,  7,  7,  8,  8,  9,  9,  10, 10, 11, 11, 12,
                                          12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
RGH_TABLE_QUALIFIER uint8_t c_tc0[52][3] = {
    {0, 0, 0},    {0, 0, 0},    {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},
    {0, 0, 0},    {0, 0, 0},    {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},
    {0, 0, 0},    {0, 0, 0},    {0, 0, 0},   {0, 0, 1},   {0, 0, 1},   {0, 0, 1},   {0, 0, 1},
    {0, 1, 1},    {0, 1, 1},    {1, 1, 1},   {1, 1, 1},   {1, 1, 1},   {1, 1, 1},   {1, 1, 2},
    {1, 1, 2},    {1, 1, 2},    {1, 1, 2},   {1, 2, 3},   {1, 2, 3},   {2, 2, 3},   {2, 2, 4},
    {2, 3, 4},    {2, 3, 4},    {3, 3, 5},   {3, 4, 6},   {3, 4, 6},   {4, 5, 7},   {4, 5, 8},
    {4, 6, 9},    {5, 7, 10},   {6, 8, 11},  {6, 8, 13},  {7, 10, 14}, {8, 11, 16}, {9, 12, 18},
    {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};
