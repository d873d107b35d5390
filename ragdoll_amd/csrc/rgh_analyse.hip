// P-frame macroblock analysis on gfx950: one wave64 per macroblock.
//
// Each function restates a piece of ragdoll's Balanced P path; file:line citations are
// relative to the ragdoll repository ("enc.rs" = crates/ragdoll-codec/src/h264/enc.rs,
// "h264.rs" = crates/ragdoll-codec/src/h264.rs, "simd/" = crates/ragdoll-simd/src/).
//
// Layout and latency: the macroblock's wave keeps its 16x16 source as 64 lanes x 4 packed
// bytes (lane l: row l >> 2, columns 4 (l & 3) .. +3) and its 8x8 chroma as one sample per lane.
// Global memory is touched in a few batched round trips: sources, neighbour decisions and
// intra edges up front; the skip-probe prediction; all integer-search candidates at once; one
// staged reference region (LDS) that serves the whole diamond search, the half-pel window and
// the quarter-pel candidates; the chroma prediction. Reductions are DPP + readlane; transform
// and quantization run one 4x4 block per lane in registers, 8x8 blocks one row or column per
// lane through LDS.
#include <stddef.h>

#include "rgh_deblock.h"
#include "rgh_device.h"
#include "rgh_tables.h"

namespace {

constexpr int ZZ4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
constexpr int ZZ8[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

constexpr int SR_W = 72;   // staged reference region: stride (68 columns used)
constexpr int SR_COLS = 68;
constexpr int SR_H = 55;   // rows
constexpr int WP_ROWS = 18;  // half-pel plane rows (simd/h264_enc.rs:710 HPEL_WIN_H)
constexpr int WP_COLS = 24;  // half-pel plane columns (HPEL_COLS)
constexpr int HR_ROWS = 23;  // unrolled horizontal-tap rows

// LDS of one macroblock's wave. The motion-search arrays (staged reference region, half-pel
// window) and the intra/residual arrays are never live together (a barrier separates the last
// prediction read from the region and the first residual write), so they share storage.
struct MbLds {
    alignas(16) uint8_t src[256], srcu[64], srcv[64];
    alignas(16) uint8_t pred[256], predu[64], predv[64];
    alignas(16) uint8_t etop[16], eleft[16], ctop[2][8], cleft[2][8];
    alignas(16) int32_t etl, ctl[2];
    alignas(16) uint32_t nb[4];  // Fast: costs of the last diamond iteration's four points
    alignas(16) uint8_t pred2[256];  // P partitions: the prediction of the shape being searched
    union {
        struct {
            alignas(16) uint8_t sr[SR_H * SR_W];
            alignas(16) int16_t hr[HR_ROWS * WP_COLS];
            alignas(16) uint8_t ph[WP_ROWS * WP_COLS], pv[WP_ROWS * WP_COLS], pc[WP_ROWS * WP_COLS];
        };
        struct {
            alignas(16) uint8_t cv[17 * 24];   // intra 4x4 reconstruction canvas (IDR), CV_W stride
            alignas(16) int16_t q4i[16][16];   // intra 4x4 levels per block (decoding order), raster
        };
        struct {
            alignas(16) int16_t coef[16][16];  // luma levels, decoding order, raster positions
            alignas(16) int16_t q8[4][64];     // 8x8 levels, raster positions
            alignas(16) int32_t t8[4][64];     // 8x8 transform intermediate
            alignas(16) int16_t craster[2][4][16];
            alignas(16) int16_t cdc_raw[2][4];
            alignas(16) int16_t cdc[2][4];
            alignas(16) int16_t cdcr[2][4];
            alignas(16) int16_t dc_raw[16], dcq[16], dcr[16];
            alignas(16) int32_t bnz[16], bbig[16], bscore[16], cscore[8], cmask[8], b8score[4];
        };
    };
};

__device__ __forceinline__ uint32_t se_bits(int d) {  // enc.rs:546
    // Length of se(v): 2 bit_length(k + 1) - 1 for code number k = 2|d| - (d > 0), with
    // k + 1 = 2d (d > 0) or 1 - 2d as a select: for a uniform d it stays a scalar select (a
    // comparison turned into an integer costs a VALU round trip there).
    const uint32_t k1 = d > 0 ? (uint32_t)(2 * d) : (uint32_t)(1 - 2 * d);
    return 2u * (uint32_t)(32 - __clz((int)k1)) - 1u;
}
__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f) {
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}

// Four bytes at any offset of an LDS row (two aligned reads and a byte align).
__device__ __forceinline__ uint32_t lds4u(const uint8_t *row, int x) {
    const uint32_t *q = (const uint32_t *)(row + (x & ~3));
    return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(x & 3));
}

// enc.rs:583 decimate_masked from the non-zero mask of the scanned levels and whether one of
// them exceeds 1 in magnitude (then the score is 9). The reference walks the set bits from the
// top, adding DECIMATE4[r] (enc.rs:88: 3, 2, 2, 1, 1, 1, 0...) for the run r of zeros below
// each (down to the next set bit, or to the start). Branch-free form: the weight is [r <= 5] +
// [r <= 2] + [r == 0], and [r <= k] holds for set bit p exactly when a set bit lies in
// p-k-1 .. p-1 or p <= k, so each term is a popcount of the mask against itself smeared k+1
// bits upwards (with the k+1 lowest bits set).
__device__ __forceinline__ uint32_t decimate_mask(uint32_t mask, int big) {
    const uint32_t t3 = mask | (mask << 1) | (mask << 2), t6 = t3 | (t3 << 3);
    const uint32_t s = (uint32_t)__builtin_popcount(mask & ((mask << 1) | 1u)) +
                       (uint32_t)__builtin_popcount(mask & ((t3 << 1) | 7u)) +
                       (uint32_t)__builtin_popcount(mask & ((t6 << 1) | 63u));
    return big ? 9u : s;
}
// enc.rs:557 decimate_score of 64 levels (8x8 table) from their zig-zag non-zero mask, in the
// same form: DECIMATE8[r] (3 x4, 2 x8, 1 x12, 0...) = [r <= 23] + [r <= 11] + [r <= 3].
__device__ __forceinline__ uint32_t decimate_mask64(uint64_t mask, int big) {
    const uint64_t t2 = mask | (mask << 1), t4 = t2 | (t2 << 2), t12 = t4 | (t4 << 4) | (t4 << 8), t24 = t12 | (t12 << 12);
    const uint32_t s = (uint32_t)__builtin_popcountll(mask & ((t4 << 1) | 0xFull)) +
                       (uint32_t)__builtin_popcountll(mask & ((t12 << 1) | 0xFFFull)) +
                       (uint32_t)__builtin_popcountll(mask & ((t24 << 1) | 0xFFFFFFull));
    return big ? 9u : s;
}

// Raster position of zig-zag scan index k (0..15) of a 4x4 block (the c_zigzag4 table, one
// nibble per index), without a memory lookup for a lane-varying k.
__device__ __forceinline__ int zz4(int k) { return (int)((0xFEB7ADC963258410ull >> (4 * k)) & 15u); }

// Value of row r (0..3) of the lane's quad (DPP quad broadcast).
template <int R>
__device__ __forceinline__ int quad_row(int v) {
    // The empty asm keeps the broadcast a v_mov_b32_dpp of its own: folded into the arithmetic
    // that consumes it (v_add_u32_dpp / v_subrev_u32_dpp with quad_perm:[R,R,R,R]), the result
    // was wrong on MI355X (the intra 4x4 transforms lost parity with the oracle).
    int r = __builtin_amdgcn_mov_dpp(v, R * 0x55, 0xF, 0xF, true);  // every source lane is valid
    __asm__ volatile("" : "+v"(r));
    return r;
}

// The y-th of four values (y in 0..3, per lane) by masks: a ternary chain on a lane-varying
// index compiles to divergent branches.
__device__ __forceinline__ int sel4(int y, int a, int b, int c, int d) {
    return (a & -(int)(y == 0)) | (b & -(int)(y == 1)) | (c & -(int)(y == 2)) | (d & -(int)(y == 3));
}

// Column pass of a 4x4 transform across the quad: output row y = m0 v0 + m1 v1 + m2 v2 + m3 v3
// over the quad's row values v0..v3 (DPP broadcasts) with the lane's multipliers, as 24-bit
// multiply-adds (every operand here stays within 24 bits: row sums of 8-bit residuals, of
// quantized levels times their scales, or of luma DC values).
// a * b + c on 24-bit signed a, b (v_mad_i32_i24). Written out because the backend cannot see the
// range of a broadcast kept behind quad_row's barrier and would pick the quarter-rate 32-bit
// multiply.
__device__ __forceinline__ int mad24(int a, int b, int c) {
    int r;
    __asm__("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
struct QuadMix {
    int m0, m1, m2, m3;
};
__device__ __forceinline__ int quad_mix(int t, const QuadMix &m) {
    const int v0 = quad_row<0>(t), v1 = quad_row<1>(t), v2 = quad_row<2>(t), v3 = quad_row<3>(t);
    return mad24(v3, m.m3, mad24(v2, m.m2, mad24(v1, m.m1, mad24(v0, m.m0, 0))));
}
// forward_dct4x4 (simd/fixed.rs:104) columns: rows (1,1,1,1), (2,1,-1,-2), (1,-1,-1,1), (1,-2,2,-1).
__device__ __forceinline__ QuadMix fdct_mix(int y) {
    return {y == 1 ? 2 : 1, sel4(y, 1, 1, -1, -2), sel4(y, 1, -1, -1, 2), sel4(y, 1, -2, 1, -1)};
}
// 4-point Hadamard with outputs (1,1,1,1), (1,1,-1,-1), (1,-1,-1,1), (1,-1,1,-1) (the luma DC
// transforms, simd/fixed.rs:318 and h264.rs:1645).
__device__ __forceinline__ QuadMix had_mix(int y) { return {1, y < 2 ? 1 : -1, (y == 0 || y == 3) ? 1 : -1, sel4(y, 1, -1, 1, -1)}; }
// The SATD's Hadamard output order (simd/h264_enc.rs:167): (1,1,1,1), (1,1,-1,-1), (1,-1,1,-1),
// (1,-1,-1,1).
__device__ __forceinline__ QuadMix satd_mix(int y) { return {1, y < 2 ? 1 : -1, (y & 1) ? -1 : 1, (y == 0 || y == 3) ? 1 : -1}; }
// inverse_idct4x4 (simd/fixed.rs:185) columns with their halved odd inputs: rows f0 + f2 + f1 +
// (f3 >> 1), f0 - f2 + (f1 >> 1) - f3, f0 - f2 - (f1 >> 1) + f3, f0 + f2 - f1 - (f3 >> 1).
struct IdctMix {
    int outer, s1, s2, s3;  // outer: rows 0 and 3
};
__device__ __forceinline__ IdctMix idct_mix(int y) {
    const int outer = y == 0 || y == 3;
    return {outer, y < 2 ? 1 : -1, outer ? 1 : -1, (y & 1) ? -1 : 1};
}
__device__ __forceinline__ int quad_idct(int t, const IdctMix &m) {
    const int f0 = quad_row<0>(t), f1 = quad_row<1>(t), f2 = quad_row<2>(t), f3 = quad_row<3>(t);
    const int a = m.outer ? f1 : f1 >> 1, b = m.outer ? f3 >> 1 : f3;
    return mad24(b, m.s3, mad24(a, m.s1, mad24(f2, m.s2, f0)));
}

// ---------------------------------------------------------------------------------------
// 4x4 transforms, one row per lane: lanes 4k..4k+3 hold rows 0..3 of block k; the column
// butterflies cross the quad with quad_row broadcasts. Same arithmetic as the block forms below.
// ---------------------------------------------------------------------------------------
// Coefficient row `row` of forward_dct4x4 (simd/fixed.rs:104) of src - pred, block at (x, y0).
__device__ __forceinline__ void fdct4_row(const uint8_t *src, const uint8_t *pred, int s, int x, int y0, int row,
                                          int *o) {
    const uint32_t a = *(const uint32_t *)&src[(y0 + row) * s + x], b = *(const uint32_t *)&pred[(y0 + row) * s + x];
    const int d0 = (int)byte_of(a, 0) - (int)byte_of(b, 0), d1 = (int)byte_of(a, 1) - (int)byte_of(b, 1);
    const int d2 = (int)byte_of(a, 2) - (int)byte_of(b, 2), d3 = (int)byte_of(a, 3) - (int)byte_of(b, 3);
    const int s03 = d0 + d3, s12 = d1 + d2, d03 = d0 - d3, d12 = d1 - d2;
    const int t[4] = {s03 + s12, (d03 << 1) + d12, s03 - s12, d03 - (d12 << 1)};
    const QuadMix m = fdct_mix(row);
#pragma unroll
    for (int c = 0; c < 4; c++) o[c] = (int16_t)quad_mix(t[c], m);
}
// quant4 on coefficient row `row`: position classes 0 / 2 (even rows), 2 / 1 (odd rows) by column.
__device__ __forceinline__ void quant4_row(int *o, int row, int qp, int intra) {
    const int shift = 15 + qp / 6, rem = qp % 6;
    const uint32_t bias = (1u << shift) / (intra ? 3u : 6u);
    const uint32_t mfe = (uint32_t)c_quant4_mf[rem][(row & 1) ? 2 : 0], mfo = (uint32_t)c_quant4_mf[rem][(row & 1) ? 1 : 2];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int v = o[c];
        uint32_t lv = (__umul24((uint32_t)iabs(v), (c & 1) ? mfo : mfe) + bias) >> shift;
        lv = lv > 32767u ? 32767u : lv;
        o[c] = v < 0 ? -(int)lv : (int)lv;
    }
}
// dequant4 on coefficient row `row`.
__device__ __forceinline__ void dequant4_row(int *c, int row, int qp) {
    const int sh = qp / 6, rem = qp % 6;
    const int se = c_dequant4_init[rem][(row & 1) ? 2 : 0] << sh, so = c_dequant4_init[rem][(row & 1) ? 1 : 2] << sh;
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = iclamp(__mul24(c[k], (k & 1) ? so : se), -32768, 32767);
}
// Row `row` of inverse_idct4x4 (simd/fixed.rs:185) of the block whose lane holds coefficient
// row c, added onto row y0 + row of a stride-s LDS block at column x with clipping.
__device__ __forceinline__ void add_idct4_row(uint8_t *dst, int s, int x, int y0, int row, const int *c) {
    const int e0 = c[0] + c[2], e1 = c[0] - c[2], e2 = (c[1] >> 1) - c[3], e3 = c[1] + (c[3] >> 1);
    const int t[4] = {e0 + e3, e1 + e2, e1 - e2, e0 - e3};
    uint32_t *p = (uint32_t *)&dst[(y0 + row) * s + x];
    const uint32_t w = *p;
    uint32_t n = 0;
    const IdctMix m = idct_mix(row);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int o = (int16_t)((quad_idct(t[k], m) + 32) >> 6);
        n |= (uint32_t)clip255((int)byte_of(w, k) + o) << (8 * k);
    }
    *p = n;
}
// Sum of a value over the lane's row of 16 lanes, in every lane of the row (the first half of
// wave_sum).
__device__ __forceinline__ uint32_t row16_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    return v;
}
// OR of a value over the lane's quad.
__device__ __forceinline__ int quad_or(int v) {
    int a = __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    __asm__ volatile("" : "+v"(a));  // kept a v_mov_b32_dpp of its own, as in quad_row
    v |= a;
    int b = __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    __asm__ volatile("" : "+v"(b));
    return v | b;
}

// ---------------------------------------------------------------------------------------
// 4x4 transforms and quantization, one block per lane, in registers.
// ---------------------------------------------------------------------------------------
// simd/fixed.rs:104 forward_dct4x4_scalar of the residual of the 4x4 block at (x, y) of two
// stride-s LDS blocks (x a multiple of 4).
__device__ __forceinline__ void sub_fdct4(const uint8_t *src, const uint8_t *pred, int s, int x, int y, int *out) {
    int t[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        uint32_t a = *(const uint32_t *)&src[(y + r) * s + x], b = *(const uint32_t *)&pred[(y + r) * s + x];
        int d0 = (int)byte_of(a, 0) - (int)byte_of(b, 0), d1 = (int)byte_of(a, 1) - (int)byte_of(b, 1);
        int d2 = (int)byte_of(a, 2) - (int)byte_of(b, 2), d3 = (int)byte_of(a, 3) - (int)byte_of(b, 3);
        int s03 = d0 + d3, s12 = d1 + d2, d03 = d0 - d3, d12 = d1 - d2;
        t[r * 4] = s03 + s12;
        t[r * 4 + 1] = (d03 << 1) + d12;
        t[r * 4 + 2] = s03 - s12;
        t[r * 4 + 3] = d03 - (d12 << 1);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int s03 = t[c] + t[c + 12], s12 = t[c + 4] + t[c + 8], d03 = t[c] - t[c + 12], d12 = t[c + 4] - t[c + 8];
        out[c] = (int16_t)(s03 + s12);
        out[c + 4] = (int16_t)((d03 << 1) + d12);
        out[c + 8] = (int16_t)(s03 - s12);
        out[c + 12] = (int16_t)(d03 - (d12 << 1));
    }
}
__device__ __forceinline__ int pos4(int i) {  // h264.rs:1204 coeff_pos_type
    int r = i >> 2, c = i & 3;
    return (!(r & 1) && !(c & 1)) ? 0 : (((r & 1) && (c & 1)) ? 1 : 2);
}
// quant_4x4_x4_scalar on one block (simd/h264_enc.rs:408; enc.rs:1731): level =
// sign(c) min((|c| mf + f) >> shift, 32767), f = 2^shift / 6 (inter) or / 3 (intra). The
// product fits 32 bits: |c| <= 32768 and mf < 2^14.
__device__ __forceinline__ void quant4(int *c, int qp, int intra) {
    const int shift = 15 + qp / 6, rem = qp % 6;
    const uint32_t bias = (1u << shift) / (intra ? 3u : 6u);
    const uint32_t mf0 = (uint32_t)c_quant4_mf[rem][0], mf1 = (uint32_t)c_quant4_mf[rem][1], mf2 = (uint32_t)c_quant4_mf[rem][2];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int v = c[i];
        uint32_t mf = pos4(i) == 0 ? mf0 : (pos4(i) == 1 ? mf1 : mf2);
        uint32_t lv = (__umul24((uint32_t)iabs(v), mf) + bias) >> shift;
        lv = lv > 32767u ? 32767u : lv;
        c[i] = v < 0 ? -(int)lv : (int)lv;
    }
}
// Zig-zag non-zero mask and "some level above 1" of a block of levels (raster).
__device__ __forceinline__ uint32_t zz_mask(const int *c, int *big) {
    uint32_t m = 0;
    int b = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        m |= (uint32_t)(c[ZZ4[i]] != 0) << i;
        b |= iabs(c[ZZ4[i]]) > 1;
    }
    *big = b;
    return m;
}
// simd/h264_enc.rs:461 dequant_4x4_x4_scalar on one block.
__device__ __forceinline__ void dequant4(int *c, int qp) {
    const int sh = qp / 6, rem = qp % 6;
    const int s0 = c_dequant4_init[rem][0] << sh, s1 = c_dequant4_init[rem][1] << sh, s2 = c_dequant4_init[rem][2] << sh;
#pragma unroll
    for (int i = 0; i < 16; i++) c[i] = iclamp(__mul24(c[i], pos4(i) == 0 ? s0 : (pos4(i) == 1 ? s1 : s2)), -32768, 32767);
}
// simd/fixed.rs:185 inverse_idct4x4_scalar added onto a stride-s LDS block with clipping
// (simd/h264_enc.rs:497 add_idct4x4_8x8_scalar for one block); x a multiple of 4.
__device__ __forceinline__ void add_idct4(uint8_t *dst, int s, int x, int y, const int *c) {
    int t[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = r * 4;
        int e0 = c[i] + c[i + 2], e1 = c[i] - c[i + 2];
        int e2 = (c[i + 1] >> 1) - c[i + 3], e3 = c[i + 1] + (c[i + 3] >> 1);
        t[i] = e0 + e3;
        t[i + 1] = e1 + e2;
        t[i + 2] = e1 - e2;
        t[i + 3] = e0 - e3;
    }
    int o[16];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int e0 = t[k] + t[k + 8], e1 = t[k] - t[k + 8];
        int e2 = (t[k + 4] >> 1) - t[k + 12], e3 = t[k + 4] + (t[k + 12] >> 1);
        o[k] = (int16_t)((e0 + e3 + 32) >> 6);
        o[4 + k] = (int16_t)((e1 + e2 + 32) >> 6);
        o[8 + k] = (int16_t)((e1 - e2 + 32) >> 6);
        o[12 + k] = (int16_t)((e0 - e3 + 32) >> 6);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        uint32_t *p = (uint32_t *)&dst[(y + r) * s + x];
        uint32_t w = *p, n = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) n |= (uint32_t)clip255((int)byte_of(w, k) + o[r * 4 + k]) << (8 * k);
        *p = n;
    }
}

// ---------------------------------------------------------------------------------------
// Reference staging and the half-pel window (simd/h264_enc.rs:704-838).
// ---------------------------------------------------------------------------------------
// Copies reference rows [oy, oy + ROWS) x columns [ox, ox + SR_COLS) (ox a multiple of 4)
// into S.sr: every lane issues all of its aligned dword loads before the first LDS store.
template <int ROWS>
__device__ void stage_region(const RghJobD &J, int ox, int oy, MbLds &S) {
    static_assert(ROWS <= SR_H, "region too tall");
    constexpr int DW = SR_COLS / 4;  // 17 dwords a row
    constexpr int N = ROWS * DW, PER = (N + 63) / 64;
    const int lane = threadIdx.x;
    const RGH_G uint8_t *base = J.ref_y + (long)oy * J.sy + ox;
    uint32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lane + 64 * k, r = i / DW, j = i - r * DW;
        v[k] = i < N ? *(const RGH_G uint32_t *)(base + (long)r * J.sy + j * 4) : 0u;
    }
    // Earlier reads of this storage (the previous region, or residual arrays sharing it) are
    // complete before it is overwritten.
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lane + 64 * k, r = i / DW, j = i - r * DW;
        if (i < N) *(uint32_t *)&S.sr[r * SR_W + j * 4] = v[k];
    }
    __syncthreads();
}

// Horizontal 6-tap sums of four consecutive outputs from the nine bytes starting at byte o of
// the three dwords w[0..2].
__device__ __forceinline__ void tap6x4(const uint32_t *w, int o, int *out) {
    const uint64_t lo = ((uint64_t)w[1] << 32) | w[0];
    const uint64_t hi = ((uint64_t)w[2] << 32) | w[1];
    int b[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int k = o + i;  // 0..11
        b[i] = k < 8 ? (int)((lo >> (8 * k)) & 0xFF) : (int)((hi >> (8 * (k - 4))) & 0xFF);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = tap6(b[i], b[i + 1], b[i + 2], b[i + 3], b[i + 4], b[i + 5]);
}

// The b (ph), h (pv) and j (pc) planes over 24 x 18 samples whose top-left integer sample
// (-1, -1 from the full-pel anchor) is region sample (wx0, wy0); the region holds rows
// wy0 - 2 .. wy0 + 20 and columns wx0 - 2 .. wx0 + 28 (simd/h264_enc.rs:765 hpel_window_scalar
// and :820 hpel_window_centre_scalar, the centre computed up front). Work items produce four
// adjacent outputs from dword LDS reads: pass 1 the unrounded horizontal taps of rows -2 .. 20
// with the b plane and the h plane; pass 2 the j plane from the kept taps.
__device__ __forceinline__ void compute_window(MbLds &S, int wx0, int wy0) {
    const int lane = threadIdx.x;
    constexpr int NA = HR_ROWS * 6, NB = WP_ROWS * 6;
#pragma unroll
    for (int it = 0; it < (NA + NB + 63) / 64; it++) {
        const int item = lane + 64 * it;
        if (item < NA) {
            const int k = item / 6, g = item - k * 6;
            const int x = wx0 + 4 * g - 2, xa = x & ~3;
            const uint8_t *row = &S.sr[(wy0 - 2 + k) * SR_W + xa];
            const uint32_t w[3] = {*(const uint32_t *)row, *(const uint32_t *)(row + 4), *(const uint32_t *)(row + 8)};
            int h[4];
            tap6x4(w, x - xa, h);
            uint32_t *hp = (uint32_t *)&S.hr[k * WP_COLS + 4 * g];
            hp[0] = ((uint32_t)(uint16_t)h[0]) | ((uint32_t)(uint16_t)h[1] << 16);
            hp[1] = ((uint32_t)(uint16_t)h[2]) | ((uint32_t)(uint16_t)h[3] << 16);
            if (k >= 2 && k < 2 + WP_ROWS) {
                uint32_t o = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) o |= sat_shr<5>(h[i] + 16) << (8 * i);
                *(uint32_t *)&S.ph[(k - 2) * WP_COLS + 4 * g] = o;
            }
        } else if (item < NA + NB) {
            const int i2 = item - NA, r = i2 / 6, g = i2 - r * 6;
            uint32_t w[6];
#pragma unroll
            for (int t = 0; t < 6; t++) w[t] = lds4u(&S.sr[(wy0 - 2 + r + t) * SR_W], wx0 + 4 * g);
            uint32_t o = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int v = tap6(byte_of(w[0], i), byte_of(w[1], i), byte_of(w[2], i), byte_of(w[3], i), byte_of(w[4], i),
                                   byte_of(w[5], i));
                o |= sat_shr<5>(v + 16) << (8 * i);
            }
            *(uint32_t *)&S.pv[r * WP_COLS + 4 * g] = o;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < (NB + 63) / 64; it++) {
        const int item = lane + 64 * it;
        if (item < NB) {
            const int r = item / 6, g = item - r * 6;
            int t[6][4];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const uint32_t *hp = (const uint32_t *)&S.hr[(r + q) * WP_COLS + 4 * g];
                const uint32_t a = hp[0], b = hp[1];
                t[q][0] = (int16_t)(a & 0xFFFF);
                t[q][1] = (int16_t)(a >> 16);
                t[q][2] = (int16_t)(b & 0xFFFF);
                t[q][3] = (int16_t)(b >> 16);
            }
            uint32_t o = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int j = tap6(t[0][i], t[1][i], t[2][i], t[3][i], t[4][i], t[5][i]);
                o |= sat_shr<10>(j + 512) << (8 * i);
            }
            *(uint32_t *)&S.pc[r * WP_COLS + 4 * g] = o;
        }
    }
    __syncthreads();
}

// Rounded byte-wise average (a + b + 1) >> 1 of four packed bytes: v_lerp_u8 with the
// round-up bit set in every byte.
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b) { return __builtin_amdgcn_lerp(a, b, 0x01010101u); }

// The lane's four prediction bytes at quarter-pel mv (relative to the macroblock at x0, y0)
// from the window anchored at full-pel (bx, by): one plane or the rounded average of two
// (enc.rs:1415 window_src, the x264 get_ref table). Integer samples come from the staged
// region whose origin is (ox, oy); mv lies within three quarter samples of the anchor.
__device__ uint32_t window_pred4(int x0, int y0, int bx, int by, int ox, int oy, int mx, int my, const MbLds &S) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    const int ix = x0 + (mx >> 2), iy = y0 + (my >> 2);
    const int wx = ix - bx + 1, wy = iy - by + 1;
    int ka, kb = -1, dxa = 0, dxb = 0, dyb = 0;
    switch ((my & 3) * 4 + (mx & 3)) {
    case 0: ka = 0; break;
    case 2: ka = 1; break;
    case 8: ka = 2; break;
    case 10: ka = 3; break;
    case 1: ka = 0; kb = 1; break;
    case 3: ka = 1; kb = 0; dxb = 1; break;
    case 4: ka = 0; kb = 2; break;
    case 12: ka = 2; kb = 0; dyb = 1; break;
    case 5: ka = 1; kb = 2; break;
    case 7: ka = 1; kb = 2; dxb = 1; break;
    case 13: ka = 2; kb = 1; dyb = 1; break;
    case 15: ka = 2; dxa = 1; kb = 1; dyb = 1; break;
    case 6: ka = 1; kb = 3; break;
    case 14: ka = 3; kb = 1; dyb = 1; break;
    case 9: ka = 2; kb = 3; break;
    default: ka = 3; kb = 2; dxb = 1; break;
    }
    auto fetch4 = [&](int k, int dx, int dy) -> uint32_t {
        if (k == 0) return lds4u(&S.sr[(iy + dy + rr - oy) * SR_W], ix + dx + c0 - ox);
        const uint8_t *pl = k == 1 ? S.ph : (k == 2 ? S.pv : S.pc);
        return lds4u(&pl[(wy + dy + rr) * WP_COLS], wx + dx + c0);
    };
    const uint32_t a = fetch4(ka, dxa, 0);
    if (kb < 0) return a;
    const uint32_t b = fetch4(kb, dxb, dyb);
    return avg4(a, b);
}

// The quarter-pel phase table of window_pred4 packed one byte per phase (my & 3) * 4 + (mx & 3):
// bits 0-1 plane a, 2-3 plane b (= a for a single plane), 4 column step of a, 5 column step
// of b, 6 row step of b; planes 0 full, 1 b (ph), 2 h (pv), 3 j (pc).
constexpr uint64_t kPhaseLo = 0x290d090821050400ull, kPhaseHi = 0x564746422b0f0e0aull;

// Byte offset in MbLds of sample (column dx, row dy) of plane k for the block at quarter-pel
// vector (mx, my), given the window anchor (bx, by) and the region origin (ox, oy), and that
// plane's row stride.
__device__ __forceinline__ int plane_offset(int k, int x0, int y0, int bx, int by, int ox, int oy, int mx, int my, int dx,
                                            int dy, int &stride) {
    const int ix = x0 + (mx >> 2) + dx, iy = y0 + (my >> 2) + dy;
    if (k == 0) {
        stride = SR_W;
        return (int)offsetof(MbLds, sr) + (iy - oy) * SR_W + ix - ox;
    }
    stride = WP_COLS;
    const int po = k == 1 ? (int)offsetof(MbLds, ph) : (k == 2 ? (int)offsetof(MbLds, pv) : (int)offsetof(MbLds, pc));
    return po + (iy - by + 1) * WP_COLS + ix - bx + 1;
}

// SADs of four quarter-pel vectors (mx[k], my[k]) (relative to the macroblock at x0, y0), each
// within +-3 quarter samples of the window anchor (bx, by), through the window: lane 16 k + r
// sums row r of vector k from its plane pair (window_pred4's table; five aligned dwords a plane),
// and each row of 16 lanes adds its vector's rows. Returns the four sums in out.
// The SAD of the lane's row's point (lane 16 k + r: row r of point k), summed over the 16-lane row.
__device__ __forceinline__ uint32_t subpel_sad_rows(int x0, int y0, int bx, int by, int ox, int oy, const int *mx,
                                                    const int *my, MbLds &S) {
    const int lane = threadIdx.x, k = lane >> 4, r = lane & 15;
    const int qx = sel4(k, mx[0], mx[1], mx[2], mx[3]), qy = sel4(k, my[0], my[1], my[2], my[3]);
    const int ph = (qy & 3) * 4 + (qx & 3);
    const uint32_t e = (uint32_t)(((ph < 8 ? kPhaseLo : kPhaseHi) >> (8 * (ph & 7))) & 0xFF);
    int sa, sb;
    const int oa = plane_offset(e & 3, x0, y0, bx, by, ox, oy, qx, qy, (e >> 4) & 1, 0, sa);
    const int ob = plane_offset((e >> 2) & 3, x0, y0, bx, by, ox, oy, qx, qy, (e >> 5) & 1, (e >> 6) & 1, sb);
    const uint8_t *base = (const uint8_t *)&S;
    const uint32_t *qa = (const uint32_t *)(base + (oa & ~3) + r * sa), *qb = (const uint32_t *)(base + (ob & ~3) + r * sb);
    const uint32_t fa = (uint32_t)(oa & 3), fb = (uint32_t)(ob & 3);
    uint32_t da[5], db[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        da[c] = qa[c];
        db[c] = qb[c];
    }
    const uint4 sw = *(const uint4 *)&S.src[r * 16];
    const uint32_t src[4] = {sw.x, sw.y, sw.z, sw.w};
    uint32_t sad = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint32_t a = __builtin_amdgcn_alignbyte(da[c + 1], da[c], fa);
        const uint32_t b = __builtin_amdgcn_alignbyte(db[c + 1], db[c], fb);
        sad = __builtin_amdgcn_sad_u8(src[c], avg4(a, b), sad);
    }
    return row16_sum(sad);
}
__device__ void subpel_sad4(int x0, int y0, int bx, int by, int ox, int oy, const int *mx, const int *my, MbLds &S,
                            uint32_t *out) {
    const uint32_t sad = subpel_sad_rows(x0, y0, bx, by, ox, oy, mx, my, S);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = (uint32_t)__builtin_amdgcn_readlane((int)sad, 16 * q);
}

// ---------------------------------------------------------------------------------------
// P partitions (RGH_FLAG_PARTITIONS): costs of a partition of W x h samples at (px, py) in the
// macroblock, in the 16-lane row layout of the diamond (lane 16 k + r: row r < h of point k).
// ---------------------------------------------------------------------------------------
// SADs of four full-sample vectors (fx[k], fy[k]) from the staged region whose origin is
// (ox, oy); every point's rows lie inside it.
template <int W>
__device__ void part_sad4(int x0, int y0, int px, int py, int h, int ox, int oy, const int *fx, const int *fy,
                          const MbLds &S, uint32_t *out) {
    const int lane = threadIdx.x, k = lane >> 4, r = lane & 15;
    const int qx = sel4(k, fx[0], fx[1], fx[2], fx[3]), qy = sel4(k, fy[0], fy[1], fy[2], fy[3]);
    uint32_t sad = 0;
    if (r < h) {
        const int lx = x0 + px + qx - ox, ly = y0 + py + qy - oy + r;
        const uint32_t *q = (const uint32_t *)&S.sr[ly * SR_W + (lx & ~3)];
        const uint32_t f = (uint32_t)(lx & 3);
        const uint32_t *sw = (const uint32_t *)&S.src[(py + r) * 16 + px];
        uint32_t d[W / 4 + 1];
#pragma unroll
        for (int c = 0; c <= W / 4; c++) d[c] = q[c];
#pragma unroll
        for (int c = 0; c < W / 4; c++) sad = __builtin_amdgcn_sad_u8(sw[c], __builtin_amdgcn_alignbyte(d[c + 1], d[c], f), sad);
    }
    sad = row16_sum(sad);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = (uint32_t)__builtin_amdgcn_readlane((int)sad, 16 * q);
}

// SADs of four quarter-sample vectors (mx[k], my[k]) of the partition through the half-sample
// window anchored at full sample (ax, ay) (the partition's full-sample result in the picture),
// each within +-3 quarter samples of it: window_pred4's plane pairs (kPhaseLo / kPhaseHi).
template <int W>
__device__ void part_subpel_sad4(int x0, int y0, int px, int py, int h, int ax, int ay, int ox, int oy, const int *mx,
                                 const int *my, MbLds &S, uint32_t *out) {
    const int lane = threadIdx.x, k = lane >> 4, r = lane & 15;
    const int qx = sel4(k, mx[0], mx[1], mx[2], mx[3]), qy = sel4(k, my[0], my[1], my[2], my[3]);
    uint32_t sad = 0;
    if (r < h) {
        const int ph = (qy & 3) * 4 + (qx & 3);
        const uint32_t e = (uint32_t)(((ph < 8 ? kPhaseLo : kPhaseHi) >> (8 * (ph & 7))) & 0xFF);
        int sa, sb;
        const int oa = plane_offset(e & 3, x0 + px, y0 + py, ax, ay, ox, oy, qx, qy, (e >> 4) & 1, 0, sa);
        const int ob = plane_offset((e >> 2) & 3, x0 + px, y0 + py, ax, ay, ox, oy, qx, qy, (e >> 5) & 1, (e >> 6) & 1, sb);
        const uint8_t *base = (const uint8_t *)&S;
        const uint32_t *qa = (const uint32_t *)(base + (oa & ~3) + r * sa), *qb = (const uint32_t *)(base + (ob & ~3) + r * sb);
        const uint32_t fa = (uint32_t)(oa & 3), fb = (uint32_t)(ob & 3);
        uint32_t da[W / 4 + 1], db[W / 4 + 1];
#pragma unroll
        for (int c = 0; c <= W / 4; c++) {
            da[c] = qa[c];
            db[c] = qb[c];
        }
        const uint32_t *sw = (const uint32_t *)&S.src[(py + r) * 16 + px];
#pragma unroll
        for (int c = 0; c < W / 4; c++) {
            const uint32_t a = __builtin_amdgcn_alignbyte(da[c + 1], da[c], fa);
            const uint32_t b = __builtin_amdgcn_alignbyte(db[c + 1], db[c], fb);
            sad = __builtin_amdgcn_sad_u8(sw[c], avg4(a, b), sad);
        }
    }
    sad = row16_sum(sad);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = (uint32_t)__builtin_amdgcn_readlane((int)sad, 16 * q);
}

__device__ __forceinline__ uint32_t ssd4(uint32_t a, uint32_t b) {
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int d = (int)byte_of(a, i) - (int)byte_of(b, i);
        s += (uint32_t)(d * d);
    }
    return s;
}

// pixel_satd_16x16 (simd/h264_enc.rs:167) of two 16-stride LDS blocks: the sum of the 4x4
// blocks' unhalved Hadamard magnitudes, halved once. One block row per lane (block lane >> 2 in
// raster order, row lane & 3): the row transform in the lane, the column transform across the
// quad, each lane summing its output row.
__device__ uint32_t satd16(const uint8_t *a, const uint8_t *b) {
    const int lane = threadIdx.x, bx = (lane >> 2) & 3, by = lane >> 4, row = lane & 3;
    const int o = (by * 4 + row) * 16 + bx * 4;
    const uint32_t wa = *(const uint32_t *)&a[o], wb = *(const uint32_t *)&b[o];
    const int d0 = (int)byte_of(wa, 0) - (int)byte_of(wb, 0), d1 = (int)byte_of(wa, 1) - (int)byte_of(wb, 1);
    const int d2 = (int)byte_of(wa, 2) - (int)byte_of(wb, 2), d3 = (int)byte_of(wa, 3) - (int)byte_of(wb, 3);
    const int s01 = d0 + d1, d01 = d0 - d1, s23 = d2 + d3, d23 = d2 - d3;
    const int t[4] = {s01 + s23, s01 - s23, d01 + d23, d01 - d23};
    const QuadMix m = satd_mix(row);
    uint32_t part = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) part += (uint32_t)iabs(quad_mix(t[c], m));
    return wave_sum(part) >> 1;
}

// The four reference chroma samples of the lane's bilinear taps for quarter-pel mv (enc.rs:1080
// mc_chroma), loaded for both planes; chroma_mc_combine applies the 1/8-sample weights
// (simd/h264_enc.rs:1279 chroma_mc_8x8_x2_scalar).
struct ChromaTaps {
    uint32_t u[4], v[4];
};
__device__ __forceinline__ ChromaTaps chroma_taps(const RghJobD &J, int mb_x, int mb_y, int mx, int my) {
    const int lane = threadIdx.x, cr = lane >> 3, cc = lane & 7;
    const long off = (long)(mb_y * 8 + (my >> 3) + cr) * J.sc + mb_x * 8 + (mx >> 3) + cc;
    ChromaTaps t;
    t.u[0] = J.ref_u[off];
    t.v[0] = J.ref_v[off];
    if ((mx | my) & 7) {
        t.u[1] = J.ref_u[off + 1];
        t.u[2] = J.ref_u[off + J.sc];
        t.u[3] = J.ref_u[off + J.sc + 1];
        t.v[1] = J.ref_v[off + 1];
        t.v[2] = J.ref_v[off + J.sc];
        t.v[3] = J.ref_v[off + J.sc + 1];
    } else {
        t.u[1] = t.u[2] = t.u[3] = t.v[1] = t.v[2] = t.v[3] = 0;
    }
    return t;
}
__device__ __forceinline__ void chroma_mc_combine(const ChromaTaps &t, int mx, int my, uint32_t *pu, uint32_t *pv) {
    const int fx = mx & 7, fy = my & 7;
    if ((fx | fy) == 0) {
        *pu = t.u[0];
        *pv = t.v[0];
        return;
    }
    const int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
    *pu = (uint32_t)((wa * (int)t.u[0] + wb * (int)t.u[1] + wc * (int)t.u[2] + wd * (int)t.u[3] + 32) >> 6);
    *pv = (uint32_t)((wa * (int)t.v[0] + wb * (int)t.v[1] + wc * (int)t.v[2] + wd * (int)t.v[3] + 32) >> 6);
}

// Chroma transform + quantization of both planes (enc.rs:1897 quant_chroma), lanes 0..31 one
// block row each (plane lane >> 4, raster 2x2 block (lane >> 2) & 3, row lane & 3), lanes 0..1
// the 2x2 DC of a plane (enc.rs:1939 quant_chroma_dc). Leaves raster levels, scanned AC masks
// and decimation scores in LDS. intra selects the rounding of quantize_block4x4.
__device__ __forceinline__ void quant_chroma(MbLds &S, int qpc, int intra) {
    const int lane = threadIdx.x;
    if (lane < 32) {
        const int p = lane >> 4, b = (lane >> 2) & 3, row = lane & 3;
        int c[4];
        fdct4_row(p ? S.srcv : S.srcu, p ? S.predv : S.predu, 8, (b & 1) * 4, (b >> 1) * 4, row, c);
        if (row == 0) S.cdc_raw[p][b] = (int16_t)c[0];
        quant4_row(c, row, qpc, intra);
        *(uint2 *)&S.craster[p][b][row * 4] =
            make_uint2(((uint32_t)c[0] & 0xFFFFu) | ((uint32_t)c[1] << 16), ((uint32_t)c[2] & 0xFFFFu) | ((uint32_t)c[3] << 16));
        // Scan positions of this row's four coefficients (byte `row` of each word), the DC's
        // bit (scan position 0) dropped by the final shift.
        const uint32_t zz[4] = {0x09030200u, 0x0A080401u, 0x0E0B0705u, 0x0F0D0C06u};
        uint32_t m = 0;
        int big = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            m |= (uint32_t)(c[k] != 0) << ((zz[k] >> (8 * row)) & 15u);
            big |= (iabs(c[k]) > 1) & (k != 0 || row != 0);
        }
        m = (uint32_t)quad_or((int)m) >> 1;
        big = quad_or(big);
        if (row == 0) {
            S.cmask[lane >> 2] = (int32_t)m;
            S.cscore[lane >> 2] = (int32_t)decimate_mask(m, big);
        }
    }
    __syncthreads();
    if (lane < 2) {
        int p = lane, d0 = S.cdc_raw[p][0], d1 = S.cdc_raw[p][1], d2 = S.cdc_raw[p][2], d3 = S.cdc_raw[p][3];
        int had[4] = {d0 + d1 + d2 + d3, d0 - d1 + d2 - d3, d0 + d1 - d2 - d3, d0 - d1 - d2 + d3};
        int qbits = 16 + qpc / 6, mf = c_quant4_mf[qpc % 6][0], f = (1 << qbits) / 3;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int l = (int)((__umul24((uint32_t)iabs(had[i]), (uint32_t)mf) + (uint32_t)f) >> qbits);
            S.cdc[p][i] = (int16_t)(had[i] < 0 ? -l : l);
        }
    }
    __syncthreads();
}

// h264.rs:1674 dequantize_idct_chroma_dc2x2 then the chroma recon (enc.rs:1990).
__device__ void recon_chroma(MbLds &S, int cbpc, int qpc) {
    const int lane = threadIdx.x;
    if (!cbpc) return;
    if (lane < 2) {
        int p = lane, qmul = (c_dequant4_init[qpc % 6][0] * 16) << (qpc / 6 + 2);
        int a = S.cdc[p][0], b = S.cdc[p][1], c = S.cdc[p][2], d = S.cdc[p][3];
        S.cdcr[p][0] = (int16_t)(((a + b + c + d) * qmul) >> 7);
        S.cdcr[p][1] = (int16_t)(((a - b + c - d) * qmul) >> 7);
        S.cdcr[p][2] = (int16_t)(((a + b - c - d) * qmul) >> 7);
        S.cdcr[p][3] = (int16_t)(((a - b - c + d) * qmul) >> 7);
    }
    __syncthreads();
    if (lane < 32) {
        const int p = lane >> 4, b = (lane >> 2) & 3, row = lane & 3;
        int blk[4] = {0, 0, 0, 0};
        if (cbpc >= 2) {
            const uint2 w = *(const uint2 *)&S.craster[p][b][row * 4];
            blk[0] = (int16_t)w.x;
            blk[1] = (int)w.x >> 16;
            blk[2] = (int16_t)w.y;
            blk[3] = (int)w.y >> 16;
            dequant4_row(blk, row, qpc);
        }
        if (row == 0) blk[0] = S.cdcr[p][b];
        add_idct4_row(p ? S.predv : S.predu, 8, (b & 1) * 4, (b >> 1) * 4, row, blk);
    }
    __syncthreads();
}

__device__ __forceinline__ int chroma_cbp(const MbLds &S) {
    uint32_t m = 0;
    int d = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) m |= (uint32_t)S.cmask[i];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int i = 0; i < 4; i++) d |= S.cdc[p][i];
    return m ? 2 : (d ? 1 : 0);
}

// 8x8 forward transform of the four 8x8 luma residuals (simd/fixed.rs:392 forward_dct8x8),
// lanes 0..31: block lane >> 3; the first pass transforms row lane & 7, the second column
// lane & 7. Returns the transformed column in col[8] (rows 0..7).
__device__ void fdct8_cols(MbLds &S, int *col) {
    const int lane = threadIdx.x;
    auto bfly = [](const int *v, int *o) {
        int a0 = v[0] + v[7], a1 = v[1] + v[6], a2 = v[2] + v[5], a3 = v[3] + v[4];
        int a4 = v[0] - v[7], a5 = v[1] - v[6], a6 = v[2] - v[5], a7 = v[3] - v[4];
        int b0 = a0 + a3, b1 = a1 + a2, b2 = a0 - a3, b3 = a1 - a2;
        int b4 = a5 + a6 + ((a4 >> 1) + a4), b5 = a4 - a7 - ((a6 >> 1) + a6);
        int b6 = a4 + a7 - ((a5 >> 1) + a5), b7 = a5 - a6 + ((a7 >> 1) + a7);
        o[0] = b0 + b1;
        o[1] = b4 + (b7 >> 2);
        o[2] = b2 + (b3 >> 1);
        o[3] = b5 + (b6 >> 2);
        o[4] = b0 - b1;
        o[5] = b6 - (b5 >> 2);
        o[6] = (b2 >> 1) - b3;
        o[7] = (b4 >> 2) - b7;
    };
    if (lane < 32) {
        int blk = lane >> 3, r = lane & 7, ox = (blk & 1) * 8, oy = (blk >> 1) * 8;
        int v[8], o[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (int)S.src[(oy + r) * 16 + ox + j] - (int)S.pred[(oy + r) * 16 + ox + j];
        bfly(v, o);
#pragma unroll
        for (int j = 0; j < 8; j++) S.t8[blk][r * 8 + j] = o[j];
    }
    __syncthreads();
    if (lane < 32) {
        int blk = lane >> 3, c = lane & 7;
        int v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = S.t8[blk][j * 8 + c];
        bfly(v, col);
#pragma unroll
        for (int j = 0; j < 8; j++) col[j] = (int16_t)col[j];
    }
}

__device__ __forceinline__ int pos8(int i) {  // h264.rs:1343 coeff8_pos_type
    int r = (i >> 3) & 3, c = i & 3;
    if (r == 0 && c == 0) return 0;
    if ((r & 1) && (c & 1)) return 1;
    if (r == 2 && c == 2) return 2;
    if ((r == 0 && (c & 1)) || ((r & 1) && c == 0)) return 3;
    if ((r == 0 && c == 2) || (r == 2 && c == 0)) return 4;
    return 5;
}

// enc.rs:2017 recon_8x8: h264.rs:1403 dequantize_block8x8, then simd/h264_idct.rs:268
// h264_idct8_add_scalar (wrapping i32), lanes 0..31 one row then one column of a coded block.
__device__ void recon8(MbLds &S, int cbpl, int qp) {
    const int lane = threadIdx.x;
    auto idct1 = [](const uint32_t *d, uint32_t *o) {
        uint32_t a0 = d[0] + d[4];
        uint32_t a1 = d[5] - d[3] - d[7] - (uint32_t)((int)d[7] >> 1);
        uint32_t a2 = d[0] - d[4];
        uint32_t a3 = d[1] + d[7] - d[3] - (uint32_t)((int)d[3] >> 1);
        uint32_t a4 = (uint32_t)((int)d[2] >> 1) - d[6];
        uint32_t a5 = d[7] - d[1] + d[5] + (uint32_t)((int)d[5] >> 1);
        uint32_t a6 = d[2] + (uint32_t)((int)d[6] >> 1);
        uint32_t a7 = d[3] + d[5] + d[1] + (uint32_t)((int)d[1] >> 1);
        uint32_t b0 = a0 + a6, b1 = a1 + (uint32_t)((int)a7 >> 2), b2 = a2 + a4;
        uint32_t b3 = a3 + (uint32_t)((int)a5 >> 2), b4 = a2 - a4;
        uint32_t b5 = (uint32_t)((int)a3 >> 2) - a5, b6 = a0 - a6, b7 = a7 - (uint32_t)((int)a1 >> 2);
        o[0] = b0 + b7;
        o[1] = b2 + b5;
        o[2] = b4 + b3;
        o[3] = b6 + b1;
        o[4] = b6 - b1;
        o[5] = b4 - b3;
        o[6] = b2 - b5;
        o[7] = b0 - b7;
    };
    const int blk = lane >> 3, k = lane & 7;
    const bool act = lane < 32 && ((cbpl >> blk) & 1);
    if (act) {
        int qd = qp / 6, rem = qp % 6;
        uint32_t d[8], o[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int i = k * 8 + j, s = c_dequant8_init[rem][pos8(i)], v;
            if (qd >= 2)
                v = S.q8[blk][i] * (s << (qd - 2));
            else {
                int rs = 2 - qd;
                v = (S.q8[blk][i] * s + (1 << (rs - 1))) >> rs;
            }
            d[j] = (uint32_t)iclamp(v, -32768, 32767);
        }
        idct1(d, o);
#pragma unroll
        for (int j = 0; j < 8; j++) S.t8[blk][k * 8 + j] = (int)o[j];
    }
    __syncthreads();
    if (act) {
        uint32_t d[8], o[8];
#pragma unroll
        for (int j = 0; j < 8; j++) d[j] = (uint32_t)S.t8[blk][j * 8 + k];
        idct1(d, o);
        int ox = (blk & 1) * 8 + k, oy = (blk >> 1) * 8;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint8_t *px = &S.pred[(oy + j) * 16 + ox];
            *px = (uint8_t)clip255((int)((uint32_t)*px + (uint32_t)((int)(o[j] + 32u) >> 6)));
        }
    }
    __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Intra 16x16 (enc.rs:1635 analyse_i16_edges, :3587 code_i16), shared by P and IDR pictures.
// ---------------------------------------------------------------------------------------
struct I16Analysis {
    uint32_t cost;
    int mode, pa, pb, pc, dc;
};

// enc.rs:1626 i16_dc from the edge samples in S (zero when a neighbour is missing).
__device__ __forceinline__ int i16_dc_of(const MbLds &S, int has_t, int has_l) {
    const uint4 te = *(const uint4 *)S.etop, le = *(const uint4 *)S.eleft;
    uint32_t st = __builtin_amdgcn_sad_u8(te.x, 0u, 0u), sl = __builtin_amdgcn_sad_u8(le.x, 0u, 0u);
    st = __builtin_amdgcn_sad_u8(te.y, 0u, st);
    st = __builtin_amdgcn_sad_u8(te.z, 0u, st);
    st = __builtin_amdgcn_sad_u8(te.w, 0u, st);
    sl = __builtin_amdgcn_sad_u8(le.y, 0u, sl);
    sl = __builtin_amdgcn_sad_u8(le.z, 0u, sl);
    sl = __builtin_amdgcn_sad_u8(le.w, 0u, sl);
    return (has_t && has_l) ? (int)((st + sl + 16) >> 5) : (has_t ? (int)((st + 8) >> 4) : (has_l ? (int)((sl + 8) >> 4) : 128));
}

// V, H and DC SATDs from one source Hadamard, plane SATD by prediction; the intra edges and
// the source are in S. One 4x4 block row per lane (block (lane >> 2) & 3, lane >> 4 in raster
// order, row lane & 3): rows in the lane, columns across the quad.
// PLANE false: no plane mode (the Fast preset, enc.rs:3322).
template <bool PLANE = true>
__device__ I16Analysis analyse_i16(MbLds &S, int mb_x, int mb_y, uint16_t &n_satd) {
    const int lane = threadIdx.x;
    const int has_t = mb_y > 0, has_l = mb_x > 0, tl = S.etl;
    // The 16 top and left edge samples as four words each (byte i: byte i & 3 of word i >> 2).
    const uint4 te = *(const uint4 *)S.etop, le = *(const uint4 *)S.eleft;
    const uint32_t tw[4] = {te.x, te.y, te.z, te.w}, lw[4] = {le.x, le.y, le.z, le.w};
    auto et = [&](int i) { return (int)byte_of(tw[i >> 2], i & 3); };
    auto el = [&](int i) { return (int)byte_of(lw[i >> 2], i & 3); };
    uint32_t st = 0, sl = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        st = __builtin_amdgcn_sad_u8(tw[k], 0u, st);
        sl = __builtin_amdgcn_sad_u8(lw[k], 0u, sl);
    }
    const int dc = (has_t && has_l) ? (int)((st + sl + 16) >> 5)
                                    : (has_t ? (int)((st + 8) >> 4) : (has_l ? (int)((sl + 8) >> 4) : 128));
    int pa = 0, pb = 0, pcc = 0;
    if (has_t && has_l) {
        int Hh = 0, Vv = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int ptop = i == 7 ? tl : et(6 - i), pl = i == 7 ? tl : el(6 - i);
            Hh += (i + 1) * (et(8 + i) - ptop);
            Vv += (i + 1) * (el(8 + i) - pl);
        }
        pa = 16 * (et(15) + el(15));
        pb = (5 * Hh + 32) >> 6;
        pcc = (5 * Vv + 32) >> 6;
    }
    // simd/h264_enc.rs:628 intra_satd_x3_16x16: V, H, DC SATDs from one source Hadamard
    // (unavailable edges read as zero, as the reference passes zero rows); the plane SATD is
    // pixel_satd_16x16 (simd/h264_enc.rs:167) of the plane prediction.
    const int bx = (lane >> 2) & 3, by = lane >> 4, row = lane & 3;
    const uint32_t w = *(const uint32_t *)&S.src[(by * 4 + row) * 16 + bx * 4];
    int hs[4], hp[4];
    {
        const int x0_ = byte_of(w, 0), x1 = byte_of(w, 1), x2 = byte_of(w, 2), x3 = byte_of(w, 3);
        const int s01 = x0_ + x1, d01 = x0_ - x1, s23 = x2 + x3, d23 = x2 - x3;
        hs[0] = s01 + s23;
        hs[1] = s01 - s23;
        hs[2] = d01 + d23;
        hs[3] = d01 - d23;
        int e[4];
        const int xy = pa + pcc * (by * 4 + row - 7) + 16;
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = (int)byte_of(w, i) - (int)sat_shr<5>(xy + pb * (bx * 4 + i - 7));
        const int p01 = e[0] + e[1], q01 = e[0] - e[1], p23 = e[2] + e[3], q23 = e[2] - e[3];
        hp[0] = p01 + p23;
        hp[1] = p01 - p23;
        hp[2] = q01 + q23;
        hp[3] = q01 - q23;
    }
    int tv[4], thr;
    {
        const uint32_t ta = bx == 0 ? te.x : (bx == 1 ? te.y : (bx == 2 ? te.z : te.w));
        const uint32_t la = by == 0 ? le.x : (by == 1 ? le.y : (by == 2 ? le.z : le.w));
        const int a0 = (int)byte_of(ta, 0), a1 = (int)byte_of(ta, 1), a2 = (int)byte_of(ta, 2), a3 = (int)byte_of(ta, 3);
        tv[0] = 4 * (a0 + a1 + a2 + a3);
        tv[1] = 4 * (a0 + a1 - a2 - a3);
        tv[2] = 4 * (a0 - a1 + a2 - a3);
        tv[3] = 4 * (a0 - a1 - a2 + a3);
        const int l0 = (int)byte_of(la, 0), l1 = (int)byte_of(la, 1), l2 = (int)byte_of(la, 2), l3 = (int)byte_of(la, 3);
        thr = 4 * sel4(row, l0 + l1 + l2 + l3, l0 + l1 - l2 - l3, l0 - l1 + l2 - l3, l0 - l1 - l2 + l3);
    }
    uint32_t part[4] = {0, 0, 0, 0};  // V, H, DC, plane
    const QuadMix sm = satd_mix(row);
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int c = (int16_t)quad_mix(hs[h], sm);
        part[0] += (uint32_t)iabs(c - (row == 0 ? tv[h] : 0));
        part[1] += (uint32_t)iabs(c - (h == 0 ? thr : 0));
        part[2] += (uint32_t)iabs(c - ((row == 0 && h == 0) ? 16 * dc : 0));
        part[3] += (uint32_t)iabs(quad_mix(hp[h], sm));
    }
    uint32_t sums[4];
    wave_sum4(part, sums);
    const uint32_t sv = sums[0] >> 1, sh = sums[1] >> 1, sd = sums[2] >> 1;
    // enc.rs:1637 analyse_i16_edges: DC, then V, then H, then plane, strict improvements.
    uint32_t ib = sd;
    int im = 2;
    if (has_t && sv < ib) { ib = sv; im = 0; }
    if (has_l && sh < ib) { ib = sh; im = 1; }
    if (PLANE && has_t && has_l) {
        const uint32_t ps = sums[3] >> 1;
        n_satd++;
        if (ps < ib) { ib = ps; im = 3; }
    }
    return {ib, im, pa, pb, pcc, dc};
}

struct ChromaIntra {
    int mode;
    uint32_t pu, pv;
};

// enc.rs:1523 analyse_chroma_intra: the chroma mode by SAD + 8 for non-DC modes (predictors
// h264.rs:1967 and enc.rs:1588) from the chroma edges in S; the lane's predicted U and V samples.
// All four modes are predicted and their SADs reduced together; unavailable modes are computed
// from whatever the edge arrays hold and then ignored. PLANE false: no plane mode (Fast); FORCED:
// the mode `forced` without analysis (Fast P pictures: the direction of the luma mode).
template <bool PLANE = true, bool FORCED = false>
__device__ ChromaIntra chroma_intra(MbLds &S, int mb_x, int mb_y, uint32_t src_u, uint32_t src_v, int forced = 0) {
    const int lane = threadIdx.x;
    const int has_t = mb_y > 0, has_l = mb_x > 0;
    const int cr = lane >> 3, cc = lane & 7;
    int smp[4][2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        // The eight top and left edge samples as two words each (byte i of the edge: byte i & 3
        // of word i >> 2); v_sad_u8 against zero sums four bytes.
        const uint2 tw = *(const uint2 *)S.ctop[p], lw = *(const uint2 *)S.cleft[p];
        auto t = [&](int i) { return (int)byte_of(i < 4 ? tw.x : tw.y, i & 3); };
        auto l = [&](int i) { return (int)byte_of(i < 4 ? lw.x : lw.y, i & 3); };
        const uint32_t t0 = __builtin_amdgcn_sad_u8(tw.x, 0u, 0u), t1 = __builtin_amdgcn_sad_u8(tw.y, 0u, 0u);
        const uint32_t l0 = __builtin_amdgcn_sad_u8(lw.x, 0u, 0u), l1 = __builtin_amdgcn_sad_u8(lw.y, 0u, 0u);
        int d00, d01, d10, d11;
        if (has_t && has_l) {
            d00 = (int)((t0 + l0 + 4) >> 3);
            d01 = (int)((t1 + 2) >> 2);
            d10 = (int)((l1 + 2) >> 2);
            d11 = (int)((t1 + l1 + 4) >> 3);
        } else if (has_t) {
            d00 = d10 = (int)((t0 + 2) >> 2);
            d01 = d11 = (int)((t1 + 2) >> 2);
        } else if (has_l) {
            d00 = d01 = (int)((l0 + 2) >> 2);
            d10 = d11 = (int)((l1 + 2) >> 2);
        } else {
            d00 = d01 = d10 = d11 = 128;
        }
        smp[0][p] = cr < 4 ? (cc < 4 ? d00 : d01) : (cc < 4 ? d10 : d11);
        smp[1][p] = (int)byte_of(cr < 4 ? lw.x : lw.y, cr & 3);
        smp[2][p] = (int)byte_of(cc < 4 ? tw.x : tw.y, cc & 3);
        int Hh = 0, Vv = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ptop = i == 3 ? S.ctl[p] : t(2 - i), pl = i == 3 ? S.ctl[p] : l(2 - i);
            Hh += (i + 1) * (t(4 + i) - ptop);
            Vv += (i + 1) * (l(4 + i) - pl);
        }
        const int a = 16 * (t(7) + l(7)), b = (17 * Hh + 16) >> 5, c = (17 * Vv + 16) >> 5;
        smp[3][p] = (int)sat_shr<5>(a + b * (cc - 3) + c * (cr - 3) + 16);
    }
    uint32_t part[4], cost[4];
#pragma unroll
    for (int m = 0; m < 4; m++) part[m] = (uint32_t)(iabs((int)src_u - smp[m][0]) + iabs((int)src_v - smp[m][1]));
    wave_sum4(part, cost);
    // First strictly cheapest available mode, DC (always available) first.
    int cmode = 0;
    uint32_t cbest = cost[0];
    if (has_l && cost[1] + 8u < cbest) {
        cbest = cost[1] + 8u;
        cmode = 1;
    }
    if (has_t && cost[2] + 8u < cbest) {
        cbest = cost[2] + 8u;
        cmode = 2;
    }
    if (PLANE && has_t && has_l && cost[3] + 8u < cbest) cmode = 3;
    if (FORCED) cmode = forced;
    const int pu = cmode == 0 ? smp[0][0] : (cmode == 1 ? smp[1][0] : (cmode == 2 ? smp[2][0] : smp[3][0]));
    const int pv = cmode == 0 ? smp[0][1] : (cmode == 1 ? smp[1][1] : (cmode == 2 ? smp[2][1] : smp[3][1]));
    return {cmode, (uint32_t)pu, (uint32_t)pv};
}

// Output stores of a macroblock, as dwords and 64-bit words (the level buffer is 8-byte aligned,
// every macroblock's levels start at a multiple of 816 bytes).
// Chroma reconstruction from S.predu / S.predv: lanes 0-31, one 4-byte half row each (plane
// lane >> 4, row (lane >> 1) & 7).
__device__ __forceinline__ void store_chroma(const MbLds &S, RGH_G uint8_t *u, RGH_G uint8_t *v, long stride, int cx0,
                                             int cy0) {
    const int lane = threadIdx.x;
    if (lane < 32) {
        const int p = lane >> 4, row = (lane >> 1) & 7, half = lane & 1;
        const uint32_t w = *(const uint32_t *)&(p ? S.predv : S.predu)[row * 8 + half * 4];
        *(RGH_G uint32_t *)&(p ? v : u)[(long)(cy0 + row) * stride + cx0 + half * 4] = w;
    }
}
__device__ __forceinline__ uint64_t pack4x16(int a, int b, int c, int d) {
    return (uint64_t)((uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)b << 16)) |
           ((uint64_t)((uint32_t)(uint16_t)c | ((uint32_t)(uint16_t)d << 16)) << 32);
}
// Luma levels [0, 256): 16 blocks (decoding order) of 16 zig-zag levels from raster-ordered LDS
// blocks; lane l stores levels 4l .. 4l + 3.
__device__ __forceinline__ void store_luma_levels(RGH_G int16_t *lv, const int16_t (*blk)[16]) {
    const int lane = threadIdx.x, b = lane >> 2, k = (lane & 3) * 4;
    ((RGH_G uint64_t *)lv)[lane] = pack4x16(blk[b][zz4(k)], blk[b][zz4(k + 1)], blk[b][zz4(k + 2)], blk[b][zz4(k + 3)]);
}
// Levels [256, 408): the luma DC (I_16x16, zig-zag), the chroma DC (U then V), and the chroma AC
// (U then V, 4 blocks of 16 zig-zag levels, slot 0 empty; zero without chroma AC).
__device__ __forceinline__ void store_dc_chroma_levels(RGH_G int16_t *lv, const MbLds &S, bool i16, int cbpc) {
    const int lane = threadIdx.x;
    if (lane < 4) {
        const int k = 4 * lane;
        ((RGH_G uint64_t *)(lv + 256))[lane] =
            i16 ? pack4x16(S.dcq[zz4(k)], S.dcq[zz4(k + 1)], S.dcq[zz4(k + 2)], S.dcq[zz4(k + 3)]) : 0ull;
    } else if (lane < 6) {
        const int p = lane - 4;
        ((RGH_G uint64_t *)(lv + 272))[p] = pack4x16(S.cdc[p][0], S.cdc[p][1], S.cdc[p][2], S.cdc[p][3]);
    } else if (lane >= 32) {
        const int l = lane - 32, p = l >> 4, b = (l >> 2) & 3, k = (l & 3) * 4;
        uint64_t w = 0;
        if (cbpc >= 2)
            w = pack4x16(k == 0 ? 0 : S.craster[p][b][zz4(k)], S.craster[p][b][zz4(k + 1)], S.craster[p][b][zz4(k + 2)],
                         S.craster[p][b][zz4(k + 3)]);
        ((RGH_G uint64_t *)(lv + 280))[l] = w;
    }
}

// The output buffers of a job, read again from its descriptor when a macroblock stores its
// results: the barrier on the descriptor's address makes the reads happen there, so these
// pointers do not hold scalar registers through the analysis.
struct JobOut {
    RGH_G uint8_t *cur_y, *cur_u, *cur_v, *pre_y, *pre_u, *pre_v;
    RGH_G int16_t *levels;
    RGH_G rgh_mb_decision *dec;
};
__device__ __forceinline__ JobOut late_outputs(const RghJob *jp) {
    const RGH_G RghJob *g = (const RGH_G RghJob *)jp;
    __asm__ volatile("" : "+s"(g));
    return {gp(g->cur_y), gp(g->cur_u), gp(g->cur_v), gp(g->pre_y), gp(g->pre_u), gp(g->pre_v), gp(g->levels), gp(g->dec)};
}

struct I16Code {
    int cbpl, cbpc, cmode;
    uint32_t pu, pv;
};

// Codes the macroblock as I_16x16 in mode A.mode: the prediction (into S.pred, then the
// reconstruction), chroma mode and prediction, luma AC / DC levels (S.coef, S.dcq), chroma
// levels and reconstruction.
template <bool CPLANE = true, bool CFORCED = false>
__device__ I16Code code_i16(MbLds &S, int mb_x, int mb_y, int qp, int qpc, const I16Analysis &A, uint32_t src_u,
                            uint32_t src_v, int cforced = 0) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    const int has_t = mb_y > 0, has_l = mb_x > 0;
    const int im = A.mode, pa = A.pa, pb = A.pb, pcc = A.pc, dc = A.dc;
    int cbpl = 0, cbpc = 0, cmode = 0;
    uint32_t pu = 0, pv = 0;
    {
        // The lane's four predicted samples (row rr, columns c0..c0 + 3); im is uniform.
        uint32_t w;
        if (im == 0) {
            w = has_t ? *(const uint32_t *)&S.etop[c0] : 0x80808080u;
        } else if (im == 1) {
            w = has_l ? (uint32_t)S.eleft[rr] * 0x01010101u : 0x80808080u;
        } else if (im == 2) {
            w = (uint32_t)dc * 0x01010101u;
        } else {
            w = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) w |= sat_shr<5>(pa + pb * (c0 + i - 7) + pcc * (rr - 7) + 16) << (8 * i);
        }
        *(uint32_t *)&S.pred[lane * 4] = w;
    }
    const ChromaIntra ci = chroma_intra<CPLANE, CFORCED>(S, mb_x, mb_y, src_u, src_v, cforced);
    cmode = ci.mode;
    pu = ci.pu;
    pv = ci.pv;
    S.predu[lane] = (uint8_t)pu;
    S.predv[lane] = (uint8_t)pv;
    __syncthreads();
    // Luma: 4x4 transforms, AC with intra rounding, DC through the 4x4 Hadamard
    // (enc.rs:3630-3660; quant_luma_dc :1959). One block row per lane: block lane >> 2 (decoding
    // order), row lane & 3.
    const int lb = lane >> 2, lrow = lane & 3, lbx = blk_x(lb) * 4, lby = blk_y(lb) * 4;
    const int lraster = (lby & ~3) + (lbx >> 2);  // raster index of block lb
    {
        int c[4];
        fdct4_row(S.src, S.pred, 16, lbx, lby, lrow, c);
        if (lrow == 0) S.dc_raw[lraster] = (int16_t)c[0];
        quant4_row(c, lrow, qp, 1);
        c[0] &= -(int)(lrow != 0);
        *(uint2 *)&S.coef[lb][lrow * 4] =
            make_uint2(((uint32_t)c[0] & 0xFFFFu) | ((uint32_t)c[1] << 16), ((uint32_t)c[2] & 0xFFFFu) | ((uint32_t)c[3] << 16));
        cbpl = __ballot((c[0] | c[1] | c[2] | c[3]) != 0) ? 15 : 0;
    }
    __syncthreads();
    {
        // simd/fixed.rs:318 hadamard4x4_scalar (with >> 1), then the DC quantizer;
        // h264.rs:1645 dequantize_idct_luma_dc4x4 for the reconstruction. Row lane & 3 of the
        // 4x4 DC matrix per lane (every quad computes it; lanes 0..3 store).
        const uint2 w = *(const uint2 *)&S.dc_raw[lrow * 4];
        const int r0 = (int16_t)w.x, r1 = (int)w.x >> 16, r2 = (int16_t)w.y, r3 = (int)w.y >> 16;
        const int s03 = r0 + r3, s12 = r1 + r2, d03 = r0 - r3, d12 = r1 - r2;
        const int t[4] = {s03 + s12, d03 + d12, s03 - s12, d03 - d12};
        const int qbits = 16 + qp / 6, mf = c_quant4_mf[qp % 6][0], f = (1 << qbits) / 3;
        int dq[4];
        const QuadMix hm = had_mix(lrow);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int had = (int16_t)(quad_mix(t[c], hm) >> 1);
            const int l = (int)((__umul24((uint32_t)iabs(had), (uint32_t)mf) + (uint32_t)f) >> qbits);
            dq[c] = had < 0 ? -l : l;
        }
        const int qmul = (c_dequant4_init[qp % 6][0] * 16) << (qp / 6 + 2);
        const int z0 = dq[0] + dq[1], z1 = dq[0] - dq[1], z2 = dq[2] - dq[3], z3 = dq[2] + dq[3];
        const int u[4] = {z0 + z3, z0 - z3, z1 - z2, z1 + z2};
        int dr[4];
#pragma unroll
        for (int c = 0; c < 4; c++) dr[c] = (int16_t)((quad_mix(u[c], hm) * qmul + 128) >> 8);
        if (lane < 4) {
            *(uint2 *)&S.dcq[lrow * 4] = make_uint2(((uint32_t)dq[0] & 0xFFFFu) | ((uint32_t)dq[1] << 16),
                                                    ((uint32_t)dq[2] & 0xFFFFu) | ((uint32_t)dq[3] << 16));
            *(uint2 *)&S.dcr[lrow * 4] = make_uint2(((uint32_t)dr[0] & 0xFFFFu) | ((uint32_t)dr[1] << 16),
                                                    ((uint32_t)dr[2] & 0xFFFFu) | ((uint32_t)dr[3] << 16));
        }
    }
    quant_chroma(S, qpc, 1);
    cbpc = chroma_cbp(S);
    {
        const uint2 w = *(const uint2 *)&S.coef[lb][lrow * 4];
        int blk[4] = {(int16_t)w.x, (int)w.x >> 16, (int16_t)w.y, (int)w.y >> 16};
        if (cbpl) dequant4_row(blk, lrow, qp);
        if (lrow == 0) blk[0] = S.dcr[lraster];
        add_idct4_row(S.pred, 16, lbx, lby, lrow, blk);
    }
    __syncthreads();
    recon_chroma(S, cbpc, qpc);
    return {cbpl, cbpc, cmode, pu, pv};
}

// Non-zero in some lane when a reference macroblock holding a sample of [px0, px1] x [py0, py1]
// (picture coordinates, clamped to the picture) is not final for a reader in launch `seq`.
// Macroblock (a, b) of the reference is final, padding included, once its deblock/pad item
// (a + 1, b + 1) is stamped by an earlier launch.
__device__ __forceinline__ int ref_box_pending(const RghJobD &J, int px0, int px1, int py0, int py1, uint32_t seq) {
    const int ax0 = iclamp(px0 >> 4, 0, J.mb_w - 1), ax1 = iclamp(px1 >> 4, 0, J.mb_w - 1);
    const int ay0 = iclamp(py0 >> 4, 0, J.mb_h - 1), ay1 = iclamp(py1 >> 4, 0, J.mb_h - 1);
    const int nx = ax1 - ax0 + 1, n = nx * (ay1 - ay0 + 1);
    int wait = 0;
    for (int i = (int)threadIdx.x; i < n; i += 64) {
        const int a = ax0 + i % nx, b = ay0 + i / nx;
        wait |= !stamp_final(J.ref_fin[(b + 1) * (J.mb_w + 1) + a + 1], J.ref_serial, seq);
    }
    return wait;
}

// ---------------------------------------------------------------------------------------
// Waiting in the persistent schedule.
// ---------------------------------------------------------------------------------------
// A producer publishes an item with its data stores completed in its XCD's L2 (s_waitcnt
// vmcnt(0)) before the stamp store; a consumer on the same XCD reads the stamp coherently (an
// agent-scope atomic load misses the L1), then invalidates its CU's L1 before reading the data.
__device__ __forceinline__ uint64_t load_stamp(const RGH_G uint64_t *p) {
    return __hip_atomic_load((const uint64_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish_stamp(RGH_G uint64_t *p, uint64_t v) {
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store((uint64_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The CU's L1 is invalidated at agent scope (buffer_inv sc1): the workgroup-scope form (sc0) does
// not invalidate it on MI355X, in or out of thread-group-split mode (stale lines of neighbouring
// macroblocks were read through it).
__device__ __forceinline__ void invalidate_l1() { __asm__ volatile("s_waitcnt vmcnt(0)\n\tbuffer_inv sc1" ::: "memory"); }

// How a macroblock's reads of the reference are made safe. Per-step launches (pw null): the
// reference's stamps must come from an earlier launch `seq`, otherwise the macroblock defers.
// Persistent schedule: it waits for stamps of items planned before its own time T; it defers
// when one is not (or its frame deferred, or the schedule gave up).
struct Gate {
    uint32_t seq;
    const RghPersist *pw;
    int T;        // this item's time
    int ref_job;  // the reference frame's job in this launch (-1: complete before the launch)
    int ref_base; // its base time
    // Per-step launches: the item was taken on the guess that its frame has not fallen behind
    // (spec). The analysis confirms the guess once its first round trip of loads is out and
    // before it uses any of them: the frame's delay (spec_delay) is 0 and the item's stamp
    // (spec_stamp) is not of this frame yet. Otherwise it sets *retry and returns false without
    // a side effect.
    int spec = 0, spec_delay = 0;
    uint64_t spec_stamp = 0;
    int *retry = nullptr;
};

// Waits until every lane's pending stamp *p carries `serial`. T_dep: the time of the item the lane
// waits for, dep_job its job. False when a pending item's frame has deferred at or before T_dep,
// or the schedule has given up (every wait then ends: the bounded sleep count sets the flag).
__device__ bool wait_stamps(bool pend, const RGH_G uint64_t *p, uint32_t serial, int T_dep, int dep_job,
                            const RghPersist *P) {
    RGH_G int32_t *deferred = gp(P->deferred);
    RGH_G uint32_t *status = gp(P->status);
    uint32_t spins = 0;
    for (;;) {
        if (pend) pend = (uint32_t)(load_stamp(p) >> 32) != serial;
        if (!__ballot(pend)) {
            if (P->debug && spins && threadIdx.x == 0) atomicAdd((uint32_t *)&status[3], spins);
            return true;
        }
        if ((spins & 7) == 7) {
            int gone = 0;
            if (pend) gone = __hip_atomic_load((int *)&deferred[dep_job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                             0x7fffffff - T_dep;
            if (__ballot(gone)) return false;
            if (__hip_atomic_load((uint32_t *)&status[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        }
        if (++spins > P->spin_limit) {
            if (threadIdx.x == 0) __hip_atomic_store((uint32_t *)&status[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Persistent schedule: waits until every reference macroblock holding a sample of
// [px0, px1] x [py0, py1] is final (deblock/pad item (a + 1, b + 1) of the reference stamped),
// then invalidates the L1; false when one cannot be waited for.
__device__ bool ref_box_wait(const RghJobD &J, int px0, int px1, int py0, int py1, const Gate &G) {
    const int ax0 = iclamp(px0 >> 4, 0, J.mb_w - 1), ax1 = iclamp(px1 >> 4, 0, J.mb_w - 1);
    const int ay0 = iclamp(py0 >> 4, 0, J.mb_h - 1), ay1 = iclamp(py1 >> 4, 0, J.mb_h - 1);
    const int nx = ax1 - ax0 + 1, n = nx * (ay1 - ay0 + 1);
    for (int i0 = 0; i0 < n; i0 += 64) {
        const int i = i0 + (int)threadIdx.x;
        const int a = ax0 + i % nx, b = ay0 + i / nx;
        const RGH_G uint64_t *p = J.ref_fin + (i < n ? (b + 1) * (J.mb_w + 1) + a + 1 : 0);
        const int T_dep = G.ref_base + (a + 1) + 2 * (b + 1) + 4;
        bool pend = i < n && (uint32_t)(load_stamp(p) >> 32) != J.ref_serial;
        if (!__ballot(pend)) continue;
        // Only items of this launch planned before this one are waited for.
        if (__ballot(pend && (G.ref_job < 0 || T_dep >= G.T))) return false;
        if (!wait_stamps(pend, p, J.ref_serial, T_dep, G.ref_job, G.pw)) return false;
    }
    invalidate_l1();
    return true;
}

// ---------------------------------------------------------------------------------------
// One macroblock (enc.rs:3011 code_p_mb and everything it calls).
// ---------------------------------------------------------------------------------------
//
// Returns false, having written nothing to global memory, when a reference sample the
// macroblock may read is not final yet for a reader in launch `seq`.
// LEARNED: the search starts from the macroblock's learned proposal (rgh_proposal: four cells
// of 4 samples and mode probabilities) instead of the reference's candidate list. Candidates
// [mvp, four proposals]; when the cheapest proposal costs more than learn_fallback percent of
// the predictor's cost, the reference's candidates (0, left, top, co-located) join and the
// diamond runs its full iteration count (the plain search), otherwise at most learn_iters
// iterations. The skip probe runs only when the skip probability reaches learn_skip_gate / 255,
// the intra 16x16 check only when the intra probability reaches learn_intra_gate / 255. Costs,
// sub-pel refinement, residual coding and reconstruction are the plain path's.
// PERSIST: the persistent schedule (the reference is waited for through G before any read of it;
// the neighbours, the co-located decision and the reference item of this macroblock were waited
// for by the caller); otherwise per-step launches (reads issued before the readiness verdict are
// discarded when the macroblock defers).
constexpr int PART_GATE = 40;   // partitions are searched above this 16x16 cost (lambda units)
constexpr int PART_RANGE = 8;   // their full-sample vectors lie within +-8 of the 16x16 one
constexpr int PART_ITERS = 8;   // diamond steps per partition

// One partition's search (RGH_FLAG_PARTITIONS) of W x ph samples at (px, py) in the macroblock,
// in the region staged at (pox, poy) around the 16x16 full-sample vector: candidates [predictor
// (pmx, pmy), 16x16 vector (m16x, m16y)] rounded to full samples and clamped into
// [lo0, hi0] x [lo1, hi1] (the second only when it differs), the first strictly cheapest of SAD +
// lambda x mvd bits; at most PART_ITERS diamond steps inside the box (enc.rs:1207's rule: the
// first of the cheapest points below the best moves); the half-sample diamond (one step) and the
// quarter-sample diamond (qpel_iters steps) within +-3 quarter samples of the full-sample result
// and inside the macroblock's vector bounds (enc.rs:1283's rule, the SAD through the window).
// Writes the partition's prediction into S.pred2 and returns its vector (x | y << 16).
template <int W>
__device__ __forceinline__ uint32_t part_search(const RghJobD &J, int x0, int y0, int px, int py, int ph, int pox,
                                                int poy, int lo0, int hi0, int lo1, int hi1, int pmx, int pmy,
                                                int m16x, int m16y, uint32_t lambda, int bmin0, int bmax0, int bmin1,
                                                int bmax1, MbLds &S) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    auto mvc = [&](int qx, int qy) { return lambda * (se_bits(qx - pmx) + se_bits(qy - pmy)); };
    const int c0x = iclamp((pmx + 2) >> 2, lo0, hi0), c0y = iclamp((pmy + 2) >> 2, lo1, hi1);
    const int c1x = iclamp((m16x + 2) >> 2, lo0, hi0), c1y = iclamp((m16y + 2) >> 2, lo1, hi1);
    int bx = c0x, by = c0y;
    uint32_t best;
    {
        const int fx[4] = {c0x, c1x, c0x, c0x}, fy[4] = {c0y, c1y, c0y, c0y};
        uint32_t sads[4];
        part_sad4<W>(x0, y0, px, py, ph, pox, poy, fx, fy, S, sads);
        best = sads[0] + mvc(4 * c0x, 4 * c0y);
        if (c1x != c0x || c1y != c0y) {
            const uint32_t c = sads[1] + mvc(4 * c1x, 4 * c1y);
            if (c < best) {
                best = c;
                bx = c1x;
                by = c1y;
            }
        }
    }
#pragma unroll 1
    for (int it = 0; it < PART_ITERS; it++) {
        const int fx[4] = {bx, bx - 1, bx + 1, bx}, fy[4] = {by - 1, by, by, by + 1};
        uint32_t sads[4];
        part_sad4<W>(x0, y0, px, py, ph, pox, poy, fx, fy, S, sads);
        uint32_t key = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ok = (int)(fx[k] >= lo0) & (int)(fx[k] <= hi0) & (int)(fy[k] >= lo1) & (int)(fy[k] <= hi1);
            key = min(key, ok ? ((sads[k] + mvc(4 * fx[k], 4 * fy[k])) << 2) | (uint32_t)k : 0xFFFFFFFFu);
        }
        if ((key >> 2) >= best) break;
        best = key >> 2;
        const int moved = (int)(key & 3u);
        bx += moved == 1 ? -1 : (moved == 2 ? 1 : 0);
        by += moved == 0 ? -1 : (moved == 3 ? 1 : 0);
    }
    // Sub-sample steps through the window anchored at the full-sample result.
    const int ax = x0 + px + bx, ay = y0 + py + by;
    compute_window(S, ax - 1 - pox, ay - 1 - poy);
    const int cx = bx * 4, cy = by * 4;
    int bmx = cx, bmy = cy;
#pragma unroll 1
    for (int pass = 0; pass < 2; pass++) {
        const int step = pass == 0 ? 2 : 1, iters = pass == 0 ? 1 : J.qpel_iters;
#pragma unroll 1
        for (int it = 0; it < iters; it++) {
            const int fx = bmx, fy = bmy;
            const int dx[4] = {0, -step, step, 0}, dy[4] = {-step, 0, 0, step};
            int mx[4], my[4], ok[4], qx[4], qy[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                mx[k] = fx + dx[k];
                my[k] = fy + dy[k];
                ok[k] = (int)(iabs(mx[k] - cx) <= 3) & (int)(iabs(my[k] - cy) <= 3) & (int)(mx[k] >= bmin0) &
                        (int)(mx[k] <= bmax0) & (int)(my[k] >= bmin1) & (int)(my[k] <= bmax1);
                qx[k] = ok[k] ? mx[k] : cx;
                qy[k] = ok[k] ? my[k] : cy;
            }
            uint32_t sads[4];
            part_subpel_sad4<W>(x0, y0, px, py, ph, ax, ay, pox, poy, qx, qy, S, sads);
            uint32_t key = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 0; k < 4; k++) key = min(key, ok[k] ? ((sads[k] + mvc(mx[k], my[k])) << 2) | (uint32_t)k : 0xFFFFFFFFu);
            if ((key >> 2) >= best) break;
            best = key >> 2;
            const int kb = (int)(key & 3u);
            bmx = fx + (kb == 1 ? -step : (kb == 2 ? step : 0));
            bmy = fy + (kb == 0 ? -step : (kb == 3 ? step : 0));
        }
    }
    if (rr >= py && rr < py + ph && c0 >= px && c0 < px + W)
        *(uint32_t *)&S.pred2[rr * 16 + c0] = window_pred4(x0, y0, ax, ay, pox, poy, bmx, bmy, S);
    return ((uint32_t)bmx & 0xFFFFu) | ((uint32_t)bmy << 16);
}

template <bool LEARNED, bool PERSIST = false, bool FAST = false, bool PARTS = false>
__device__ __forceinline__ bool analyse_mb(const RghJobD &J, const RghJob *jp, int mb_x, int mb_y, MbLds &S,
                                           const Gate &G) {
    const int lane = threadIdx.x;
    const int rr = lane >> 2, c0 = (lane & 3) * 4;
    const int mb_w = J.mb_w, idx = mb_y * mb_w + mb_x;
    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8;
    const int qp = J.qp, qpc = chroma_qp_of(qp, J.cqo);
    const uint32_t lambda = c_lambda[qp];
    const int W = mb_w * 16, H = J.mb_h * 16;

    // ---- Round trip 1: source, neighbour motion, co-located vector, intra edges. ----
    // Source macroblock, edge-replicated past the display size (enc.rs:389 MbSrc::load,
    // h264.rs:3660 extract_mb_planes).
    uint32_t srcw = 0;
    {
        const int sy = min(y0 + rr, J.height - 1), sx = x0 + c0;
        const RGH_G uint8_t *row = J.src_y + (long)sy * J.src_sy;
        if (sx + 3 < J.width && (((uintptr_t)(row + sx)) & 3) == 0) {
            srcw = *(const RGH_G uint32_t *)(row + sx);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) srcw |= (uint32_t)row[min(sx + i, J.width - 1)] << (8 * i);
        }
    }
    uint32_t src_u, src_v;
    {
        const int cw = (J.width + 1) / 2, ch = (J.height + 1) / 2;
        const int csy = min(cy0 + (lane >> 3), ch - 1), csx = min(cx0 + (lane & 7), cw - 1);
        src_u = J.src_u[(long)csy * J.src_su + csx];
        src_v = J.src_v[(long)csy * J.src_sv + csx];
    }
    // The learned proposal (16 B, uniform): words 0-1 the four cells, word 2 the probabilities.
    uint32_t pw0 = 0, pw1 = 0, pw2 = 0;
    if constexpr (LEARNED) {
        const RGH_G uint32_t *pp = (const RGH_G uint32_t *)&J.prop[idx];
        pw0 = pp[0];
        pw1 = pp[1];
        pw2 = pp[2];
    }
    // Neighbour motion (left, top, top-right else top-left decisions of this frame, and the
    // co-located decision of the reference frame) and intra edges (enc.rs:1457 i16_edges, :1476:
    // the undeblocked current picture) in one round trip: every lane issues one aligned dword
    // load before any value is used. Lanes 0-3 read decision words 0 (mb_type) and 2 (mv);
    // lanes 0-15 the left luma column, 16-31 the left chroma columns, 32-35 the top luma row,
    // 36-39 the top chroma rows, 40-42 the top-left corners. Unavailable neighbours read a valid
    // stand-in address and are discarded.
    const bool has_l = mb_x > 0, has_t = mb_y > 0, has_c = has_t && (mb_x + 1 < mb_w || has_l);
    const int ci = mb_x + 1 < mb_w ? idx - mb_w + 1 : idx - mb_w - 1;
    uint32_t dw0 = 0, dw2 = 0, dw5 = 0;
    const bool has_tr = has_t && mb_x + 1 < mb_w;
    if (lane < (FAST || PARTS ? 5 : 4)) {
        // Lane 4: the top-left decision (the Fast preset's skip-probe gate; with partitions the
        // top-left quadrant 3 vector, and lane 2 reads the top-right record only).
        const int di = lane == 0 ? (has_l ? idx - 1 : idx)
                                 : lane == 1 ? (has_t ? idx - mb_w : idx)
                                             : lane == 4 ? (has_l && has_t ? idx - mb_w - 1 : idx)
                                                         : (PARTS ? (has_tr ? idx - mb_w + 1 : idx) : (has_c ? ci : idx));
        const RGH_G uint32_t *dp = (const RGH_G uint32_t *)(lane == 3 ? &J.prev[idx] : &J.dec[di]);
        dw0 = dp[0];
        if constexpr (PARTS) {
            // Records in the quadrant layout (words 2-5): the blocks next to this macroblock are
            // the left one's quadrants 1 and 3, the top one's 2 and 3, the top-right one's 2, the
            // top-left one's 3 and the co-located one's 0 (the reference's prev_mb_mv).
            const uint64_t qa = *(const RGH_G uint64_t *)(dp + 2), qb = *(const RGH_G uint64_t *)(dp + 4);
            const uint32_t q0 = (uint32_t)qa, q1 = (uint32_t)(qa >> 32), q2 = (uint32_t)qb, q3 = (uint32_t)(qb >> 32);
            dw2 = lane == 0 ? q1 : (lane == 1 || lane == 2) ? q2 : lane == 3 ? q0 : q3;
            dw5 = q3;
        } else {
            dw2 = dp[2];
        }
    }
    uint32_t ew = 0;
    int eshift = 0;
    if (lane < 43) {
        const RGH_G uint8_t *ea;
        if (lane < 16) {
            ea = J.cur_y + (long)(y0 + lane) * J.sy + x0 - 1;
        } else if (lane < 32) {
            const RGH_G uint8_t *cp = ((lane >> 3) & 1) ? J.cur_v : J.cur_u;
            ea = cp + (long)(cy0 + (lane & 7)) * J.sc + cx0 - 1;
        } else if (lane < 36) {
            ea = J.cur_y + (long)(y0 - 1) * J.sy + x0 + 4 * (lane - 32);
        } else if (lane < 40) {
            const RGH_G uint8_t *cp = ((lane - 36) >> 1) ? J.cur_v : J.cur_u;
            ea = cp + (long)(cy0 - 1) * J.sc + cx0 + 4 * ((lane - 36) & 1);
        } else {
            const RGH_G uint8_t *cp = lane == 40 ? J.cur_y : (lane == 41 ? J.cur_u : J.cur_v);
            ea = cp + (long)((lane == 40 ? y0 : cy0) - 1) * (lane == 40 ? J.sy : J.sc) + (lane == 40 ? x0 : cx0) - 1;
        }
        eshift = 8 * (int)((uintptr_t)ea & 3);
        ew = *(const RGH_G uint32_t *)(ea - ((uintptr_t)ea & 3));
    }
    if (G.spec && (G.spec_delay != 0 || (uint32_t)(G.spec_stamp >> 32) == J.serial)) {
        *G.retry = 1;
        return false;
    }
    // Intra records (I_16x16, I_4x4, I_PCM; with partitions 5-7 are inter).
    auto intra_rec = [](uint32_t w0) { return PARTS ? (w0 & 0xFFu) - 2u <= 2u : (w0 & 0xFFu) >= 2u; };
    int lref = -2, tref = -2, dref = -2, ax = 0, ay = 0, bx_ = 0, by_ = 0, cxv = 0, cyv = 0;
    int nskip = 1;  // Fast: a P_Skip among the left, top, top-left and top-right neighbours
    {
        const uint32_t l0 = __builtin_amdgcn_readlane(dw0, 0), l2 = __builtin_amdgcn_readlane(dw2, 0);
        const uint32_t t0 = __builtin_amdgcn_readlane(dw0, 1), t2 = __builtin_amdgcn_readlane(dw2, 1);
        // With partitions lane 2 holds the top-right record only; C falls back to D (lane 4).
        const int cl = PARTS && mb_x + 1 >= mb_w ? 4 : 2;
        const uint32_t c0w = __builtin_amdgcn_readlane(dw0, cl), c2 = __builtin_amdgcn_readlane(dw2, cl);
        if constexpr (FAST) {
            const uint32_t tl0 = __builtin_amdgcn_readlane(dw0, 4);
            nskip = (int)(has_l && (l0 & 0xFF) == 0) | (int)(has_t && (t0 & 0xFF) == 0) |
                    (int)(has_t && mb_x + 1 < mb_w && (c0w & 0xFF) == 0) | (int)(has_l && has_t && (tl0 & 0xFF) == 0);
        }
        // Intra neighbours have no motion (an I_4x4 record's mv words hold its modes).
        if (has_l) {
            lref = intra_rec(l0) ? -1 : 0;
            ax = lref ? 0 : (int16_t)(l2 & 0xFFFF);
            ay = lref ? 0 : (int16_t)(l2 >> 16);
        }
        if (has_t) {
            tref = intra_rec(t0) ? -1 : 0;
            bx_ = tref ? 0 : (int16_t)(t2 & 0xFFFF);
            by_ = tref ? 0 : (int16_t)(t2 >> 16);
        }
        if (has_c) {
            dref = intra_rec(c0w) ? -1 : 0;
            cxv = dref ? 0 : (int16_t)(c2 & 0xFFFF);
            cyv = dref ? 0 : (int16_t)(c2 >> 16);
        }
    }
    const uint32_t col0 = __builtin_amdgcn_readlane(dw0, 3), col2 = __builtin_amdgcn_readlane(dw2, 3);
    const bool col_intra = intra_rec(col0);
    const int colx = col_intra ? 0 : (int16_t)(col2 & 0xFFFF), coly = col_intra ? 0 : (int16_t)(col2 >> 16);
    // With partitions: the neighbours' quadrant vectors that partition predictors read (packed
    // x | y << 16; zero for intra) and their reference indices (-2 outside the picture, -1 intra).
    uint32_t nq_l1 = 0, nq_l3 = 0, nq_t2 = 0, nq_t3 = 0, nq_tr = 0, nq_tl = 0;
    int nr_tr = -2, nr_tl = -2;
    if constexpr (PARTS) {
        nq_l1 = ((uint32_t)ax & 0xFFFFu) | ((uint32_t)ay << 16);
        nq_t2 = ((uint32_t)bx_ & 0xFFFFu) | ((uint32_t)by_ << 16);
        nq_l3 = lref == 0 ? __builtin_amdgcn_readlane(dw5, 0) : 0u;
        nq_t3 = tref == 0 ? __builtin_amdgcn_readlane(dw5, 1) : 0u;
        if (has_tr) {
            nr_tr = intra_rec(__builtin_amdgcn_readlane(dw0, 2)) ? -1 : 0;
            nq_tr = nr_tr == 0 ? __builtin_amdgcn_readlane(dw2, 2) : 0u;
        }
        if (has_l && has_t) {
            nr_tl = intra_rec(__builtin_amdgcn_readlane(dw0, 4)) ? -1 : 0;
            nq_tl = nr_tl == 0 ? __builtin_amdgcn_readlane(dw2, 4) : 0u;
        }
    }
    {
        const uint32_t eb = (ew >> eshift) & 0xFF;
        if (lane < 16)
            S.eleft[lane] = has_l ? (uint8_t)eb : 0;
        else if (lane < 32)
            S.cleft[(lane >> 3) & 1][lane & 7] = has_l ? (uint8_t)eb : 0;
        else if (lane < 36)
            *(uint32_t *)&S.etop[4 * (lane - 32)] = has_t ? ew : 0u;
        else if (lane < 40)
            *(uint32_t *)&S.ctop[(lane - 36) >> 1][4 * ((lane - 36) & 1)] = has_t ? ew : 0u;
        else if (lane == 40)
            S.etl = has_l && has_t ? (int)eb : 128;
        else if (lane < 43)
            S.ctl[lane - 41] = has_l && has_t ? (int)eb : 128;
    }
    *(uint32_t *)&S.src[lane * 4] = srcw;
    S.srcu[lane] = (uint8_t)src_u;
    S.srcv[lane] = (uint8_t)src_v;

    // Motion vector predictor (h264.rs:3577 pred_motion_partition, 16x16, ref 0).
    auto med3 = [](int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); };
    int mvpx, mvpy;
    {
        const int match = (lref == 0) + (tref == 0) + (dref == 0);
        if (match == 1) {
            mvpx = lref == 0 ? ax : (tref == 0 ? bx_ : cxv);
            mvpy = lref == 0 ? ay : (tref == 0 ? by_ : cyv);
        } else if (match == 0 && tref == -2 && dref == -2 && lref != -2) {
            mvpx = ax;
            mvpy = ay;
        } else {
            mvpx = med3(ax, bx_, cxv);
            mvpy = med3(ay, by_, cyv);
        }
    }
    // The P_Skip vector (enc.rs:3029-3044, 8.4.1.1).
    int pskx = mvpx, psky = mvpy;
    if (mb_x == 0 || mb_y == 0 || (lref == 0 && ax == 0 && ay == 0) || (tref == 0 && bx_ == 0 && by_ == 0)) pskx = psky = 0;

    // Vector bounds (enc.rs:1104 MvBounds::new).
    const int bmin0 = -4 * (x0 + RGH_MV_MARGIN), bmin1 = max(-4 * (y0 + RGH_MV_MARGIN), -2048);
    const int bmax0 = 4 * (W - 16 - x0 + RGH_MV_MARGIN), bmax1 = min(4 * (H - 16 - y0 + RGH_MV_MARGIN), 2044);
    auto in_bounds = [&](int mx, int my) { return (int)(mx >= bmin0) & (int)(mx <= bmax0) & (int)(my >= bmin1) & (int)(my <= bmax1); };
    auto mv_cost = [&](int mx, int my) { return lambda * (se_bits(mx - mvpx) + se_bits(my - mvpy)); };

    // Integer-search candidates (enc.rs:1185 full_pel, :3081-3095): [mvp, 0, left, top,
    // co-located], each rounded to full-pel, clamped and deduplicated. Fixed slots keep the
    // reference's list order; absent neighbours are not live. Learned: [mvp, proposals 0-3],
    // then the reference's list (its fallback) in slots 5-8.
    const int imn0 = (bmin0 + 3) >> 2, imn1 = (bmin1 + 3) >> 2, imx0 = bmax0 >> 2, imx1 = bmax1 >> 2;
    constexpr int NC = LEARNED ? 9 : 5;
    auto pcell = [&](int k, int comp) { return (int)(int8_t)(((k < 2 ? pw0 : pw1) >> (16 * (k & 1) + 8 * comp)) & 0xFF) * 16; };
    int cvx[NC], cvy[NC], present[NC];
    if constexpr (LEARNED) {
        cvx[0] = mvpx;
        cvy[0] = mvpy;
        present[0] = 1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cvx[1 + k] = pcell(k, 0);
            cvy[1 + k] = pcell(k, 1);
            present[1 + k] = 1;
        }
        const int ex[4] = {0, ax, bx_, colx}, ey[4] = {0, ay, by_, coly}, ep[4] = {1, mb_x > 0, mb_y > 0, 1};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cvx[5 + k] = ex[k];
            cvy[5 + k] = ey[k];
            present[5 + k] = ep[k];
        }
    } else {
        const int ex[5] = {mvpx, 0, ax, bx_, colx}, ey[5] = {mvpy, 0, ay, by_, coly}, ep[5] = {1, 1, mb_x > 0, mb_y > 0, 1};
#pragma unroll
        for (int k = 0; k < 5; k++) {
            cvx[k] = ex[k];
            cvy[k] = ey[k];
            present[k] = ep[k];
        }
    }
    int fxs[NC], fys[NC], live[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) {
        fxs[k] = iclamp((cvx[k] + 2) >> 2, imn0, imx0);
        fys[k] = iclamp((cvy[k] + 2) >> 2, imn1, imx1);
        live[k] = present[k];
#pragma unroll
        for (int j = 0; j < k; j++) live[k] &= !(live[j] && fxs[j] == fxs[k] && fys[j] == fys[k]);
    }

    // ---- Round trip 2: readiness stamps, candidate rows and the skip probe's samples. ----
    // Readiness of the reference (when it may still be in flight): while the diamond stays
    // within 16 samples of the centre c of the staged region (a candidate, or the point the walk
    // re-staged at), every reference read stays within [c - 22, c + 68] x [c - 19, c + 36] (the
    // staged 68x55 search region, the skip-probe window, the half-pel window and the chroma taps
    // around a vector at most 16 samples from c). This box around every candidate is checked
    // here; a walk that re-stages checks the box of its new centre then. The zero candidate keeps
    // this macroblock's co-located decision inside. The candidate rows and probe samples are
    // loaded before the verdict (loads only; nothing is written before it).
    int wait = 0;
    if (J.ref_fin) {
        int lx = fxs[0], hx = fxs[0], ly = fys[0], hy = fys[0];
#pragma unroll
        for (int k = 1; k < NC; k++) {
            lx = min(lx, fxs[k]);
            hx = max(hx, fxs[k]);
            ly = min(ly, fys[k]);
            hy = max(hy, fys[k]);
        }
        if constexpr (PERSIST) {
            if (!ref_box_wait(J, x0 + lx - 22, x0 + hx + 68, y0 + ly - 19, y0 + hy + 36, G)) return false;
        } else {
            wait = ref_box_pending(J, x0 + lx - 22, x0 + hx + 68, y0 + ly - 19, y0 + hy + 36, G.seq);
        }
    }
    uint32_t rw[NC];
    {
        const RGH_G uint8_t *rb = J.ref_y + (long)(y0 + rr) * J.sy + x0 + c0;
#pragma unroll
        for (int k = 0; k < NC; k++) rw[k] = live[k] ? ld4u(rb + (long)fys[k] * J.sy + fxs[k]) : 0;
    }
    const int p_skip = (int)(pw2 & 0xFF), p_intra = (int)((pw2 >> 16) & 0xFF);
    // The Fast preset probes only next to a skipped macroblock (enc.rs:3050-3063).
    const bool probe = in_bounds(pskx, psky) && (!LEARNED || p_skip >= J.learn_skip_gate) && (!FAST || nskip);
    const int int_y = ((pskx | psky) & 3) == 0;
    uint32_t predw = 0, pu = 0, pv = 0;
    ChromaTaps ct{};
    if (probe) {
        ct = chroma_taps(J, mb_x, mb_y, pskx, psky);
        if (int_y) predw = ld4u(J.ref_y + (long)(y0 + (psky >> 2) + rr) * J.sy + x0 + (pskx >> 2) + c0);
    }
    if (__ballot(wait)) return false;

    rgh_mb_decision D;
    memset(&D, 0, sizeof(D));
    D.mvp[0] = (int16_t)mvpx;
    D.mvp[1] = (int16_t)mvpy;
    D.qp = (uint8_t)qp;
    int mb_type = -1, mvx = 0, mvy = 0;
    int cbpl = 0, cbpc = 0, use8 = 0, i16_mode = 0, cmode = 0;
    int shape = 0;                    // partitions: 0 16x16, 1 16x8, 2 8x16, 3 8x8
    uint32_t qv[4] = {0, 0, 0, 0};    // ... and the chosen shape's quadrant vectors (x | y << 16)

    // ---- Skip probe (enc.rs:2906 try_p_skip; Balanced probes every macroblock). ----
    if (probe) {
        if (!int_y) {
            // Spec luma MC at the skip vector through the window anchored at its full-pel part
            // (the window table equals h264_luma_mc: enc.rs:4030 test).
            const int fbx = x0 + (pskx >> 2), fby = y0 + (psky >> 2);
            const int ox = (fbx - 3) & ~3, oy = fby - 3;
            stage_region<23>(J, ox, oy, S);
            compute_window(S, fbx - 1 - ox, fby - 1 - oy);
            predw = window_pred4(x0, y0, fbx, fby, ox, oy, pskx, psky, S);
        }
        chroma_mc_combine(ct, pskx, psky, &pu, &pv);
        // The luma SAD (below 2^16) with the U SAD (below 2^14) in its upper half, and the V SAD,
        // reduced together; the SSDs only when a test needs them.
        const uint32_t du = (uint32_t)iabs((int)src_u - (int)pu), dv = (uint32_t)iabs((int)src_v - (int)pv);
        const uint32_t in2[2] = {__builtin_amdgcn_sad_u8(srcw, predw, 0) | (du << 16), dv};
        uint32_t ps[2];
        wave_sums<2>(in2, ps);
        const uint32_t sad = ps[0] & 0xFFFFu;
        if (sad < J.sb_probe) {
            int luma_zero = sad <= J.sb_luma || wave_sum(ssd4(srcw, predw)) <= J.sb_luma_ssd;
            if (!luma_zero) {
                // enc.rs:2163 probe_skip_luma: no level above 1 and a total score below 6.
                *(uint32_t *)&S.pred[lane * 4] = predw;
                __syncthreads();
                // One block row per lane (block lane >> 2, row lane & 3).
                const int lb = lane >> 2, lrow = lane & 3;
                int c[4];
                fdct4_row(S.src, S.pred, 16, blk_x(lb) * 4, blk_y(lb) * 4, lrow, c);
                quant4_row(c, lrow, qp, 0);
                const uint32_t zz[4] = {0x09030200u, 0x0A080401u, 0x0E0B0705u, 0x0F0D0C06u};
                uint32_t m = 0;
                int big = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    m |= (uint32_t)(c[k] != 0) << ((zz[k] >> (8 * lrow)) & 15u);
                    big |= iabs(c[k]) > 1;
                }
                m = (uint32_t)quad_or((int)m);
                const int bigq = quad_or(big);
                const uint32_t score = lrow == 0 ? decimate_mask(m, bigq) : 0u;
                luma_zero = __ballot(big) == 0 && wave_sum(score) < 6;
            }
            if (luma_zero) {
                int chroma_zero = (ps[0] >> 16) <= J.sb_chroma || wave_sum(du * du) <= J.sb_chroma_ssd;
                if (chroma_zero) chroma_zero = ps[1] <= J.sb_chroma || wave_sum(dv * dv) <= J.sb_chroma_ssd;
                if (!chroma_zero) {
                    // enc.rs:2186 probe_skip_chroma: no DC level and AC scores below 7.
                    S.predu[lane] = (uint8_t)pu;
                    S.predv[lane] = (uint8_t)pv;
                    __syncthreads();
                    quant_chroma(S, qpc, 0);
                    int dcnz = 0;
#pragma unroll
                    for (int p = 0; p < 2; p++)
#pragma unroll
                        for (int i = 0; i < 4; i++) dcnz |= S.cdc[p][i];
                    const int su = S.cscore[0] + S.cscore[1] + S.cscore[2] + S.cscore[3];
                    const int sv = S.cscore[4] + S.cscore[5] + S.cscore[6] + S.cscore[7];
                    chroma_zero = !dcnz && su < 7 && sv < 7;
                }
                if (chroma_zero) {
                    mb_type = 0;
                    D.flags |= 1;
                    mvx = pskx;
                    mvy = psky;
                }
            }
        }
    }

    if constexpr (LEARNED) {
        if (in_bounds(pskx, psky) && !probe) D.flags |= 4;
    }
    if (mb_type < 0) {
        // ---- Integer candidates: the first strictly cheapest wins. ----
        int bestx = 0, besty = 0;
        uint32_t best_cost = 0xFFFFFFFFu, bestw = 0;
        int iters = J.me_iters;
        bool quick = false;  // learned: an intra-likely macroblock keeps its full-pel candidate
        // Every candidate's SAD, two to a reduction (each below 2^16).
        uint32_t csad[NC];
        {
            constexpr int NP = (NC + 1) / 2;
            uint32_t in[NP], out[NP];
#pragma unroll
            for (int k = 0; k < NP; k++)
                in[k] = __builtin_amdgcn_sad_u8(srcw, rw[2 * k], 0) |
                        (2 * k + 1 < NC ? __builtin_amdgcn_sad_u8(srcw, rw[2 * k + 1 < NC ? 2 * k + 1 : 0], 0) << 16 : 0u);
            wave_sums<NP>(in, out);
#pragma unroll
            for (int k = 0; k < NC; k++) csad[k] = (k & 1) ? out[k >> 1] >> 16 : out[k >> 1] & 0xFFFFu;
        }
        auto take = [&](int k) {
            if (!live[k]) return;
            const uint32_t cst = csad[k] + mv_cost(fxs[k] * 4, fys[k] * 4);
            D.n_sad++;
            if (cst < best_cost) {
                best_cost = cst;
                bestx = fxs[k];
                besty = fys[k];
                bestw = rw[k];
            }
        };
        if constexpr (LEARNED) {
            // The predictor (slot 0) is always live; a proposal equal to it is not evaluated twice.
            take(0);
            const uint32_t c_mvp = best_cost;
            best_cost = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 1; k < 5; k++) take(k);
            const uint32_t c_prop = best_cost == 0xFFFFFFFFu ? c_mvp : best_cost;  // all equal to the predictor
            if (c_mvp <= best_cost) {
                best_cost = c_mvp;
                bestx = fxs[0];
                besty = fys[0];
                bestw = rw[0];
            }
            quick = p_intra >= J.learn_intra_first;
            if (quick) {
                D.flags |= 16;
            } else if ((uint64_t)c_prop * 100u > (uint64_t)c_mvp * (uint32_t)J.learn_fallback) {
                D.flags |= 2;
#pragma unroll
                for (int k = 5; k < NC; k++) take(k);
            } else {
                iters = min(iters, J.learn_iters);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 5; k++) take(k);
        }

        uint32_t me_cost = best_cost;
        if (quick) {
            predw = bestw;
            mvx = bestx * 4;
            mvy = besty * 4;
        } else {
        // ---- Round trip 4: the reference region around the best candidate, in LDS. ----
        // Columns [bx - 19 (aligned down), +68) and rows [by - 19, +55) hold every 16x16 block
        // within 17 samples of the centre (bx, by) and the half-pel window of any of them.
        int scx = bestx, scy = besty;  // centre of the staged region (full-pel vector)
        int ox = (x0 + scx - 19) & ~3, oy = y0 + scy - 19;
        stage_region<SR_H>(J, ox, oy, S);
        // Small diamond (enc.rs:1207-1272): the first of the cheapest in-bounds points below the
        // best moves, until none does or the iterations run out. An iteration starts at most 15
        // samples from the staged centre, so its points lie within 16 of it, inside the region
        // (columns 3..38 + 16, rows 3..35 + 16), and the walk ends within 16 of it: the four SADs
        // read LDS without a branch, so their loads issue together; out-of-bounds points are
        // computed and ignored. The reference's 16 iterations never leave the first region; a
        // longer walk (me_range > 16) stages a new region around its best point once that point is
        // 16 samples from the centre, after checking that the reference macroblocks its box
        // reaches are final (else the macroblock defers, having written nothing).
        // Lane 16 k + r computes row r of point k's SAD (k: up, left, right, down) from the source
        // row in registers, and each row of 16 lanes sums one point.
        const int dk = lane >> 4, drow = lane & 15;
        const int ddx = dk == 1 ? -1 : (dk == 2 ? 1 : 0), ddy = dk == 0 ? -1 : (dk == 3 ? 1 : 0);
        const uint4 srow = *(const uint4 *)&S.src[drow * 16];
        // Fast: the four points' costs of the last iteration (enc.rs:1207 nb), in LDS.
        if (FAST && lane < 4) S.nb[lane] = 0xFFFFFFFFu;
        for (int it = 0; it < iters; it++) {
            if (iabs(bestx - scx) >= 16 || iabs(besty - scy) >= 16) {
                if (J.ref_fin) {
                    if constexpr (PERSIST) {
                        if (!ref_box_wait(J, x0 + bestx - 22, x0 + bestx + 68, y0 + besty - 19, y0 + besty + 36, G))
                            return false;
                    } else if (__ballot(ref_box_pending(J, x0 + bestx - 22, x0 + bestx + 68, y0 + besty - 19,
                                                        y0 + besty + 36, G.seq))) {
                        return false;
                    }
                }
                scx = bestx;
                scy = besty;
                ox = (x0 + scx - 19) & ~3;
                oy = y0 + scy - 19;
                stage_region<SR_H>(J, ox, oy, S);
            }
            // The centre is in range (the walk starts at a clamped candidate and moves only to
            // points in range), so each point has one bound to check.
            const int ok0 = (int)(besty > imn1), ok1 = (int)(bestx > imn0), ok2 = (int)(bestx < imx0),
                      ok3 = (int)(besty < imx1);
            uint32_t sad;
            {
                const int lx = x0 + bestx + ddx - ox, ly = y0 + besty + ddy - oy + drow;
                const uint32_t *q = (const uint32_t *)&S.sr[ly * SR_W + (lx & ~3)];
                const uint32_t f = (uint32_t)(lx & 3);
                const uint32_t w0 = q[0], w1 = q[1], w2 = q[2], w3 = q[3], w4 = q[4];
                sad = __builtin_amdgcn_sad_u8(srow.x, __builtin_amdgcn_alignbyte(w1, w0, f), 0u);
                sad = __builtin_amdgcn_sad_u8(srow.y, __builtin_amdgcn_alignbyte(w2, w1, f), sad);
                sad = __builtin_amdgcn_sad_u8(srow.z, __builtin_amdgcn_alignbyte(w3, w2, f), sad);
                sad = __builtin_amdgcn_sad_u8(srow.w, __builtin_amdgcn_alignbyte(w4, w3, f), sad);
                sad = row16_sum(sad);
            }
            // The first strictly cheapest in-range point: the least (cost, k) key (costs stay
            // below 2^30). Each 16-lane row forms its point's key (the vector cost needs only the
            // centre, so it is ready before the SAD), then the rows' keys meet by row broadcasts
            // in lane 63.
            uint32_t key, cst0 = 0, cst1 = 0, cst2 = 0, cst3 = 0;
            if constexpr (!FAST && !LEARNED) {  // (the learned kernels keep the scalar form: fewer spills there)
                const int okk = dk == 0 ? ok0 : (dk == 1 ? ok1 : (dk == 2 ? ok2 : ok3));
                const uint32_t mvk = lambda * (se_bits(4 * (bestx + ddx) - mvpx) + se_bits(4 * (besty + ddy) - mvpy));
                uint32_t kk = okk ? ((sad + mvk) << 2) | (uint32_t)dk : 0xFFFFFFFFu;
                kk = min(kk, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)kk, 0x142, 0xA, 0xF, false));  // row_bcast:15
                kk = min(kk, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)kk, 0x143, 0xC, 0xF, false));  // row_bcast:31
                key = (uint32_t)__builtin_amdgcn_readlane((int)kk, 63);
            } else {
                uint32_t sums[4];
#pragma unroll
                for (int k = 0; k < 4; k++) sums[k] = (uint32_t)__builtin_amdgcn_readlane((int)sad, 16 * k);
                // Vector costs from the three distinct lengths per component.
                const uint32_t ex0 = se_bits(4 * bestx - mvpx), exm = se_bits(4 * bestx - 4 - mvpx),
                               exp_ = se_bits(4 * bestx + 4 - mvpx);
                const uint32_t ey0 = se_bits(4 * besty - mvpy), eym = se_bits(4 * besty - 4 - mvpy),
                               eyp = se_bits(4 * besty + 4 - mvpy);
                cst0 = sums[0] + lambda * (ex0 + eym);
                cst1 = sums[1] + lambda * (exm + ey0);
                cst2 = sums[2] + lambda * (exp_ + ey0);
                cst3 = sums[3] + lambda * (ex0 + eyp);
                key = ok0 ? cst0 << 2 : 0xFFFFFFFFu;
                key = min(key, ok1 ? (cst1 << 2) | 1u : 0xFFFFFFFFu);
                key = min(key, ok2 ? (cst2 << 2) | 2u : 0xFFFFFFFFu);
                key = min(key, ok3 ? (cst3 << 2) | 3u : 0xFFFFFFFFu);
            }
            const int nok = ok0 + ok1 + ok2 + ok3;
            if (FAST && lane < 4) {
                const int k = lane;
                S.nb[k] = (k == 0 ? ok0 : k == 1 ? ok1 : k == 2 ? ok2 : ok3)
                              ? (k == 0 ? cst0 : k == 1 ? cst1 : k == 2 ? cst2 : cst3)
                              : 0xFFFFFFFFu;
            }
            D.n_sad += nok;
            if ((key >> 2) >= best_cost) break;
            best_cost = key >> 2;
            // The move (up, left, right, down) as nibble tables of offset + 1: scalar arithmetic
            // (a comparison turned into an integer costs a VALU round trip).
            const uint32_t sh = 4u * (key & 3u);
            bestx += (int)((0x1201u >> sh) & 15u) - 1;
            besty += (int)((0x2110u >> sh) & 15u) - 1;
        }

        // Sub-pel refinement unless the full-pel residual already quantizes to nothing
        // (enc.rs:3097-3117; enc.rs:1283 sub_pel: half-pel, then quarter-pel diamonds).
        me_cost = best_cost;
        const int fbx = x0 + bestx, fby = y0 + besty;
        int rox = ox, roy = oy;
        if (fbx - 3 < rox || fbx + 29 > rox + SR_COLS || fby - 3 < roy || fby + 20 > roy + SR_H) {
            rox = (fbx - 3) & ~3;
            roy = fby - 3;
            stage_region<23>(J, rox, roy, S);
        }
        if (best_cost - mv_cost(bestx * 4, besty * 4) <= J.sb_luma) {
            predw = lds4u(&S.sr[(fby - roy + rr) * SR_W], fbx - rox + c0);
            mvx = bestx * 4;
            mvy = besty * 4;
        } else if (FAST) {
            // enc.rs:1341 half_pel_toward: on each axis the half sample toward the cheaper
            // full-pel neighbour, estimated by the SAD against the average of its two integer
            // neighbours (out of bounds: no estimate); the better estimate (vertical on a tie),
            // if below the full-pel cost, is interpolated (the window's 6-tap planes) and taken if
            // its cost is below the full-pel cost.
            __syncthreads();
            const uint32_t n0 = S.nb[0], n1 = S.nb[1], n2 = S.nb[2], n3 = S.nb[3];
            const bool up = n0 < n3, left = n1 < n2;
            const int cx = bestx * 4, cy = besty * 4;
            const int vx = cx, vy = cy + (up ? -2 : 2), hx = cx + (left ? -2 : 2), hy = cy;
            const int ry = fby - roy + rr, rx = fbx - rox + c0;
            const uint32_t a = lds4u(&S.sr[ry * SR_W], rx);
            const uint32_t bv = lds4u(&S.sr[(ry + (up ? -1 : 1)) * SR_W], rx);
            const uint32_t bh = lds4u(&S.sr[ry * SR_W], rx + (left ? -1 : 1));
            const uint32_t sw = *(const uint32_t *)&S.src[lane * 4];
            const uint32_t in2[2] = {__builtin_amdgcn_sad_u8(sw, avg4(a, bv), 0u), __builtin_amdgcn_sad_u8(sw, avg4(a, bh), 0u)};
            uint32_t e2[2];
            wave_sums<2>(in2, e2);
            const int okv = in_bounds(vx, vy), okh = in_bounds(hx, hy);
            const uint32_t ev = okv ? e2[0] + mv_cost(vx, vy) : 0xFFFFFFFFu;
            const uint32_t eh = okh ? e2[1] + mv_cost(hx, hy) : 0xFFFFFFFFu;
            D.n_subpel += okv + okh;
            const bool vert = ev <= eh;
            const int hmx = vert ? vx : hx, hmy = vert ? vy : hy;
            bool took = false;
            if ((vert ? ev : eh) < me_cost) {
                compute_window(S, fbx - 1 - rox, fby - 1 - roy);
                D.n_hpel++;
                D.n_subpel++;
                const int qx[4] = {hmx, hmx, hmx, hmx}, qy[4] = {hmy, hmy, hmy, hmy};
                uint32_t sads[4];
                subpel_sad4(x0, y0, fbx, fby, rox, roy, qx, qy, S, sads);
                const uint32_t c = sads[0] + mv_cost(hmx, hmy);
                if (c < me_cost) {
                    me_cost = c;
                    predw = window_pred4(x0, y0, fbx, fby, rox, roy, hmx, hmy, S);
                    mvx = hmx;
                    mvy = hmy;
                    took = true;
                }
            }
            if (!took) {
                predw = a;
                mvx = cx;
                mvy = cy;
            }
        } else {
            const int cx = bestx * 4, cy = besty * 4;
            compute_window(S, fbx - 1 - rox, fby - 1 - roy);
            D.n_hpel++;
            // The half-pel diamond, then the quarter-pel diamond (costs, bounds and strict-< ties as
            // enc.rs:1283 sub_pel), each step's four SADs computed together. Every point lies
            // within +-3 quarter samples of the centre (cx, cy), inside the window; points out of
            // range are computed at the centre and ignored.
            int bmx = cx, bmy = cy;
            // The centre of each step is in range (the walk starts at the full-sample result and
            // moves only to points in range), so each point has one bound to check: the nearer of
            // the window's and the vector bounds' on its side.
            const int lo_x = max(cx - 3, bmin0), hi_x = min(cx + 3, bmax0);
            const int lo_y = max(cy - 3, bmin1), hi_y = min(cy + 3, bmax1);
            for (int pass = 0; pass < 2; pass++) {
                const int step = pass == 0 ? 2 : 1, iters = pass == 0 ? 1 : J.qpel_iters;
                for (int it = 0; it < iters; it++) {
                    const int fx = bmx, fy = bmy;
                    const int ok0 = (int)(fy - step >= lo_y), ok1 = (int)(fx - step >= lo_x),
                              ok2 = (int)(fx + step <= hi_x), ok3 = (int)(fy + step <= hi_y);
                    // Points out of range are computed at the centre and ignored.
                    const int qx[4] = {ok0 ? fx : cx, ok1 ? fx - step : cx, ok2 ? fx + step : cx, ok3 ? fx : cx};
                    const int qy[4] = {ok0 ? fy - step : cy, ok1 ? fy : cy, ok2 ? fy : cy, ok3 ? fy + step : cy};
                    // The first strictly cheapest point below me_cost: the least (cost, k) key
                    // over the in-range points (costs stay below 2^30), without branches.
                    uint32_t key;
                    if constexpr (!LEARNED) {
                        // Each 16-lane row forms its point's key (as the diamond does).
                        const uint32_t sadr = subpel_sad_rows(x0, y0, fbx, fby, rox, roy, qx, qy, S);
                        const int okk = dk == 0 ? ok0 : (dk == 1 ? ok1 : (dk == 2 ? ok2 : ok3));
                        const uint32_t mvk = lambda * (se_bits(fx + ddx * step - mvpx) + se_bits(fy + ddy * step - mvpy));
                        uint32_t kk = okk ? ((sadr + mvk) << 2) | (uint32_t)dk : 0xFFFFFFFFu;
                        kk = min(kk, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)kk, 0x142, 0xA, 0xF, false));
                        kk = min(kk, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)kk, 0x143, 0xC, 0xF, false));
                        key = (uint32_t)__builtin_amdgcn_readlane((int)kk, 63);
                    } else {
                        uint32_t sads[4];
                        subpel_sad4(x0, y0, fbx, fby, rox, roy, qx, qy, S, sads);
                        // Vector costs from the three distinct lengths per component.
                        const uint32_t ex0 = se_bits(fx - mvpx), exm = se_bits(fx - step - mvpx), exp_ = se_bits(fx + step - mvpx);
                        const uint32_t ey0 = se_bits(fy - mvpy), eym = se_bits(fy - step - mvpy), eyp = se_bits(fy + step - mvpy);
                        const uint32_t c0 = sads[0] + lambda * (ex0 + eym), c1 = sads[1] + lambda * (exm + ey0),
                                       c2 = sads[2] + lambda * (exp_ + ey0), c3 = sads[3] + lambda * (ex0 + eyp);
                        key = ok0 ? c0 << 2 : 0xFFFFFFFFu;
                        key = min(key, ok1 ? (c1 << 2) | 1u : 0xFFFFFFFFu);
                        key = min(key, ok2 ? (c2 << 2) | 2u : 0xFFFFFFFFu);
                        key = min(key, ok3 ? (c3 << 2) | 3u : 0xFFFFFFFFu);
                    }
                    D.n_subpel += ok0 + ok1 + ok2 + ok3;
                    if ((key >> 2) >= me_cost) break;
                    me_cost = key >> 2;
                    // The move (up, left, right, down) from nibble tables of offset + 1.
                    const uint32_t sh = 4u * (key & 3u);
                    bmx = fx + ((int)((0x1201u >> sh) & 15u) - 1) * step;
                    bmy = fy + ((int)((0x2110u >> sh) & 15u) - 1) * step;
                }
            }
            predw = window_pred4(x0, y0, fbx, fby, rox, roy, bmx, bmy, S);
            mvx = bmx;
            mvy = bmy;
        }
        }
        D.cost = me_cost;
        // ---- Round trip 5 (overlapped with the SATD): chroma taps at the final vector. ----
        ChromaTaps ct = chroma_taps(J, mb_x, mb_y, mvx, mvy);
        *(uint32_t *)&S.pred[lane * 4] = predw;
        __syncthreads();

        // Intra 16x16 when the inter prediction is poor (enc.rs:3119-3152). The Fast preset first
        // needs the best of the V / H / DC SADs within 1.75 times the inter SAD
        // (simd/h264_enc.rs:612 intra_sad_x3_16x16; missing edges read as zero).
        bool try_intra = true;
        if constexpr (FAST) {
            const uint32_t inter_sad = me_cost - mv_cost(mvx, mvy);
            const int dcv = i16_dc_of(S, mb_y > 0, mb_x > 0);
            const uint32_t sw = *(const uint32_t *)&S.src[lane * 4];
            const uint32_t tw = *(const uint32_t *)&S.etop[c0], lw = (uint32_t)S.eleft[rr] * 0x01010101u;
            const uint32_t in2[2] = {__builtin_amdgcn_sad_u8(sw, tw, 0u) | (__builtin_amdgcn_sad_u8(sw, lw, 0u) << 16),
                                     __builtin_amdgcn_sad_u8(sw, (uint32_t)dcv * 0x01010101u, 0u)};
            uint32_t o2[2];
            wave_sums<2>(in2, o2);
            uint32_t best = o2[1];
            if (mb_y > 0) best = min(best, o2[0] & 0xFFFFu);
            if (mb_x > 0) best = min(best, o2[0] >> 16);
            try_intra = 4u * best < 7u * inter_sad;
        }
        uint32_t inter_satd = 0;
        if (try_intra) {
            inter_satd = satd16(S.src, S.pred);
            D.n_satd++;
        }
        uint32_t inter_cost = inter_satd + lambda * (se_bits(mvx - mvpx) + se_bits(mvy - mvpy) + 1);
        if constexpr (PARTS) {
            // The 16x8, 8x16 and 8x8 shapes when the 16x16 cost exceeds PART_GATE lambda: each
            // partition searched in decoding order (part_search) around the 16x16 full-sample
            // vector (bestx, besty), all inside one staged region; a shape costs the 16x16 SATD
            // of its prediction + lambda x (mvd bits + mb_type bits: 3, 3, 3 + 4); the first
            // strictly cheapest of 16x16, 16x8, 8x16, 8x8 wins.
            const uint32_t v16 = ((uint32_t)mvx & 0xFFFFu) | ((uint32_t)mvy << 16);
            qv[0] = qv[1] = qv[2] = qv[3] = v16;
            if (inter_cost > (uint32_t)PART_GATE * lambda) {
                // Region: columns [x0 + bestx - 14 (aligned down), +68), rows [y0 + besty - 11,
                // +47): every partition vector within +-8 samples, its window and chroma taps.
                const int pox = (x0 + bestx - PART_RANGE - 3) & ~3, poy = y0 + besty - PART_RANGE - 3;
                if (J.ref_fin && __ballot(ref_box_pending(J, pox, pox + SR_COLS, poy, poy + 47, G.seq))) return false;
                stage_region<47>(J, pox, poy, S);
                const int lo0 = max(imn0, bestx - PART_RANGE), hi0 = min(imx0, bestx + PART_RANGE);
                const int lo1 = max(imn1, besty - PART_RANGE), hi1 = min(imx1, besty + PART_RANGE);
                // The 4x4 block (x, y) (4x4 units, relative to this macroblock) a predictor reads:
                // a neighbour's quadrant vector, or a decided quadrant of this shape (cq, done).
                auto nbk = [&](int x, int y, const uint32_t *cq, int done, int &r) -> uint32_t {
                    if (y < 0) {
                        if (x < 0) {
                            r = nr_tl;
                            return nq_tl;
                        }
                        if (x < 4) {
                            r = tref;
                            return x >= 2 ? nq_t3 : nq_t2;
                        }
                        r = nr_tr;
                        return nq_tr;
                    }
                    if (x < 0) {
                        r = lref;
                        return y >= 2 ? nq_l3 : nq_l1;
                    }
                    if (x >= 4) {
                        r = -2;
                        return 0u;
                    }
                    const int k = (y >> 1) * 2 + (x >> 1);
                    r = ((done >> k) & 1) ? 0 : -2;
                    return r == 0 ? cq[k] : 0u;
                };
                auto vx = [](uint32_t v) { return (int)(int16_t)(v & 0xFFFFu); };
                auto vy = [](uint32_t v) { return (int)(int16_t)(v >> 16); };
                uint32_t best_cost = inter_cost;
#pragma unroll 1
                for (int sh = 1; sh < 4; sh++) {
                    uint32_t cq[4] = {0, 0, 0, 0};
                    int done = 0;
                    uint32_t bits = sh == 3 ? 7u : 3u;
                    const int np = sh == 3 ? 4 : 2;
#pragma unroll 1
                    for (int p = 0; p < np; p++) {
                        const int px = sh == 2 ? 8 * p : (sh == 3 ? 8 * (p & 1) : 0);
                        const int py = sh == 1 ? 8 * p : (sh == 3 ? 8 * (p >> 1) : 0);
                        const int pw = sh == 1 ? 16 : 8, ph = sh == 2 ? 16 : 8;
                        // 8.4.1.3: A, B, C (else D), B and C from A when only A is available, the
                        // directional rules of 16x8 / 8x16, else the single match or the median.
                        const int bx4 = px >> 2, by4 = py >> 2, pw4 = pw >> 2;
                        int ra, rb, rc;
                        uint32_t va = nbk(bx4 - 1, by4, cq, done, ra), vb = nbk(bx4, by4 - 1, cq, done, rb);
                        uint32_t vc = nbk(bx4 + pw4, by4 - 1, cq, done, rc);
                        if (rc == -2) vc = nbk(bx4 - 1, by4 - 1, cq, done, rc);
                        if (rb == -2 && rc == -2 && ra != -2) {
                            vb = vc = va;
                            rb = rc = ra;
                        }
                        int dirk = -1;
                        if (sh == 1) dirk = p == 0 ? 1 : 0;
                        if (sh == 2) dirk = p == 0 ? 0 : 2;
                        uint32_t pv;
                        if (dirk >= 0 && (dirk == 0 ? ra : dirk == 1 ? rb : rc) == 0) {
                            pv = dirk == 0 ? va : dirk == 1 ? vb : vc;
                        } else {
                            const int hits = (ra == 0) + (rb == 0) + (rc == 0);
                            if (hits == 1) {
                                pv = ra == 0 ? va : (rb == 0 ? vb : vc);
                            } else {
                                const int mx_ = max(min(vx(va), vx(vb)), min(max(vx(va), vx(vb)), vx(vc)));
                                const int my_ = max(min(vy(va), vy(vb)), min(max(vy(va), vy(vb)), vy(vc)));
                                pv = ((uint32_t)mx_ & 0xFFFFu) | ((uint32_t)my_ << 16);
                            }
                        }
                        const int pmx = vx(pv), pmy = vy(pv);
                        const uint32_t mv = pw == 16 ? part_search<16>(J, x0, y0, px, py, ph, pox, poy, lo0, hi0, lo1, hi1,
                                                                        pmx, pmy, mvx, mvy, lambda, bmin0, bmax0, bmin1, bmax1, S)
                                                     : part_search<8>(J, x0, y0, px, py, ph, pox, poy, lo0, hi0, lo1, hi1,
                                                                       pmx, pmy, mvx, mvy, lambda, bmin0, bmax0, bmin1, bmax1, S);
                        bits += se_bits(vx(mv) - pmx) + se_bits(vy(mv) - pmy);
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int kx = (k & 1) * 8, ky = (k >> 1) * 8;
                            if (kx >= px && kx < px + pw && ky >= py && ky < py + ph) {
                                cq[k] = mv;
                                done |= 1 << k;
                            }
                        }
                    }
                    __syncthreads();
                    const uint32_t c = satd16(S.src, S.pred2) + lambda * bits;
                    if (c < best_cost) {
                        best_cost = c;
                        shape = sh;
                        *(uint32_t *)&S.pred[lane * 4] = *(const uint32_t *)&S.pred2[lane * 4];
#pragma unroll
                        for (int k = 0; k < 4; k++) qv[k] = cq[k];
                    }
                    __syncthreads();
                }
                inter_cost = best_cost;
            }
        }
        bool try_i16 = inter_satd > 24 * lambda;
        if constexpr (LEARNED) {
            if (try_i16 && p_intra < J.learn_intra_gate) {
                try_i16 = false;
                D.flags |= 8;
            }
        }
        if (try_i16) {
            const I16Analysis ia = analyse_i16<!FAST>(S, mb_x, mb_y, D.n_satd);
            // The Fast preset asks intra to win by twice the margin.
            if (ia.cost + lambda * (FAST ? 16u : 8u) < inter_cost) {
                // enc.rs:3587 code_i16 in a P slice (Fast: chroma in the direction of the luma mode).
                mb_type = 2;
                i16_mode = ia.mode;
                mvx = mvy = 0;
                const int from_luma = ia.mode == 0 ? 2 : (ia.mode == 1 ? 1 : (ia.mode == 2 ? 0 : 3));
                const I16Code ic = code_i16<true, FAST>(S, mb_x, mb_y, qp, qpc, ia, src_u, src_v, from_luma);
                cbpl = ic.cbpl;
                cbpc = ic.cbpc;
                cmode = ic.cmode;
                pu = ic.pu;
                pv = ic.pv;
            }
        }

        if (mb_type < 0) {
            // P_L0_16x16 residual (enc.rs:3154-3191), or P_Skip when it vanishes at the skip vector.
            if (PARTS && shape) {
                // Each chroma 4x4 block at its quadrant's vector (enc/inter.rs:229 predict_inter).
                const int k = ((lane >> 5) << 1) | ((lane >> 2) & 1);
                const uint32_t v = k == 0 ? qv[0] : k == 1 ? qv[1] : k == 2 ? qv[2] : qv[3];
                const int lmx = (int)(int16_t)(v & 0xFFFFu), lmy = (int)(int16_t)(v >> 16);
                const ChromaTaps t = chroma_taps(J, mb_x, mb_y, lmx, lmy);
                chroma_mc_combine(t, lmx, lmy, &pu, &pv);
            } else {
                chroma_mc_combine(ct, mvx, mvy, &pu, &pv);
            }
            S.predu[lane] = (uint8_t)pu;
            S.predv[lane] = (uint8_t)pv;
            // enc.rs:1810 quant_luma_4x4_inter: decimated 8x8 quadrants, then the whole MB. One
            // block row per lane: block lane >> 2 (decoding order), row lane & 3.
            const int lb = lane >> 2, lrow = lane & 3;
            int q4[4];
            {
                fdct4_row(S.src, S.pred, 16, blk_x(lb) * 4, blk_y(lb) * 4, lrow, q4);
                quant4_row(q4, lrow, qp, 0);
                // Scan positions of the row's coefficients (byte lrow of each word).
                const uint32_t zz[4] = {0x09030200u, 0x0A080401u, 0x0E0B0705u, 0x0F0D0C06u};
                uint32_t m = 0;
                int big = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    m |= (uint32_t)(q4[k] != 0) << ((zz[k] >> (8 * lrow)) & 15u);
                    big |= iabs(q4[k]) > 1;
                }
                m = (uint32_t)quad_or((int)m);
                big = quad_or(big);
                if (lrow == 0) {
                    S.bnz[lb] = m != 0;
                    S.bbig[lb] = big;
                    S.bscore[lb] = (int32_t)decimate_mask(m, big);
                }
            }
            __syncthreads();
            int keep = 0;
            {
                uint32_t total = 0;
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const int nz = S.bnz[4 * g] | S.bnz[4 * g + 1] | S.bnz[4 * g + 2] | S.bnz[4 * g + 3];
                    const int big = S.bbig[4 * g] | S.bbig[4 * g + 1] | S.bbig[4 * g + 2] | S.bbig[4 * g + 3];
                    if (!nz) continue;
                    if (big) {
                        keep |= 1 << g;
                        total += 9;
                        continue;
                    }
                    const uint32_t sc = (uint32_t)(S.bscore[4 * g] + S.bscore[4 * g + 1] + S.bscore[4 * g + 2] + S.bscore[4 * g + 3]);
                    if (sc >= 4) keep |= 1 << g;
                    total += sc;
                }
                if (total < 6) keep = 0;
            }
            cbpl = keep;
            {
                const int kept = -(int)((keep >> (lb >> 2)) & 1);
#pragma unroll
                for (int k = 0; k < 4; k++) q4[k] &= kept;
                *(uint2 *)&S.coef[lb][lrow * 4] = make_uint2(((uint32_t)q4[0] & 0xFFFFu) | ((uint32_t)q4[1] << 16),
                                                             ((uint32_t)q4[2] & 0xFFFFu) | ((uint32_t)q4[3] << 16));
            }
            __syncthreads();
            if (J.t8x8 && cbpl) {
                // enc.rs:2041 quant_luma_8x8_inter; the 8x8 transform wins when it leaves no
                // more non-zero levels (enc.rs:3170-3179).
                int col[8];
                fdct8_cols(S, col);
                if (lane < 32) {
                    const int blk = lane >> 3, c = lane & 7, rem = qp % 6, shift = 16 + qp / 6;
                    const uint32_t bias = (1u << shift) / 6u;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int i = j * 8 + c, v = col[j];
                        uint32_t lv = ((uint32_t)iabs(v) * (uint32_t)c_quant8_mf[rem][pos8(i)] + bias) >> shift;
                        lv = lv > 32767u ? 32767u : lv;
                        S.q8[blk][i] = (int16_t)(v < 0 ? -(int)lv : (int)lv);
                    }
                }
                __syncthreads();
                if (lane < 4) {
                    uint64_t m = 0;
                    int big = 0;
#pragma unroll
                    for (int i = 0; i < 64; i++) {
                        const int v = S.q8[lane][ZZ8[i]];
                        m |= (uint64_t)(v != 0) << i;
                        big |= iabs(v) > 1;
                    }
                    S.b8score[lane] = (int32_t)(m ? decimate_mask64(m, big) : 0);
                }
                __syncthreads();
                uint32_t total = 0;
                int cbp8 = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    total += (uint32_t)S.b8score[b];
                    if (S.b8score[b] >= 4) cbp8 |= 1 << b;
                }
                if (total < 6) cbp8 = 0;
                uint32_t n4 = (uint32_t)((q4[0] != 0) + (q4[1] != 0) + (q4[2] != 0) + (q4[3] != 0)), n8 = 0;
                if (lane < 32 && ((cbp8 >> (lane >> 3)) & 1)) {
#pragma unroll
                    for (int j = 0; j < 8; j++) n8 += S.q8[lane >> 3][j * 8 + (lane & 7)] != 0;
                }
                n4 = wave_sum(n4);
                n8 = wave_sum(n8);
                if (n8 <= n4) {
                    use8 = 1;
                    cbpl = cbp8;
                    if (lane < 32 && !((cbp8 >> (lane >> 3)) & 1)) {
#pragma unroll
                        for (int j = 0; j < 8; j++) S.q8[lane >> 3][j * 8 + (lane & 7)] = 0;
                    }
                }
                __syncthreads();
            }
            // enc.rs:2084 quant_chroma_inter: a plane's AC goes when it scores below 7.
            quant_chroma(S, qpc, 0);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int any = S.cmask[4 * p] | S.cmask[4 * p + 1] | S.cmask[4 * p + 2] | S.cmask[4 * p + 3];
                const int sc = S.cscore[4 * p] + S.cscore[4 * p + 1] + S.cscore[4 * p + 2] + S.cscore[4 * p + 3];
                if (any && sc < 7) {
                    __syncthreads();
                    if (lane < 4) {
                        S.cmask[4 * p + lane] = 0;
#pragma unroll
                        for (int i = 0; i < 16; i++) S.craster[p][lane][i] = 0;
                    }
                    __syncthreads();
                }
            }
            cbpc = chroma_cbp(S);
            if (shape == 0 && cbpl == 0 && cbpc == 0 && mvx == pskx && mvy == psky) {
                mb_type = 0;
            } else {
                mb_type = shape == 0 ? 1 : (shape == 1 ? 6 : (shape == 2 ? 7 : 5));
                // Reconstruction (enc.rs:3282-3306).
                if (use8) {
                    recon8(S, cbpl, qp);
                } else {
                    // Blocks outside the coded quadrants hold zeros and add nothing.
                    const uint2 w = *(const uint2 *)&S.coef[lb][lrow * 4];
                    int blk[4] = {(int16_t)w.x, (int)w.x >> 16, (int16_t)w.y, (int)w.y >> 16};
                    dequant4_row(blk, lrow, qp);
                    add_idct4_row(S.pred, 16, blk_x(lb) * 4, blk_y(lb) * 4, lrow, blk);
                    __syncthreads();
                }
                recon_chroma(S, cbpc, qpc);
            }
        }
    }

    // ---- Outputs: reconstruction, decision record and levels. ----
    if (mb_type == 0) {
        // P_Skip: the prediction is the reconstruction.
        *(uint32_t *)&S.pred[lane * 4] = predw;
        S.predu[lane] = (uint8_t)pu;
        S.predv[lane] = (uint8_t)pv;
        cbpl = cbpc = 0;
    }
    __syncthreads();
    const JobOut O = late_outputs(jp);
    *(uint32_t *)&O.cur_y[(long)(y0 + rr) * J.sy + x0 + c0] = *(const uint32_t *)&S.pred[lane * 4];
    store_chroma(S, O.cur_u, O.cur_v, J.sc, cx0, cy0);
    if (O.pre_y) {
        *(uint32_t *)&O.pre_y[(long)(y0 + rr) * W + x0 + c0] = *(const uint32_t *)&S.pred[lane * 4];
        store_chroma(S, O.pre_u, O.pre_v, W / 2, cx0, cy0);
    }

    RGH_G int16_t *lv = O.levels + (long)idx * RGH_LEVELS_PER_MB;
    uint32_t nnz = 0;
    const bool inter_coded = PARTS ? (mb_type == 1 || mb_type >= 5) : mb_type == 1;
    if (inter_coded || mb_type == 2) {
        if (inter_coded && use8) {
            for (int i = lane; i < 256; i += 64) lv[i] = S.q8[i >> 6][c_zigzag8[i & 63]];
            nnz = (cbpl & 1 ? 0x0033u : 0) | (cbpl & 2 ? 0x00CCu : 0) | (cbpl & 4 ? 0x3300u : 0) | (cbpl & 8 ? 0xCC00u : 0);
        } else {
            store_luma_levels(lv, S.coef);
            int nzb = 0;
            if (lane < 16 && !(mb_type == 2 && !cbpl)) {
#pragma unroll
                for (int i = 0; i < 16; i++) nzb |= S.coef[lane][i] != 0;
            }
            const uint64_t bl = __ballot(nzb);
#pragma unroll
            for (int b = 0; b < 16; b++)
                if ((bl >> b) & 1) nnz |= 1u << (blk_y(b) * 4 + blk_x(b));
        }
        store_dc_chroma_levels(lv, S, mb_type == 2, cbpc);
    } else {
        if (lane < RGH_LEVELS_PER_MB / 8) {  // 816 bytes, 16 a lane
            ((RGH_G uint64_t *)lv)[2 * lane] = 0;
            ((RGH_G uint64_t *)lv)[2 * lane + 1] = 0;
        }
    }
    if (lane == 0) {
        D.mb_type = (uint8_t)mb_type;
        D.i16_mode = (uint8_t)(mb_type == 2 ? i16_mode : 0);
        D.chroma_mode = (uint8_t)(mb_type == 2 ? cmode : 0);
        D.cbp_luma = (uint8_t)cbpl;
        D.cbp_chroma = (uint8_t)cbpc;
        D.transform_8x8 = (uint8_t)(inter_coded && use8 && cbpl);
        D.mv[0] = (int16_t)mvx;
        D.mv[1] = (int16_t)mvy;
        D.nnz = (uint16_t)nnz;
        if constexpr (PARTS) {
            // The quadrant layout: words 2-5 the quadrant vectors (zero for intra), word 6 zero.
            uint32_t w[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                w[k] = mb_type == 2 ? 0u
                       : mb_type == 0 ? (((uint32_t)pskx & 0xFFFFu) | ((uint32_t)psky << 16))
                                      : (shape ? qv[k] : (((uint32_t)mvx & 0xFFFFu) | ((uint32_t)mvy << 16)));
            memcpy(D.mv, &w[0], 4);
            memcpy(D.mvp, &w[1], 4);
            memcpy(&D.cost, &w[2], 4);
            D.n_sad = (uint16_t)(w[3] & 0xFFFFu);
            D.n_subpel = (uint16_t)(w[3] >> 16);
            D.n_satd = D.n_hpel = 0;
        }
        store_dec(&O.dec[idx], D);
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// IDR pictures: enc.rs:3311 code_intra — I_16x16 analysis, then I_4x4 (enc.rs:3333 analyse_i4)
// when the 16x16 SATD exceeds 40 lambda, coded by code_i4 (:3441) or code_i16 (:3587).
// ---------------------------------------------------------------------------------------
constexpr int CV_W = 24;  // canvas: row 0 = the top edge, rows 1-16 the macroblock; column 3 = the
                          // left edge, 4-19 the macroblock, 20-23 the top-right edge

// Intra 4x4 prediction taps (8.3.1.2, enc.rs:1489 pred4), for the lane of mode m = lane / 4 and
// row y = lane % 4: one byte per pixel x, (kind << 5) | i over the edge line G, where G[0..3] =
// p[-1, 3..0], G[4] = p[-1, -1], G[5..12] = p[0..7, -1] (the top-right substituted), G[13] =
// p[7, -1], G[14..21] = p[-1, 0..3] then p[-1, 3] repeated. Kinds: 0 G[i]; 1 (G[i] + G[i + 1] + 1)
// >> 1; 2 (G[i - 1] + 2 G[i] + G[i + 1] + 2) >> 2; 3 the DC value. Every prediction of the
// standard reduces to one of these per pixel.
__constant__ uint32_t c_i4_taps[36] = {
    0x08070605, 0x08070605, 0x08070605, 0x08070605, 0x03030303, 0x02020202, 0x01010101, 0x00000000,
    0x60606060, 0x60606060, 0x60606060, 0x60606060, 0x49484746, 0x4a494847, 0x4b4a4948, 0x4c4b4a49,
    0x47464544, 0x46454443, 0x45444342, 0x44434241, 0x27262524, 0x47464544, 0x26252443, 0x46454442,
    0x46454423, 0x44234322, 0x43224221, 0x42214120, 0x28272625, 0x49484746, 0x29282726, 0x4a494847,
    0x502f4f2e, 0x5130502f, 0x52315130, 0x53325231};

__device__ __forceinline__ int intra4_available(int mode, bool has_t, bool has_l) {  // h264.rs:3273
    return (mode == 0 || mode == 3 || mode == 7) ? has_t : ((mode == 1 || mode == 8) ? has_l : (mode == 2 ? 1 : has_t && has_l));
}

// Mode of block (bx, by) (0..3 each) of a packed decoding-order mode word.
__device__ __forceinline__ int mode_at(uint64_t packed, int bx, int by) {
    const int b = (bx & 1) | ((by & 1) << 1) | ((bx >> 1) << 2) | ((by >> 1) << 3);
    return (int)((packed >> (4 * b)) & 15);
}

// Canvas offset, relative to the block's top-left edge sample (row py, column px), of edge-line
// entry i (see c_i4_taps), with (tr) or without the top-right samples.
__device__ __forceinline__ int g4_offset(int i, bool tr) {
    if (i < 4) return (4 - i) * CV_W + 3;
    if (i == 4) return 3;
    if (i < 14) {
        const int k = min(i - 5, 7);
        return 4 + ((k < 4 || tr) ? k : 3);
    }
    return (1 + min(i - 14, 3)) * CV_W + 3;
}

// enc.rs:3333 analyse_i4 on the canvas (edges in row 0 / column 3), lane-parallel: per block (in
// decoding order) lanes 4m..4m+3 predict rows 0-3 of mode m; the SATD (simd/h264_enc.rs:196:
// sum of the Hadamard coefficients, halved) runs its column pass across the quad; cost = SATD +
// lambda * (1 for the predicted mode — min(left, top), DC at the picture's edges — else 4); the
// first cheapest available mode wins. Every quad also runs the forward transform (rows in the
// lane, columns across the quad), the intra quantizer, the dequantizer and the inverse
// transform (rows, then columns, then (v + 32) >> 6) before the winner is known; the winner's
// levels are kept and its reconstruction written to the canvas before the next block. Returns
// false as soon as the running cost (from 24 lambda) reaches budget. lmodes / tmodes: the left
// and top macroblocks' packed modes (all DC when they are not I_4x4).
__device__ bool analyse_i4(MbLds &S, int mb_x, int mb_y, int mb_w, int qp, uint32_t lambda, uint32_t budget,
                           uint64_t lmodes, uint64_t tmodes, uint64_t &modes, int &cbpl) {
    const int lane = threadIdx.x, m = lane >> 2, y = lane & 3;
    const uint32_t taps = c_i4_taps[lane < 36 ? lane : 0];
    // Canvas offsets of the three edge entries around each pixel's tap (bytes 3x .. 3x + 2),
    // with and without the top-right samples.
    uint32_t off_tr[3] = {0, 0, 0}, off_nt[3] = {0, 0, 0};
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const int i = (int)((taps >> (8 * x)) & 31);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int e = max(i - 1 + k, 0), slot = 3 * x + k;
            off_tr[slot >> 2] |= (uint32_t)g4_offset(e, true) << (8 * (slot & 3));
            off_nt[slot >> 2] |= (uint32_t)g4_offset(e, false) << (8 * (slot & 3));
        }
    }
    // Quantizer and dequantizer of the lane's row: position classes 0 / 2 / 0 / 2 (even rows) or
    // 2 / 1 / 2 / 1 (odd rows) across the columns.
    const int qshift = 15 + qp / 6, qrem = qp % 6, dsh = qp / 6;
    const uint32_t qbias = (1u << qshift) / 3u;
    const uint32_t mfe = (uint32_t)c_quant4_mf[qrem][(y & 1) ? 2 : 0], mfo = (uint32_t)c_quant4_mf[qrem][(y & 1) ? 1 : 2];
    const int dqe = c_dequant4_init[qrem][(y & 1) ? 2 : 0] << dsh, dqo = c_dequant4_init[qrem][(y & 1) ? 1 : 2] << dsh;
    const QuadMix fm = fdct_mix(y);
    const IdctMix im = idct_mix(y);
    uint32_t total = 24 * lambda;
    modes = 0;
    cbpl = 0;
    for (int b = 0; b < 16; b++) {
        const int bx = blk_x(b), by = blk_y(b), px = 4 * bx, py = 4 * by;
        const bool has_t = mb_y * 16 + py > 0, has_l = mb_x * 16 + px > 0;
        const bool tr = !(b == 3 || b == 7 || b == 11 || b == 13 || b == 15) && (b != 5 || mb_x + 1 < mb_w);
        const int lm = bx ? mode_at(modes, bx - 1, by) : mode_at(lmodes, 3, by);
        const int tm = by ? mode_at(modes, bx, by - 1) : mode_at(tmodes, bx, 3);
        const int pm = (mb_x * 4 + bx == 0 || mb_y * 4 + by == 0) ? 2 : min(lm, tm);
        const uint8_t *cb = &S.cv[py * CV_W + px];
        int dc;
        {
            const uint32_t tw = *(const uint32_t *)(cb + 4);
            const int st = (int)__builtin_amdgcn_sad_u8(tw, 0u, 0u);
            const int sl = cb[CV_W + 3] + cb[2 * CV_W + 3] + cb[3 * CV_W + 3] + cb[4 * CV_W + 3];
            dc = has_t && has_l ? (st + sl + 4) >> 3 : (has_t ? (st + 2) >> 2 : (has_l ? (sl + 2) >> 2 : 128));
        }
        // Prediction and residual of the lane's row.
        const uint32_t srow = *(const uint32_t *)&S.src[(py + y) * 16 + px];
        int p[4], d[4];
#pragma unroll
        for (int x = 0; x < 4; x++) {
            const int k = (int)((taps >> (8 * x + 5)) & 7);
            int g[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int slot = 3 * x + j;
                const uint32_t w = tr ? off_tr[slot >> 2] : off_nt[slot >> 2];
                g[j] = cb[(w >> (8 * (slot & 3))) & 0xFF];
            }
            // Every lane computes the three predictor forms and picks its own with masks: a
            // ternary chain here compiles to divergent branches around the loads.
            const int f1 = (g[1] + g[2] + 1) >> 1, f2 = (g[0] + 2 * g[1] + g[2] + 2) >> 2;
            const int v = (g[1] & -(int)(k == 0)) | (f1 & -(int)(k == 1)) | (f2 & -(int)(k == 2)) | (dc & -(int)(k > 2));
            p[x] = v;
            d[x] = (int)byte_of(srow, x) - v;
        }
        // SATD: row Hadamard in the lane, column butterflies across the quad.
        uint32_t sa;
        {
            const int s03 = d[0] + d[3], s12 = d[1] + d[2], d03 = d[0] - d[3], d12 = d[1] - d[2];
            int h[4] = {s03 + s12, d03 + d12, s03 - s12, d03 - d12};
            int acc = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int q1 = __builtin_amdgcn_update_dpp(0, h[c], 0xB1, 0xF, 0xF, false);  // row y ^ 1
                const int a = (y & 1) ? q1 - h[c] : h[c] + q1;
                const int q2 = __builtin_amdgcn_update_dpp(0, a, 0x4E, 0xF, 0xF, false);     // row y ^ 2
                acc += iabs((y & 2) ? q2 - a : a + q2);
            }
            acc += __builtin_amdgcn_update_dpp(0, acc, 0xB1, 0xF, 0xF, false);
            acc += __builtin_amdgcn_update_dpp(0, acc, 0x4E, 0xF, 0xF, false);
            sa = (uint32_t)acc >> 1;
        }
        const uint32_t cost = (lane < 36 && intra4_available(m, has_t, has_l)) ? sa + lambda * (m == pm ? 1u : 4u)
                                                                                 : 0xFFFFFFFFu;
        // Forward transform (rows, then columns across the quad) and intra quantization.
        int lv[4];
        {
            const int s03 = d[0] + d[3], s12 = d[1] + d[2], d03 = d[0] - d[3], d12 = d[1] - d[2];
            const int t[4] = {s03 + s12, (d03 << 1) + d12, s03 - s12, d03 - (d12 << 1)};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int cv = (int16_t)quad_mix(t[c], fm);
                uint32_t q = (__umul24((uint32_t)iabs(cv), (c & 1) ? mfo : mfe) + qbias) >> qshift;
                q = q > 32767u ? 32767u : q;
                lv[c] = cv < 0 ? -(int)q : (int)q;
            }
        }
        // Dequantization (simd/h264_enc.rs:461) and inverse transform (simd/fixed.rs:185) added to
        // the prediction; all-zero levels leave the prediction.
        uint32_t rec = 0;
        {
            int c[4];
#pragma unroll
            for (int x = 0; x < 4; x++) c[x] = iclamp(__mul24(lv[x], (x & 1) ? dqo : dqe), -32768, 32767);
            const int e0 = c[0] + c[2], e1 = c[0] - c[2], e2 = (c[1] >> 1) - c[3], e3 = c[1] + (c[3] >> 1);
            const int t[4] = {e0 + e3, e1 + e2, e1 - e2, e0 - e3};
#pragma unroll
            for (int k = 0; k < 4; k++) rec |= (uint32_t)clip255(p[k] + (int16_t)((quad_idct(t[k], im) + 32) >> 6)) << (8 * k);
        }
        // The first cheapest mode: the least (cost, mode) key over the quads (costs stay below
        // 2^27; unavailable modes and lanes 36-63 hold the largest key), reduced by DPP within
        // each row of 16 lanes (four quads), then over the three rows that hold modes.
        uint32_t key = cost == 0xFFFFFFFFu ? 0xFFFFFFFFu : (cost << 4) | (uint32_t)m;
        key = min(key, (uint32_t)__builtin_amdgcn_mov_dpp((int)key, 0x141, 0xF, 0xF, true));  // row_half_mirror
        key = min(key, (uint32_t)__builtin_amdgcn_mov_dpp((int)key, 0x140, 0xF, 0xF, true));  // row_mirror
        key = min(min((uint32_t)__builtin_amdgcn_readlane((int)key, 0), (uint32_t)__builtin_amdgcn_readlane((int)key, 16)),
                  (uint32_t)__builtin_amdgcn_readlane((int)key, 32));
        const uint32_t best = key == 0xFFFFFFFFu ? 0xFFFFFFFFu : key >> 4;
        const int bm = key == 0xFFFFFFFFu ? 2 : (int)(key & 15u);
        total += best;
        if (total >= budget) return false;
        modes |= (uint64_t)bm << (4 * b);
        const bool mine = m == bm;
        if (__ballot(mine && (lv[0] | lv[1] | lv[2] | lv[3]) != 0)) cbpl |= 1 << (b >> 2);
        if (mine) {
            *(uint2 *)&S.q4i[b][y * 4] = make_uint2((uint32_t)(uint16_t)lv[0] | ((uint32_t)(uint16_t)lv[1] << 16),
                                                    (uint32_t)(uint16_t)lv[2] | ((uint32_t)(uint16_t)lv[3] << 16));
            *(uint32_t *)&S.cv[(py + 1 + y) * CV_W + px + 4] = rec;
        }
        __syncthreads();
    }
    return true;
}

// One IDR macroblock: decision record, levels and reconstruction. FAST: the Fast preset's IDR
// analysis (no plane modes, no intra 4x4; enc.rs:3322-3330).
template <bool FAST = false>
__device__ __forceinline__ void analyse_intra_mb(const RghJobD &J, const RghJob *jp, int mb_x, int mb_y, MbLds &S) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    const int mb_w = J.mb_w, idx = mb_y * mb_w + mb_x;
    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8;
    const int qp = J.qp, qpc = chroma_qp_of(qp, J.cqo);
    const uint32_t lambda = c_lambda[qp];
    const int W = mb_w * 16;
    const bool has_l = mb_x > 0, has_t = mb_y > 0;

    // Round trip: source, the left / top decisions (their intra 4x4 modes), the edges
    // (luma top row with the top-right four samples, left column, corners; chroma edges).
    uint32_t srcw = 0;
    {
        const int sy = min(y0 + rr, J.height - 1), sx = x0 + c0;
        const RGH_G uint8_t *row = J.src_y + (long)sy * J.src_sy;
        if (sx + 3 < J.width && (((uintptr_t)(row + sx)) & 3) == 0) {
            srcw = *(const RGH_G uint32_t *)(row + sx);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) srcw |= (uint32_t)row[min(sx + i, J.width - 1)] << (8 * i);
        }
    }
    uint32_t src_u, src_v;
    {
        const int cw = (J.width + 1) / 2, ch = (J.height + 1) / 2;
        const int csy = min(cy0 + (lane >> 3), ch - 1), csx = min(cx0 + (lane & 7), cw - 1);
        src_u = J.src_u[(long)csy * J.src_su + csx];
        src_v = J.src_v[(long)csy * J.src_sv + csx];
    }
    uint32_t dw0 = 0, dw2 = 0, dw3 = 0;
    if (lane < 2) {
        const int di = lane == 0 ? (has_l ? idx - 1 : idx) : (has_t ? idx - mb_w : idx);
        const RGH_G uint32_t *dp = (const RGH_G uint32_t *)&J.dec[di];
        dw0 = dp[0];
        dw2 = dp[2];
        dw3 = dp[3];
    }
    uint32_t ew = 0;
    int eshift = 0;
    if (lane < 44) {
        const RGH_G uint8_t *ea;
        if (lane < 16) {
            ea = J.cur_y + (long)(y0 + lane) * J.sy + x0 - 1;
        } else if (lane < 32) {
            const RGH_G uint8_t *cp = ((lane >> 3) & 1) ? J.cur_v : J.cur_u;
            ea = cp + (long)(cy0 + (lane & 7)) * J.sc + cx0 - 1;
        } else if (lane < 37) {  // top luma row and the top-right dword
            ea = J.cur_y + (long)(y0 - 1) * J.sy + x0 + 4 * (lane - 32);
        } else if (lane < 41) {
            const RGH_G uint8_t *cp = ((lane - 37) >> 1) ? J.cur_v : J.cur_u;
            ea = cp + (long)(cy0 - 1) * J.sc + cx0 + 4 * ((lane - 37) & 1);
        } else {
            const RGH_G uint8_t *cp = lane == 41 ? J.cur_y : (lane == 42 ? J.cur_u : J.cur_v);
            ea = cp + (long)((lane == 41 ? y0 : cy0) - 1) * (lane == 41 ? J.sy : J.sc) + (lane == 41 ? x0 : cx0) - 1;
        }
        eshift = 8 * (int)((uintptr_t)ea & 3);
        ew = *(const RGH_G uint32_t *)(ea - ((uintptr_t)ea & 3));
    }
    const bool has_tr = has_t && mb_x + 1 < mb_w;
    {
        const uint32_t eb = (ew >> eshift) & 0xFF;
        if (lane < 16) {
            S.eleft[lane] = has_l ? (uint8_t)eb : 0;
        } else if (lane < 32) {
            S.cleft[(lane >> 3) & 1][lane & 7] = has_l ? (uint8_t)eb : 0;
        } else if (lane < 36) {
            *(uint32_t *)&S.etop[4 * (lane - 32)] = has_t ? ew : 0u;
        } else if (lane < 37) {
        } else if (lane < 41) {
            *(uint32_t *)&S.ctop[(lane - 37) >> 1][4 * ((lane - 37) & 1)] = has_t ? ew : 0u;
        } else if (lane == 41) {
            S.etl = has_l && has_t ? (int)eb : 128;
        } else if (lane < 44) {
            S.ctl[lane - 42] = has_l && has_t ? (int)eb : 128;
        }
    }
    *(uint32_t *)&S.src[lane * 4] = srcw;
    S.srcu[lane] = (uint8_t)src_u;
    S.srcv[lane] = (uint8_t)src_v;
    // Intra 4x4 modes of the neighbours (DC for anything but I_4x4).
    const uint32_t l0 = __builtin_amdgcn_readlane(dw0, 0), t0 = __builtin_amdgcn_readlane(dw0, 1);
    const uint64_t lpack = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(dw3, 0) << 32) | (uint32_t)__builtin_amdgcn_readlane(dw2, 0);
    const uint64_t tpack = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(dw3, 1) << 32) | (uint32_t)__builtin_amdgcn_readlane(dw2, 1);
    const uint64_t all_dc = 0x2222222222222222ull;
    const uint64_t lmodes = has_l && (l0 & 0xFF) == 3 ? lpack : all_dc, tmodes = has_t && (t0 & 0xFF) == 3 ? tpack : all_dc;
    __syncthreads();

    const uint32_t tr_word = has_tr ? (uint32_t)__builtin_amdgcn_readlane((int)ew, 36) : 0u;

    uint16_t n_satd = 0;
    const I16Analysis ia = analyse_i16<!FAST>(S, mb_x, mb_y, n_satd);
    int mb_type = 2, cbpl = 0, cbpc = 0, cmode = 0;
    uint32_t pu = 0, pv = 0;
    uint64_t modes = 0;
    bool i4 = false;
    if (!FAST && ia.cost > 40 * lambda) {
        // The canvas shares storage with the 16x16 analysis' plane prediction: filled now.
        __syncthreads();
        if (lane < 16)
            S.cv[(1 + lane) * CV_W + 3] = S.eleft[lane];
        else if (lane < 20)
            *(uint32_t *)&S.cv[4 + 4 * (lane - 16)] = *(const uint32_t *)&S.etop[4 * (lane - 16)];
        else if (lane == 20)
            *(uint32_t *)&S.cv[20] = tr_word;
        else if (lane == 21)
            S.cv[3] = (uint8_t)S.etl;
        __syncthreads();
        i4 = analyse_i4(S, mb_x, mb_y, mb_w, qp, lambda, ia.cost + 4 * lambda, lmodes, tmodes, modes, cbpl);
    }
    if (i4) {
        // enc.rs:3441 code_i4: the canvas holds the luma reconstruction.
        mb_type = 3;
        *(uint32_t *)&S.pred[lane * 4] = *(const uint32_t *)&S.cv[(1 + rr) * CV_W + 4 + c0];
        const ChromaIntra ci = chroma_intra(S, mb_x, mb_y, src_u, src_v);
        cmode = ci.mode;
        pu = ci.pu;
        pv = ci.pv;
        S.predu[lane] = (uint8_t)pu;
        S.predv[lane] = (uint8_t)pv;
        __syncthreads();
        quant_chroma(S, qpc, 1);
        cbpc = chroma_cbp(S);
        recon_chroma(S, cbpc, qpc);
    } else {
        cbpl = 0;
        const I16Code ic = code_i16<!FAST>(S, mb_x, mb_y, qp, qpc, ia, src_u, src_v);
        cbpl = ic.cbpl;
        cbpc = ic.cbpc;
        cmode = ic.cmode;
    }
    __syncthreads();
    const JobOut O = late_outputs(jp);
    *(RGH_G uint32_t *)&O.cur_y[(long)(y0 + rr) * J.sy + x0 + c0] = *(const uint32_t *)&S.pred[lane * 4];
    store_chroma(S, O.cur_u, O.cur_v, J.sc, cx0, cy0);
    if (O.pre_y) {
        *(RGH_G uint32_t *)&O.pre_y[(long)(y0 + rr) * W + x0 + c0] = *(const uint32_t *)&S.pred[lane * 4];
        store_chroma(S, O.pre_u, O.pre_v, W / 2, cx0, cy0);
    }
    RGH_G int16_t *lv = O.levels + (long)idx * RGH_LEVELS_PER_MB;
    uint32_t nnz = 0;
    {
        int nzb = 0;
        if (mb_type == 3) {
            store_luma_levels(lv, S.q4i);
            if (lane < 16)
#pragma unroll
                for (int i = 0; i < 16; i++) nzb |= S.q4i[lane][i] != 0;
        } else {
            store_luma_levels(lv, S.coef);
            if (lane < 16 && cbpl)
#pragma unroll
                for (int i = 0; i < 16; i++) nzb |= S.coef[lane][i] != 0;
        }
        const uint64_t bl = __ballot(nzb);
#pragma unroll
        for (int b = 0; b < 16; b++)
            if ((bl >> b) & 1) nnz |= 1u << (blk_y(b) * 4 + blk_x(b));
    }
    store_dc_chroma_levels(lv, S, mb_type == 2, cbpc);
    if (lane == 0) {
        rgh_mb_decision D;
        memset(&D, 0, sizeof(D));
        D.mb_type = (uint8_t)mb_type;
        D.i16_mode = (uint8_t)(mb_type == 2 ? ia.mode : 0);
        D.chroma_mode = (uint8_t)cmode;
        D.cbp_luma = (uint8_t)cbpl;
        D.cbp_chroma = (uint8_t)cbpc;
        D.qp = (uint8_t)qp;
        if (mb_type == 3) __builtin_memcpy(D.mv, &modes, sizeof(modes));
        D.nnz = (uint16_t)nnz;
        store_dec(&O.dec[idx], D);
    }
}

#include "rgh_compact.h"

union StepLds {
    MbLds a;
    DbLds d;
};
union CompactStepLds {
    CompactLds a;
    DbLds d;
};
template <int MODE>
struct StepLdsOf {
    using T = StepLds;
};
template <>
struct StepLdsOf<2> {
    using T = CompactStepLds;
};

// Waves per SIMD the Balanced P kernels are register-bounded for (4: at most 128 VGPRs).
#ifndef RGH_P_WAVES
#define RGH_P_WAVES 4
#endif

// One launch of a frame schedule, sequence number seq0 + rseq (seq0 from the job descriptors).
// The grid has 2 F rows of workgroups for the launch's F frame jobs j0 .. j0 + F - 1: row r < F
// analyses the macroblocks of frame job j0 + r's current wavefront step s = launch - base -
// delay, and row F + r runs that frame's deblock/pad items (x, y), 0 <= x <= mb_w,
// 0 <= y <= mb_h, of step s - 4 (item (x, y) deblocks macroblock (x, y) when it exists and pads
// macroblock (x - 1, y - 1)); the analyses, the longer items, are dispatched first. Column bx
// takes the item on row band_item_row(bx, first row of the step, band phase). The lag of 4 steps
// lets every analysis that reads undeblocked samples of (x, y) or of the samples its edges modify
// ((x + 1, y), (x - 1, y + 1), (x, y + 1), (x + 1, y + 1)) finish first.
//
// A macroblock whose reference samples are not final, or the first step of a frame whose picture
// slot is still read by a later frame, defers: it raises the frame's delay for the next launch
// and the whole frame repeats the step then (items already done skip themselves by their
// stamps). Nothing waits inside a launch.
//
// IDR jobs run the same schedule in their own kernel (rgh_k_intra_step), and Compact-preset P
// pictures in theirs (rgh_k_compact_step), so that each analysis has its own register and LDS
// budget; Balanced P pictures searched from learned proposals run rgh_k_learned_step. A launch
// that holds both IDR and Balanced P pictures runs them in one kernel (rgh_k_mixed_step, or
// rgh_k_mixed_learned_step with learned proposals), each block taking its picture's analysis,
// so that neither waits for the other. Fast-preset P pictures run their own pair (rgh_k_fast_step,
// rgh_k_mixed_fast_step, and rgh_k_intra_fast_step for its IDR pictures): the preset's branches
// stay out of the Balanced kernels' code. MODE: 0 Balanced P, 1 IDR, 2 Compact P, 3 Balanced P
// with learned proposals, 4 IDR and Balanced P, 5 IDR and Balanced P with learned proposals,
// 6 Fast P, 7 IDR and Fast P, 8 Fast IDR, 9 Balanced P with partitions (rgh_k_parts_step), 10 IDR
// and Balanced P with partitions (rgh_k_mixed_parts_step).
template <int MODE>
__device__ __forceinline__ void frame_step(const RghJob *__restrict__ jobs, int j0, int launch, uint32_t rseq) {
    const int nf = (int)(gridDim.y >> 1);
    const bool deb = (int)blockIdx.y >= nf;  // a deblock/pad row
    const int fr = deb ? (int)blockIdx.y - nf : (int)blockIdx.y;
    const RghJobD J = device_view(jobs[j0 + fr]);
    const uint32_t seq = J.seq0 + rseq;  // this launch's sequence number
    RGH_G RghFrameCtl *ctl = J.ctl;
    const int lane = threadIdx.x;
    const int delay = ctl->delay[launch & 1];
    const int s0 = launch - J.base;  // the frame's step unless it has fallen behind
    const int bx = (int)blockIdx.x, h = band_phase((uint64_t)J.src_y);
    __shared__ typename StepLdsOf<MODE>::T S;
    if (!deb) {
        // The item is first taken on the guess that the frame has not fallen behind (delay 0, the
        // common case): its stamp and the analysis's first round trip of loads go out together
        // with the delay, and the analysis confirms the guess before it uses any of them
        // (Gate::spec). Otherwise, and for the analyses without that check (IDR pictures, the
        // Compact preset), the item is taken from the delay read.
        constexpr bool GATED = MODE == 0 || MODE == 3 || MODE == 4 || MODE == 5 || MODE == 6 || MODE == 7 ||
                               MODE == 9 || MODE == 10;
        bool spec = GATED && !J.intra;
        int retry = 0;
#pragma unroll 1
        for (int pass = 0; pass < 2; pass++) {
            const int s = spec ? s0 : s0 - delay;
            bool valid = s >= 0 && s < J.steps;
            int y = 0;
            if (valid) {
                const int a = s - J.mb_w + 1;
                y = band_item_row(bx, a > 0 ? (a + 1) >> 1 : 0, h);
                valid = y <= min(J.mb_h - 1, s >> 1);
            }
            if (spec && (!valid || (s == 0 && J.wait_analysed))) {
                spec = false;
                continue;
            }
            if (!valid) break;
            const int x = s - 2 * y, m = y * J.mb_w + x;
            const uint64_t st = J.ana[m];
            if (!spec && (uint32_t)(st >> 32) == J.serial) break;  // analysed in an earlier launch
            bool ok = true;
            if (s == 0 && J.wait_analysed) ok = __atomic_load_n(J.wait_analysed, __ATOMIC_RELAXED) >= J.wait_total;
            if (ok) {
                const RghJob *jp = jobs + j0 + fr;
                const Gate G{seq, nullptr, 0, -1, 0, spec ? 1 : 0, delay, st, &retry};
                if constexpr (MODE == 1 || MODE == 8) {
                    analyse_intra_mb<MODE == 8>(J, jp, x, y, S.a);
                } else if constexpr (MODE == 4 || MODE == 5 || MODE == 7) {
                    if (J.intra)
                        analyse_intra_mb<MODE == 7>(J, jp, x, y, S.a);
                    else
                        ok = analyse_mb<MODE == 5, false, MODE == 7>(J, jp, x, y, S.a, G);
                } else if constexpr (MODE == 9) {
                    ok = analyse_mb<false, false, false, true>(J, jp, x, y, S.a, G);
                } else if constexpr (MODE == 10) {
                    if (J.intra)
                        analyse_intra_mb<false>(J, jp, x, y, S.a);
                    else
                        ok = analyse_mb<false, false, false, true>(J, jp, x, y, S.a, G);
                } else if constexpr (MODE == 2)
                    ok = analyse_compact_mb(J, x, y, S.a, seq);
                else
                    ok = analyse_mb<MODE == 3, false, MODE == 6>(J, jp, x, y, S.a, G);
            }
            if (retry) {  // the frame has fallen behind, or the item was done in an earlier launch
                spec = false;
                retry = 0;
                continue;
            }
            if (lane == 0) {
                // The stamp and progress buffers, read again from the descriptor (as late_outputs).
                const RGH_G RghJob *g = (const RGH_G RghJob *)(jobs + j0 + fr);
                __asm__ volatile("" : "+s"(g));
                RGH_G RghFrameCtl *lctl = gp(g->ctl);
                if (ok) {
                    gp(g->ana)[m] = make_stamp(g->serial, seq);
                    atomicAdd((int *)&lctl->analysed, 1);
                } else {
                    atomicMax((int *)&lctl->delay[(launch + 1) & 1], delay + 1);
                }
            }
            break;
        }
        // The delay carries over to the next launch (raised there by deferrals of this one).
        if (bx == 0 && lane == 0 && delay > 0) atomicMax((int *)&ctl->delay[(launch + 1) & 1], delay);
    } else {
        const int sd = s0 - delay - 4;
        if (sd < 0 || sd > J.steps + 2) return;
        const int a = sd - J.mb_w;
        const int y = band_item_row(bx, a > 0 ? (a + 1) >> 1 : 0, h);
        if (y > min(J.mb_h, sd >> 1)) return;
        const int x = sd - 2 * y, m = y * (J.mb_w + 1) + x;
        if ((uint32_t)(J.fin[m] >> 32) == J.serial) return;  // done in an earlier launch
        if constexpr (MODE >= 4) {
            // The mixed kernels inline both analyses; the deblock and pad items stay inline too.
            if (x < J.mb_w && y < J.mb_h && J.deblock) deblock_mb_inline(J, x, y, S.d);
            if (x > 0 && y > 0) pad_mb_inline(J, x - 1, y - 1);
        } else {
            if (x < J.mb_w && y < J.mb_h && J.deblock) deblock_mb(J, x, y, S.d);
            if (x > 0 && y > 0) pad_mb(J, x - 1, y - 1);
        }
        if (lane == 0) J.fin[m] = make_stamp(J.serial, seq);
    }
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_frame_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                   uint32_t seq) {
    frame_step<0>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_intra_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                   uint32_t seq) {
    frame_step<1>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_mixed_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                   uint32_t seq) {
    frame_step<4>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_mixed_learned_step(const RghJob *__restrict__ jobs, int j0,
                                                                           int launch, uint32_t seq) {
    frame_step<5>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_intra_fast_step(const RghJob *__restrict__ jobs, int j0,
                                                                        int launch, uint32_t seq) {
    frame_step<8>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_fast_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                  uint32_t seq) {
    frame_step<6>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_mixed_fast_step(const RghJob *__restrict__ jobs, int j0,
                                                                        int launch, uint32_t seq) {
    frame_step<7>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_compact_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                     uint32_t seq) {
    frame_step<2>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_parts_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                   uint32_t seq) {
    frame_step<9>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_mixed_parts_step(const RghJob *__restrict__ jobs, int j0,
                                                                         int launch, uint32_t seq) {
    frame_step<10>(jobs, j0, launch, seq);
}

extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RGH_P_WAVES))) void rgh_k_learned_step(const RghJob *__restrict__ jobs, int j0, int launch,
                                                                     uint32_t seq) {
    frame_step<3>(jobs, j0, launch, seq);
}

// Fills n 64-bit words (stamps) with v.
extern "C" __global__ __launch_bounds__(256) void rgh_k_fill_u64(uint64_t *p, int n, uint64_t v) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------
// The persistent schedule: one launch runs every item of a call.
// ---------------------------------------------------------------------------------------

// The XCC (XCD) the calling wave runs on.
__device__ __forceinline__ int xcc_id() {
    uint32_t v;
    __asm__ volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & 15u);
}

// Sets bit xcc_id() of *mask for every workgroup: which XCC ids the device's waves report.
extern "C" __global__ __launch_bounds__(64) void rgh_k_xcc_probe(uint32_t *mask) {
    if (threadIdx.x == 0) atomicOr(mask, 1u << xcc_id());
}

// Recounts the analysed macroblocks of job blockIdx.x (stamps of its serial) into its progress
// record, for the per-step schedule's picture-slot gate after a persistent launch.
extern "C" __global__ __launch_bounds__(64) void rgh_k_recount(const RghJob *__restrict__ jobs) {
    const RghJobD J = device_view(jobs[blockIdx.x]);
    const int n = J.mb_w * J.mb_h;
    uint32_t c = 0;
    for (int m = (int)threadIdx.x; m < n; m += 64) c += (uint32_t)(J.ana[m] >> 32) == J.serial;
    c = wave_sum(c);
    if (threadIdx.x == 0) J.ctl->analysed = (int32_t)c;
}

// Analysis macroblocks of a frame at wavefront step s: rows [y0, y0 + n) (n <= 0: none).
__device__ __forceinline__ void step_rows(int s, int mb_w, int mb_h, int steps, int &y0, int &n) {
    const int a = s - mb_w + 1;
    y0 = a > 0 ? (a + 1) >> 1 : 0;
    n = (s < 0 || s >= steps) ? 0 : min(mb_h - 1, s >> 1) - y0 + 1;
}
// Deblock/pad items ((mb_w + 1) x (mb_h + 1) grid) at step sd: rows [y0, y0 + n).
__device__ __forceinline__ void item_rows(int sd, int mb_w, int mb_h, int steps, int &y0, int &n) {
    const int a = sd - mb_w;
    y0 = a > 0 ? (a + 1) >> 1 : 0;
    n = (sd < 0 || sd > steps + 2) ? 0 : min(mb_h, sd >> 1) - y0 + 1;
}

__device__ __forceinline__ void defer_item(const RghPersist *P, int j, int T) {
    if (threadIdx.x == 0) {
        atomicMax((int *)&gp(P->deferred)[j], 0x7fffffff - T);
        atomicAdd((uint32_t *)&gp(P->status)[1], 1u);
    }
}

// One item (job j, time T, index r among the job's items of that time): an analysis macroblock,
// else a deblock/pad item. It waits for the items it reads (all planned before T): an analysis
// for its left, top, top-right and top-left neighbours (decisions and undeblocked edge samples)
// and, in a P picture, for the reference's item (x + 1, y + 1) (the co-located decision and the
// previous picture's neighbouring decisions); a deblock/pad item (x, y) for the analyses that read
// samples it modifies ((x, y), (x + 1, y), (x - 1, y + 1), (x, y + 1), (x + 1, y + 1)) and the
// items (x - 1, y), (x, y - 1), (x + 1, y - 1) that modify samples it reads (the reference's
// raster deblocking order). The first step of a frame also waits until the frame that last read
// its picture slot has been analysed.
template <int MODE>
__device__ __forceinline__ void persist_item(const RghJob *__restrict__ jobs, const RghPersist *P, int j, int T, int r,
                                             uint32_t seq, StepLds &S) {
    const uint64_t t_item0 = __builtin_amdgcn_s_memtime();
    const RghJobD J = device_view(jobs[j]);
    const RghJob *jp = jobs + j;
    const int lane = threadIdx.x, s = T - J.base;
    int ya, na;
    step_rows(s, J.mb_w, J.mb_h, J.steps, ya, na);
    if (r < na) {
        const int y = ya + r, x = s - 2 * y, m = y * J.mb_w + x;
        if ((uint32_t)(load_stamp(&J.ana[m]) >> 32) == J.serial) return;  // done by an earlier launch
        const int ref_base = J.ref_job >= 0 ? jobs[J.ref_job].base : 0;
        bool has = false;
        const RGH_G uint64_t *p = J.ana;
        int T_dep = 0, dj = j;
        uint32_t ser = J.serial;
        if (lane == 0 && x > 0) {
            p = &J.ana[m - 1];
            T_dep = T - 1;
            has = true;
        } else if (lane == 1 && y > 0) {
            p = &J.ana[m - J.mb_w];
            T_dep = T - 2;
            has = true;
        } else if (lane == 2 && y > 0 && x + 1 < J.mb_w) {
            p = &J.ana[m - J.mb_w + 1];
            T_dep = T - 1;
            has = true;
        } else if (lane == 3 && y > 0 && x > 0) {
            p = &J.ana[m - J.mb_w - 1];
            T_dep = T - 3;
            has = true;
        } else if (lane == 4 && !J.intra) {
            p = &J.ref_fin[(y + 1) * (J.mb_w + 1) + x + 1];
            T_dep = ref_base + (x + 1) + 2 * (y + 1) + 4;
            ser = J.ref_serial;
            dj = J.ref_job;
            has = true;
        }
        bool pend = has && (uint32_t)(load_stamp(p) >> 32) != ser;
        if (__ballot(pend && (dj < 0 || T_dep >= T)) || !wait_stamps(pend, p, ser, T_dep, dj, P)) {
            defer_item(P, j, T);
            return;
        }
        if (s == 0 && J.wait_job >= 0) {
            // The picture slot is free once the frame that last read it is analysed (planned to
            // finish before this frame starts): its last macroblock's analysis waited, through its
            // neighbours, for every other one. (A gating frame outside this launch is complete.)
            const RghJob *wjp = jobs + J.wait_job;
            const RGH_G uint64_t *wp = gp(wjp->ana) + (wjp->mb_w * wjp->mb_h - 1);
            const int wT = wjp->base + wjp->steps - 1;
            const uint32_t wser = wjp->serial;
            const bool wpend = lane == 0 && (uint32_t)(load_stamp(wp) >> 32) != wser;
            if (!wait_stamps(wpend, wp, wser, wT, J.wait_job, P)) {
                defer_item(P, j, T);
                return;
            }
        }
        invalidate_l1();
        const uint64_t t_run0 = __builtin_amdgcn_s_memtime();
        bool ok = true;
        if (J.intra)
            analyse_intra_mb<MODE == 2>(J, jp, x, y, S.a);
        else
            ok = analyse_mb<MODE == 1, true, MODE == 2>(J, jp, x, y, S.a, Gate{seq, P, T, J.ref_job, ref_base});
        if (!ok) {
            defer_item(P, j, T);
            return;
        }
        if (lane == 0) {
            const RGH_G RghJob *g = (const RGH_G RghJob *)jp;
            __asm__ volatile("" : "+s"(g));
            publish_stamp(&gp(g->ana)[m], make_stamp(g->serial, seq));
            if (P->debug) {
                const uint64_t t_run1 = __builtin_amdgcn_s_memtime();
                atomicAdd((uint32_t *)&gp(P->status)[2], 1u);
                atomicAdd((uint32_t *)&gp(P->status)[4], (uint32_t)((t_run0 - t_item0) >> 8));
                atomicAdd((uint32_t *)&gp(P->status)[5], (uint32_t)((t_run1 - t_run0) >> 8));
                atomicAdd((uint32_t *)&gp(P->status)[6], (uint32_t)((t_run1 - t_run0) >> 8));
                atomicAdd((uint32_t *)&gp(P->status)[7], 1u);
            }
        }
        return;
    }
    const int sd = s - 4;
    int yd, nd;
    item_rows(sd, J.mb_w, J.mb_h, J.steps, yd, nd);
    const int y = yd + (r - na), x = sd - 2 * y, iw = J.mb_w + 1, m = y * iw + x;
    if ((uint32_t)(load_stamp(&J.fin[m]) >> 32) == J.serial) return;
    bool has = false;
    const RGH_G uint64_t *p = J.ana;
    int T_dep = 0;
    const bool in_x = x < J.mb_w, in_y = y < J.mb_h, nx_ = x + 1 < J.mb_w, ny_ = y + 1 < J.mb_h;
    if (lane == 0 && in_x && in_y) {
        p = &J.ana[y * J.mb_w + x];
        T_dep = T - 4;
        has = true;
    } else if (lane == 1 && nx_ && in_y) {
        p = &J.ana[y * J.mb_w + x + 1];
        T_dep = T - 3;
        has = true;
    } else if (lane == 2 && x > 0 && ny_) {
        p = &J.ana[(y + 1) * J.mb_w + x - 1];
        T_dep = T - 3;
        has = true;
    } else if (lane == 3 && in_x && ny_) {
        p = &J.ana[(y + 1) * J.mb_w + x];
        T_dep = T - 2;
        has = true;
    } else if (lane == 4 && nx_ && ny_) {
        p = &J.ana[(y + 1) * J.mb_w + x + 1];
        T_dep = T - 1;
        has = true;
    } else if (lane == 5 && x > 0) {
        p = &J.fin[m - 1];
        T_dep = T - 1;
        has = true;
    } else if (lane == 6 && y > 0) {
        p = &J.fin[m - iw];
        T_dep = T - 2;
        has = true;
    } else if (lane == 7 && y > 0 && x + 1 <= J.mb_w) {
        p = &J.fin[m - iw + 1];
        T_dep = T - 1;
        has = true;
    }
    bool pend = has && (uint32_t)(load_stamp(p) >> 32) != J.serial;
    if (!wait_stamps(pend, p, J.serial, T_dep, j, P)) {
        defer_item(P, j, T);
        return;
    }
    invalidate_l1();
    const uint64_t t_run0 = __builtin_amdgcn_s_memtime();
    if (in_x && in_y && J.deblock) deblock_mb_inline(J, x, y, S.d);
    if (x > 0 && y > 0) pad_mb_inline(J, x - 1, y - 1);
    if (lane == 0) {
        publish_stamp(&J.fin[m], make_stamp(J.serial, seq));
        if (P->debug) {
            const uint64_t t_run1 = __builtin_amdgcn_s_memtime();
            atomicAdd((uint32_t *)&gp(P->status)[2], 1u);
            atomicAdd((uint32_t *)&gp(P->status)[4], (uint32_t)((t_run0 - t_item0) >> 8));
            atomicAdd((uint32_t *)&gp(P->status)[5], (uint32_t)((t_run1 - t_run0) >> 8));
        }
    }
}

// MODE 0: Balanced P and IDR pictures; 1: Balanced P searched from learned proposals, and IDR;
// 2: Fast P and IDR.
// Each workgroup (one wave) runs one ticket of the queue of the XCD it runs on. The grid is dealt
// round-robin over the n_xcc XCDs in workgroup order (from whichever XCD the dispatcher's turn
// starts at), so each group of n_xcc consecutive workgroups puts one on every XCD and workgroup b
// takes ticket b / n_xcc there, in dispatch order: no shared counter (a device-scope atomic on one
// address costs about a microsecond per operation in series, more than an item). The workgroup
// claims its ticket in a word of its own; a ticket claimed twice (a dispatch that is not
// round-robin) flags the launch (status 2) and the host finishes the call with the per-step
// schedule. Every item an item waits for (an earlier ticket of its queue) belongs to a workgroup
// dispatched before it and still resident: the earliest unfinished item never waits, and the
// schedule always advances. (A loop over tickets inside one workgroup would keep the analysis's
// loop-invariant lane constants live across the whole body, several times the registers of a
// single item.) Then the time slot by binary search over the queue's slot table, and the job by
// binary search over the slot's segments.
template <int MODE>
__device__ __forceinline__ void persist(const RghJob *__restrict__ jobs, const RghPersist *__restrict__ P, uint32_t seq) {
    __shared__ StepLds S;
    const int lane = threadIdx.x;
    const int q = P->queue_of_xcc[xcc_id()];
    if (q < 0) return;
    const RghQueue Q = P->q[q];
    const uint32_t t = blockIdx.x / (uint32_t)P->n_xcc;
    if (t >= Q.total) return;
    // Every lane takes part (lane 0 sets the word): no divergent branch around the atomic.
    if (__builtin_amdgcn_readfirstlane(atomicOr(P->claim + Q.claim0 + t, lane == 0 ? 1u : 0u))) {
        if (lane == 0) __hip_atomic_store(&P->status[0], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    int lo = 0, hi = (int)Q.n_slots - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P->slot_tl[Q.slot0 + mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    int e0 = (int)P->slot_seg[Q.slot0 + lo], e1 = (int)P->slot_seg[Q.slot0 + lo + 1] - 1;
    while (e0 < e1) {
        const int mid = (e0 + e1 + 1) >> 1;
        if (P->seg_tl[mid] <= t)
            e0 = mid;
        else
            e1 = mid - 1;
    }
    if (P->jitter) {
        // Test hook: a pseudo-random delay, so that items finish out of ticket order.
        const uint32_t h = (t * 2654435761u) ^ (t >> 7);
        for (uint32_t k = h % P->jitter; k > 0; k--) __builtin_amdgcn_s_sleep(1);
    }
    // The decoded values are uniform; readfirstlane states it (they address the job descriptor
    // through scalar registers).
    const int j = __builtin_amdgcn_readfirstlane((int)P->seg_job[e0]);
    const int r = __builtin_amdgcn_readfirstlane((int)(t - P->seg_tl[e0]));
    persist_item<MODE>(jobs, P, j, __builtin_amdgcn_readfirstlane(Q.t0 + lo), r, seq, S);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_persist(const RghJob *__restrict__ jobs,
                                                                const RghPersist *__restrict__ P, uint32_t seq) {
    persist<0>(jobs, P, seq);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_persist_learned(const RghJob *__restrict__ jobs,
                                                                        const RghPersist *__restrict__ P, uint32_t seq) {
    persist<1>(jobs, P, seq);
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_persist_fast(const RghJob *__restrict__ jobs,
                                                                     const RghPersist *__restrict__ P, uint32_t seq) {
    persist<2>(jobs, P, seq);
}
