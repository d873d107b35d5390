// P-frame macroblock analysis on gfx950: one wave64 per macroblock, macroblocks of one
// anti-diagonal wavefront step (t = x + 2y) of every stream of a batch per launch.
//
// Each function restates a piece of ragdoll's Balanced P path; file:line citations are
// relative to the ragdoll repository ("enc.rs" = crates/ragdoll-codec/src/h264/enc.rs,
// "h264.rs" = crates/ragdoll-codec/src/h264.rs, "simd/" = crates/ragdoll-simd/src/).
//
// Lane layouts: a 16x16 luma block is 64 lanes x 4 packed bytes (lane l holds row l >> 2,
// columns 4 (l & 3) .. +3); 8x8 chroma is one sample per lane (row l >> 3, column l & 7);
// transform and quantization work runs one 4x4 block per lane (16 luma, 8 chroma lanes) on
// blocks staged in LDS, in the reference's own arithmetic.
#include "rgh_device.h"
#include "rgh_tables.h"

namespace {

constexpr int WS_ROWS = 23;   // staged reference rows of a half-pel window
constexpr int WS_STRIDE = 32; // staged reference columns
constexpr int WP_ROWS = 18;   // half-pel plane rows (simd/h264_enc.rs:710 HPEL_WIN_H)
constexpr int WP_COLS = 24;   // half-pel plane columns (HPEL_COLS)

struct MbLds {
    uint8_t src[256], srcu[64], srcv[64];
    uint8_t pred[256], predu[64], predv[64];
    uint8_t tmp[256];
    uint8_t ws[WS_ROWS * WS_STRIDE];
    int16_t hr[WS_ROWS * WP_COLS];
    uint8_t ph[WP_ROWS * WP_COLS], pv[WP_ROWS * WP_COLS], pc[WP_ROWS * WP_COLS];
    int16_t coef[16][16];    // luma levels, decoding order, raster positions
    int16_t q8[4][64];       // 8x8-transform levels, raster positions
    int16_t craster[2][4][16];
    int16_t cdc_raw[2][4];
    int16_t cdc[2][4];
    int16_t cdcr[2][4];
    int16_t dc_raw[16], dcq[16], dcr[16];
    int32_t bnz[16], bbig[16], bscore[16], cscore[8], cmask[8], b8score[4];
    uint8_t etop[16], eleft[16], ctop[2][8], cleft[2][8];
};

__device__ __forceinline__ uint32_t se_bits(int d) {  // enc.rs:546
    uint32_t v = (uint32_t)iabs(d) * 2 - (uint32_t)(d > 0);
    return 2u * (uint32_t)(32 - __clz((int)(v + 1))) - 1u;
}

__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f) {
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}

// ---------------------------------------------------------------------------------------
// 4x4 transforms and quantization, one block per lane.
// ---------------------------------------------------------------------------------------
// simd/fixed.rs:104 forward_dct4x4_scalar.
__device__ void fdct4(const int *d, int16_t *out) {
    int t[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int i = r * 4;
        int s03 = d[i] + d[i + 3], s12 = d[i + 1] + d[i + 2], d03 = d[i] - d[i + 3], d12 = d[i + 1] - d[i + 2];
        t[i] = s03 + s12;
        t[i + 1] = (d03 << 1) + d12;
        t[i + 2] = s03 - s12;
        t[i + 3] = d03 - (d12 << 1);
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
// Residual of the 4x4 block at (x, y) of two 16- or 8-stride blocks, transformed.
__device__ void sub_dct4(const uint8_t *src, const uint8_t *pred, int stride, int x, int y, int16_t *out) {
    int d[16];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) d[r * 4 + c] = (int)src[(y + r) * stride + x + c] - (int)pred[(y + r) * stride + x + c];
    fdct4(d, out);
}
// simd/fixed.rs:185 inverse_idct4x4_scalar, added onto a block with clipping
// (simd/h264_enc.rs:497 add_idct4x4_8x8_scalar for one block).
__device__ void add_idct4(uint8_t *dst, int stride, const int16_t *c) {
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
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int e0 = t[k] + t[k + 8], e1 = t[k] - t[k + 8];
        int e2 = (t[k + 4] >> 1) - t[k + 12], e3 = t[k + 4] + (t[k + 12] >> 1);
        int o[4] = {(int16_t)((e0 + e3 + 32) >> 6), (int16_t)((e1 + e2 + 32) >> 6), (int16_t)((e1 - e2 + 32) >> 6),
                    (int16_t)((e0 - e3 + 32) >> 6)};
#pragma unroll
        for (int r = 0; r < 4; r++) dst[r * stride + k] = (uint8_t)clip255(dst[r * stride + k] + o[r]);
    }
}
__device__ __forceinline__ int coeff_pos_type(int i) {  // h264.rs:1204
    int r = i >> 2, c = i & 3;
    return (!(r & 1) && !(c & 1)) ? 0 : (((r & 1) && (c & 1)) ? 1 : 2);
}
__device__ __forceinline__ int coeff8_pos_type(int i) {  // h264.rs:1343
    int r = (i >> 3) & 3, c = i & 3;
    if (r == 0 && c == 0) return 0;
    if ((r & 1) && (c & 1)) return 1;
    if (r == 2 && c == 2) return 2;
    if ((r == 0 && (c & 1)) || ((r & 1) && c == 0)) return 3;
    if ((r == 0 && c == 2) || (r == 2 && c == 0)) return 4;
    return 5;
}
// quant_4x4_x4_scalar on one block (simd/h264_enc.rs:408; enc.rs:1731 QUANT4): level =
// sign(c) min((|c| mf + bias) >> shift, 32767) with f = 2^shift / 6 (inter) or / 3 (intra).
__device__ void quant4(int16_t *c, int qp, int intra, int *any, int *big) {
    int shift = 15 + qp / 6;
    uint32_t bias = (1u << shift) / (intra ? 3u : 6u);
    int a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int v = c[i];
        uint64_t lv = ((uint64_t)iabs(v) * (uint32_t)c_quant4_mf[qp % 6][coeff_pos_type(i)] + bias) >> shift;
        if (lv > 32767) lv = 32767;
        c[i] = (int16_t)(v < 0 ? -(int)lv : (int)lv);
        a |= lv != 0;
        b |= lv > 1;
    }
    *any = a;
    *big = b;
}
// simd/h264_enc.rs:461 dequant_4x4_x4_scalar on one block.
__device__ void dequant4(int16_t *c, int qp) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int s = c_dequant4_init[qp % 6][coeff_pos_type(i)] << (qp / 6);
        c[i] = (int16_t)iclamp(c[i] * s, -32768, 32767);
    }
}
// enc.rs:583 decimate_masked over n scanned levels with their non-zero mask.
__device__ uint32_t decimate_masked(const int16_t *sc, uint32_t mask) {
    uint32_t score = 0, m = mask;
    while (m) {
        int p = 31 - __clz((int)m);
        if (iabs(sc[p]) > 1) return 9;
        m ^= 1u << p;
        int below = m ? 31 - __clz((int)m) : -1;
        score += c_decimate4[p - below - 1];
    }
    return score;
}

// ---------------------------------------------------------------------------------------
// Half-pel window (simd/h264_enc.rs:704-838): the b, h and j planes over 24 x 18 samples
// around a full-pel position, from a 23 x 32 staged copy of the reference.
// ---------------------------------------------------------------------------------------
__device__ void compute_window(const RghJob &J, int bx, int by, MbLds &S) {
    const int lane = threadIdx.x;
    // ws[k][j] = ref(bx - 3 + j, by - 3 + k)
    const uint8_t *base = J.ref_y + (long)(by - 3) * J.sy + (bx - 3);
    for (int i = lane; i < WS_ROWS * 8; i += 64) {
        int k = i >> 3, j4 = (i & 7) * 4;
        *(uint32_t *)&S.ws[k * WS_STRIDE + j4] = ld4u(base + (long)k * J.sy + j4);
    }
    __syncthreads();
    // Unrounded horizontal taps of rows -2 .. 20 (hr) and the vertical half-pel plane.
    for (int i = lane; i < WS_ROWS * WP_COLS; i += 64) {
        int k = i / WP_COLS, c = i - k * WP_COLS;
        const uint8_t *p = &S.ws[k * WS_STRIDE + c];
        S.hr[i] = (int16_t)tap6(p[0], p[1], p[2], p[3], p[4], p[5]);
    }
    for (int i = lane; i < WP_ROWS * WP_COLS; i += 64) {
        int r = i / WP_COLS, c = i - r * WP_COLS;
        const uint8_t *p = &S.ws[r * WS_STRIDE + c + 2];
        int v = tap6(p[0], p[WS_STRIDE], p[2 * WS_STRIDE], p[3 * WS_STRIDE], p[4 * WS_STRIDE], p[5 * WS_STRIDE]);
        S.pv[i] = (uint8_t)clip255((v + 16) >> 5);
    }
    __syncthreads();
    for (int i = lane; i < WP_ROWS * WP_COLS; i += 64) {
        int r = i / WP_COLS, c = i - r * WP_COLS;
        const int16_t *t = &S.hr[r * WP_COLS + c];
        S.ph[i] = (uint8_t)clip255((t[2 * WP_COLS] + 16) >> 5);
        int j = tap6(t[0], t[WP_COLS], t[2 * WP_COLS], t[3 * WP_COLS], t[4 * WP_COLS], t[5 * WP_COLS]);
        S.pc[i] = (uint8_t)clip255((j + 512) >> 10);
    }
    __syncthreads();
}

// The lane's four prediction bytes at quarter-pel mv from the window anchored at full-pel
// (bx, by): one window/reference plane or the rounded average of two (enc.rs:1415 window_src,
// the x264 get_ref table). mv lies within three quarter samples of the anchor.
__device__ uint32_t window_pred4(int x0, int y0, int bx, int by, int mx, int my, const MbLds &S) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    int ix = x0 + (mx >> 2), iy = y0 + (my >> 2);
    int wx = ix - bx + 1, wy = iy - by + 1;
    // Source descriptors: 0 = integer samples (staged), 1 = b (h plane), 2 = h (v plane), 3 = j.
    int ka, kb = -1, dxa = 0, dya = 0, dxb = 0, dyb = 0;
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
    auto fetch = [&](int k, int dx, int dy, int c) -> uint32_t {
        if (k == 0) return S.ws[(wy + 2 + dy + rr) * WS_STRIDE + wx + 2 + dx + c];
        const uint8_t *pl = k == 1 ? S.ph : (k == 2 ? S.pv : S.pc);
        return pl[(wy + dy + rr) * WP_COLS + wx + dx + c];
    };
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t a = fetch(ka, dxa, dya, c0 + i);
        if (kb >= 0) a = (a + fetch(kb, dxb, dyb, c0 + i) + 1) >> 1;
        w |= a << (8 * i);
    }
    return w;
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

// pixel_satd_16x16 (simd/h264_enc.rs:167) of two 16-stride LDS blocks: lanes 0..15 each sum one
// 4x4 block's unhalved Hadamard magnitudes; the total is halved once.
__device__ uint32_t satd16(const uint8_t *a, const uint8_t *b) {
    const int lane = threadIdx.x;
    uint32_t part = 0;
    if (lane < 16) {
        int bx = (lane & 3) * 4, by = (lane >> 2) * 4, d[4][4];
#pragma unroll
        for (int y = 0; y < 4; y++) {
#pragma unroll
            for (int x = 0; x < 4; x++) d[y][x] = (int)a[(by + y) * 16 + bx + x] - (int)b[(by + y) * 16 + bx + x];
            int s01 = d[y][0] + d[y][1], d01 = d[y][0] - d[y][1], s23 = d[y][2] + d[y][3], d23 = d[y][2] - d[y][3];
            d[y][0] = s01 + s23;
            d[y][1] = s01 - s23;
            d[y][2] = d01 + d23;
            d[y][3] = d01 - d23;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int s01 = d[0][c] + d[1][c], d01 = d[0][c] - d[1][c], s23 = d[2][c] + d[3][c], d23 = d[2][c] - d[3][c];
            part += (uint32_t)(iabs(s01 + s23) + iabs(s01 - s23) + iabs(d01 + d23) + iabs(d01 - d23));
        }
    }
    return wave_sum(part) >> 1;
}

// Chroma bilinear motion compensation of the lane's sample for both planes (enc.rs:1080
// mc_chroma; simd/h264_enc.rs:1279 chroma_mc_8x8_x2_scalar).
__device__ void mc_chroma_lane(const RghJob &J, int mb_x, int mb_y, int mx, int my, uint32_t *pu, uint32_t *pv) {
    const int lane = threadIdx.x, cr = lane >> 3, cc = lane & 7;
    long off = (long)(mb_y * 8 + (my >> 3) + cr) * J.sc + mb_x * 8 + (mx >> 3) + cc;
    int fx = mx & 7, fy = my & 7;
    if ((fx | fy) == 0) {
        *pu = J.ref_u[off];
        *pv = J.ref_v[off];
        return;
    }
    int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
    const uint8_t *u = J.ref_u + off, *v = J.ref_v + off;
    *pu = (uint32_t)((wa * u[0] + wb * u[1] + wc * u[J.sc] + wd * u[J.sc + 1] + 32) >> 6);
    *pv = (uint32_t)((wa * v[0] + wb * v[1] + wc * v[J.sc] + wd * v[J.sc + 1] + 32) >> 6);
}

// Chroma transform + quantization of both planes (enc.rs:1897 quant_chroma), lanes 0..7 one
// block each (plane lane >> 2, block lane & 3, raster 2x2), lanes 0..1 the 2x2 DC of a plane
// (enc.rs:1939 quant_chroma_dc). Leaves raster levels, scanned AC masks and decimation scores
// in LDS. intra selects the rounding of quantize_block4x4.
__device__ void quant_chroma(MbLds &S, int qpc, int intra) {
    const int lane = threadIdx.x;
    if (lane < 8) {
        int p = lane >> 2, b = lane & 3;
        int16_t c[16];
        sub_dct4(p ? S.srcv : S.srcu, p ? S.predv : S.predu, 8, (b & 1) * 4, (b >> 1) * 4, c);
        S.cdc_raw[p][b] = c[0];
        int any, big;
        quant4(c, qpc, intra, &any, &big);
        int16_t zz[16];
        uint32_t m = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            zz[i] = c[c_zigzag4[i]];
            m |= (uint32_t)(zz[i] != 0) << i;
        }
#pragma unroll
        for (int i = 0; i < 16; i++) S.craster[p][b][i] = c[i];
        S.cmask[lane] = (int32_t)(m >> 1);
        S.cscore[lane] = (int32_t)decimate_masked(zz + 1, m >> 1);
    }
    __syncthreads();
    if (lane < 2) {
        int p = lane, d0 = S.cdc_raw[p][0], d1 = S.cdc_raw[p][1], d2 = S.cdc_raw[p][2], d3 = S.cdc_raw[p][3];
        int had[4] = {d0 + d1 + d2 + d3, d0 - d1 + d2 - d3, d0 + d1 - d2 - d3, d0 - d1 - d2 + d3};
        int qbits = 16 + qpc / 6, mf = c_quant4_mf[qpc % 6][0], f = (1 << qbits) / 3;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int l = (iabs(had[i]) * mf + f) >> qbits;
            S.cdc[p][i] = (int16_t)(had[i] < 0 ? -l : l);
        }
    }
    __syncthreads();
}

// h264.rs:1674 dequantize_idct_chroma_dc2x2 then the chroma recon (enc.rs:1990): lanes 0..7
// add their block's inverse transform onto the prediction in LDS.
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
    if (lane < 8) {
        int p = lane >> 2, b = lane & 3;
        int16_t blk[16];
#pragma unroll
        for (int i = 0; i < 16; i++) blk[i] = cbpc >= 2 ? S.craster[p][b][i] : 0;
        if (cbpc >= 2) dequant4(blk, qpc);
        blk[0] = S.cdcr[p][b];
        add_idct4((p ? S.predv : S.predu) + (b >> 1) * 32 + (b & 1) * 4, 8, blk);
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

// Intra 16x16 / chroma prediction sample (h264.rs:1715 predict_intra_16x16_scalar, enc.rs:1559
// pred16; h264.rs:1967 predict_intra_chroma_8x8_scalar, enc.rs:1588 pred_chroma).
__device__ int pred16_sample(int mode, const uint8_t *t, const uint8_t *l, int has_t, int has_l, int tl, int dc,
                             int pb, int pc_, int pa, int x, int y) {
    if (mode == 0) return has_t ? t[x] : 128;
    if (mode == 1) return has_l ? l[y] : 128;
    if (mode == 2) return dc;
    return clip255((pa + pb * (x - 7) + pc_ * (y - 7) + 16) >> 5);
}

struct Edges16 {
    int has_t, has_l, tl;
};

}  // namespace

// ---------------------------------------------------------------------------------------
// One macroblock (enc.rs:3011 code_p_mb and everything it calls).
// ---------------------------------------------------------------------------------------
__device__ void analyse_mb(const RghJob &J, int mb_x, int mb_y, MbLds &S) {
    const int lane = threadIdx.x;
    const int rr = lane >> 2, c0 = (lane & 3) * 4;
    const int mb_w = J.mb_w, idx = mb_y * mb_w + mb_x;
    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8;
    const int qp = J.qp, qpc = c_chroma_qp[qp];
    const uint32_t lambda = c_lambda[qp];
    const int W = mb_w * 16, H = J.mb_h * 16;

    // Source macroblock, edge-replicated past the display size (enc.rs:389 MbSrc::load,
    // h264.rs:3660 extract_mb_planes).
    uint32_t srcw = 0;
    {
        int sy = min(y0 + rr, J.height - 1);
#pragma unroll
        for (int i = 0; i < 4; i++) srcw |= (uint32_t)J.src_y[(long)sy * J.src_sy + min(x0 + c0 + i, J.width - 1)] << (8 * i);
        *(uint32_t *)&S.src[lane * 4] = srcw;
        int cw = (J.width + 1) / 2, ch = (J.height + 1) / 2;
        int csy = min(cy0 + (lane >> 3), ch - 1), csx = min(cx0 + (lane & 7), cw - 1);
        S.srcu[lane] = J.src_u[(long)csy * J.src_su + csx];
        S.srcv[lane] = J.src_v[(long)csy * J.src_sv + csx];
    }
    const uint32_t src_u = S.srcu[lane], src_v = S.srcv[lane];  // (read after the stores below)

    // Neighbour motion (h264.rs:3577 pred_motion_partition for one 16x16 partition, ref 0).
    int lref = -2, tref = -2, dref = -2, ax = 0, ay = 0, bx_ = 0, by_ = 0, cxv = 0, cyv = 0;
    if (mb_x > 0) {
        const rgh_mb_decision &d = J.dec[idx - 1];
        lref = d.mb_type >= 2 ? -1 : 0;
        ax = d.mv[0];
        ay = d.mv[1];
    }
    if (mb_y > 0) {
        const rgh_mb_decision &d = J.dec[idx - mb_w];
        tref = d.mb_type >= 2 ? -1 : 0;
        bx_ = d.mv[0];
        by_ = d.mv[1];
        const rgh_mb_decision *c = nullptr;
        if (mb_x + 1 < mb_w)
            c = &J.dec[idx - mb_w + 1];
        else if (mb_x > 0)
            c = &J.dec[idx - mb_w - 1];
        if (c) {
            dref = c->mb_type >= 2 ? -1 : 0;
            cxv = c->mv[0];
            cyv = c->mv[1];
        }
    }
    auto med3 = [](int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); };
    int mvpx, mvpy;
    {
        int match = (lref == 0) + (tref == 0) + (dref == 0);
        if (match > 1) {
            mvpx = med3(ax, bx_, cxv);
            mvpy = med3(ay, by_, cyv);
        } else if (match == 1) {
            mvpx = lref == 0 ? ax : (tref == 0 ? bx_ : cxv);
            mvpy = lref == 0 ? ay : (tref == 0 ? by_ : cyv);
        } else if (tref == -2 && dref == -2 && lref != -2) {
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
    auto in_bounds = [&](int mx, int my) { return mx >= bmin0 && mx <= bmax0 && my >= bmin1 && my <= bmax1; };
    auto mv_cost = [&](int mx, int my) { return lambda * se_bits(mx - mvpx) + lambda * se_bits(my - mvpy); };

    rgh_mb_decision D;
    memset(&D, 0, sizeof(D));
    D.mvp[0] = (int16_t)mvpx;
    D.mvp[1] = (int16_t)mvpy;
    D.qp = (uint8_t)qp;
    int mb_type = -1;          // decided type
    int mvx = 0, mvy = 0;      // final vector
    uint32_t predw = 0;        // the lane's luma prediction bytes
    uint32_t pu = 0, pv = 0;   // the lane's chroma prediction samples
    int cbpl = 0, cbpc = 0, use8 = 0, i16_mode = 0, cmode = 0;

    // 1. Skip probe (enc.rs:2906 try_p_skip; Balanced probes every macroblock).
    if (in_bounds(pskx, psky)) {
        int int_y = ((pskx | psky) & 3) == 0, int_c = ((pskx | psky) & 7) == 0;
        if (int_y) {
            predw = ld4u(J.ref_y + (long)(y0 + (psky >> 2) + rr) * J.sy + x0 + (pskx >> 2) + c0);
        } else {
            // Spec luma MC at the skip vector, through the window anchored at its full-pel part
            // (the window table equals h264_luma_mc, enc.rs:4030 test).
            int fbx = x0 + (pskx >> 2), fby = y0 + (psky >> 2);
            compute_window(J, fbx, fby, S);
            predw = window_pred4(x0, y0, fbx, fby, pskx, psky, S);
        }
        uint32_t sad = wave_sum(__builtin_amdgcn_sad_u8(srcw, predw, 0));
        if (sad < J.sb_probe) {
            int luma_zero = sad <= J.sb_luma || wave_sum(ssd4(srcw, predw)) <= J.sb_luma_ssd;
            if (!luma_zero) {
                // enc.rs:2163 probe_skip_luma: no level above 1 and a total score below 6.
                *(uint32_t *)&S.pred[lane * 4] = predw;
                __syncthreads();
                int big = 0;
                uint32_t score = 0;
                if (lane < 16) {
                    int16_t c[16];
                    sub_dct4(S.src, S.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, c);
                    int any;
                    quant4(c, qp, 0, &any, &big);
                    int16_t zz[16];
                    uint32_t m = 0;
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        zz[i] = c[c_zigzag4[i]];
                        m |= (uint32_t)(zz[i] != 0) << i;
                    }
                    score = decimate_masked(zz, m);
                }
                uint64_t bigs = __ballot(big);
                uint32_t total = wave_sum(score);
                luma_zero = bigs == 0 && total < 6;
            }
            if (luma_zero) {
                if (int_c) {
                    long off = (long)(cy0 + (psky >> 3) + (lane >> 3)) * J.sc + cx0 + (pskx >> 3) + (lane & 7);
                    pu = J.ref_u[off];
                    pv = J.ref_v[off];
                } else {
                    mc_chroma_lane(J, mb_x, mb_y, pskx, psky, &pu, &pv);
                }
                uint32_t du = (uint32_t)iabs((int)src_u - (int)pu), dv = (uint32_t)iabs((int)src_v - (int)pv);
                int chroma_zero = wave_sum(du) <= J.sb_chroma || wave_sum(du * du) <= J.sb_chroma_ssd;
                if (chroma_zero) chroma_zero = wave_sum(dv) <= J.sb_chroma || wave_sum(dv * dv) <= J.sb_chroma_ssd;
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
                    int su = S.cscore[0] + S.cscore[1] + S.cscore[2] + S.cscore[3];
                    int sv = S.cscore[4] + S.cscore[5] + S.cscore[6] + S.cscore[7];
                    chroma_zero = !dcnz && su < 7 && sv < 7;
                }
                if (chroma_zero) {
                    mb_type = 0;
                    D.flags = 1;
                    mvx = pskx;
                    mvy = psky;
                }
            }
        }
    }

    if (mb_type < 0) {
        // 2. Integer search (enc.rs:1185 full_pel) from [mvp, 0, left, top, co-located].
        int cand[5][2], nc = 0;
        cand[nc][0] = mvpx; cand[nc++][1] = mvpy;
        cand[nc][0] = 0; cand[nc++][1] = 0;
        if (mb_x > 0) { cand[nc][0] = ax; cand[nc++][1] = ay; }
        if (mb_y > 0) { cand[nc][0] = bx_; cand[nc++][1] = by_; }
        cand[nc][0] = J.prev[idx].mv[0];
        cand[nc++][1] = J.prev[idx].mv[1];
        const int imn0 = (bmin0 + 3) >> 2, imn1 = (bmin1 + 3) >> 2, imx0 = bmax0 >> 2, imx1 = bmax1 >> 2;
        const uint8_t *rb = J.ref_y + (long)(y0 + rr) * J.sy + x0 + c0;
        auto sad_at = [&](int fx, int fy) {
            return wave_sum(__builtin_amdgcn_sad_u8(srcw, ld4u(rb + (long)fy * J.sy + fx), 0));
        };
        int bestx = 0, besty = 0;
        uint32_t best_cost = 0xFFFFFFFFu;
        int seen[5][2];
        for (int k = 0; k < nc; k++) {
            int fx = iclamp((cand[k][0] + 2) >> 2, imn0, imx0), fy = iclamp((cand[k][1] + 2) >> 2, imn1, imx1);
            int dup = 0;
            for (int j = 0; j < k; j++) dup |= seen[j][0] == fx && seen[j][1] == fy;
            seen[k][0] = dup ? 0x7FFFFFFF : fx;
            seen[k][1] = dup ? 0x7FFFFFFF : fy;
            if (dup) continue;
            uint32_t cst = sad_at(fx, fy) + mv_cost(fx * 4, fy * 4);
            D.n_sad++;
            if (cst < best_cost) {
                best_cost = cst;
                bestx = fx;
                besty = fy;
            }
        }
        // Small diamond: the first of the cheapest in-bounds points below the best moves.
        for (int it = 0; it < J.me_iters; it++) {
            const int px[4] = {bestx, bestx - 1, bestx + 1, bestx};
            const int py[4] = {besty - 1, besty, besty, besty + 1};
            int moved = -1;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (px[k] < imn0 || px[k] > imx0 || py[k] < imn1 || py[k] > imx1) continue;
                uint32_t cst = sad_at(px[k], py[k]) + mv_cost(px[k] * 4, py[k] * 4);
                D.n_sad++;
                if (cst < best_cost) {
                    best_cost = cst;
                    moved = k;
                }
            }
            if (moved < 0) break;
            bestx = px[moved];
            besty = py[moved];
        }
        // 3. Sub-pel refinement unless the full-pel residual already quantizes to nothing
        // (enc.rs:3097-3117; enc.rs:1283 sub_pel).
        uint32_t me_cost = best_cost;
        if (best_cost - mv_cost(bestx * 4, besty * 4) <= J.sb_luma) {
            predw = ld4u(rb + (long)besty * J.sy + bestx);
            mvx = bestx * 4;
            mvy = besty * 4;
        } else {
            int cx = bestx * 4, cy = besty * 4, fbx = x0 + bestx, fby = y0 + besty;
            compute_window(J, fbx, fby, S);
            D.n_hpel++;
            int bmx = cx, bmy = cy;
            for (int pass = 0; pass < 2; pass++) {
                int step = pass == 0 ? 2 : 1, iters = pass == 0 ? 1 : J.qpel_iters;
                for (int it = 0; it < iters; it++) {
                    int fx = bmx, fy = bmy;
                    const int dx[4] = {0, -step, step, 0}, dy[4] = {-step, 0, 0, step};
                    for (int k = 0; k < 4; k++) {
                        int mx = fx + dx[k], my = fy + dy[k];
                        if (iabs(mx - cx) > 3 || iabs(my - cy) > 3 || !in_bounds(mx, my)) continue;
                        uint32_t w = window_pred4(x0, y0, fbx, fby, mx, my, S);
                        uint32_t cst = wave_sum(__builtin_amdgcn_sad_u8(srcw, w, 0)) + mv_cost(mx, my);
                        D.n_subpel++;
                        if (cst < me_cost) {
                            me_cost = cst;
                            bmx = mx;
                            bmy = my;
                        }
                    }
                    if (bmx == fx && bmy == fy) break;
                }
            }
            predw = window_pred4(x0, y0, fbx, fby, bmx, bmy, S);
            mvx = bmx;
            mvy = bmy;
        }
        D.cost = me_cost;
        *(uint32_t *)&S.pred[lane * 4] = predw;
        __syncthreads();

        // 4. Intra 16x16 when the inter prediction is poor (enc.rs:3119-3152).
        uint32_t inter_satd = satd16(S.src, S.pred);
        D.n_satd++;
        uint32_t inter_cost = inter_satd + lambda * (se_bits(mvx - mvpx) + se_bits(mvy - mvpy) + 1);
        if (inter_satd > 24 * lambda) {
            // Edges from the undeblocked current picture (enc.rs:1457 i16_edges).
            Edges16 E;
            E.has_t = mb_y > 0;
            E.has_l = mb_x > 0;
            E.tl = (mb_x > 0 && mb_y > 0) ? J.cur_y[(long)(y0 - 1) * J.sy + x0 - 1] : 128;
            if (lane < 16) {
                S.etop[lane] = E.has_t ? J.cur_y[(long)(y0 - 1) * J.sy + x0 + lane] : 0;
                S.eleft[lane] = E.has_l ? J.cur_y[(long)(y0 + lane) * J.sy + x0 - 1] : 0;
            }
            __syncthreads();
            uint32_t st = 0, sl = 0;
            for (int i = 0; i < 16; i++) {
                st += S.etop[i];
                sl += S.eleft[i];
            }
            int dc = (E.has_t && E.has_l) ? (int)((st + sl + 16) >> 5)
                                          : (E.has_t ? (int)((st + 8) >> 4) : (E.has_l ? (int)((sl + 8) >> 4) : 128));
            // simd/h264_enc.rs:628 intra_satd_x3_16x16: V, H, DC SATDs from one source Hadamard.
            uint32_t sv = 0, sh = 0, sd = 0;
            if (lane < 16) {
                int bx = lane & 3, by = lane >> 2, rows[4][4];
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    const uint8_t *x = &S.src[(by * 4 + y) * 16 + bx * 4];
                    int s01 = x[0] + x[1], d01 = x[0] - x[1], s23 = x[2] + x[3], d23 = x[2] - x[3];
                    rows[y][0] = s01 + s23;
                    rows[y][1] = s01 - s23;
                    rows[y][2] = d01 + d23;
                    rows[y][3] = d01 - d23;
                }
                int tv[4], th[4];
                {
                    const uint8_t *x = &S.etop[bx * 4];
                    int s01 = x[0] + x[1], d01 = x[0] - x[1], s23 = x[2] + x[3], d23 = x[2] - x[3];
                    tv[0] = 4 * (s01 + s23);
                    tv[1] = 4 * (s01 - s23);
                    tv[2] = 4 * (d01 + d23);
                    tv[3] = 4 * (d01 - d23);
                    x = &S.eleft[by * 4];
                    s01 = x[0] + x[1];
                    d01 = x[0] - x[1];
                    s23 = x[2] + x[3];
                    d23 = x[2] - x[3];
                    th[0] = 4 * (s01 + s23);
                    th[1] = 4 * (s01 - s23);
                    th[2] = 4 * (d01 + d23);
                    th[3] = 4 * (d01 - d23);
                }
#pragma unroll
                for (int h = 0; h < 4; h++) {
                    int s01 = rows[0][h] + rows[1][h], d01 = rows[0][h] - rows[1][h];
                    int s23 = rows[2][h] + rows[3][h], d23 = rows[2][h] - rows[3][h];
                    int o[4] = {(int16_t)(s01 + s23), (int16_t)(s01 - s23), (int16_t)(d01 + d23), (int16_t)(d01 - d23)};
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        int c = o[v], k = v * 4 + h;
                        int pvv = v == 0 ? tv[h] : 0, phh = h == 0 ? th[v] : 0, pdd = k == 0 ? 16 * dc : 0;
                        sv += (uint32_t)iabs(c - pvv);
                        sh += (uint32_t)iabs(c - phh);
                        sd += (uint32_t)iabs(c - pdd);
                    }
                }
            }
            sv = wave_sum(sv) >> 1;
            sh = wave_sum(sh) >> 1;
            sd = wave_sum(sd) >> 1;
            // enc.rs:1637 analyse_i16_edges: DC, then V, then H, then plane, strict improvements.
            uint32_t ib = sd;
            int im = 2;
            if (E.has_t && sv < ib) { ib = sv; im = 0; }
            if (E.has_l && sh < ib) { ib = sh; im = 1; }
            int pa = 0, pb = 0, pcc = 0;
            if (E.has_t && E.has_l) {
                int Hh = 0, Vv = 0;
                for (int i = 0; i < 8; i++) {
                    int ptop = i == 7 ? E.tl : S.etop[6 - i], pl = i == 7 ? E.tl : S.eleft[6 - i];
                    Hh += (i + 1) * (S.etop[8 + i] - ptop);
                    Vv += (i + 1) * (S.eleft[8 + i] - pl);
                }
                pa = 16 * (S.etop[15] + S.eleft[15]);
                pb = (5 * Hh + 32) >> 6;
                pcc = (5 * Vv + 32) >> 6;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    S.tmp[lane * 4 + i] = (uint8_t)pred16_sample(3, S.etop, S.eleft, 1, 1, E.tl, dc, pb, pcc, pa, c0 + i, rr);
                __syncthreads();
                uint32_t ps = satd16(S.src, S.tmp);
                D.n_satd++;
                if (ps < ib) { ib = ps; im = 3; }
            }
            if (ib + lambda * 8 < inter_cost) {
                // enc.rs:3587 code_i16 in a P slice.
                mb_type = 2;
                i16_mode = im;
                mvx = mvy = 0;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    S.pred[lane * 4 + i] = (uint8_t)pred16_sample(im, S.etop, S.eleft, E.has_t, E.has_l, E.tl, dc, pb, pcc, pa, c0 + i, rr);
                // Chroma mode by SAD + 8 for non-DC modes (enc.rs:1523 analyse_chroma_intra).
                if (lane < 8) {
                    for (int p = 0; p < 2; p++) {
                        const uint8_t *cp = p ? J.cur_v : J.cur_u;
                        S.ctop[p][lane] = mb_y > 0 ? cp[(long)(cy0 - 1) * J.sc + cx0 + lane] : 0;
                        S.cleft[p][lane] = mb_x > 0 ? cp[(long)(cy0 + lane) * J.sc + cx0 - 1] : 0;
                    }
                }
                int ctl[2];
                for (int p = 0; p < 2; p++) {
                    const uint8_t *cp = p ? J.cur_v : J.cur_u;
                    ctl[p] = (mb_x > 0 && mb_y > 0) ? cp[(long)(cy0 - 1) * J.sc + cx0 - 1] : 128;
                }
                __syncthreads();
                const int cr = lane >> 3, cc = lane & 7;
                uint32_t cbest = 0xFFFFFFFFu;
                uint32_t bpu = 0, bpv = 0;
                for (int mode = 0; mode < 4; mode++) {
                    int ok = mode == 1 ? mb_x > 0 : (mode == 2 ? mb_y > 0 : (mode == 3 ? (mb_x > 0 && mb_y > 0) : 1));
                    if (!ok) continue;
                    uint32_t smp[2];
                    for (int p = 0; p < 2; p++) {
                        const uint8_t *t = S.ctop[p], *l = S.cleft[p];
                        int ht = mb_y > 0, hl = mb_x > 0, v;
                        if (mode == 0) {
                            uint32_t t0 = 0, t1 = 0, l0 = 0, l1 = 0;
                            for (int i = 0; i < 4; i++) {
                                t0 += ht ? t[i] : 128;
                                t1 += ht ? t[4 + i] : 128;
                                l0 += hl ? l[i] : 128;
                                l1 += hl ? l[4 + i] : 128;
                            }
                            int d00, d01, d10, d11;
                            if (ht && hl) {
                                d00 = (int)((t0 + l0 + 4) >> 3);
                                d01 = (int)((t1 + 2) >> 2);
                                d10 = (int)((l1 + 2) >> 2);
                                d11 = (int)((t1 + l1 + 4) >> 3);
                            } else if (ht) {
                                d00 = d10 = (int)((t0 + 2) >> 2);
                                d01 = d11 = (int)((t1 + 2) >> 2);
                            } else if (hl) {
                                d00 = d01 = (int)((l0 + 2) >> 2);
                                d10 = d11 = (int)((l1 + 2) >> 2);
                            } else {
                                d00 = d01 = d10 = d11 = 128;
                            }
                            v = cr < 4 ? (cc < 4 ? d00 : d01) : (cc < 4 ? d10 : d11);
                        } else if (mode == 1) {
                            v = l[cr];
                        } else if (mode == 2) {
                            v = t[cc];
                        } else {
                            int Hh = 0, Vv = 0;
                            for (int i = 0; i < 4; i++) {
                                int ptop = i == 3 ? ctl[p] : t[2 - i], pl = i == 3 ? ctl[p] : l[2 - i];
                                Hh += (i + 1) * (t[4 + i] - ptop);
                                Vv += (i + 1) * (l[4 + i] - pl);
                            }
                            int a = 16 * (t[7] + l[7]), b = (17 * Hh + 16) >> 5, c = (17 * Vv + 16) >> 5;
                            v = clip255((a + b * (cc - 3) + c * (cr - 3) + 16) >> 5);
                        }
                        smp[p] = (uint32_t)v;
                    }
                    uint32_t cost = wave_sum((uint32_t)iabs((int)src_u - (int)smp[0]) + (uint32_t)iabs((int)src_v - (int)smp[1])) +
                                    (mode ? 8u : 0u);
                    if (cost < cbest) {
                        cbest = cost;
                        cmode = mode;
                        bpu = smp[0];
                        bpv = smp[1];
                    }
                }
                pu = bpu;
                pv = bpv;
                S.predu[lane] = (uint8_t)pu;
                S.predv[lane] = (uint8_t)pv;
                __syncthreads();
                // Luma: 4x4 transforms, AC quantized with intra rounding, DC through the 4x4
                // Hadamard (enc.rs:3630-3660; quant_luma_dc :1959).
                int acnz = 0;
                if (lane < 16) {
                    int16_t c[16];
                    sub_dct4(S.src, S.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, c);
                    const int z2r[16] = {0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15};
                    S.dc_raw[z2r[lane]] = c[0];
                    int any, big;
                    quant4(c, qp, 1, &any, &big);
                    c[0] = 0;
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        S.coef[lane][i] = c[i];
                        acnz |= c[i] != 0;
                    }
                }
                uint64_t acm = __ballot(acnz);
                cbpl = acm ? 15 : 0;
                __syncthreads();
                if (lane == 0) {
                    // simd/fixed.rs:318 hadamard4x4_scalar (with >> 1), then the DC quantizer.
                    int t[16], had[16];
                    for (int r = 0; r < 4; r++) {
                        int i = r * 4;
                        int s03 = S.dc_raw[i] + S.dc_raw[i + 3], s12 = S.dc_raw[i + 1] + S.dc_raw[i + 2];
                        int d03 = S.dc_raw[i] - S.dc_raw[i + 3], d12 = S.dc_raw[i + 1] - S.dc_raw[i + 2];
                        t[i] = s03 + s12;
                        t[i + 1] = d03 + d12;
                        t[i + 2] = s03 - s12;
                        t[i + 3] = d03 - d12;
                    }
                    for (int c = 0; c < 4; c++) {
                        int s03 = t[c] + t[c + 12], s12 = t[c + 4] + t[c + 8], d03 = t[c] - t[c + 12], d12 = t[c + 4] - t[c + 8];
                        had[c] = (int16_t)((s03 + s12) >> 1);
                        had[c + 4] = (int16_t)((d03 + d12) >> 1);
                        had[c + 8] = (int16_t)((s03 - s12) >> 1);
                        had[c + 12] = (int16_t)((d03 - d12) >> 1);
                    }
                    int qbits = 16 + qp / 6, mf = c_quant4_mf[qp % 6][0], f = (1 << qbits) / 3;
                    for (int i = 0; i < 16; i++) {
                        int l = (iabs(had[i]) * mf + f) >> qbits;
                        S.dcq[i] = (int16_t)(had[i] < 0 ? -l : l);
                    }
                    // h264.rs:1645 dequantize_idct_luma_dc4x4
                    int qmul = (c_dequant4_init[qp % 6][0] * 16) << (qp / 6 + 2);
                    for (int i = 0; i < 4; i++) {
                        int z0 = S.dcq[4 * i] + S.dcq[4 * i + 1], z1 = S.dcq[4 * i] - S.dcq[4 * i + 1];
                        int z2 = S.dcq[4 * i + 2] - S.dcq[4 * i + 3], z3 = S.dcq[4 * i + 2] + S.dcq[4 * i + 3];
                        t[4 * i] = z0 + z3;
                        t[4 * i + 1] = z0 - z3;
                        t[4 * i + 2] = z1 - z2;
                        t[4 * i + 3] = z1 + z2;
                    }
                    for (int i = 0; i < 4; i++) {
                        int z0 = t[i] + t[8 + i], z1 = t[i] - t[8 + i], z2 = t[4 + i] - t[12 + i], z3 = t[4 + i] + t[12 + i];
                        S.dcr[i] = (int16_t)(((z0 + z3) * qmul + 128) >> 8);
                        S.dcr[4 + i] = (int16_t)(((z1 + z2) * qmul + 128) >> 8);
                        S.dcr[8 + i] = (int16_t)(((z1 - z2) * qmul + 128) >> 8);
                        S.dcr[12 + i] = (int16_t)(((z0 - z3) * qmul + 128) >> 8);
                    }
                }
                quant_chroma(S, qpc, 1);
                cbpc = chroma_cbp(S);
                if (lane < 16) {
                    const int z2r[16] = {0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15};
                    int16_t blk[16];
                    int any = 0;
#pragma unroll
                    for (int i = 0; i < 16; i++) blk[i] = S.coef[lane][i];
                    if (cbpl) dequant4(blk, qp);
                    blk[0] = S.dcr[z2r[lane]];
#pragma unroll
                    for (int i = 0; i < 16; i++) any |= blk[i];
                    if (any) add_idct4(&S.pred[blk_y(lane) * 64 + blk_x(lane) * 4], 16, blk);
                }
                __syncthreads();
                recon_chroma(S, cbpc, qpc);
            }
        }

        if (mb_type < 0) {
            // 5. P_L0_16x16 residual (enc.rs:3154-3191), or P_Skip when it vanishes at the skip vector.
            mc_chroma_lane(J, mb_x, mb_y, mvx, mvy, &pu, &pv);
            S.predu[lane] = (uint8_t)pu;
            S.predv[lane] = (uint8_t)pv;
            // enc.rs:1810 quant_luma_4x4_inter
            if (lane < 16) {
                int16_t c[16];
                sub_dct4(S.src, S.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, c);
                int any, big;
                quant4(c, qp, 0, &any, &big);
                int16_t zz[16];
                uint32_t m = 0;
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    S.coef[lane][i] = c[i];
                    zz[i] = c[c_zigzag4[i]];
                    m |= (uint32_t)(zz[i] != 0) << i;
                }
                S.bnz[lane] = any;
                S.bbig[lane] = big;
                S.bscore[lane] = (int32_t)decimate_masked(zz, m);
            }
            __syncthreads();
            int keep = 0;
            {
                uint32_t total = 0;
                for (int g = 0; g < 4; g++) {
                    int nz = S.bnz[4 * g] | S.bnz[4 * g + 1] | S.bnz[4 * g + 2] | S.bnz[4 * g + 3];
                    int big = S.bbig[4 * g] | S.bbig[4 * g + 1] | S.bbig[4 * g + 2] | S.bbig[4 * g + 3];
                    if (!nz) continue;
                    if (big) {
                        keep |= 1 << g;
                        total += 9;
                        continue;
                    }
                    uint32_t sc = (uint32_t)(S.bscore[4 * g] + S.bscore[4 * g + 1] + S.bscore[4 * g + 2] + S.bscore[4 * g + 3]);
                    if (sc >= 4) keep |= 1 << g;
                    total += sc;
                }
                if (total < 6) keep = 0;
            }
            cbpl = keep;
            __syncthreads();
            if (lane < 16 && !((keep >> (lane >> 2)) & 1))
#pragma unroll
                for (int i = 0; i < 16; i++) S.coef[lane][i] = 0;
            __syncthreads();
            if (J.t8x8 && cbpl) {
                // enc.rs:2041 quant_luma_8x8_inter; the 8x8 transform wins when it leaves no
                // more non-zero levels (enc.rs:3170-3179).
                if (lane < 4) {
                    int ox = (lane & 1) * 8, oy = (lane >> 1) * 8;
                    int t[64], o[64];
                    for (int y = 0; y < 8; y++)
                        for (int x = 0; x < 8; x++)
                            t[y * 8 + x] = (int)S.src[(oy + y) * 16 + ox + x] - (int)S.pred[(oy + y) * 16 + ox + x];
                    // simd/fixed.rs:392 forward_dct8x8
                    for (int pass = 0; pass < 2; pass++) {
                        for (int k = 0; k < 8; k++) {
                            int v[8];
                            for (int j = 0; j < 8; j++) v[j] = pass == 0 ? t[k * 8 + j] : o[j * 8 + k];
                            int a0 = v[0] + v[7], a1 = v[1] + v[6], a2 = v[2] + v[5], a3 = v[3] + v[4];
                            int a4 = v[0] - v[7], a5 = v[1] - v[6], a6 = v[2] - v[5], a7 = v[3] - v[4];
                            int b0 = a0 + a3, b1 = a1 + a2, b2 = a0 - a3, b3 = a1 - a2;
                            int b4 = a5 + a6 + ((a4 >> 1) + a4), b5 = a4 - a7 - ((a6 >> 1) + a6);
                            int b6 = a4 + a7 - ((a5 >> 1) + a5), b7 = a5 - a6 + ((a7 >> 1) + a7);
                            int r8[8] = {b0 + b1, b4 + (b7 >> 2), b2 + (b3 >> 1), b5 + (b6 >> 2),
                                         b0 - b1, b6 - (b5 >> 2), (b2 >> 1) - b3, (b4 >> 2) - b7};
                            for (int j = 0; j < 8; j++) {
                                if (pass == 0)
                                    o[k * 8 + j] = r8[j];
                                else
                                    t[j * 8 + k] = (int16_t)r8[j];
                            }
                        }
                    }
                    int shift = 16 + qp / 6;
                    uint32_t bias = (1u << shift) / 6u;
                    int any = 0, big = 0;
                    for (int i = 0; i < 64; i++) {
                        int v = t[i];
                        uint64_t lv = ((uint64_t)iabs(v) * (uint32_t)c_quant8_mf[qp % 6][coeff8_pos_type(i)] + bias) >> shift;
                        if (lv > 32767) lv = 32767;
                        any |= lv != 0;
                        big |= lv > 1;
                        S.q8[lane][i] = (int16_t)(v < 0 ? -(int)lv : (int)lv);
                    }
                    uint32_t score = 0;
                    if (big)
                        score = 9;
                    else if (any) {
                        // enc.rs:557 decimate_score with the 8x8 table
                        int n = 64;
                        while (n > 0 && S.q8[lane][c_zigzag8[n - 1]] == 0) n--;
                        while (n > 0) {
                            n--;
                            int run = 0;
                            while (n > 0 && S.q8[lane][c_zigzag8[n - 1]] == 0) {
                                n--;
                                run++;
                            }
                            score += c_decimate8[run];
                        }
                    }
                    S.b8score[lane] = (int32_t)score;
                }
                __syncthreads();
                uint32_t total = 0;
                int cbp8 = 0;
                for (int b = 0; b < 4; b++) {
                    total += (uint32_t)S.b8score[b];
                    if (S.b8score[b] >= 4) cbp8 |= 1 << b;
                }
                if (total < 6) cbp8 = 0;
                uint32_t n4 = 0, n8 = 0;
                if (lane < 16)
                    for (int i = 0; i < 16; i++) n4 += S.coef[lane][i] != 0;
                if (lane < 4 && ((cbp8 >> lane) & 1))
                    for (int i = 0; i < 64; i++) n8 += S.q8[lane][i] != 0;
                n4 = wave_sum(n4);
                n8 = wave_sum(n8);
                __syncthreads();
                if (n8 <= n4) {
                    use8 = 1;
                    cbpl = cbp8;
                    if (lane < 4 && !((cbp8 >> lane) & 1))
                        for (int i = 0; i < 64; i++) S.q8[lane][i] = 0;
                }
                __syncthreads();
            }
            // enc.rs:2084 quant_chroma_inter: a plane's AC goes when it scores below 7.
            quant_chroma(S, qpc, 0);
            for (int p = 0; p < 2; p++) {
                int any = S.cmask[4 * p] | S.cmask[4 * p + 1] | S.cmask[4 * p + 2] | S.cmask[4 * p + 3];
                int sc = S.cscore[4 * p] + S.cscore[4 * p + 1] + S.cscore[4 * p + 2] + S.cscore[4 * p + 3];
                if (any && sc < 7) {
                    __syncthreads();
                    if (lane < 4) {
                        S.cmask[4 * p + lane] = 0;
                        for (int i = 0; i < 16; i++) S.craster[p][lane][i] = 0;
                    }
                    __syncthreads();
                }
            }
            cbpc = chroma_cbp(S);
            if (cbpl == 0 && cbpc == 0 && mvx == pskx && mvy == psky) {
                mb_type = 0;
            } else {
                mb_type = 1;
                // Reconstruction (enc.rs:3282-3306).
                if (use8) {
                    if (lane < 4 && ((cbpl >> lane) & 1)) {
                        // enc.rs:2017 recon_8x8: h264.rs:1403 dequantize_block8x8, then
                        // simd/h264_idct.rs:268 h264_idct8_add_scalar.
                        int qd = qp / 6, d[64];
                        for (int i = 0; i < 64; i++) {
                            int s = c_dequant8_init[qp % 6][coeff8_pos_type(i)], v;
                            if (qd >= 2)
                                v = S.q8[lane][i] * (s << (qd - 2));
                            else {
                                int rs = 2 - qd;
                                v = (S.q8[lane][i] * s + (1 << (rs - 1))) >> rs;
                            }
                            d[i] = iclamp(v, -32768, 32767);
                        }
                        int g[64];
                        for (int pass = 0; pass < 2; pass++) {
                            for (int k = 0; k < 8; k++) {
                                uint32_t v[8];
                                for (int j = 0; j < 8; j++) v[j] = (uint32_t)(pass == 0 ? d[k * 8 + j] : g[j * 8 + k]);
                                uint32_t a0 = v[0] + v[4];
                                uint32_t a1 = v[5] - v[3] - v[7] - (uint32_t)((int)v[7] >> 1);
                                uint32_t a2 = v[0] - v[4];
                                uint32_t a3 = v[1] + v[7] - v[3] - (uint32_t)((int)v[3] >> 1);
                                uint32_t a4 = (uint32_t)((int)v[2] >> 1) - v[6];
                                uint32_t a5 = v[7] - v[1] + v[5] + (uint32_t)((int)v[5] >> 1);
                                uint32_t a6 = v[2] + (uint32_t)((int)v[6] >> 1);
                                uint32_t a7 = v[3] + v[5] + v[1] + (uint32_t)((int)v[1] >> 1);
                                uint32_t b0 = a0 + a6, b1 = a1 + (uint32_t)((int)a7 >> 2), b2 = a2 + a4;
                                uint32_t b3 = a3 + (uint32_t)((int)a5 >> 2), b4 = a2 - a4;
                                uint32_t b5 = (uint32_t)((int)a3 >> 2) - a5, b6 = a0 - a6, b7 = a7 - (uint32_t)((int)a1 >> 2);
                                uint32_t o8[8] = {b0 + b7, b2 + b5, b4 + b3, b6 + b1, b6 - b1, b4 - b3, b2 - b5, b0 - b7};
                                for (int j = 0; j < 8; j++) {
                                    if (pass == 0)
                                        g[k * 8 + j] = (int)o8[j];
                                    else {
                                        uint8_t *px = &S.pred[((lane >> 1) * 8 + j) * 16 + (lane & 1) * 8 + k];
                                        *px = (uint8_t)clip255((int)((uint32_t)*px + (uint32_t)((int)(o8[j] + 32u) >> 6)));
                                    }
                                }
                            }
                        }
                    }
                } else if (lane < 16 && ((cbpl >> (lane >> 2)) & 1)) {
                    int16_t blk[16];
#pragma unroll
                    for (int i = 0; i < 16; i++) blk[i] = S.coef[lane][i];
                    dequant4(blk, qp);
                    add_idct4(&S.pred[blk_y(lane) * 64 + blk_x(lane) * 4], 16, blk);
                }
                __syncthreads();
                recon_chroma(S, cbpc, qpc);
            }
        }
    }

    // 6. Outputs: reconstruction, decision record and levels.
    __syncthreads();
    if (mb_type == 0) {
        // P_Skip: the prediction is the reconstruction (store the chroma prediction).
        if (D.flags) {
            // probe skip: chroma prediction is in registers, luma in predw
            S.predu[lane] = (uint8_t)pu;
            S.predv[lane] = (uint8_t)pv;
            *(uint32_t *)&S.pred[lane * 4] = predw;
        }
        __syncthreads();
        cbpl = cbpc = 0;
    }
    *(uint32_t *)&J.cur_y[(long)(y0 + rr) * J.sy + x0 + c0] = *(const uint32_t *)&S.pred[lane * 4];
    J.cur_u[(long)(cy0 + (lane >> 3)) * J.sc + cx0 + (lane & 7)] = S.predu[lane];
    J.cur_v[(long)(cy0 + (lane >> 3)) * J.sc + cx0 + (lane & 7)] = S.predv[lane];

    int16_t *lv = J.levels + (long)idx * RGH_LEVELS_PER_MB;
    uint32_t nnz = 0;
    if (mb_type == 1 || mb_type == 2) {
        if (mb_type == 1 && use8) {
            for (int i = lane; i < 256; i += 64) lv[i] = S.q8[i >> 6][c_zigzag8[i & 63]];
            nnz = (cbpl & 1 ? 0x0033u : 0) | (cbpl & 2 ? 0x00CCu : 0) | (cbpl & 4 ? 0x3300u : 0) | (cbpl & 8 ? 0xCC00u : 0);
        } else {
            for (int i = lane; i < 256; i += 64) lv[i] = S.coef[i >> 4][c_zigzag4[i & 15]];
            int nzb = 0;
            if (lane < 16) {
                for (int i = 0; i < 16; i++) nzb |= S.coef[lane][i] != 0;
                if (mb_type == 2 && !cbpl) nzb = 0;
            }
            uint64_t bl = __ballot(nzb);
            for (int b = 0; b < 16; b++)
                if ((bl >> b) & 1) nnz |= 1u << (blk_y(b) * 4 + blk_x(b));
        }
        if (lane < 16) lv[256 + lane] = mb_type == 2 ? S.dcq[c_zigzag4[lane]] : 0;
        if (lane < 8) lv[272 + lane] = S.cdc[lane >> 2][lane & 3];
        for (int i = lane; i < 128; i += 64) {
            int p = i >> 6, b = (i >> 4) & 3, k = i & 15;
            lv[280 + i] = (k == 0 || cbpc < 2) ? 0 : S.craster[p][b][c_zigzag4[k]];
        }
    } else {
        for (int i = lane; i < RGH_LEVELS_PER_MB; i += 64) lv[i] = 0;
    }
    if (lane == 0) {
        D.mb_type = (uint8_t)mb_type;
        D.i16_mode = (uint8_t)(mb_type == 2 ? i16_mode : 0);
        D.chroma_mode = (uint8_t)(mb_type == 2 ? cmode : 0);
        D.cbp_luma = (uint8_t)cbpl;
        D.cbp_chroma = (uint8_t)cbpc;
        D.transform_8x8 = (uint8_t)(mb_type == 1 && use8 && cbpl);
        D.mv[0] = (int16_t)mvx;
        D.mv[1] = (int16_t)mvy;
        D.nnz = (uint16_t)nnz;
        J.dec[idx] = D;
    }
}

extern "C" __global__ __launch_bounds__(64) void rgh_k_analyse_step(const RghJob *__restrict__ jobs, int t) {
    const RghJob J = jobs[blockIdx.y];
    if (t >= J.steps) return;
    int a = t - J.mb_w + 1;
    int ylo = a > 0 ? (a + 1) >> 1 : 0, yhi = min(J.mb_h - 1, t >> 1);
    int y = ylo + (int)blockIdx.x;
    if (y > yhi) return;
    __shared__ MbLds S;
    analyse_mb(J, t - 2 * y, y, S);
}
