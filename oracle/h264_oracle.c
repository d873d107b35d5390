/*
 * CPU oracle: plain-C restatement of ragdoll's Balanced H.264 encoder. TEST INFRASTRUCTURE
 * ONLY — see h264_oracle.h for the scope and the pinning. Each function cites the reference
 * file:line it restates (paths relative to the ragdoll repository root; "enc.rs" is
 * crates/ragdoll-codec/src/h264/enc.rs, "h264.rs" is crates/ragdoll-codec/src/h264.rs and
 * "simd/" is crates/ragdoll-simd/src/).
 */
#include "h264_oracle.h"
#include "h264_tables.h"

#include <stdlib.h>
#include <string.h>

#define PAD_Y 48 /* enc.rs:56 */
#define PAD_C 16 /* enc.rs:58 */
#define MV_MARGIN 28 /* enc.rs:61 */
#define MV_COST_RANGE 4096 /* enc.rs:1134 */
#define WIN_W 32 /* simd/h264_enc.rs:705 HPEL_WIN_W */
#define WIN_COLS 24 /* HPEL_COLS */
#define WIN_H 18 /* HPEL_WIN_H */
#define WIN_HR_ROWS (WIN_H + 5)

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline uint8_t clip_u8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }
static inline int iabs(int v) { return v < 0 ? -v : v; }
static inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
static inline int pop32(uint32_t v) { return __builtin_popcount(v); }

/* ------------------------------------------------------------------------------------------ */
/* Bit writer: MSB-first (enc.rs:2674-2774 Bits; ragdoll-core bitstream.rs BitWriter).          */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    uint8_t *buf;
    long len, cap;
    uint64_t acc;
    int n;
} bits_t;

static void bits_init(bits_t *b) {
    b->cap = 1 << 16;
    b->buf = (uint8_t *)malloc((size_t)b->cap);
    b->len = 0;
    b->acc = 0;
    b->n = 0;
}
static void bits_free(bits_t *b) { free(b->buf); }
static void bits_push(bits_t *b, uint8_t v) {
    if (b->len == b->cap) {
        b->cap *= 2;
        b->buf = (uint8_t *)realloc(b->buf, (size_t)b->cap);
    }
    b->buf[b->len++] = v;
}
/* enc.rs:2699 put32: low `len` bits of val, len <= 32 */
static void put32(bits_t *b, uint32_t val, int len) {
    if (len == 0) return;
    b->acc = (b->acc << len) | (uint64_t)(len == 32 ? val : (val & ((1u << len) - 1)));
    b->n += len;
    while (b->n >= 8) {
        b->n -= 8;
        bits_push(b, (uint8_t)(b->acc >> b->n));
    }
}
static void put_bits(bits_t *b, uint64_t val, int len) {
    if (len > 32) {
        put32(b, (uint32_t)(val >> 32), len - 32);
        put32(b, (uint32_t)val, 32);
    } else {
        put32(b, (uint32_t)val, len);
    }
}
static void put_ue(bits_t *b, uint64_t v) { /* enc.rs:2730 */
    uint64_t code = v + 1;
    int len = 64 - __builtin_clzll(code);
    put_bits(b, 0, len - 1);
    put_bits(b, code, len);
}
static void put_se(bits_t *b, int64_t v) { put_ue(b, v > 0 ? (uint64_t)(2 * v - 1) : (uint64_t)(-2 * v)); }
static long bits_written(const bits_t *b) { return b->len * 8 + b->n; }
static void bits_align_zero(bits_t *b) {
    if (b->n & 7) put32(b, 0, 8 - (b->n & 7));
}
static void bits_trailing(bits_t *b) { /* write_rbsp_trailing_bits */
    put32(b, 1, 1);
    bits_align_zero(b);
}

/* ------------------------------------------------------------------------------------------ */
/* Padded planes (enc.rs:95-189).                                                              */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    uint8_t *buf;
    int stride, w, h, pad;
} plane_t;

static void plane_alloc(plane_t *p, int w, int h, int pad) {
    p->stride = (w + 2 * pad + 31) & ~31;
    p->w = w;
    p->h = h;
    p->pad = pad;
    p->buf = (uint8_t *)calloc((size_t)p->stride * (size_t)(h + 2 * pad), 1);
}
static inline long pat(const plane_t *p, int x, int y) {
    return (long)(y + p->pad) * p->stride + x + p->pad;
}
/* enc.rs:164 extend_borders */
static void plane_extend(plane_t *p) {
    int P = p->pad, s = p->stride, w = p->w, h = p->h;
    for (int y = P; y < P + h; y++) {
        uint8_t *row = p->buf + (long)y * s;
        memset(row, row[P], (size_t)P);
        memset(row + P + w, row[P + w - 1], (size_t)(s - P - w));
    }
    for (int y = 0; y < P; y++) {
        memcpy(p->buf + (long)y * s, p->buf + (long)P * s, (size_t)s);
        memcpy(p->buf + (long)(P + h + y) * s, p->buf + (long)(P + h - 1) * s, (size_t)s);
    }
}
typedef struct {
    plane_t y, u, v;
} pic_t;

/* ------------------------------------------------------------------------------------------ */
/* Pixel kernels (scalar twins in simd/).                                                      */
/* ------------------------------------------------------------------------------------------ */
static uint32_t sad_wh(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h) {
    uint32_t s = 0; /* simd/h264_enc.rs:68 sad_scalar */
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) s += (uint32_t)iabs(a[y * sa + x] - b[y * sb + x]);
    return s;
}
static uint32_t ssd_n(const uint8_t *a, int sa, const uint8_t *b, int sb, int n) {
    uint32_t s = 0; /* simd/h264_enc.rs:105 pixel_ssd_scalar */
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            int d = a[y * sa + x] - b[y * sb + x];
            s += (uint32_t)(d * d);
        }
    return s;
}
/* simd/metrics.rs:705 satd_4x4_u8_scalar (Σ|H d Hᵀ| >> 1) */
static uint32_t satd4(const uint8_t *a, int sa, const uint8_t *b, int sb) {
    int t[4][4];
    for (int y = 0; y < 4; y++) {
        int d0 = a[y * sa] - b[y * sb], d1 = a[y * sa + 1] - b[y * sb + 1];
        int d2 = a[y * sa + 2] - b[y * sb + 2], d3 = a[y * sa + 3] - b[y * sb + 3];
        int s03 = d0 + d3, s12 = d1 + d2, d03 = d0 - d3, d12 = d1 - d2;
        t[y][0] = s03 + s12;
        t[y][1] = d03 + d12;
        t[y][2] = s03 - s12;
        t[y][3] = d03 - d12;
    }
    uint32_t sum = 0;
    for (int x = 0; x < 4; x++) {
        int s03 = t[0][x] + t[3][x], s12 = t[1][x] + t[2][x];
        int d03 = t[0][x] - t[3][x], d12 = t[1][x] - t[2][x];
        sum += (uint32_t)(iabs(s03 + s12) + iabs(d03 + d12) + iabs(s03 - s12) + iabs(d03 - d12));
    }
    return sum >> 1;
}
/* simd/h264_enc.rs:167 pixel_satd_16x16_scalar: sum of unhalved 4x4 SATDs, halved once */
static uint32_t satd16(const uint8_t *a, int sa, const uint8_t *b, int sb) {
    uint32_t sum = 0;
    for (int by = 0; by < 4; by++)
        for (int bx = 0; bx < 4; bx++) {
            int d[4][4];
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++)
                    d[y][x] = a[(by * 4 + y) * sa + bx * 4 + x] - b[(by * 4 + y) * sb + bx * 4 + x];
            for (int y = 0; y < 4; y++) {
                int s01 = d[y][0] + d[y][1], d01 = d[y][0] - d[y][1];
                int s23 = d[y][2] + d[y][3], d23 = d[y][2] - d[y][3];
                d[y][0] = s01 + s23;
                d[y][1] = s01 - s23;
                d[y][2] = d01 + d23;
                d[y][3] = d01 - d23;
            }
            for (int c = 0; c < 4; c++) {
                int s01 = d[0][c] + d[1][c], d01 = d[0][c] - d[1][c];
                int s23 = d[2][c] + d[3][c], d23 = d[2][c] - d[3][c];
                sum += (uint32_t)(iabs(s01 + s23) + iabs(s01 - s23) + iabs(d01 + d23) + iabs(d01 - d23));
            }
        }
    return sum >> 1;
}
/* simd/fixed.rs:104 forward_dct4x4_scalar */
static void fdct4(const int16_t *blk, int16_t *out) {
    int t[16];
    for (int r = 0; r < 4; r++) {
        int i = r * 4;
        int s03 = blk[i] + blk[i + 3], s12 = blk[i + 1] + blk[i + 2];
        int d03 = blk[i] - blk[i + 3], d12 = blk[i + 1] - blk[i + 2];
        t[i] = s03 + s12;
        t[i + 1] = (d03 << 1) + d12;
        t[i + 2] = s03 - s12;
        t[i + 3] = d03 - (d12 << 1);
    }
    for (int c = 0; c < 4; c++) {
        int s03 = t[c] + t[c + 12], s12 = t[c + 4] + t[c + 8];
        int d03 = t[c] - t[c + 12], d12 = t[c + 4] - t[c + 8];
        out[c] = (int16_t)(s03 + s12);
        out[c + 4] = (int16_t)((d03 << 1) + d12);
        out[c + 8] = (int16_t)(s03 - s12);
        out[c + 12] = (int16_t)(d03 - (d12 << 1));
    }
}
/* simd/fixed.rs:185 inverse_idct4x4_scalar (with +32 >> 6) */
static void idct4(const int16_t *c, int16_t *out) {
    int t[16];
    for (int r = 0; r < 4; r++) {
        int i = r * 4;
        int e0 = c[i] + c[i + 2], e1 = c[i] - c[i + 2];
        int e2 = (c[i + 1] >> 1) - c[i + 3], e3 = c[i + 1] + (c[i + 3] >> 1);
        t[i] = e0 + e3;
        t[i + 1] = e1 + e2;
        t[i + 2] = e1 - e2;
        t[i + 3] = e0 - e3;
    }
    for (int k = 0; k < 4; k++) {
        int e0 = t[k] + t[k + 8], e1 = t[k] - t[k + 8];
        int e2 = (t[k + 4] >> 1) - t[k + 12], e3 = t[k + 4] + (t[k + 12] >> 1);
        out[k] = (int16_t)((e0 + e3 + 32) >> 6);
        out[k + 4] = (int16_t)((e1 + e2 + 32) >> 6);
        out[k + 8] = (int16_t)((e1 - e2 + 32) >> 6);
        out[k + 12] = (int16_t)((e0 - e3 + 32) >> 6);
    }
}
/* simd/fixed.rs:318 hadamard4x4_scalar (luma DC, with the final >> 1) */
static void hadamard4_dc(const int16_t *b, int16_t *out) {
    int t[16];
    for (int r = 0; r < 4; r++) {
        int i = r * 4;
        int s03 = b[i] + b[i + 3], s12 = b[i + 1] + b[i + 2];
        int d03 = b[i] - b[i + 3], d12 = b[i + 1] - b[i + 2];
        t[i] = s03 + s12;
        t[i + 1] = d03 + d12;
        t[i + 2] = s03 - s12;
        t[i + 3] = d03 - d12;
    }
    for (int c = 0; c < 4; c++) {
        int s03 = t[c] + t[c + 12], s12 = t[c + 4] + t[c + 8];
        int d03 = t[c] - t[c + 12], d12 = t[c + 4] - t[c + 8];
        out[c] = (int16_t)((s03 + s12) >> 1);
        out[c + 4] = (int16_t)((d03 + d12) >> 1);
        out[c + 8] = (int16_t)((s03 - s12) >> 1);
        out[c + 12] = (int16_t)((d03 - d12) >> 1);
    }
}
/* simd/h264_enc.rs:313 sub_dct4x4_scalar */
static void sub_dct4(const uint8_t *src, int ss, const uint8_t *pred, int ps, int16_t *out) {
    int16_t d[16];
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) d[y * 4 + x] = (int16_t)(src[y * ss + x] - pred[y * ps + x]);
    fdct4(d, out);
}
/* simd/h264_enc.rs:297 sub_dct4x4_8x8_scalar: four 4x4 blocks of an 8x8, raster order */
static void sub_dct4_8x8(const uint8_t *src, int ss, const uint8_t *pred, int ps, int16_t out[4][16]) {
    for (int b = 0; b < 4; b++)
        sub_dct4(src + (b >> 1) * 4 * ss + (b & 1) * 4, ss, pred + (b >> 1) * 4 * ps + (b & 1) * 4, ps,
                 out[b]);
}
/* simd/fixed.rs:392 forward_dct8x8 */
static void fdct8(const int16_t *blk, int16_t *out) {
    int t[64];
    for (int pass = 0; pass < 2; pass++) {
        for (int k = 0; k < 8; k++) {
            int v[8];
            for (int j = 0; j < 8; j++) v[j] = pass == 0 ? blk[k * 8 + j] : t[j * 8 + k];
            int a0 = v[0] + v[7], a1 = v[1] + v[6], a2 = v[2] + v[5], a3 = v[3] + v[4];
            int a4 = v[0] - v[7], a5 = v[1] - v[6], a6 = v[2] - v[5], a7 = v[3] - v[4];
            int b0 = a0 + a3, b1 = a1 + a2, b2 = a0 - a3, b3 = a1 - a2;
            int b4 = a5 + a6 + ((a4 >> 1) + a4);
            int b5 = a4 - a7 - ((a6 >> 1) + a6);
            int b6 = a4 + a7 - ((a5 >> 1) + a5);
            int b7 = a5 - a6 + ((a7 >> 1) + a7);
            int o[8] = {b0 + b1, b4 + (b7 >> 2), b2 + (b3 >> 1), b5 + (b6 >> 2),
                        b0 - b1, b6 - (b5 >> 2), (b2 >> 1) - b3, (b4 >> 2) - b7};
            for (int j = 0; j < 8; j++) {
                if (pass == 0)
                    t[k * 8 + j] = o[j];
                else
                    out[j * 8 + k] = (int16_t)o[j];
            }
        }
    }
}
/* simd/h264_idct.rs:226 idct8_1d (wrapping i32) */
static void idct8_1d(const int32_t *d, int32_t *o) {
    uint32_t a0 = (uint32_t)d[0] + (uint32_t)d[4];
    uint32_t a1 = (uint32_t)d[5] - (uint32_t)d[3] - (uint32_t)d[7] - (uint32_t)(d[7] >> 1);
    uint32_t a2 = (uint32_t)d[0] - (uint32_t)d[4];
    uint32_t a3 = (uint32_t)d[1] + (uint32_t)d[7] - (uint32_t)d[3] - (uint32_t)(d[3] >> 1);
    uint32_t a4 = (uint32_t)(d[2] >> 1) - (uint32_t)d[6];
    uint32_t a5 = (uint32_t)d[7] - (uint32_t)d[1] + (uint32_t)d[5] + (uint32_t)(d[5] >> 1);
    uint32_t a6 = (uint32_t)d[2] + (uint32_t)(d[6] >> 1);
    uint32_t a7 = (uint32_t)d[3] + (uint32_t)d[5] + (uint32_t)d[1] + (uint32_t)(d[1] >> 1);
    uint32_t b0 = a0 + a6, b1 = a1 + (uint32_t)((int32_t)a7 >> 2), b2 = a2 + a4;
    uint32_t b3 = a3 + (uint32_t)((int32_t)a5 >> 2), b4 = a2 - a4;
    uint32_t b5 = (uint32_t)((int32_t)a3 >> 2) - a5, b6 = a0 - a6;
    uint32_t b7 = a7 - (uint32_t)((int32_t)a1 >> 2);
    o[0] = (int32_t)(b0 + b7);
    o[1] = (int32_t)(b2 + b5);
    o[2] = (int32_t)(b4 + b3);
    o[3] = (int32_t)(b6 + b1);
    o[4] = (int32_t)(b6 - b1);
    o[5] = (int32_t)(b4 - b3);
    o[6] = (int32_t)(b2 - b5);
    o[7] = (int32_t)(b0 - b7);
}
/* simd/h264_idct.rs:268 h264_idct8_add_scalar */
static void idct8_add(uint8_t *dst, int stride, const int32_t *d) {
    int32_t g[64], col[8], out[8];
    for (int i = 0; i < 8; i++) idct8_1d(d + i * 8, g + i * 8);
    for (int j = 0; j < 8; j++) {
        for (int i = 0; i < 8; i++) col[i] = g[i * 8 + j];
        idct8_1d(col, out);
        for (int i = 0; i < 8; i++) {
            uint8_t *p = dst + i * stride + j;
            *p = clip_u8((int32_t)((uint32_t)*p + (uint32_t)((int32_t)((uint32_t)out[i] + 32u) >> 6)));
        }
    }
}
/* simd/h264_enc.rs:374-430 Quant4x4 and quant_4x4_x4_scalar */
typedef struct {
    uint16_t mf[16];
    uint32_t bias, shift;
} quant4_t;
static int coeff_pos_type(int i) { /* h264.rs:1204 */
    int r = i >> 2, c = i & 3;
    if (!(r & 1) && !(c & 1)) return 0;
    if ((r & 1) && (c & 1)) return 1;
    return 2;
}
static int coeff8_pos_type(int i) { /* h264.rs:1343 */
    int r = (i >> 3) & 3, c = (i & 7) & 3;
    if (r == 0 && c == 0) return 0;
    if ((r & 1) && (c & 1)) return 1;
    if (r == 2 && c == 2) return 2;
    if ((r == 0 && (c & 1)) || ((r & 1) && c == 0)) return 3;
    if ((r == 0 && c == 2) || (r == 2 && c == 0)) return 4;
    return 5;
}
static quant4_t QUANT4[52][2];
static int32_t DEQUANT4[52][16];
static uint16_t QUANT8_MF16[6][64];
static int32_t DEQUANT8_POS[6][64];
static void init_tables(void) {
    static int done = 0;
    if (done) return;
    for (int qp = 0; qp < 52; qp++) { /* enc.rs:1731 QUANT4 / :1761 DEQUANT4 */
        uint32_t shift = 15 + (uint32_t)(qp / 6);
        for (int k = 0; k < 2; k++) {
            for (int i = 0; i < 16; i++) QUANT4[qp][k].mf[i] = (uint16_t)T_QUANT4_MF[qp % 6][coeff_pos_type(i)];
            QUANT4[qp][k].shift = shift;
            QUANT4[qp][k].bias = (1u << shift) / (k ? 3 : 6);
        }
        for (int i = 0; i < 16; i++) DEQUANT4[qp][i] = T_DEQUANT4_INIT[qp % 6][coeff_pos_type(i)] << (qp / 6);
    }
    for (int r = 0; r < 6; r++)
        for (int i = 0; i < 64; i++) { /* h264.rs:1357 QUANT8_POS */
            QUANT8_MF16[r][i] = (uint16_t)T_QUANT8_MF[r][coeff8_pos_type(i)];
            DEQUANT8_POS[r][i] = T_DEQUANT8_INIT[r][coeff8_pos_type(i)];
        }
    done = 1;
}
static const quant4_t *quant4(int qp, int intra) { return &QUANT4[imin(qp, 51)][intra ? 1 : 0]; }
/* quant_4x4_x4_scalar: bit b = block b non-zero, bit 4+b = a level above 1 */
static uint32_t quant_4x4_x4(int16_t blocks[4][16], const quant4_t *q) {
    uint32_t nz = 0;
    for (int b = 0; b < 4; b++) {
        int any = 0, big = 0;
        for (int i = 0; i < 16; i++) {
            int c = blocks[b][i];
            uint64_t lv = ((uint64_t)iabs(c) * q->mf[i] + q->bias) >> q->shift;
            if (lv > 32767) lv = 32767;
            blocks[b][i] = (int16_t)(c < 0 ? -(int)lv : (int)lv);
            any |= lv != 0;
            big |= lv > 1;
        }
        if (any) nz |= 1u << b;
        if (big) nz |= 0x10u << b;
    }
    return nz;
}
/* simd/h264_enc.rs:461 dequant_4x4_x4_scalar */
static void dequant_4x4(int16_t *blk, const int32_t *scale) {
    for (int i = 0; i < 16; i++) blk[i] = (int16_t)iclamp(blk[i] * scale[i], -32768, 32767);
}
/* simd/h264_enc.rs:497 add_idct4x4_8x8_scalar: blocks are the 4 quadrants of an 8x8 */
static void add_idct4_8x8(uint8_t *dst, int ds, int16_t blocks[4][16]) {
    for (int b = 0; b < 4; b++) {
        int16_t res[16];
        idct4(blocks[b], res);
        uint8_t *o = dst + (b >> 1) * 4 * ds + (b & 1) * 4;
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) o[y * ds + x] = clip_u8(o[y * ds + x] + res[y * 4 + x]);
    }
}
/* simd/fixed.rs:288 add4x4_idct_u8 (DC-only fast path is exact) */
static void add_idct4(uint8_t *dst, int ds, const int16_t *c) {
    int16_t res[16];
    idct4(c, res);
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) dst[y * ds + x] = clip_u8(dst[y * ds + x] + res[y * 4 + x]);
}
static void zigzag4(const int16_t *raster, int16_t *zz) {
    for (int i = 0; i < 16; i++) zz[i] = raster[T_ZIGZAG4[i]];
}
static uint32_t nz_mask(const int16_t *c, int n) {
    uint32_t m = 0;
    for (int i = 0; i < n; i++) m |= (uint32_t)(c[i] != 0) << i;
    return m;
}
/* enc.rs:583 decimate_masked (and :557 decimate_score with the 4x4 table) */
static uint32_t decimate_masked(const int16_t *scanned, uint32_t mask) {
    uint32_t score = 0, m = mask;
    while (m) {
        int p = 31 - clz32(m);
        if (iabs(scanned[p]) > 1) return 9;
        m ^= 1u << p;
        int below = m ? 31 - clz32(m) : -1;
        score += T_DECIMATE4[p - below - 1];
    }
    return score;
}
/* enc.rs:557 decimate_score (8x8 table for 64 levels) */
static uint32_t decimate_score(const int16_t *scanned, int n) {
    const uint8_t *table = n == 64 ? T_DECIMATE8 : T_DECIMATE4;
    int idx = n;
    while (idx > 0 && scanned[idx - 1] == 0) idx--;
    uint32_t score = 0;
    while (idx > 0) {
        idx--;
        if (iabs(scanned[idx]) > 1) return 9;
        int run = 0;
        while (idx > 0 && scanned[idx - 1] == 0) {
            idx--;
            run++;
        }
        score += table[run];
    }
    return score;
}

/* simd/h264_mc.rs:129 h264_luma_mc_scalar: spec 8.4.2.2.1 at quarter-pel phase (mx, my) */
static int tap6(int a, int b, int c, int d, int e, int f) { return a - 5 * b + 20 * c + 20 * d - 5 * e + f; }
static void luma_mc(uint8_t *dst, int ds, const uint8_t *src, int ss, int w, int h, int mx, int my) {
#define AT(r, c) ((int)src[(long)(r) * ss + (c)])
#define HRAW(r, c) tap6(AT(r, (c)-2), AT(r, (c)-1), AT(r, c), AT(r, (c) + 1), AT(r, (c) + 2), AT(r, (c) + 3))
#define G(r, c) AT(r, c)
#define B(r, c) clip_u8((HRAW(r, c) + 16) >> 5)
#define V(r, c) clip_u8((tap6(AT((r)-2, c), AT((r)-1, c), AT(r, c), AT((r) + 1, c), AT((r) + 2, c), AT((r) + 3, c)) + 16) >> 5)
#define J(r, c)                                                                                     \
    clip_u8((tap6(HRAW((r)-2, c), HRAW((r)-1, c), HRAW(r, c), HRAW((r) + 1, c), HRAW((r) + 2, c), \
                  HRAW((r) + 3, c)) +                                                             \
             512) >>                                                                              \
            10)
#define AVG(a, b) (uint8_t)(((a) + (b) + 1) >> 1)
    mx &= 3;
    my &= 3;
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            uint8_t v;
            switch (my * 4 + mx) {
            case 0: v = (uint8_t)G(r, c); break;
            case 1: v = AVG(G(r, c), B(r, c)); break;
            case 2: v = B(r, c); break;
            case 3: v = AVG(B(r, c), G(r, c + 1)); break;
            case 4: v = AVG(G(r, c), V(r, c)); break;
            case 5: v = AVG(B(r, c), V(r, c)); break;
            case 6: v = AVG(B(r, c), J(r, c)); break;
            case 7: v = AVG(B(r, c), V(r, c + 1)); break;
            case 8: v = V(r, c); break;
            case 9: v = AVG(V(r, c), J(r, c)); break;
            case 10: v = J(r, c); break;
            case 11: v = AVG(J(r, c), V(r, c + 1)); break;
            case 12: v = AVG(V(r, c), G(r + 1, c)); break;
            case 13: v = AVG(V(r, c), B(r + 1, c)); break;
            case 14: v = AVG(J(r, c), B(r + 1, c)); break;
            default: v = AVG(V(r, c + 1), B(r + 1, c)); break;
            }
            dst[r * ds + c] = v;
        }
#undef AT
#undef HRAW
#undef G
#undef B
#undef V
#undef J
#undef AVG
}
/* simd/h264_enc.rs:1279 chroma_mc_8x8_x2_scalar (one plane) */
static void chroma_mc8(uint8_t *out, const uint8_t *src, int ss, int fx, int fy) {
    int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            const uint8_t *s0 = src + (long)r * ss, *s1 = src + (long)(r + 1) * ss;
            int x = wa * s0[c] + wb * s0[c + 1] + wc * s1[c] + wd * s1[c + 1];
            out[r * 8 + c] = (uint8_t)((x + 32) >> 6);
        }
}

/* simd/h264_enc.rs:704-838 HpelWindow: b (h), h (v) and j (c) planes over 24 x 18 */
typedef struct {
    uint8_t h[WIN_W * WIN_H], v[WIN_W * WIN_H], c[WIN_W * WIN_H];
    int16_t hr[WIN_HR_ROWS * WIN_COLS];
} hpel_win_t;
static void hpel_window(const uint8_t *src, int ss, hpel_win_t *w) { /* :765 */
#define AT(r, c) ((int)src[(long)(r) * ss + (c)])
    for (int k = 0; k < WIN_HR_ROWS; k++) {
        int r = k - 2;
        for (int c = 0; c < WIN_COLS; c++) {
            int hh = tap6(AT(r, c - 2), AT(r, c - 1), AT(r, c), AT(r, c + 1), AT(r, c + 2), AT(r, c + 3));
            w->hr[k * WIN_COLS + c] = (int16_t)hh;
            if (r >= 0 && r < WIN_H) {
                int i = r * WIN_W + c;
                w->h[i] = clip_u8((hh + 16) >> 5);
                int vv = tap6(AT(r - 2, c), AT(r - 1, c), AT(r, c), AT(r + 1, c), AT(r + 2, c), AT(r + 3, c));
                w->v[i] = clip_u8((vv + 16) >> 5);
            }
        }
    }
#undef AT
}
static void hpel_window_centre(hpel_win_t *w) { /* :820 */
    for (int r = 0; r < WIN_H; r++)
        for (int c = 0; c < WIN_COLS; c++) {
            const int16_t *t = w->hr + r * WIN_COLS + c;
            int j = t[0] - 5 * t[WIN_COLS] + 20 * t[2 * WIN_COLS] + 20 * t[3 * WIN_COLS] - 5 * t[4 * WIN_COLS] +
                    t[5 * WIN_COLS];
            w->c[r * WIN_W + c] = clip_u8((j + 512) >> 10);
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Intra predictors.                                                                           */
/* ------------------------------------------------------------------------------------------ */
/* h264.rs:1715 predict_intra_16x16_scalar, plus enc.rs:1559 pred16 plane path (identical) */
static void pred16(int mode, const uint8_t *top, const uint8_t *left, int tl, uint8_t *dst) {
    uint8_t t[16], l[16];
    for (int i = 0; i < 16; i++) {
        t[i] = top ? top[i] : 128;
        l[i] = left ? left[i] : 128;
    }
    if (mode == 0) {
        for (int y = 0; y < 16; y++) memcpy(dst + y * 16, t, 16);
    } else if (mode == 1) {
        for (int y = 0; y < 16; y++) memset(dst + y * 16, l[y], 16);
    } else if (mode == 2) {
        int dc;
        uint32_t st = 0, sl = 0;
        for (int i = 0; i < 16; i++) {
            st += t[i];
            sl += l[i];
        }
        if (top && left)
            dc = (int)((st + sl + 16) >> 5);
        else if (top)
            dc = (int)((st + 8) >> 4);
        else if (left)
            dc = (int)((sl + 8) >> 4);
        else
            dc = 128;
        memset(dst, dc, 256);
    } else {
        int H = 0, Vv = 0;
        for (int i = 0; i < 8; i++) {
            int pt = i == 7 ? tl : t[6 - i], pl = i == 7 ? tl : l[6 - i];
            H += (i + 1) * (t[8 + i] - pt);
            Vv += (i + 1) * (l[8 + i] - pl);
        }
        int a = 16 * (t[15] + l[15]), b = (5 * H + 32) >> 6, c = (5 * Vv + 32) >> 6;
        for (int y = 0; y < 16; y++)
            for (int x = 0; x < 16; x++) dst[y * 16 + x] = clip_u8((a + b * (x - 7) + c * (y - 7) + 16) >> 5);
    }
}
/* h264.rs:1967 predict_intra_chroma_8x8_scalar and enc.rs:1588 pred_chroma (plane) */
static void pred_chroma(int mode, const uint8_t *top, const uint8_t *left, int tl, uint8_t *dst) {
    uint8_t t[8], l[8];
    for (int i = 0; i < 8; i++) {
        t[i] = top ? top[i] : 128;
        l[i] = left ? left[i] : 128;
    }
    if (mode == 0) {
        uint32_t st0 = t[0] + t[1] + t[2] + t[3], st1 = t[4] + t[5] + t[6] + t[7];
        uint32_t sl0 = l[0] + l[1] + l[2] + l[3], sl1 = l[4] + l[5] + l[6] + l[7];
        int d00, d01, d10, d11;
        if (top && left) {
            d00 = (int)((st0 + sl0 + 4) >> 3);
            d01 = (int)((st1 + 2) >> 2);
            d10 = (int)((sl1 + 2) >> 2);
            d11 = (int)((st1 + sl1 + 4) >> 3);
        } else if (top) {
            d00 = d10 = (int)((st0 + 2) >> 2);
            d01 = d11 = (int)((st1 + 2) >> 2);
        } else if (left) {
            d00 = d01 = (int)((sl0 + 2) >> 2);
            d10 = d11 = (int)((sl1 + 2) >> 2);
        } else {
            d00 = d01 = d10 = d11 = 128;
        }
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++)
                dst[y * 8 + x] = (uint8_t)(y < 4 ? (x < 4 ? d00 : d01) : (x < 4 ? d10 : d11));
    } else if (mode == 1) {
        for (int y = 0; y < 8; y++) memset(dst + y * 8, l[y], 8);
    } else if (mode == 2) {
        for (int y = 0; y < 8; y++) memcpy(dst + y * 8, t, 8);
    } else {
        int H = 0, Vv = 0;
        for (int i = 0; i < 4; i++) {
            int pt = i == 3 ? tl : t[2 - i], pl = i == 3 ? tl : l[2 - i];
            H += (i + 1) * (t[4 + i] - pt);
            Vv += (i + 1) * (l[4 + i] - pl);
        }
        int a = 16 * (t[7] + l[7]), b = (17 * H + 16) >> 5, c = (17 * Vv + 16) >> 5;
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) dst[y * 8 + x] = clip_u8((a + b * (x - 3) + c * (y - 3) + 16) >> 5);
    }
}
/* simd/h264_dsp.rs:799 predict_intra_4x4_u8_scalar (modes 0..8) */
static void pred4(int mode, const uint8_t *top8, const uint8_t *left4, int tl, uint8_t *dst, int ds) {
    uint8_t t[8], l[4];
    for (int i = 0; i < 8; i++) t[i] = top8 ? top8[i] : 128;
    for (int i = 0; i < 4; i++) l[i] = left4 ? left4[i] : 128;
    uint32_t A = t[0], B = t[1], C = t[2], D = t[3], E = t[4], F = t[5], Gg = t[6], Hh = t[7];
    uint32_t I = l[0], Jj = l[1], K = l[2], L = l[3], M = (uint32_t)tl;
#define A2(x, y) (uint8_t)(((x) + (y) + 1) >> 1)
#define A3(x, y, z) (uint8_t)(((x) + 2 * (y) + (z) + 2) >> 2)
    uint8_t p[16];
    switch (mode) {
    case 0:
        for (int y = 0; y < 4; y++) memcpy(p + y * 4, t, 4);
        break;
    case 1:
        for (int y = 0; y < 4; y++) memset(p + y * 4, l[y], 4);
        break;
    case 2: {
        int dc;
        if (top8 && left4)
            dc = (int)((A + B + C + D + I + Jj + K + L + 4) >> 3);
        else if (top8)
            dc = (int)((A + B + C + D + 2) >> 2);
        else if (left4)
            dc = (int)((I + Jj + K + L + 2) >> 2);
        else
            dc = 128;
        memset(p, dc, 16);
    } break;
    case 3: {
        uint8_t q[7] = {A3(A, B, C), A3(B, C, D), A3(C, D, E), A3(D, E, F), A3(E, F, Gg), A3(F, Gg, Hh), A3(Gg, Hh, Hh)};
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) p[y * 4 + x] = q[x + y];
    } break;
    case 4: {
        /* index by (x - y): -3..3 */
        uint8_t q[7] = {A3(L, K, Jj), A3(K, Jj, I), A3(Jj, I, M), A3(I, M, A), A3(M, A, B), A3(A, B, C), A3(B, C, D)};
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) p[y * 4 + x] = q[x - y + 3];
    } break;
    case 5: {
        uint8_t r0[4] = {A2(M, A), A2(A, B), A2(B, C), A2(C, D)};
        uint8_t r1[4] = {A3(I, M, A), A3(M, A, B), A3(A, B, C), A3(B, C, D)};
        uint8_t r2[4] = {A3(Jj, I, M), r0[0], r0[1], r0[2]};
        uint8_t r3[4] = {A3(K, Jj, I), r1[0], r1[1], r1[2]};
        memcpy(p, r0, 4);
        memcpy(p + 4, r1, 4);
        memcpy(p + 8, r2, 4);
        memcpy(p + 12, r3, 4);
    } break;
    case 6: {
        uint8_t r0[4] = {A2(M, I), A3(I, M, A), A3(M, A, B), A3(A, B, C)};
        uint8_t r1[4] = {A2(I, Jj), A3(M, I, Jj), r0[0], r0[1]};
        uint8_t r2[4] = {A2(Jj, K), A3(I, Jj, K), r1[0], r1[1]};
        uint8_t r3[4] = {A2(K, L), A3(Jj, K, L), r2[0], r2[1]};
        memcpy(p, r0, 4);
        memcpy(p + 4, r1, 4);
        memcpy(p + 8, r2, 4);
        memcpy(p + 12, r3, 4);
    } break;
    case 7: {
        uint8_t r0[4] = {A2(A, B), A2(B, C), A2(C, D), A2(D, E)};
        uint8_t r1[4] = {A3(A, B, C), A3(B, C, D), A3(C, D, E), A3(D, E, F)};
        uint8_t r2[4] = {r0[1], r0[2], r0[3], A2(E, F)};
        uint8_t r3[4] = {r1[1], r1[2], r1[3], A3(E, F, Gg)};
        memcpy(p, r0, 4);
        memcpy(p + 4, r1, 4);
        memcpy(p + 8, r2, 4);
        memcpy(p + 12, r3, 4);
    } break;
    default: {
        uint8_t p0 = A2(I, Jj), p1 = A3(I, Jj, K), p2 = A2(Jj, K), p3 = A3(Jj, K, L), p4 = A2(K, L),
                p5 = A3(K, L, L), p6 = (uint8_t)L;
        uint8_t r[16] = {p0, p1, p2, p3, p2, p3, p4, p5, p4, p5, p6, p6, p6, p6, p6, p6};
        memcpy(p, r, 16);
    } break;
    }
#undef A2
#undef A3
    for (int y = 0; y < 4; y++) memcpy(dst + y * ds, p + y * 4, 4);
}
static int intra4_mode_available(int mode, int has_top, int has_left) { /* h264.rs:3273 */
    switch (mode) {
    case 0: case 3: case 7: return has_top;
    case 1: case 8: return has_left;
    case 2: return 1;
    default: return has_top && has_left;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* CAVLC block coder (h264.rs:2215 encode_cavlc_block; bit-identical to enc.rs:867).           */
/* ------------------------------------------------------------------------------------------ */
static int nc_class(int nc) { return nc < 2 ? 0 : (nc < 4 ? 1 : (nc < 8 ? 2 : 3)); }
static void cavlc_block(bits_t *w, const int16_t *sc, int maxc, int nc, int chroma_dc) {
    int tbl = nc_class(nc);
    int last = -1;
    for (int i = maxc - 1; i >= 0; i--)
        if (sc[i]) {
            last = i;
            break;
        }
    if (last < 0) {
        if (chroma_dc)
            put32(w, T_CDC_TOKEN_BITS[0], T_CDC_TOKEN_LEN[0]);
        else
            put32(w, T_COEFF_TOKEN_BITS[tbl][0], T_COEFF_TOKEN_LEN[tbl][0]);
        return;
    }
    int levels[16], runs[16], nlev = 0, total_zeros = 0, cur_run = 0;
    for (int i = last; i >= 0; i--) {
        if (sc[i]) {
            if (nlev > 0) {
                runs[nlev - 1] = cur_run;
                total_zeros += cur_run;
            }
            levels[nlev++] = sc[i];
            cur_run = 0;
        } else {
            cur_run++;
        }
    }
    runs[nlev - 1] = cur_run;
    total_zeros += cur_run;
    int t1 = 0;
    for (int i = 0; i < nlev && i < 3; i++) {
        if (iabs(levels[i]) == 1)
            t1++;
        else
            break;
    }
    int tok = nlev * 4 + t1;
    if (chroma_dc)
        put32(w, T_CDC_TOKEN_BITS[tok], T_CDC_TOKEN_LEN[tok]);
    else
        put32(w, T_COEFF_TOKEN_BITS[tbl][tok], T_COEFF_TOKEN_LEN[tbl][tok]);
    for (int i = 0; i < t1; i++) put32(w, levels[i] < 0, 1);
    uint32_t sfx = (nlev > 10 && t1 < 3) ? 1 : 0;
    for (int idx = t1; idx < nlev; idx++) {
        int lv = levels[idx];
        int code = lv > 0 ? 2 * lv - 2 : -2 * lv - 1;
        if (idx == t1 && t1 < 3) code -= 2;
        uint32_t cu = (uint32_t)code;
        if (cu < (14u << sfx)) {
            uint32_t prefix = cu >> sfx;
            put_bits(w, 0, (int)prefix);
            put32(w, 1, 1);
            if (sfx) put32(w, cu & ((1u << sfx) - 1), (int)sfx);
        } else if (sfx == 0 && cu < 30) {
            put32(w, 0, 14);
            put32(w, 1, 1);
            put32(w, cu - 14, 4);
        } else if (sfx > 0 && cu < (15u << sfx)) {
            put32(w, 0, 14);
            put32(w, 1, 1);
            put32(w, cu - (14u << sfx), (int)sfx);
        } else {
            uint32_t esc = cu - (15u << sfx);
            if (sfx == 0) esc -= 15;
            put32(w, 0, 15);
            put32(w, 1, 1);
            put32(w, esc & 0xFFF, 12);
        }
        if (sfx == 0) sfx = 1;
        if ((uint32_t)iabs(lv) > (3u << (sfx - 1)) && sfx < 6) sfx++;
    }
    if (nlev < maxc) {
        if (chroma_dc)
            put32(w, T_CDC_TZ_BITS[nlev - 1][total_zeros], T_CDC_TZ_LEN[nlev - 1][total_zeros]);
        else
            put32(w, T_TOTAL_ZEROS_BITS[nlev - 1][total_zeros], T_TOTAL_ZEROS_LEN[nlev - 1][total_zeros]);
    }
    int zl = total_zeros;
    for (int i = 0; i < nlev - 1 && zl > 0; i++) {
        int vi = imin(zl - 1, 6);
        put32(w, T_RUN_BITS[vi][runs[i]], T_RUN_LEN[vi][runs[i]]);
        zl -= runs[i];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Encoder state (enc.rs:193-360, h264.rs:2996-3165).                                          */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int16_t dc[2][4];
    int16_t ac[2][4][15];
    uint32_t ac_masks[2][4];
    int16_t raster[2][4][16];
} chroma_levels_t;

typedef struct {
    uint32_t luma, luma_ssd, chroma, chroma_ssd, probe;
    int qp;
} skip_bounds_t;

struct ora_enc {
    int width, height, mb_w, mb_h, fps_num, fps_den, qp, profile, keyint, me_iters;
    int cabac, t8x8;
    unsigned frame_count;
    int has_ref, last_idr;
    pic_t cur, refp;
    uint8_t *luma_nnz, *chroma_nnz[2], *intra4x4_modes;
    int8_t *ref_grid;
    int16_t (*mv_grid)[2];
    int16_t (*prev_mb_mv)[2];
    uint8_t *mb_is_intra, *mb_is_i16, *mb_is_skip, *mb_cbp_luma, *mb_cbp_chroma, *mb_qps, *mb_8x8;
    uint16_t *mv_costs; /* 2*MV_COST_RANGE+1 entries at the slice QP */
    int mv_costs_qp;
    ora_mb_decision *dec;
    int16_t *levels;
    uint8_t *pre_y, *pre_u, *pre_v; /* the last picture before deblocking, coded size */
    /* slice state */
    int is_idr, skip_run, cur_qp;
    skip_bounds_t sb;
    hpel_win_t win;
    /* Compact preset (enc/compact.rs, enc/inter.rs): chroma QP offset 1, fine inter rounding,
     * half-sample planes of the reference, the 16x16 search vectors of this and the previous
     * picture (mvr), the Compact motion vector costs. */
    int compact, cqo;
    int fast; /* H264Speed::Fast (enc.rs:503): half-pel motion, intra 16x16 only, no loop filter */
    int parts; /* Balanced with P partitions 16x8 / 8x16 / 8x8 (ora_set_partitions) */
    plane_t hh, hv, hc;
    int16_t (*mvr)[2], (*prev_mvr)[2];
    uint16_t *mv_costs_c;
    int mv_costs_c_qp;
};

int ora_mb_w(const ora_enc *e) { return e->mb_w; }
int ora_mb_h(const ora_enc *e) { return e->mb_h; }
int ora_last_was_idr(const ora_enc *e) { return e->last_idr; }

/* enc.rs:540 chroma_qp with the picture parameter set's chroma_qp_index_offset */
static int chroma_qp_o(const ora_enc *e, int qp) { return T_CHROMA_QP[iclamp(qp + e->cqo, 0, 51)]; }
#define chroma_qp(qp) chroma_qp_o(e, (qp))
static uint32_t se_bits(int d) { /* enc.rs:546 */
    uint32_t v = (uint32_t)iabs(d) * 2 - (uint32_t)(d > 0);
    return 2 * (uint32_t)(32 - clz32(v + 1)) - 1;
}

ora_enc *ora_create(int width, int height, int fps_num, int fps_den, int qp, int profile, int keyint) {
    init_tables();
    if (width <= 0 || height <= 0 || (width & 1) || (height & 1)) return NULL;
    ora_enc *e = (ora_enc *)calloc(1, sizeof(ora_enc));
    e->width = width;
    e->height = height;
    e->mb_w = (width + 15) / 16;
    e->mb_h = (height + 15) / 16;
    e->fps_num = fps_num;
    e->fps_den = fps_den;
    e->qp = iclamp(qp, 0, 51);
    e->profile = profile;
    e->keyint = keyint > 0 ? keyint : 30;
    e->me_iters = 16;
    e->cabac = profile >= 1;
    e->t8x8 = profile == 2;
    int W = e->mb_w * 16, H = e->mb_h * 16, n = e->mb_w * e->mb_h, n4 = n * 16;
    plane_alloc(&e->cur.y, W, H, PAD_Y);
    plane_alloc(&e->cur.u, W / 2, H / 2, PAD_C);
    plane_alloc(&e->cur.v, W / 2, H / 2, PAD_C);
    plane_alloc(&e->refp.y, W, H, PAD_Y);
    plane_alloc(&e->refp.u, W / 2, H / 2, PAD_C);
    plane_alloc(&e->refp.v, W / 2, H / 2, PAD_C);
    e->luma_nnz = (uint8_t *)calloc((size_t)n4, 1);
    e->chroma_nnz[0] = (uint8_t *)calloc((size_t)n * 4, 1);
    e->chroma_nnz[1] = (uint8_t *)calloc((size_t)n * 4, 1);
    e->intra4x4_modes = (uint8_t *)calloc((size_t)n4, 1);
    e->ref_grid = (int8_t *)malloc((size_t)n4);
    memset(e->ref_grid, -2, (size_t)n4);
    e->mv_grid = calloc((size_t)n4, sizeof(*e->mv_grid));
    e->prev_mb_mv = calloc((size_t)n, sizeof(*e->prev_mb_mv));
    e->mb_is_intra = calloc((size_t)n, 1);
    e->mb_is_i16 = calloc((size_t)n, 1);
    e->mb_is_skip = calloc((size_t)n, 1);
    e->mb_cbp_luma = calloc((size_t)n, 1);
    e->mb_cbp_chroma = calloc((size_t)n, 1);
    e->mb_qps = calloc((size_t)n, 1);
    e->mb_8x8 = calloc((size_t)n, 1);
    e->mv_costs = (uint16_t *)malloc(sizeof(uint16_t) * (2 * MV_COST_RANGE + 1));
    e->mv_costs_qp = -1;
    e->dec = calloc((size_t)n, sizeof(ora_mb_decision));
    e->levels = calloc((size_t)n * ORA_LEVELS_PER_MB, sizeof(int16_t));
    e->pre_y = malloc((size_t)W * H);
    e->pre_u = malloc((size_t)W * H / 4);
    e->pre_v = malloc((size_t)W * H / 4);
    e->mvr = calloc((size_t)n, sizeof(*e->mvr));
    e->prev_mvr = calloc((size_t)n, sizeof(*e->prev_mvr));
    e->mv_costs_c = (uint16_t *)malloc(sizeof(uint16_t) * (2 * MV_COST_RANGE + 1));
    e->mv_costs_c_qp = -1;
    return e;
}

/* H264Encoder::with_speed(Compact) (h264.rs:3066-3098): chroma QP offset 1; with CAVLC no B
 * pictures. Returns -1 for the CABAC profiles, whose Compact configuration codes B pictures. */
int ora_set_compact(ora_enc *e, int on) {
    if (on && (e->cabac || e->fast || e->parts)) return -1;
    e->compact = on != 0;
    e->cqo = on ? 1 : 0;
    if (on && !e->hh.buf) {
        int W = e->mb_w * 16, H = e->mb_h * 16;
        plane_alloc(&e->hh, W, H, PAD_Y);
        plane_alloc(&e->hv, W, H, PAD_Y);
        plane_alloc(&e->hc, W, H, PAD_Y);
    }
    return 0;
}

/* H264Encoder::with_speed(Fast) (h264.rs:3032-3065 with enc.rs:2270): the Fast gates of
 * code_p_mb (enc.rs:3050-3152), half_pel_toward (enc.rs:1341), no intra 4x4 and no plane modes
 * (enc.rs:3322, :3457), chroma from the luma mode in P (enc.rs:3605) and no loop filter
 * (enc.rs:2334). Not combinable with Compact. */
int ora_set_fast(ora_enc *e, int on) {
    if (on && (e->compact || e->parts)) return -1;
    e->fast = on != 0;
    return 0;
}

/* The partition extension of the Balanced preset (16x8, 8x16 and 8x8 P partitions; see
 * parts_analyse). Not combinable with Compact or Fast. */
int ora_set_partitions(ora_enc *e, int on) {
    if (on && (e->compact || e->fast)) return -1;
    e->parts = on != 0;
    return 0;
}

void ora_destroy(ora_enc *e) {
    if (!e) return;
    free(e->cur.y.buf);
    free(e->cur.u.buf);
    free(e->cur.v.buf);
    free(e->refp.y.buf);
    free(e->refp.u.buf);
    free(e->refp.v.buf);
    free(e->luma_nnz);
    free(e->chroma_nnz[0]);
    free(e->chroma_nnz[1]);
    free(e->intra4x4_modes);
    free(e->ref_grid);
    free(e->mv_grid);
    free(e->prev_mb_mv);
    free(e->mb_is_intra);
    free(e->mb_is_i16);
    free(e->mb_is_skip);
    free(e->mb_cbp_luma);
    free(e->mb_cbp_chroma);
    free(e->mb_qps);
    free(e->mb_8x8);
    free(e->mv_costs);
    free(e->dec);
    free(e->levels);
    free(e->pre_y);
    free(e->pre_u);
    free(e->pre_v);
    free(e->mvr);
    free(e->prev_mvr);
    free(e->mv_costs_c);
    free(e->hh.buf);
    free(e->hv.buf);
    free(e->hc.buf);
    free(e);
}

void ora_set_me_range(ora_enc *e, int iters) { e->me_iters = iters; }

/* enc.rs:266 mv_costs */
static void ensure_mv_costs(ora_enc *e, int qp) {
    if (e->mv_costs_qp == qp) return;
    uint32_t lambda = T_LAMBDA[qp];
    for (int d = -MV_COST_RANGE; d <= MV_COST_RANGE; d++)
        e->mv_costs[d + MV_COST_RANGE] = (uint16_t)(lambda * se_bits(d));
    e->mv_costs_qp = qp;
}
static uint32_t mvd_cost(const ora_enc *e, uint32_t lambda, int d) { /* enc.rs:1154 */
    if (d >= -MV_COST_RANGE && d <= MV_COST_RANGE) return e->mv_costs[d + MV_COST_RANGE];
    return lambda * se_bits(d);
}

/* h264.rs:3554 fill_grid_4x4 over a whole macroblock */
static void fill_ref(ora_enc *e, int mb_x, int mb_y, int8_t v) {
    int gs = e->mb_w * 4;
    for (int dy = 0; dy < 4; dy++) memset(e->ref_grid + (mb_y * 4 + dy) * gs + mb_x * 4, v, 4);
}
static void fill_mv(ora_enc *e, int mb_x, int mb_y, int16_t mx, int16_t my) {
    int gs = e->mb_w * 4;
    for (int dy = 0; dy < 4; dy++)
        for (int dx = 0; dx < 4; dx++) {
            e->mv_grid[(mb_y * 4 + dy) * gs + mb_x * 4 + dx][0] = mx;
            e->mv_grid[(mb_y * 4 + dy) * gs + mb_x * 4 + dx][1] = my;
        }
}
static void set_intra_motion(ora_enc *e, int mb_x, int mb_y) { /* enc.rs:364 */
    fill_ref(e, mb_x, mb_y, -1);
    fill_mv(e, mb_x, mb_y, 0, 0);
}

static int16_t median3(int16_t a, int16_t b, int16_t c) { /* h264.rs:3570 */
    int16_t mx = a > b ? a : b, mn = a < b ? a : b;
    int16_t t = mn > c ? mn : c;
    return mx < t ? mx : t;
}
/* h264.rs:3577 pred_motion_partition / enc/inter.rs:104 neighbours + :143 mvp_from, for the
 * partition of reference 0 whose top-left 4x4 block is (gx, gy), pw blocks wide, directional
 * rule 0 (16x16 and 8x8 partitions) */
static void pred_mv_part(const ora_enc *e, int gx, int gy, int pw, int16_t mvp[2]) {
    int gs = e->mb_w * 4;
    int lref = -2, tref = -2, dref = -2;
    int16_t a[2] = {0, 0}, b[2] = {0, 0}, c[2] = {0, 0};
    if (gx > 0) {
        int i = gy * gs + gx - 1;
        lref = e->ref_grid[i];
        a[0] = e->mv_grid[i][0];
        a[1] = e->mv_grid[i][1];
    }
    if (gy > 0) {
        int i = (gy - 1) * gs + gx;
        tref = e->ref_grid[i];
        b[0] = e->mv_grid[i][0];
        b[1] = e->mv_grid[i][1];
    }
    int bx = gx & 3, by = gy & 3, trx = gx + pw;
    int tr_avail = gy > 0 && trx < gs &&
                   !((by == 1 && (bx + pw == 2 || bx + pw == 4)) || (by == 2 && bx + pw == 4) ||
                     (by == 3 && (bx + pw == 2 || bx + pw == 4)));
    if (tr_avail && e->ref_grid[(gy - 1) * gs + trx] != -2) {
        int i = (gy - 1) * gs + trx;
        dref = e->ref_grid[i];
        c[0] = e->mv_grid[i][0];
        c[1] = e->mv_grid[i][1];
    } else if (gy > 0 && gx > 0) {
        int i = (gy - 1) * gs + gx - 1;
        dref = e->ref_grid[i];
        c[0] = e->mv_grid[i][0];
        c[1] = e->mv_grid[i][1];
    }
    int match = (lref == 0) + (tref == 0) + (dref == 0);
    if (match > 1) {
        mvp[0] = median3(a[0], b[0], c[0]);
        mvp[1] = median3(a[1], b[1], c[1]);
    } else if (match == 1) {
        const int16_t *s = lref == 0 ? a : (tref == 0 ? b : c);
        mvp[0] = s[0];
        mvp[1] = s[1];
    } else if (tref == -2 && dref == -2 && lref != -2) {
        mvp[0] = a[0];
        mvp[1] = a[1];
    } else {
        mvp[0] = median3(a[0], b[0], c[0]);
        mvp[1] = median3(a[1], b[1], c[1]);
    }
}
static void pred_mv(const ora_enc *e, int gx, int gy, int16_t mvp[2]) { pred_mv_part(e, gx, gy, 4, mvp); }

/* enc.rs:389 MbSrc::load + h264.rs:3660 extract_mb_planes (edge replication) */
typedef struct {
    uint8_t y[256], u[64], v[64];
} mbsrc_t;
typedef struct {
    const uint8_t *y, *u, *v;
    int sy, su, sv, w, h;
} frame_t;
static void load_mb(const frame_t *f, int mb_x, int mb_y, mbsrc_t *s) {
    int x0 = mb_x * 16, y0 = mb_y * 16, cw = (f->w + 1) / 2, ch = (f->h + 1) / 2;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++)
            s->y[r * 16 + c] = f->y[(long)imin(y0 + r, f->h - 1) * f->sy + imin(x0 + c, f->w - 1)];
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            int sy = imin(y0 / 2 + r, ch - 1), sx = imin(x0 / 2 + c, cw - 1);
            s->u[r * 8 + c] = f->u[(long)sy * f->su + sx];
            s->v[r * 8 + c] = f->v[(long)sy * f->sv + sx];
        }
}
static void store_block(plane_t *p, int x0, int y0, const uint8_t *src, int n) {
    for (int r = 0; r < n; r++) memcpy(p->buf + pat(p, x0, y0 + r), src + r * n, (size_t)n);
}
static void load_block(const plane_t *p, long off, uint8_t *dst, int n) {
    for (int r = 0; r < n; r++) memcpy(dst + r * n, p->buf + off + (long)r * p->stride, (size_t)n);
}
static void store_mb(ora_enc *e, int mb_x, int mb_y, const uint8_t *y, const uint8_t *u, const uint8_t *v) {
    store_block(&e->cur.y, mb_x * 16, mb_y * 16, y, 16);
    store_block(&e->cur.u, mb_x * 8, mb_y * 8, u, 8);
    store_block(&e->cur.v, mb_x * 8, mb_y * 8, v, 8);
}

/* enc.rs:1056 mc_luma */
static void mc_luma_mb(const plane_t *r, int x0, int y0, const int16_t mv[2], uint8_t *dst) {
    long off = pat(r, x0 + (mv[0] >> 2), y0 + (mv[1] >> 2));
    if (((mv[0] | mv[1]) & 3) == 0) {
        load_block(r, off, dst, 16);
        return;
    }
    luma_mc(dst, 16, r->buf + off, r->stride, 16, 16, mv[0] & 3, mv[1] & 3);
}
/* enc.rs:1080 mc_chroma */
static void mc_chroma_mb(const pic_t *p, int cx0, int cy0, const int16_t mv[2], uint8_t *u, uint8_t *v) {
    long off = pat(&p->u, cx0 + (mv[0] >> 3), cy0 + (mv[1] >> 3));
    int fx = mv[0] & 7, fy = mv[1] & 7;
    if ((fx | fy) == 0) {
        load_block(&p->u, off, u, 8);
        load_block(&p->v, off, v, 8);
        return;
    }
    chroma_mc8(u, p->u.buf + off, p->u.stride, fx, fy);
    chroma_mc8(v, p->v.buf + off, p->v.stride, fx, fy);
}

/* enc.rs:1784 zero_sad_bound, :2122 zero_ssd_bound, :2131 SkipBounds::new */
static uint32_t zero_sad_bound(const quant4_t *q) {
    static const uint32_t W[4] = {1, 2, 1, 2};
    uint32_t room = (1u << q->shift) - q->bias - 1, m = UINT32_MAX;
    for (int i = 0; i < 16; i++) {
        uint32_t v = room / (q->mf[i] * W[i >> 2] * W[i & 3]);
        if (v < m) m = v;
    }
    return m;
}
static uint32_t zero_ssd_bound(const quant4_t *q) {
    static const uint32_t N[4] = {4, 10, 4, 10};
    uint32_t room = (1u << q->shift) - q->bias - 1, m = UINT32_MAX;
    for (int i = 0; i < 16; i++) {
        uint32_t lim = room / q->mf[i];
        uint32_t v = lim * lim / (N[i >> 2] * N[i & 3]);
        if (v < m) m = v;
    }
    return m;
}
static void skip_bounds_init(skip_bounds_t *s, int qp, int qpc, int fine) {
    const quant4_t *q = quant4(qp, fine), *qc = quant4(qpc, fine);
    uint32_t dc_shift = 16 + (uint32_t)qpc / 6;
    uint32_t dc_room = (1u << dc_shift) - (1u << dc_shift) / 3 - 1;
    uint32_t dc = dc_room / (uint32_t)T_QUANT4_MF[qpc % 6][0];
    s->qp = qp;
    s->luma = zero_sad_bound(q);
    s->luma_ssd = zero_ssd_bound(q);
    uint32_t c = zero_sad_bound(qc);
    s->chroma = c < dc ? c : dc;
    uint32_t cs = zero_ssd_bound(qc), dd = dc * dc / 64;
    s->chroma_ssd = cs < dd ? cs : dd;
    s->probe = 16 * (((1u << q->shift) - q->bias) / q->mf[0] + 1);
}

/* enc.rs:1790 dct_luma: sixteen 4x4 blocks in decoding order */
static void dct_luma(const uint8_t *src, const uint8_t *pred, int16_t out[16][16]) {
    for (int i8 = 0; i8 < 4; i8++) {
        int o = (i8 >> 1) * 128 + (i8 & 1) * 8;
        sub_dct4_8x8(src + o, 16, pred + o, 16, &out[i8 * 4]);
    }
}

/* enc.rs:2163 probe_skip_luma */
static int probe_skip_luma(const mbsrc_t *s, const uint8_t *py, int qp, int fine) {
    int16_t c[16][16];
    dct_luma(s->y, py, c);
    const quant4_t *q = quant4(qp, fine);
    uint32_t total = 0;
    for (int i8 = 0; i8 < 4; i8++) {
        uint32_t nz = quant_4x4_x4(&c[i8 * 4], q);
        if (!nz) continue;
        if (nz >> 4) return 0;
        for (int k = 0; k < 4; k++) {
            int16_t zz[16];
            zigzag4(c[i8 * 4 + k], zz);
            total += decimate_masked(zz, nz_mask(zz, 16));
            if (total >= 6) return 0;
        }
    }
    return 1;
}

/* enc.rs:1939 quant_chroma_dc (intra-style f = 2^qbits / 3) */
static void quant_chroma_dc(const int16_t *dc, int qpc, int16_t *out) {
    int d0 = dc[0], d1 = dc[1], d2 = dc[2], d3 = dc[3];
    int had[4] = {d0 + d1 + d2 + d3, d0 - d1 + d2 - d3, d0 + d1 - d2 - d3, d0 - d1 - d2 + d3};
    qpc = imin(qpc, 51);
    int qbits = 16 + qpc / 6, mf = T_QUANT4_MF[qpc % 6][0], f = (1 << qbits) / 3;
    for (int i = 0; i < 4; i++) {
        int l = (iabs(had[i]) * mf + f) >> qbits;
        out[i] = (int16_t)(had[i] < 0 ? -l : l);
    }
}
/* enc.rs:1897 quant_chroma */
static void quant_chroma(const mbsrc_t *s, const uint8_t *pu, const uint8_t *pv, int qpc, int intra,
                         chroma_levels_t *out) {
    const quant4_t *q = quant4(qpc, intra);
    for (int p = 0; p < 2; p++) {
        const uint8_t *src = p ? s->v : s->u, *pred = p ? pv : pu;
        int16_t c[4][16];
        sub_dct4_8x8(src, 8, pred, 8, c);
        int16_t dc_raw[4] = {c[0][0], c[1][0], c[2][0], c[3][0]};
        if (quant_4x4_x4(c, q)) {
            for (int b = 0; b < 4; b++) {
                int16_t zz[16];
                zigzag4(c[b], zz);
                out->ac_masks[p][b] = nz_mask(zz, 16) >> 1;
                memcpy(out->ac[p][b], zz + 1, 15 * sizeof(int16_t));
                memcpy(out->raster[p][b], c[b], 16 * sizeof(int16_t));
            }
        } else {
            memset(out->ac_masks[p], 0, sizeof(out->ac_masks[p]));
            memset(out->ac[p], 0, sizeof(out->ac[p]));
            memset(out->raster[p], 0, sizeof(out->raster[p]));
        }
        quant_chroma_dc(dc_raw, qpc, out->dc[p]);
    }
}
static int chroma_cbp(const chroma_levels_t *c) { /* enc.rs:1867 */
    uint32_t m = 0;
    int d = 0;
    for (int p = 0; p < 2; p++)
        for (int b = 0; b < 4; b++) {
            m |= c->ac_masks[p][b];
            d |= c->dc[p][b];
        }
    return m ? 2 : (d ? 1 : 0);
}
static uint32_t chroma_ac_score(const chroma_levels_t *c, int p) {
    uint32_t s = 0;
    for (int b = 0; b < 4; b++) s += decimate_masked(c->ac[p][b], c->ac_masks[p][b]);
    return s;
}
/* enc.rs:2186 probe_skip_chroma */
static int probe_skip_chroma(const mbsrc_t *s, const uint8_t *pu, const uint8_t *pv, int qpc, int fine) {
    chroma_levels_t c;
    quant_chroma(s, pu, pv, qpc, fine, &c);
    for (int p = 0; p < 2; p++)
        for (int b = 0; b < 4; b++)
            if (c.dc[p][b]) return 0;
    return chroma_ac_score(&c, 0) < 7 && chroma_ac_score(&c, 1) < 7;
}
/* enc.rs:2084 quant_chroma_inter */
static void quant_chroma_inter(const mbsrc_t *s, const uint8_t *pu, const uint8_t *pv, int qpc,
                               chroma_levels_t *c, int fine) {
    quant_chroma(s, pu, pv, qpc, fine, c);
    for (int p = 0; p < 2; p++) {
        uint32_t any = c->ac_masks[p][0] | c->ac_masks[p][1] | c->ac_masks[p][2] | c->ac_masks[p][3];
        if (any && chroma_ac_score(c, p) < 7) {
            memset(c->ac[p], 0, sizeof(c->ac[p]));
            memset(c->ac_masks[p], 0, sizeof(c->ac_masks[p]));
            memset(c->raster[p], 0, sizeof(c->raster[p]));
        }
    }
}
/* enc.rs:1810 quant_luma_4x4_inter */
static int quant_luma_4x4_inter(const uint8_t *src, const uint8_t *pred, int qp, int16_t q4[16][16], int fine) {
    dct_luma(src, pred, q4);
    const quant4_t *q = quant4(qp, fine);
    int cbp = 0;
    uint32_t total = 0;
    for (int i8 = 0; i8 < 4; i8++) {
        int16_t (*g)[16] = &q4[i8 * 4];
        uint32_t nz = quant_4x4_x4(g, q);
        if (!nz) continue;
        if (nz >> 4) {
            cbp |= 1 << i8;
            total += 9;
            continue;
        }
        uint32_t score = 0;
        for (int k = 0; k < 4; k++) {
            int16_t zz[16];
            zigzag4(g[k], zz);
            score += decimate_masked(zz, nz_mask(zz, 16));
        }
        if (score < 4)
            memset(g, 0, sizeof(int16_t) * 64);
        else
            cbp |= 1 << i8;
        total += score;
    }
    if (total < 6 && cbp) {
        memset(q4, 0, sizeof(int16_t) * 256);
        cbp = 0;
    }
    return cbp;
}
/* enc.rs:2041 quant_luma_8x8_inter (x264 quant_8x8 + sub8x8_dct) */
static int quant_luma_8x8_inter(const uint8_t *src, const uint8_t *pred, int qp, int16_t q8[4][64]) {
    qp = imin(qp, 51);
    uint32_t shift = 16 + (uint32_t)(qp / 6), bias = (1u << shift) / 6;
    const uint16_t *mf = QUANT8_MF16[qp % 6];
    int cbp = 0;
    uint32_t total = 0;
    for (int b = 0; b < 4; b++) {
        int o = (b >> 1) * 128 + (b & 1) * 8;
        int16_t d[64];
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) d[y * 8 + x] = (int16_t)(src[o + y * 16 + x] - pred[o + y * 16 + x]);
        fdct8(d, q8[b]);
        int any = 0, big = 0;
        for (int i = 0; i < 64; i++) { /* simd/h264_dct8.rs:84 quant_8x8_scalar */
            int c = q8[b][i];
            uint64_t lv = ((uint64_t)iabs(c) * mf[i] + bias) >> shift;
            if (lv > 32767) lv = 32767;
            any |= lv != 0;
            big |= lv > 1;
            q8[b][i] = (int16_t)(c < 0 ? -(int)lv : (int)lv);
        }
        uint32_t score = 0;
        if (big)
            score = 9;
        else if (any) {
            int16_t zz[64];
            for (int i = 0; i < 64; i++) zz[i] = q8[b][T_ZIGZAG8[i]];
            score = decimate_score(zz, 64);
        }
        total += score;
        if (score >= 4)
            cbp |= 1 << b;
        else
            memset(q8[b], 0, sizeof(q8[b]));
    }
    if (total < 6 && cbp) {
        memset(q8, 0, sizeof(int16_t) * 256);
        cbp = 0;
    }
    return cbp;
}

/* h264.rs:1645 dequantize_idct_luma_dc4x4 */
static void dequant_idct_luma_dc(const int16_t *lv, int qp, int16_t *out) {
    qp = imin(qp, 51);
    int qmul = (T_DEQUANT4_INIT[qp % 6][0] * 16) << (qp / 6 + 2);
    int t[16];
    for (int i = 0; i < 4; i++) {
        int z0 = lv[4 * i] + lv[4 * i + 1], z1 = lv[4 * i] - lv[4 * i + 1];
        int z2 = lv[4 * i + 2] - lv[4 * i + 3], z3 = lv[4 * i + 2] + lv[4 * i + 3];
        t[4 * i] = z0 + z3;
        t[4 * i + 1] = z0 - z3;
        t[4 * i + 2] = z1 - z2;
        t[4 * i + 3] = z1 + z2;
    }
    for (int i = 0; i < 4; i++) {
        int z0 = t[i] + t[8 + i], z1 = t[i] - t[8 + i], z2 = t[4 + i] - t[12 + i], z3 = t[4 + i] + t[12 + i];
        out[i] = (int16_t)(((z0 + z3) * qmul + 128) >> 8);
        out[4 + i] = (int16_t)(((z1 + z2) * qmul + 128) >> 8);
        out[8 + i] = (int16_t)(((z1 - z2) * qmul + 128) >> 8);
        out[12 + i] = (int16_t)(((z0 - z3) * qmul + 128) >> 8);
    }
}
/* h264.rs:1674 dequantize_idct_chroma_dc2x2 */
static void dequant_idct_chroma_dc(const int16_t *lv, int qpc, int16_t *out) {
    qpc = imin(qpc, 51);
    int qmul = (T_DEQUANT4_INIT[qpc % 6][0] * 16) << (qpc / 6 + 2);
    int a = lv[0], b = lv[1], c = lv[2], d = lv[3];
    out[0] = (int16_t)(((a + b + c + d) * qmul) >> 7);
    out[1] = (int16_t)(((a - b + c - d) * qmul) >> 7);
    out[2] = (int16_t)(((a + b - c - d) * qmul) >> 7);
    out[3] = (int16_t)(((a - b - c + d) * qmul) >> 7);
}
/* enc.rs:1976 recon_luma_4x4 */
static void recon_luma_4x4(uint8_t *pred, int16_t q4[16][16], int cbp, int qp) {
    const int32_t *scale = DEQUANT4[imin(qp, 51)];
    for (int i8 = 0; i8 < 4; i8++) {
        if (!(cbp & (1 << i8))) continue;
        int16_t g[4][16];
        memcpy(g, q4[i8 * 4], sizeof(g));
        for (int k = 0; k < 4; k++) dequant_4x4(g[k], scale);
        add_idct4_8x8(pred + (i8 >> 1) * 128 + (i8 & 1) * 8, 16, g);
    }
}
/* h264.rs:1403 dequantize_block8x8 (flat scaling lists: LevelScale8x8 = 16 x normAdjust8x8) */
static void dequant8(const int16_t *q8, int qp, int32_t *d) {
    qp = imin(qp, 51);
    int qd = qp / 6;
    const int32_t *scale = DEQUANT8_POS[qp % 6];
    for (int i = 0; i < 64; i++) {
        int v;
        if (qd >= 2)
            v = q8[i] * (scale[i] << (qd - 2));
        else {
            int rs = 2 - qd;
            v = (q8[i] * scale[i] + (1 << (rs - 1))) >> rs;
        }
        d[i] = iclamp(v, -32768, 32767);
    }
}
/* enc.rs:2017 recon_8x8 */
static void recon_8x8(uint8_t *dst, const int16_t *q8, int qp) {
    int32_t d[64];
    dequant8(q8, qp, d);
    idct8_add(dst, 16, d);
}
/* enc.rs:1990 recon_chroma */
static void recon_chroma(uint8_t *pu, uint8_t *pv, const chroma_levels_t *c, int cbpc, int qpc) {
    if (!cbpc) return;
    const int32_t *scale = DEQUANT4[imin(qpc, 51)];
    for (int p = 0; p < 2; p++) {
        int16_t dcr[4], blocks[4][16];
        dequant_idct_chroma_dc(c->dc[p], qpc, dcr);
        memset(blocks, 0, sizeof(blocks));
        if (cbpc >= 2) {
            memcpy(blocks, c->raster[p], sizeof(blocks));
            for (int b = 0; b < 4; b++) dequant_4x4(blocks[b], scale);
        }
        for (int b = 0; b < 4; b++) blocks[b][0] = dcr[b];
        add_idct4_8x8(p ? pv : pu, 8, blocks);
    }
}

/* enc.rs:432 blk_xy: column and row of 4x4 block b (decoding order) */
static void blk_xy(int b, int *x, int *y) {
    *x = (b & 1) + ((b >> 1) & 2);
    *y = ((b >> 1) & 1) + ((b >> 2) & 2);
}
static const int Z_TO_RASTER[16] = {0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15};

/* ------------------------------------------------------------------------------------------ */
/* CAVLC macroblock syntax (enc.rs:779-1054, 2779-3306).                                       */
/* ------------------------------------------------------------------------------------------ */
/* enc.rs:793 NnzCache: nC of block b from the left / top counts, 0x80 = unavailable */
typedef struct {
    uint8_t c[32];
} nnz_cache_t;
static int NNZ_AT(int b) {
    int x, y;
    blk_xy(b, &x, &y);
    return (y + 1) * 5 + x + 1;
}
static void nnz_load(const ora_enc *e, int mb_x, int mb_y, nnz_cache_t *c) {
    int gs = e->mb_w * 4, gx = mb_x * 4, gy = mb_y * 4;
    memset(c->c, 0x80, 32);
    if (mb_y > 0) memcpy(c->c + 1, e->luma_nnz + (gy - 1) * gs + gx, 4);
    if (mb_x > 0)
        for (int y = 0; y < 4; y++) c->c[(y + 1) * 5] = e->luma_nnz[(gy + y) * gs + gx - 1];
}
static int nnz_nc(const nnz_cache_t *c, int b) {
    int i = NNZ_AT(b);
    uint32_t s = (uint32_t)c->c[i - 1] + c->c[i - 5];
    return (int)((s < 0x80 ? (s + 1) >> 1 : s) & 0x7F);
}
static void nnz_store(ora_enc *e, int mb_x, int mb_y, const nnz_cache_t *c) {
    int gs = e->mb_w * 4;
    for (int y = 0; y < 4; y++) memcpy(e->luma_nnz + (mb_y * 4 + y) * gs + mb_x * 4, c->c + (y + 1) * 5 + 1, 4);
}
static int predict_chroma_nc(const uint8_t *nnz, int cgs, int cgx, int cgy) { /* h264.rs:2119 */
    int hl = cgx > 0, ht = cgy > 0;
    int l = hl ? nnz[cgy * cgs + cgx - 1] : 0, t = ht ? nnz[(cgy - 1) * cgs + cgx] : 0;
    if (hl && ht) return (l + t + 1) >> 1;
    if (hl) return l;
    if (ht) return t;
    return 0;
}
/* enc.rs:1009 write_cavlc_chroma */
static void write_cavlc_chroma(ora_enc *e, bits_t *w, int mb_x, int mb_y, int cbpc, const chroma_levels_t *c) {
    if (cbpc >= 1)
        for (int p = 0; p < 2; p++) cavlc_block(w, c->dc[p], 4, 0, 1);
    if (cbpc >= 2) {
        int cgs = e->mb_w * 2;
        for (int p = 0; p < 2; p++)
            for (int b = 0; b < 4; b++) {
                int cgx = mb_x * 2 + (b & 1), cgy = mb_y * 2 + (b >> 1);
                int nc = predict_chroma_nc(e->chroma_nnz[p], cgs, cgx, cgy);
                cavlc_block(w, c->ac[p][b], 15, nc, 0);
                e->chroma_nnz[p][cgy * cgs + cgx] = (uint8_t)pop32(c->ac_masks[p][b]);
            }
    }
}
/* enc.rs:976 write_cavlc_residuals */
static void write_cavlc_residuals(ora_enc *e, bits_t *w, int mb_x, int mb_y, int cbpl, int cbpc,
                                  int16_t q4[16][16], const chroma_levels_t *c) {
    nnz_cache_t cache;
    nnz_load(e, mb_x, mb_y, &cache);
    for (int i8 = 0; i8 < 4; i8++) {
        if (!(cbpl & (1 << i8))) {
            for (int k = 0; k < 4; k++) cache.c[NNZ_AT(i8 * 4 + k)] = 0;
            continue;
        }
        for (int k = 0; k < 4; k++) {
            int b = i8 * 4 + k;
            int16_t zz[16];
            zigzag4(q4[b], zz);
            uint32_t m = nz_mask(zz, 16);
            if (w) cavlc_block(w, zz, 16, nnz_nc(&cache, b), 0);
            cache.c[NNZ_AT(b)] = (uint8_t)pop32(m);
        }
    }
    nnz_store(e, mb_x, mb_y, &cache);
    if (w) write_cavlc_chroma(e, w, mb_x, mb_y, cbpc, c);
}
static uint32_t inter_cbp_golomb(int cbp) {
    for (int i = 0; i < 48; i++)
        if (T_GOLOMB_TO_INTER_CBP[i] == cbp) return (uint32_t)i;
    return 0;
}
static uint32_t intra4_cbp_golomb(int cbp) {
    for (int i = 0; i < 48; i++)
        if (T_GOLOMB_TO_INTRA4X4_CBP[i] == cbp) return (uint32_t)i;
    return 3;
}
/* The CABAC configurations record non-zero counts like encode_cabac_cbp_and_residuals
 * (h264.rs:3418-3470): a coded 8x8 block marks its four 4x4 entries 1. */
static void cabac_luma_nnz(ora_enc *e, int mb_x, int mb_y, int cbpl, int cbpc, int16_t q4[16][16],
                           int16_t q8[4][64], int use8) {
    if (!cbpl && !cbpc) return;
    int gs = e->mb_w * 4;
    if (use8) {
        for (int i8 = 0; i8 < 4; i8++) {
            int v = (cbpl >> i8) & 1;
            int bx = (i8 & 1) * 2, by = (i8 >> 1) * 2;
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) e->luma_nnz[(mb_y * 4 + by + dy) * gs + mb_x * 4 + bx + dx] = (uint8_t)v;
        }
    } else {
        for (int b = 0; b < 16; b++) {
            int bx, by;
            blk_xy(b, &bx, &by);
            int n = 0;
            if (cbpl & (1 << (b >> 2)))
                for (int i = 0; i < 16; i++) n += q4[b][i] != 0;
            e->luma_nnz[(mb_y * 4 + by) * gs + mb_x * 4 + bx] = (uint8_t)n;
        }
    }
    (void)q8;
}

/* Level record of a macroblock (layout of ORA_LV_*). */
static void rec_luma4(int16_t *lv, int16_t q4[16][16]) {
    for (int b = 0; b < 16; b++) zigzag4(q4[b], lv + ORA_LV_LUMA + b * 16);
}
static void rec_chroma(int16_t *lv, const chroma_levels_t *c) {
    for (int p = 0; p < 2; p++) {
        memcpy(lv + ORA_LV_CHROMA_DC + p * 4, c->dc[p], 4 * sizeof(int16_t));
        for (int b = 0; b < 4; b++) {
            int16_t *o = lv + ORA_LV_CHROMA_AC + (p * 4 + b) * 16;
            o[0] = 0;
            memcpy(o + 1, c->ac[p][b], 15 * sizeof(int16_t));
        }
    }
}

/* enc.rs:2781 begin_coded_mb (CAVLC) */
static void begin_coded_mb(ora_enc *e, bits_t *w) {
    if (e->is_idr) return;
    if (w) put_ue(w, (uint64_t)e->skip_run);
    e->skip_run = 0;
}
/* enc.rs:2874 code_skip */
static void code_skip(ora_enc *e, int mb_x, int mb_y, const int16_t mv[2]) {
    int idx = mb_y * e->mb_w + mb_x;
    e->skip_run++;
    fill_ref(e, mb_x, mb_y, 0);
    fill_mv(e, mb_x, mb_y, mv[0], mv[1]);
    e->mb_is_skip[idx] = 1;
    e->mb_qps[idx] = (uint8_t)e->cur_qp;
}

/* enc.rs:1457 i16_edges / :1476 chroma_edges from the (undeblocked) current picture */
typedef struct {
    uint8_t top[16], left[16];
    int has_top, has_left, tl;
} edges_t;
static void get_edges(const plane_t *p, int mb_x, int mb_y, int n, edges_t *e) {
    int x0 = mb_x * n, y0 = mb_y * n;
    e->has_top = mb_y > 0;
    e->has_left = mb_x > 0;
    if (e->has_top) memcpy(e->top, p->buf + pat(p, x0, y0 - 1), (size_t)n);
    if (e->has_left)
        for (int r = 0; r < n; r++) e->left[r] = p->buf[pat(p, x0 - 1, y0 + r)];
    e->tl = (mb_x > 0 && mb_y > 0) ? p->buf[pat(p, x0 - 1, y0 - 1)] : 128;
}
/* simd/h264_enc.rs:628 intra_satd_x3_16x16_scalar: V, H, DC SATDs from one source Hadamard */
static void intra_satd_x3_16(const uint8_t *src, const uint8_t *top, const uint8_t *left, int dc, uint32_t out[3]) {
    int tv[4][4], th[4][4];
    for (int b = 0; b < 4; b++) {
        const uint8_t *x = top + b * 4, *y = left + b * 4;
        int s01 = x[0] + x[1], d01 = x[0] - x[1], s23 = x[2] + x[3], d23 = x[2] - x[3];
        tv[b][0] = 4 * (s01 + s23);
        tv[b][1] = 4 * (s01 - s23);
        tv[b][2] = 4 * (d01 + d23);
        tv[b][3] = 4 * (d01 - d23);
        s01 = y[0] + y[1];
        d01 = y[0] - y[1];
        s23 = y[2] + y[3];
        d23 = y[2] - y[3];
        th[b][0] = 4 * (s01 + s23);
        th[b][1] = 4 * (s01 - s23);
        th[b][2] = 4 * (d01 + d23);
        th[b][3] = 4 * (d01 - d23);
    }
    uint32_t sum[3] = {0, 0, 0};
    for (int by = 0; by < 4; by++)
        for (int bx = 0; bx < 4; bx++) {
            int rows[4][4], o[16];
            for (int y = 0; y < 4; y++) {
                const uint8_t *x = src + (by * 4 + y) * 16 + bx * 4;
                int s01 = x[0] + x[1], d01 = x[0] - x[1], s23 = x[2] + x[3], d23 = x[2] - x[3];
                rows[y][0] = s01 + s23;
                rows[y][1] = s01 - s23;
                rows[y][2] = d01 + d23;
                rows[y][3] = d01 - d23;
            }
            for (int h = 0; h < 4; h++) {
                int s01 = rows[0][h] + rows[1][h], d01 = rows[0][h] - rows[1][h];
                int s23 = rows[2][h] + rows[3][h], d23 = rows[2][h] - rows[3][h];
                o[0 * 4 + h] = s01 + s23;
                o[1 * 4 + h] = s01 - s23;
                o[2 * 4 + h] = d01 + d23;
                o[3 * 4 + h] = d01 - d23;
            }
            for (int k = 0; k < 16; k++) {
                int v = k / 4, h = k % 4, c = (int16_t)o[k];
                int pv = v == 0 ? tv[bx][h] : 0, ph = h == 0 ? th[by][v] : 0, pd = k == 0 ? 16 * dc : 0;
                sum[0] += (uint32_t)iabs(c - pv);
                sum[1] += (uint32_t)iabs(c - ph);
                sum[2] += (uint32_t)iabs(c - pd);
            }
        }
    for (int i = 0; i < 3; i++) out[i] = sum[i] >> 1;
}
static int i16_dc(const edges_t *ed) { /* enc.rs:1624 */
    uint32_t st = 0, sl = 0;
    for (int i = 0; i < 16; i++) {
        st += ed->top[i];
        sl += ed->left[i];
    }
    if (ed->has_top && ed->has_left) return (int)((st + sl + 16) >> 5);
    if (ed->has_top) return (int)((st + 8) >> 4);
    if (ed->has_left) return (int)((sl + 8) >> 4);
    return 128;
}
/* enc.rs:1637 analyse_i16_edges (the plane mode is tried unless `plane` is 0: Fast) */
static void analyse_i16(const edges_t *ed, const uint8_t *src, uint32_t *satd_out, int *mode_out, uint16_t *n_satd,
                        int plane) {
    static const uint8_t zero[16] = {0};
    int dc = i16_dc(ed);
    uint32_t vhd[3];
    intra_satd_x3_16(src, ed->has_top ? ed->top : zero, ed->has_left ? ed->left : zero, dc, vhd);
    uint32_t best = vhd[2];
    int mode = 2;
    if (ed->has_top && vhd[0] < best) {
        best = vhd[0];
        mode = 0;
    }
    if (ed->has_left && vhd[1] < best) {
        best = vhd[1];
        mode = 1;
    }
    if (plane && ed->has_top && ed->has_left) {
        uint8_t pred[256];
        pred16(3, ed->top, ed->left, ed->tl, pred);
        uint32_t c = satd16(src, 16, pred, 16);
        if (n_satd) (*n_satd)++;
        if (c < best) {
            best = c;
            mode = 3;
        }
    }
    *satd_out = best;
    *mode_out = mode;
}
/* enc.rs:1523 analyse_chroma_intra (the plane mode only when `plane`) */
static int analyse_chroma_intra(const ora_enc *e, int mb_x, int mb_y, const mbsrc_t *s, uint8_t *pu, uint8_t *pv,
                                int plane) {
    edges_t eu, ev;
    get_edges(&e->cur.u, mb_x, mb_y, 8, &eu);
    get_edges(&e->cur.v, mb_x, mb_y, 8, &ev);
    uint32_t best = UINT32_MAX;
    int bm = 0;
    for (int mode = 0; mode < 4; mode++) {
        int ok = mode == 1 ? eu.has_left : mode == 2 ? eu.has_top : mode == 3 ? (plane && eu.has_top && eu.has_left) : 1;
        if (!ok) continue;
        uint8_t a[64], b[64];
        pred_chroma(mode, eu.has_top ? eu.top : NULL, eu.has_left ? eu.left : NULL, eu.tl, a);
        pred_chroma(mode, ev.has_top ? ev.top : NULL, ev.has_left ? ev.left : NULL, ev.tl, b);
        uint32_t cost = sad_wh(s->u, 8, a, 8, 8, 8) + sad_wh(s->v, 8, b, 8, 8, 8) + (mode ? 8 : 0);
        if (cost < best) {
            best = cost;
            bm = mode;
            memcpy(pu, a, 64);
            memcpy(pv, b, 64);
        }
    }
    return bm;
}

/* enc.rs:3587 code_i16 */
static void code_i16(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const mbsrc_t *s, int mode) {
    int idx = mb_y * e->mb_w + mb_x;
    ora_mb_decision *d = &e->dec[idx];
    int16_t *lv = e->levels + (long)idx * ORA_LEVELS_PER_MB;
    edges_t ed;
    get_edges(&e->cur.y, mb_x, mb_y, 16, &ed);
    uint8_t pred[256], pu[64], pv[64];
    pred16(mode, ed.has_top ? ed.top : NULL, ed.has_left ? ed.left : NULL, ed.tl, pred);
    int cmode;
    if (e->fast && !e->is_idr) { /* enc.rs:3605: chroma predicted in the direction of the luma mode */
        static const int from_luma[4] = {2, 1, 0, 3};
        cmode = from_luma[mode & 3];
        edges_t eu, ev;
        get_edges(&e->cur.u, mb_x, mb_y, 8, &eu);
        get_edges(&e->cur.v, mb_x, mb_y, 8, &ev);
        pred_chroma(cmode, eu.has_top ? eu.top : NULL, eu.has_left ? eu.left : NULL, eu.tl, pu);
        pred_chroma(cmode, ev.has_top ? ev.top : NULL, ev.has_left ? ev.left : NULL, ev.tl, pv);
    } else {
        cmode = analyse_chroma_intra(e, mb_x, mb_y, s, pu, pv, !e->fast);
    }
    e->mb_is_intra[idx] = 1;
    e->mb_is_i16[idx] = 1;
    set_intra_motion(e, mb_x, mb_y);
    int qp_delta = qp - e->cur_qp, eff_qp = iclamp(e->cur_qp + qp_delta, 0, 51), eff_qpc = chroma_qp(eff_qp);
    int16_t ac[16][16], dc_raw[16], ac_zz[16][16];
    uint32_t ac_masks[16];
    memset(ac_zz, 0, sizeof(ac_zz));
    memset(ac_masks, 0, sizeof(ac_masks));
    dct_luma(s->y, pred, ac);
    for (int b = 0; b < 16; b++) dc_raw[Z_TO_RASTER[b]] = ac[b][0];
    const quant4_t *q = quant4(eff_qp, 1);
    for (int i8 = 0; i8 < 4; i8++) {
        if (quant_4x4_x4(&ac[i8 * 4], q)) {
            for (int k = 0; k < 4; k++) ac[i8 * 4 + k][0] = 0;
            for (int k = 0; k < 4; k++) {
                zigzag4(ac[i8 * 4 + k], ac_zz[i8 * 4 + k]);
                ac_masks[i8 * 4 + k] = nz_mask(ac_zz[i8 * 4 + k], 16) >> 1;
            }
        }
    }
    uint32_t any_ac = 0;
    for (int b = 0; b < 16; b++) any_ac |= ac_masks[b];
    /* h264.rs:1245 quantize_luma_dc via enc.rs:1959 quant_luma_dc */
    int16_t had[16], dc_q[16], dc_sc[16];
    hadamard4_dc(dc_raw, had);
    {
        int qbits = 16 + eff_qp / 6, mf = T_QUANT4_MF[eff_qp % 6][0], f = (1 << qbits) / 3;
        for (int i = 0; i < 16; i++) {
            int l = (iabs(had[i]) * mf + f) >> qbits;
            dc_q[i] = (int16_t)(had[i] < 0 ? -l : l);
        }
    }
    zigzag4(dc_q, dc_sc);
    int cbpl = any_ac ? 15 : 0;
    chroma_levels_t c;
    quant_chroma(s, pu, pv, eff_qpc, 1, &c);
    int cbpc = chroma_cbp(&c);
    int mb_type = 1 + mode + 4 * cbpc + (cbpl ? 12 : 0);
    e->mb_cbp_luma[idx] = (uint8_t)cbpl;
    e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    if (w) {
        put_ue(w, (uint64_t)((e->is_idr ? 0 : 5) + mb_type));
        put_ue(w, (uint64_t)cmode);
        put_se(w, qp_delta);
    }
    if (e->cabac) {
        /* enc.rs:3679-3706: the AC counts are recorded only when cbp_luma is set */
        if (cbpl) {
            int gs = e->mb_w * 4;
            for (int b = 0; b < 16; b++) {
                int bx, by;
                blk_xy(b, &bx, &by);
                e->luma_nnz[(mb_y * 4 + by) * gs + mb_x * 4 + bx] = (uint8_t)pop32(ac_masks[b]);
            }
        }
    } else {
        nnz_cache_t cache;
        nnz_load(e, mb_x, mb_y, &cache);
        cavlc_block(w, dc_sc, 16, nnz_nc(&cache, 0), 0);
        if (cbpl) {
            for (int b = 0; b < 16; b++) {
                cavlc_block(w, ac_zz[b] + 1, 15, nnz_nc(&cache, b), 0);
                cache.c[NNZ_AT(b)] = (uint8_t)pop32(ac_masks[b]);
            }
            nnz_store(e, mb_x, mb_y, &cache);
        }
        write_cavlc_chroma(e, w, mb_x, mb_y, cbpc, &c);
    }
    e->cur_qp = eff_qp;
    e->mb_qps[idx] = (uint8_t)eff_qp;
    /* levels record */
    memset(lv, 0, ORA_LEVELS_PER_MB * sizeof(int16_t));
    for (int b = 0; b < 16; b++) memcpy(lv + ORA_LV_LUMA + b * 16, ac_zz[b], 16 * sizeof(int16_t));
    memcpy(lv + ORA_LV_LUMA_DC, dc_sc, 16 * sizeof(int16_t));
    rec_chroma(lv, &c);
    d->mb_type = 2;
    d->i16_mode = (uint8_t)mode;
    d->chroma_mode = (uint8_t)cmode;
    d->cbp_luma = (uint8_t)cbpl;
    d->cbp_chroma = (uint8_t)cbpc;
    d->qp = (uint8_t)eff_qp;
    /* reconstruction */
    int16_t dcr[16];
    dequant_idct_luma_dc(dc_q, eff_qp, dcr);
    const int32_t *scale = DEQUANT4[eff_qp];
    for (int i8 = 0; i8 < 4; i8++) {
        int16_t (*g)[16] = &ac[i8 * 4];
        if (cbpl)
            for (int k = 0; k < 4; k++) dequant_4x4(g[k], scale);
        int any = 0;
        for (int k = 0; k < 4; k++) {
            g[k][0] = dcr[Z_TO_RASTER[i8 * 4 + k]];
            for (int i = 0; i < 16; i++) any |= g[k][i];
        }
        if (any) add_idct4_8x8(pred + (i8 >> 1) * 128 + (i8 & 1) * 8, 16, g);
    }
    recon_chroma(pu, pv, &c, cbpc, eff_qpc);
    store_mb(e, mb_x, mb_y, pred, pu, pv);
}

/* enc.rs:1183 i4_edges + h264.rs:3283 pred_intra4x4_mode */
static int pred_i4_mode(const ora_enc *e, int gx, int gy) {
    if (gx == 0 || gy == 0) return 2;
    int gs = e->mb_w * 4;
    int a = e->intra4x4_modes[gy * gs + gx - 1], b = e->intra4x4_modes[(gy - 1) * gs + gx];
    return a < b ? a : b;
}
/* enc.rs:1700 intra_satd_x3_4x4 from the source block's Hadamard */
static void intra_satd_x3_4(const int16_t *had, const uint8_t *top, const uint8_t *left, uint32_t out[3]) {
    int total = 0;
    for (int i = 0; i < 16; i++) total += iabs(had[i]);
    int dc;
    int st = top ? top[0] + top[1] + top[2] + top[3] : 0, sl = left ? left[0] + left[1] + left[2] + left[3] : 0;
    if (top && left)
        dc = (st + sl + 4) >> 3;
    else if (top)
        dc = (st + 2) >> 2;
    else if (left)
        dc = (sl + 2) >> 2;
    else
        dc = 128;
    int sd = total - iabs(had[0]) + iabs(had[0] - 16 * dc);
    int sv = 0, sh = 0;
    if (top) {
        int x0 = top[0], x1 = top[1], x2 = top[2], x3 = top[3];
        int s01 = x0 + x1, d01 = x0 - x1, s23 = x2 + x3, d23 = x2 - x3;
        int p[4] = {4 * (s01 + s23), 4 * (s01 - s23), 4 * (d01 + d23), 4 * (d01 - d23)};
        sv = total;
        for (int h = 0; h < 4; h++) sv += iabs(had[h] - p[h]) - iabs(had[h]);
    }
    if (left) {
        int x0 = left[0], x1 = left[1], x2 = left[2], x3 = left[3];
        int s01 = x0 + x1, d01 = x0 - x1, s23 = x2 + x3, d23 = x2 - x3;
        int p[4] = {4 * (s01 + s23), 4 * (s01 - s23), 4 * (d01 + d23), 4 * (d01 - d23)};
        sh = total;
        for (int v = 0; v < 4; v++) sh += iabs(had[4 * v] - p[v]) - iabs(had[4 * v]);
    }
    out[0] = (uint32_t)(sv >> 1);
    out[1] = (uint32_t)(sh >> 1);
    out[2] = (uint32_t)(sd >> 1);
}
/* simd/h264_enc.rs:355 hadamard4x4_16x4_scalar on one 4x4 block (o[4v + h]) */
static void hadamard4_src(const uint8_t *src, int ss, int16_t *o) {
    int rows[4][4];
    for (int y = 0; y < 4; y++) {
        const uint8_t *x = src + y * ss;
        int s01 = x[0] + x[1], d01 = x[0] - x[1], s23 = x[2] + x[3], d23 = x[2] - x[3];
        rows[y][0] = s01 + s23;
        rows[y][1] = s01 - s23;
        rows[y][2] = d01 + d23;
        rows[y][3] = d01 - d23;
    }
    for (int h = 0; h < 4; h++) {
        int s01 = rows[0][h] + rows[1][h], d01 = rows[0][h] - rows[1][h];
        int s23 = rows[2][h] + rows[3][h], d23 = rows[2][h] - rows[3][h];
        o[0 + h] = (int16_t)(s01 + s23);
        o[4 + h] = (int16_t)(s01 - s23);
        o[8 + h] = (int16_t)(d01 + d23);
        o[12 + h] = (int16_t)(d01 - d23);
    }
}
typedef struct {
    uint8_t modes[16];
    int16_t q4[16][16];
    int cbp_luma;
    uint8_t recon[256];
} i4_choice_t;
/* enc.rs:3333 analyse_i4: returns 1 with the choice when it beats budget */
static int analyse_i4(ora_enc *e, int mb_x, int mb_y, int qp, const mbsrc_t *s, uint32_t budget, i4_choice_t *ch) {
    uint32_t lambda = T_LAMBDA[qp];
    int gs = e->mb_w * 4;
    uint8_t saved[16];
    for (int b = 0; b < 16; b++) {
        int bx, by;
        blk_xy(b, &bx, &by);
        saved[b] = e->intra4x4_modes[(mb_y * 4 + by) * gs + mb_x * 4 + bx];
        ch->modes[b] = 2;
    }
    memset(ch->q4, 0, sizeof(ch->q4));
    ch->cbp_luma = 0;
    int16_t had[16][16];
    for (int by = 0; by < 4; by++)
        for (int bx = 0; bx < 4; bx++) hadamard4_src(s->y + by * 64 + bx * 4, 16, had[by * 4 + bx]);
    const quant4_t *quant = quant4(qp, 1);
    const int32_t *scale = DEQUANT4[qp];
    uint32_t total = lambda * 24;
    plane_t *py = &e->cur.y;
    for (int b = 0; b < 16; b++) {
        int bx, by;
        blk_xy(b, &bx, &by);
        int px = mb_x * 16 + bx * 4, pyy = mb_y * 16 + by * 4, gx = mb_x * 4 + bx, gy = mb_y * 4 + by;
        uint8_t top[8], left[4];
        int has_top = pyy > 0, has_left = px > 0;
        if (has_top) {
            const uint8_t *r = py->buf + pat(py, px, pyy - 1);
            memcpy(top, r, 4);
            int tr_avail = !(b == 3 || b == 7 || b == 11 || b == 13 || b == 15) && (b != 5 || mb_x + 1 < e->mb_w);
            if (tr_avail)
                memcpy(top + 4, r + 4, 4);
            else
                memset(top + 4, top[3], 4);
        }
        if (has_left)
            for (int r = 0; r < 4; r++) left[r] = py->buf[pat(py, px - 1, pyy + r)];
        int tl = (px > 0 && pyy > 0) ? py->buf[pat(py, px - 1, pyy - 1)] : 128;
        int pm = pred_i4_mode(e, gx, gy);
        int so = by * 64 + bx * 4;
        uint32_t sat[3];
        intra_satd_x3_4(had[by * 4 + bx], has_top ? top : NULL, has_left ? left : NULL, sat);
        uint32_t best = UINT32_MAX;
        int bm = 2;
        for (int m = 0; m < 3; m++) {
            int ok = m == 0 ? has_top : m == 1 ? has_left : 1;
            uint32_t c = sat[m] + lambda * (m == pm ? 1 : 4);
            if (ok && c < best) {
                best = c;
                bm = m;
            }
        }
        uint8_t dirs[6][16];
        for (int m = 3; m <= 8; m++) {
            if (!intra4_mode_available(m, has_top, has_left)) continue;
            pred4(m, has_top ? top : NULL, has_left ? left : NULL, tl, dirs[m - 3], 4);
            uint32_t c = satd4(s->y + so, 16, dirs[m - 3], 4) + lambda * (m == pm ? 1 : 4);
            if (c < best) {
                best = c;
                bm = m;
            }
        }
        uint8_t pred[16];
        if (bm >= 3)
            memcpy(pred, dirs[bm - 3], 16);
        else
            pred4(bm, has_top ? top : NULL, has_left ? left : NULL, tl, pred, 4);
        total += best;
        if (total >= budget) {
            for (int k = 0; k < 16; k++) {
                int kx, ky;
                blk_xy(k, &kx, &ky);
                e->intra4x4_modes[(mb_y * 4 + ky) * gs + mb_x * 4 + kx] = saved[k];
            }
            return 0;
        }
        ch->modes[b] = (uint8_t)bm;
        e->intra4x4_modes[gy * gs + gx] = (uint8_t)bm;
        int16_t cf[16], q[16];
        sub_dct4(s->y + so, 16, pred, 4, cf);
        int nz = 0;
        for (int i = 0; i < 16; i++) { /* enc.rs:1704 quant_one */
            int l = (int)(((uint32_t)iabs(cf[i]) * quant->mf[i] + quant->bias) >> quant->shift);
            q[i] = (int16_t)(cf[i] < 0 ? -l : l);
            nz |= q[i] != 0;
        }
        uint8_t rec[16];
        memcpy(rec, pred, 16);
        if (nz) {
            ch->cbp_luma |= 1 << (b >> 2);
            int16_t dq[16];
            for (int i = 0; i < 16; i++) dq[i] = (int16_t)iclamp(q[i] * scale[i], -32768, 32767);
            add_idct4(rec, 4, dq);
        }
        memcpy(ch->q4[b], q, sizeof(q));
        for (int r = 0; r < 4; r++) {
            memcpy(py->buf + pat(py, px, pyy + r), rec + r * 4, 4);
            memcpy(ch->recon + so + r * 16, rec + r * 4, 4);
        }
    }
    return 1;
}
/* enc.rs:3441 code_i4 */
static void code_i4(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const mbsrc_t *s, i4_choice_t *i4) {
    int idx = mb_y * e->mb_w + mb_x, gs = e->mb_w * 4;
    ora_mb_decision *d = &e->dec[idx];
    int16_t *lv = e->levels + (long)idx * ORA_LEVELS_PER_MB;
    uint8_t pu[64], pv[64];
    int cmode = analyse_chroma_intra(e, mb_x, mb_y, s, pu, pv, !e->fast);
    int qpc = chroma_qp(qp);
    chroma_levels_t c;
    quant_chroma(s, pu, pv, qpc, 1, &c);
    int cbpc = chroma_cbp(&c);
    int cbp = i4->cbp_luma | (cbpc << 4);
    int qp_delta = cbp ? qp - e->cur_qp : 0, eff_qp = iclamp(e->cur_qp + qp_delta, 0, 51);
    int eff_qpc = chroma_qp(eff_qp);
    e->mb_is_intra[idx] = 1;
    set_intra_motion(e, mb_x, mb_y);
    for (int b = 0; b < 16; b++) {
        int bx, by;
        blk_xy(b, &bx, &by);
        e->intra4x4_modes[(mb_y * 4 + by) * gs + mb_x * 4 + bx] = 2;
    }
    if (w) {
        put_ue(w, e->is_idr ? 0 : 5);
        if (e->t8x8) put32(w, 0, 1);
    }
    for (int b = 0; b < 16; b++) {
        int bx, by;
        blk_xy(b, &bx, &by);
        int gx = mb_x * 4 + bx, gy = mb_y * 4 + by, m = i4->modes[b];
        int pm = pred_i4_mode(e, gx, gy);
        if (w) {
            if (m == pm)
                put32(w, 1, 1);
            else {
                put32(w, 0, 1);
                put32(w, (uint32_t)(m < pm ? m : m - 1), 3);
            }
        }
        e->intra4x4_modes[gy * gs + gx] = (uint8_t)m;
    }
    if (e->cabac) {
        cabac_luma_nnz(e, mb_x, mb_y, i4->cbp_luma, cbpc, i4->q4, NULL, 0);
        e->mb_cbp_luma[idx] = (uint8_t)i4->cbp_luma;
        e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    } else {
        if (w) {
            put_ue(w, (uint64_t)cmode);
            put_ue(w, intra4_cbp_golomb(cbp));
        }
        if (cbp) {
            if (w) put_se(w, qp_delta);
            write_cavlc_residuals(e, w, mb_x, mb_y, i4->cbp_luma, cbpc, i4->q4, &c);
        }
        e->mb_cbp_luma[idx] = (uint8_t)i4->cbp_luma;
        e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    }
    e->cur_qp = eff_qp;
    e->mb_qps[idx] = (uint8_t)eff_qp;
    memset(lv, 0, ORA_LEVELS_PER_MB * sizeof(int16_t));
    rec_luma4(lv, i4->q4);
    rec_chroma(lv, &c);
    d->mb_type = 3;
    {
        /* The record's mv/mvp words carry the 16 prediction modes, 4 bits each (block b in bits
         * 4b..4b+3, decoding order); an intra macroblock has no motion. */
        uint64_t packed = 0;
        for (int b = 0; b < 16; b++) packed |= (uint64_t)i4->modes[b] << (4 * b);
        memcpy(d->mv, &packed, sizeof(packed));
    }
    d->chroma_mode = (uint8_t)cmode;
    d->cbp_luma = (uint8_t)i4->cbp_luma;
    d->cbp_chroma = (uint8_t)cbpc;
    d->qp = (uint8_t)eff_qp;
    recon_chroma(pu, pv, &c, cbpc, eff_qpc);
    store_mb(e, mb_x, mb_y, i4->recon, pu, pv);
}
/* enc.rs:3311 code_intra */
static void code_intra(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const mbsrc_t *s) {
    uint32_t lambda = T_LAMBDA[qp];
    edges_t ed;
    get_edges(&e->cur.y, mb_x, mb_y, 16, &ed);
    uint32_t i16_satd;
    int i16_mode;
    analyse_i16(&ed, s->y, &i16_satd, &i16_mode, NULL, !e->fast);
    if (!e->fast && i16_satd > 40 * lambda) {
        i4_choice_t ch;
        if (analyse_i4(e, mb_x, mb_y, qp, s, i16_satd + lambda * 4, &ch)) {
            code_i4(e, w, mb_x, mb_y, qp, s, &ch);
            return;
        }
    }
    code_i16(e, w, mb_x, mb_y, qp, s, i16_mode);
}

/* ------------------------------------------------------------------------------------------ */
/* Balanced P macroblock (enc.rs:2906-3306, 1097-1455).                                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int min[2], max[2];
} mvb_t;
static void mvb_init(mvb_t *b, int x0, int y0, int w, int h) { /* enc.rs:1104 */
    b->min[0] = -4 * (x0 + MV_MARGIN);
    b->min[1] = imax(-4 * (y0 + MV_MARGIN), -2048);
    b->max[0] = 4 * (w - 16 - x0 + MV_MARGIN);
    b->max[1] = imin(4 * (h - 16 - y0 + MV_MARGIN), 2044);
}
static int mvb_contains(const mvb_t *b, int mx, int my) {
    return mx >= b->min[0] && mx <= b->max[0] && my >= b->min[1] && my <= b->max[1];
}

typedef struct {
    const plane_t *r;
    const uint8_t *src;
    int x0, y0;
    int16_t mvp[2];
    uint32_t lambda;
    const ora_enc *e;
    mvb_t b;
    ora_mb_decision *d;
} ms_t;
static uint32_t ms_mv_cost(const ms_t *m, int mx, int my) {
    return mvd_cost(m->e, m->lambda, mx - m->mvp[0]) + mvd_cost(m->e, m->lambda, my - m->mvp[1]);
}
static uint32_t ms_cost_at(const ms_t *m, int fx, int fy) { /* enc.rs:1176 */
    long off = pat(m->r, m->x0 + fx, m->y0 + fy);
    m->d->n_sad++;
    return sad_wh(m->src, 16, m->r->buf + off, m->r->stride, 16, 16) + ms_mv_cost(m, fx * 4, fy * 4);
}
/* enc.rs:1185 full_pel; nb (may be NULL): the costs of the last iteration's four points (up,
 * left, right, down; UINT32_MAX outside the range or when no iteration ran), which the Fast
 * preset's half-pel step reads. */
static void ms_full_pel(const ms_t *m, const int16_t (*cands)[2], int nc, int iters, int best[2], uint32_t *best_cost,
                        uint32_t nb[4]) {
    if (nb) nb[0] = nb[1] = nb[2] = nb[3] = UINT32_MAX;
    int imn[2] = {(m->b.min[0] + 3) >> 2, (m->b.min[1] + 3) >> 2};
    int imx[2] = {m->b.max[0] >> 2, m->b.max[1] >> 2};
    best[0] = best[1] = 0;
    *best_cost = UINT32_MAX;
    int seen[8][2];
    for (int k = 0; k < nc && k < 8; k++) {
        int mx = iclamp((cands[k][0] + 2) >> 2, imn[0], imx[0]);
        int my = iclamp((cands[k][1] + 2) >> 2, imn[1], imx[1]);
        seen[k][0] = seen[k][1] = INT32_MAX;
        int dup = 0;
        for (int j = 0; j < k; j++)
            if (seen[j][0] == mx && seen[j][1] == my) dup = 1;
        if (dup) continue;
        seen[k][0] = mx;
        seen[k][1] = my;
        uint32_t c = ms_cost_at(m, mx, my);
        if (c < *best_cost) {
            *best_cost = c;
            best[0] = mx;
            best[1] = my;
        }
    }
    for (int it = 0; it < iters; it++) {
        int pts[4][2] = {{best[0], best[1] - 1}, {best[0] - 1, best[1]}, {best[0] + 1, best[1]}, {best[0], best[1] + 1}};
        int moved = -1;
        /* Both branches of the reference pick the first of the cheapest points below the
         * current best (sad_x4 with index tie-break, or sequential strict compares). */
        for (int k = 0; k < 4; k++) {
            if (nb) nb[k] = UINT32_MAX;
            if (pts[k][0] < imn[0] || pts[k][0] > imx[0] || pts[k][1] < imn[1] || pts[k][1] > imx[1]) continue;
            uint32_t c = ms_cost_at(m, pts[k][0], pts[k][1]);
            if (nb) nb[k] = c;
            if (c < *best_cost) {
                *best_cost = c;
                moved = k;
            }
        }
        if (moved < 0) break;
        best[0] = pts[moved][0];
        best[1] = pts[moved][1];
    }
}
/* enc.rs:1415 window_src: the one or two window/reference blocks of quarter-pel mv */
typedef struct {
    const uint8_t *a, *b;
    int sa, sb;
} winsrc_t;
static winsrc_t ms_window_src(const ms_t *m, const hpel_win_t *w, int bx, int by, int mx, int my) {
    int ix = m->x0 + (mx >> 2), iy = m->y0 + (my >> 2);
    int wx = ix - bx + 1, wy = iy - by + 1;
    const plane_t *r = m->r;
#define GP(dx, dy) (r->buf + pat(r, ix + (dx), iy + (dy)))
#define WP(p, dx, dy) ((p) + (wy + (dy)) * WIN_W + wx + (dx))
    int sg = r->stride, sw = WIN_W;
    winsrc_t s = {NULL, NULL, 0, 0};
    switch ((my & 3) * 4 + (mx & 3)) {
    case 0: s.a = GP(0, 0); s.sa = sg; break;
    case 2: s.a = WP(w->h, 0, 0); s.sa = sw; break;
    case 8: s.a = WP(w->v, 0, 0); s.sa = sw; break;
    case 10: s.a = WP(w->c, 0, 0); s.sa = sw; break;
    case 1: s.a = GP(0, 0); s.sa = sg; s.b = WP(w->h, 0, 0); s.sb = sw; break;
    case 3: s.a = WP(w->h, 0, 0); s.sa = sw; s.b = GP(1, 0); s.sb = sg; break;
    case 4: s.a = GP(0, 0); s.sa = sg; s.b = WP(w->v, 0, 0); s.sb = sw; break;
    case 12: s.a = WP(w->v, 0, 0); s.sa = sw; s.b = GP(0, 1); s.sb = sg; break;
    case 5: s.a = WP(w->h, 0, 0); s.sa = sw; s.b = WP(w->v, 0, 0); s.sb = sw; break;
    case 7: s.a = WP(w->h, 0, 0); s.sa = sw; s.b = WP(w->v, 1, 0); s.sb = sw; break;
    case 13: s.a = WP(w->v, 0, 0); s.sa = sw; s.b = WP(w->h, 0, 1); s.sb = sw; break;
    case 15: s.a = WP(w->v, 1, 0); s.sa = sw; s.b = WP(w->h, 0, 1); s.sb = sw; break;
    case 6: s.a = WP(w->h, 0, 0); s.sa = sw; s.b = WP(w->c, 0, 0); s.sb = sw; break;
    case 14: s.a = WP(w->c, 0, 0); s.sa = sw; s.b = WP(w->h, 0, 1); s.sb = sw; break;
    case 9: s.a = WP(w->v, 0, 0); s.sa = sw; s.b = WP(w->c, 0, 0); s.sb = sw; break;
    default: s.a = WP(w->c, 0, 0); s.sa = sw; s.b = WP(w->v, 1, 0); s.sb = sw; break;
    }
#undef GP
#undef WP
    return s;
}
static void winsrc_pred(const winsrc_t *s, uint8_t *dst) {
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++)
            dst[r * 16 + c] = s->b ? (uint8_t)((s->a[r * s->sa + c] + s->b[r * s->sb + c] + 1) >> 1) : s->a[r * s->sa + c];
}
/* enc.rs:1283 sub_pel (Balanced: half-pel 1 iteration, quarter-pel qpel_iters) */
static void ms_sub_pel(const ms_t *m, hpel_win_t *w, int full[2], uint32_t full_cost, int qpel_iters, int16_t mv_out[2],
                       uint32_t *cost_out, uint8_t *pred) {
    int cx = full[0] * 4, cy = full[1] * 4;
    int bx = m->x0 + full[0], by = m->y0 + full[1];
    hpel_window(m->r->buf + pat(m->r, bx - 1, by - 1), m->r->stride, w);
    m->d->n_hpel++;
    /* The reference interpolates the centre plane lazily; computing it up front is the same. */
    hpel_window_centre(w);
    int best[2] = {cx, cy};
    uint32_t best_cost = full_cost;
    for (int pass = 0; pass < 2; pass++) {
        int step = pass == 0 ? 2 : 1, iters = pass == 0 ? 1 : qpel_iters;
        for (int it = 0; it < iters; it++) {
            int from[2] = {best[0], best[1]};
            int ds[4][2] = {{0, -step}, {-step, 0}, {step, 0}, {0, step}};
            for (int k = 0; k < 4; k++) {
                int mx = from[0] + ds[k][0], my = from[1] + ds[k][1];
                if (iabs(mx - cx) > 3 || iabs(my - cy) > 3 || !mvb_contains(&m->b, mx, my)) continue;
                winsrc_t s = ms_window_src(m, w, bx, by, mx, my);
                uint8_t p[256];
                winsrc_pred(&s, p);
                m->d->n_subpel++;
                uint32_t c = sad_wh(m->src, 16, p, 16, 16, 16) + ms_mv_cost(m, mx, my);
                if (c < best_cost) {
                    best_cost = c;
                    best[0] = mx;
                    best[1] = my;
                }
            }
            if (best[0] == from[0] && best[1] == from[1]) break;
        }
    }
    winsrc_t s = ms_window_src(m, w, bx, by, best[0], best[1]);
    winsrc_pred(&s, pred);
    mv_out[0] = (int16_t)best[0];
    mv_out[1] = (int16_t)best[1];
    *cost_out = best_cost;
}

/* enc.rs:1341 half_pel_toward (Fast): on each axis only the half sample toward the cheaper
 * full-pel neighbour is a candidate, estimated by the SAD against the average of its two integer
 * neighbours (enc.rs:1380 half_estimate; UINT32_MAX outside the bounds); the better estimate (the
 * vertical one on a tie), if below the full-pel cost, is interpolated with the 6-tap filter
 * (simd/h264_enc.rs:1345 hpel_16x16) and taken if its SAD cost is below the full-pel cost.
 * Counts each estimate and the interpolated candidate as sub-pel evaluations, the
 * interpolation as a half-pel window. */
static void ms_half_pel_toward(const ms_t *m, const int full[2], uint32_t full_cost, const uint32_t nb[4],
                               int16_t mv_out[2], uint32_t *cost_out, uint8_t *pred) {
    const plane_t *r = m->r;
    int cx = full[0] * 4, cy = full[1] * 4;
    int up = nb[0] < nb[3], left = nb[1] < nb[2];
    long off = pat(r, m->x0 + full[0], m->y0 + full[1]), st = r->stride;
    int mvv[2] = {cx, cy + (up ? -2 : 2)}, mvh[2] = {cx + (left ? -2 : 2), cy};
    uint32_t est[2];
    for (int k = 0; k < 2; k++) {
        const int *mv = k == 0 ? mvv : mvh;
        est[k] = UINT32_MAX;
        if (!mvb_contains(&m->b, mv[0], mv[1])) continue;
        long o2 = k == 0 ? (up ? off - st : off + st) : (left ? off - 1 : off + 1);
        uint32_t sad = 0;
        for (int y = 0; y < 16; y++)
            for (int x = 0; x < 16; x++) {
                int a = r->buf[off + y * st + x], b = r->buf[o2 + y * st + x];
                sad += (uint32_t)iabs((int)m->src[y * 16 + x] - ((a + b + 1) >> 1));
            }
        m->d->n_subpel++;
        est[k] = sad + ms_mv_cost(m, mv[0], mv[1]);
    }
    int vertical = est[0] <= est[1];
    const int *mv = vertical ? mvv : mvh;
    int back = vertical ? up : left;
    if ((vertical ? est[0] : est[1]) < full_cost) {
        uint8_t block[256];
        long base = off - (back ? (vertical ? st : 1) : 0);
        for (int y = 0; y < 16; y++)
            for (int x = 0; x < 16; x++) {
                const uint8_t *q = r->buf + base + y * st + x;
                long step = vertical ? st : 1;
                int v = tap6(q[-2 * step], q[-step], q[0], q[step], q[2 * step], q[3 * step]);
                block[y * 16 + x] = clip_u8((v + 16) >> 5);
            }
        m->d->n_subpel++;
        m->d->n_hpel++;
        uint32_t cost = sad_wh(m->src, 16, block, 16, 16, 16) + ms_mv_cost(m, mv[0], mv[1]);
        if (cost < full_cost) {
            memcpy(pred, block, 256);
            mv_out[0] = (int16_t)mv[0];
            mv_out[1] = (int16_t)mv[1];
            *cost_out = cost;
            return;
        }
    }
    load_block(r, off, pred, 16);
    mv_out[0] = (int16_t)cx;
    mv_out[1] = (int16_t)cy;
    *cost_out = full_cost;
}

/* enc.rs:2906 try_p_skip (probe enabled for Balanced) */
static int try_p_skip(ora_enc *e, int mb_x, int mb_y, int qp, int qpc, const mbsrc_t *s, const int16_t pskip[2],
                      int probe, uint8_t *py, uint8_t *pu, uint8_t *pv) {
    if (!probe) return 0;
    int x0 = mb_x * 16, y0 = mb_y * 16;
    int int_y = ((pskip[0] | pskip[1]) & 3) == 0, int_c = ((pskip[0] | pskip[1]) & 7) == 0;
    const plane_t *ry = &e->refp.y, *ru = &e->refp.u, *rv = &e->refp.v;
    long yo = pat(ry, x0 + (pskip[0] >> 2), y0 + (pskip[1] >> 2));
    long co = pat(ru, mb_x * 8 + (pskip[0] >> 3), mb_y * 8 + (pskip[1] >> 3));
    uint32_t sad;
    if (int_y)
        sad = sad_wh(s->y, 16, ry->buf + yo, ry->stride, 16, 16);
    else {
        mc_luma_mb(ry, x0, y0, pskip, py);
        sad = sad_wh(s->y, 16, py, 16, 16, 16);
    }
    const skip_bounds_t *z = &e->sb;
    if (sad >= z->probe) return 0;
    int luma_zero = sad <= z->luma ||
                    (int_y ? ssd_n(s->y, 16, ry->buf + yo, ry->stride, 16) : ssd_n(s->y, 16, py, 16, 16)) <= z->luma_ssd;
    if (!luma_zero) {
        if (int_y) load_block(ry, yo, py, 16);
        luma_zero = probe_skip_luma(s, py, qp, e->compact);
    }
    if (!luma_zero) return 0;
    if (!int_c) mc_chroma_mb(&e->refp, mb_x * 8, mb_y * 8, pskip, pu, pv);
    int chroma_zero = 1;
    for (int p = 0; p < 2 && chroma_zero; p++) {
        const uint8_t *src = p ? s->v : s->u;
        const uint8_t *b = int_c ? (p ? rv->buf + co : ru->buf + co) : (p ? pv : pu);
        int st = int_c ? ru->stride : 8;
        uint32_t sd = sad_wh(src, 8, b, st, 8, 8);
        chroma_zero = sd <= z->chroma || ssd_n(src, 8, b, st, 8) <= z->chroma_ssd;
    }
    if (!chroma_zero) {
        if (int_c) {
            load_block(ru, co, pu, 8);
            load_block(rv, co, pv, 8);
        }
        chroma_zero = probe_skip_chroma(s, pu, pv, qpc, e->compact);
    }
    if (!chroma_zero) return 0;
    code_skip(e, mb_x, mb_y, pskip);
    uint8_t ty[256], tu[64], tv[64];
    if (int_y)
        load_block(ry, yo, ty, 16);
    else
        memcpy(ty, py, 256);
    if (int_c) {
        load_block(ru, co, tu, 8);
        load_block(rv, co, tv, 8);
    } else {
        memcpy(tu, pu, 64);
        memcpy(tv, pv, 64);
    }
    store_mb(e, mb_x, mb_y, ty, tu, tv);
    return 1;
}

/* enc.rs:3196 code_inter */
static void code_inter(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const int16_t mv[2], const int16_t mvp[2],
                       int cbpl, int cbpc, int16_t q4[16][16], int16_t q8[4][64], int use8,
                       const chroma_levels_t *c, uint8_t *py, uint8_t *pu, uint8_t *pv) {
    int idx = mb_y * e->mb_w + mb_x;
    int cbp = cbpl | (cbpc << 4);
    int mvd0 = mv[0] - mvp[0], mvd1 = mv[1] - mvp[1];
    int qp_delta = cbp ? qp - e->cur_qp : 0, eff_qp = iclamp(e->cur_qp + qp_delta, 0, 51);
    int eff_qpc = chroma_qp(eff_qp);
    if (use8 && cbpl) e->mb_8x8[idx] = 1;
    if (e->cabac) {
        cabac_luma_nnz(e, mb_x, mb_y, cbpl, cbpc, q4, q8, use8);
        e->mb_cbp_luma[idx] = (uint8_t)cbpl;
        e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    } else {
        if (w) {
            put_ue(w, 0);
            put_se(w, mvd0);
            put_se(w, mvd1);
            put_ue(w, inter_cbp_golomb(cbp));
        }
        if (cbp) {
            if (w) {
                if (e->t8x8) put32(w, 0, 1);
                put_se(w, qp_delta);
            }
            write_cavlc_residuals(e, w, mb_x, mb_y, cbpl, cbpc, q4, c);
        }
        e->mb_cbp_luma[idx] = (uint8_t)cbpl;
        e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    }
    e->cur_qp = eff_qp;
    e->mb_qps[idx] = (uint8_t)eff_qp;
    if (use8) {
        for (int i8 = 0; i8 < 4; i8++)
            if (cbpl & (1 << i8)) recon_8x8(py + (i8 >> 1) * 128 + (i8 & 1) * 8, q8[i8], eff_qp);
    } else {
        recon_luma_4x4(py, q4, cbpl, eff_qp);
    }
    recon_chroma(pu, pv, c, cbpc, eff_qpc);
    fill_ref(e, mb_x, mb_y, 0);
    fill_mv(e, mb_x, mb_y, mv[0], mv[1]);
    store_mb(e, mb_x, mb_y, py, pu, pv);
}

/* ------------------------------------------------------------------------------------------ */
/* P partitions 16x8 / 8x16 / 8x8 on top of the Balanced analysis (ora_set_partitions).         */
/* The reference's Balanced preset codes P_L0_16x16 only (enc.rs:3008-3191); its Compact preset */
/* adds 8x8 partitions (enc/inter.rs:651-652). This extension is defined here, by the oracle:   */
/* when the 16x16 inter cost exceeds PART_GATE lambda, each shape is searched partition by      */
/* partition in decoding order: the ITU-T H.264 8.4.1.3 predictor (directional for 16x8 and     */
/* 8x16), candidates [predictor, 16x16 vector] rounded to full samples, a small diamond of at   */
/* most PART_ITERS steps inside +-PART_RANGE samples of the 16x16 full-sample vector (and the    */
/* macroblock's vector bounds), then the half-sample and quarter-sample diamonds of enc.rs:1283  */
/* (SAD of the spec luma MC, +-3 quarter samples around the full-sample result). A shape costs   */
/* the 16x16 SATD of its assembled prediction plus lambda x (its mvd bits and mb_type bits:      */
/* ue(1) 16x8, ue(2) 8x16, ue(3) + 4 x ue(0) 8x8); the cheapest of 16x16, 16x8, 8x16, 8x8 wins   */
/* (strict <, in that order). Records use the quadrant layout of set_compact_record.            */
/* ------------------------------------------------------------------------------------------ */
#define PART_GATE 40
#define PART_RANGE 8
#define PART_ITERS 8
static void fill_part(ora_enc *e, int gx, int gy, int w4, int h4, int8_t r, int mx, int my);
static void set_compact_record(ora_mb_decision *d, const int q[4][2], const int16_t mvr[2]);

/* shape 0 16x16, 1 16x8, 2 8x16, 3 8x8: partition p's rectangle in samples within the macroblock */
static void part_rect(int shape, int p, int *bx, int *by, int *w, int *h) {
    *bx = shape == 2 ? 8 * p : (shape == 3 ? 8 * (p & 1) : 0);
    *by = shape == 1 ? 8 * p : (shape == 3 ? 8 * (p >> 1) : 0);
    *w = shape <= 1 ? 16 : 8;
    *h = (shape == 0 || shape == 2) ? 16 : 8;
}
static const int PART_COUNT[4] = {1, 2, 2, 4};
static const uint32_t PART_TYPE_BITS[4] = {1, 3, 3, 7};
static const uint8_t PART_MB_TYPE[4] = {1, 6, 7, 5};

/* ITU-T H.264 8.4.1.3 (with 6.4.11.7 for the neighbours) for reference 0: the partition whose
 * top-left 4x4 block is (gx, gy) (picture 4x4 units), pw blocks wide, partition `part` of
 * `shape`; `done` marks the 4x4 blocks (bit 4 y + x) of this macroblock already decided. */
static void pred_mv_shape(const ora_enc *e, int gx, int gy, int pw, int shape, int part, uint32_t done, int16_t mvp[2]) {
    int gs = e->mb_w * 4, gh = e->mb_h * 4, mx0 = gx & ~3, my0 = gy & ~3;
    /* availability and contents of the 4x4 block at picture position (x, y) */
    int av[3], ref[3];
    int16_t mv[3][2];
    const int pos[4][2] = {{gx - 1, gy}, {gx, gy - 1}, {gx + pw, gy - 1}, {gx - 1, gy - 1}};
    int cav = 0, cref = -1;
    int16_t cmv[2] = {0, 0};
    for (int n = 0; n < 4; n++) {
        int x = pos[n][0], y = pos[n][1], a;
        if (x < 0 || y < 0 || x >= gs || y >= gh)
            a = 0;
        else if (y >= my0 && y < my0 + 4 && x >= mx0 && x < mx0 + 4)
            a = (int)((done >> (4 * (y - my0) + (x - mx0))) & 1); /* this macroblock: decided yet? */
        else if (y >= my0 && x >= mx0)
            a = 0; /* right of this macroblock in its rows (or below it): not decoded */
        else
            a = 1; /* a macroblock row above, or the macroblock to the left */
        int r = -1;
        int16_t v[2] = {0, 0};
        if (a) {
            long i = (long)y * gs + x;
            r = e->ref_grid[i];
            if (r == 0) {
                v[0] = e->mv_grid[i][0];
                v[1] = e->mv_grid[i][1];
            } else {
                r = -1;
            }
        }
        if (n < 2) {
            av[n] = a;
            ref[n] = r;
            mv[n][0] = v[0];
            mv[n][1] = v[1];
        } else if (n == 2) {
            cav = a;
            cref = r;
            cmv[0] = v[0];
            cmv[1] = v[1];
        } else if (!cav) { /* 8.4.1.3.2: C replaced by D */
            cav = a;
            cref = r;
            cmv[0] = v[0];
            cmv[1] = v[1];
        }
    }
    av[2] = cav;
    ref[2] = cref;
    mv[2][0] = cmv[0];
    mv[2][1] = cmv[1];
    if (!av[1] && !av[2] && av[0]) { /* 8.4.1.3.1: B and C from A */
        for (int n = 1; n < 3; n++) {
            ref[n] = ref[0];
            mv[n][0] = mv[0][0];
            mv[n][1] = mv[0][1];
        }
    }
    int dir = -1;
    if (shape == 1) dir = part == 0 ? 1 : 0;
    if (shape == 2) dir = part == 0 ? 0 : 2;
    if (dir >= 0 && ref[dir] == 0) {
        mvp[0] = mv[dir][0];
        mvp[1] = mv[dir][1];
        return;
    }
    int n0 = (ref[0] == 0) + (ref[1] == 0) + (ref[2] == 0);
    if (n0 == 1) {
        int k = ref[0] == 0 ? 0 : (ref[1] == 0 ? 1 : 2);
        mvp[0] = mv[k][0];
        mvp[1] = mv[k][1];
        return;
    }
    mvp[0] = median3(mv[0][0], mv[1][0], mv[2][0]);
    mvp[1] = median3(mv[0][1], mv[1][1], mv[2][1]);
}

/* One partition's search (see above). src: the macroblock's luma, stride 16; (bx, by, w, h): the
 * partition; F: the 16x16 full-sample vector; pred: the macroblock's prediction (stride 16), whose
 * partition area receives the result's prediction. Returns the SAD cost. */
static uint32_t part_search(ora_enc *e, const uint8_t *src, int x0, int y0, const mvb_t *b, int bx, int by, int w,
                            int h, const int16_t mvp[2], const int16_t mv16[2], const int F[2], uint32_t lambda,
                            int16_t out[2], uint8_t *pred) {
    const plane_t *r = &e->refp.y;
    int lo[2], hi[2];
    for (int k = 0; k < 2; k++) {
        lo[k] = imax((b->min[k] + 3) >> 2, F[k] - PART_RANGE);
        hi[k] = imin(b->max[k] >> 2, F[k] + PART_RANGE);
    }
#define PCOST(fx, fy)                                                                                                 \
    (sad_wh(src + by * 16 + bx, 16, r->buf + pat(r, x0 + bx + (fx), y0 + by + (fy)), r->stride, w, h) +            \
     mvd_cost(e, lambda, 4 * (fx)-mvp[0]) + mvd_cost(e, lambda, 4 * (fy)-mvp[1]))
    int cand[2][2];
    const int16_t *cs[2] = {mvp, mv16};
    for (int k = 0; k < 2; k++) {
        cand[k][0] = iclamp((cs[k][0] + 2) >> 2, lo[0], hi[0]);
        cand[k][1] = iclamp((cs[k][1] + 2) >> 2, lo[1], hi[1]);
    }
    int best[2] = {cand[0][0], cand[0][1]};
    uint32_t best_cost = PCOST(best[0], best[1]);
    if (cand[1][0] != cand[0][0] || cand[1][1] != cand[0][1]) {
        uint32_t c = PCOST(cand[1][0], cand[1][1]);
        if (c < best_cost) {
            best_cost = c;
            best[0] = cand[1][0];
            best[1] = cand[1][1];
        }
    }
    for (int it = 0; it < PART_ITERS; it++) {
        int pts[4][2] = {{best[0], best[1] - 1}, {best[0] - 1, best[1]}, {best[0] + 1, best[1]}, {best[0], best[1] + 1}};
        int moved = -1;
        for (int k = 0; k < 4; k++) {
            if (pts[k][0] < lo[0] || pts[k][0] > hi[0] || pts[k][1] < lo[1] || pts[k][1] > hi[1]) continue;
            uint32_t c = PCOST(pts[k][0], pts[k][1]);
            if (c < best_cost) {
                best_cost = c;
                moved = k;
            }
        }
        if (moved < 0) break;
        best[0] = pts[moved][0];
        best[1] = pts[moved][1];
    }
#undef PCOST
    /* half-sample then quarter-sample diamonds around the full-sample result */
    int cx = best[0] * 4, cy = best[1] * 4, bm[2] = {cx, cy};
    uint8_t p[256];
    for (int pass = 0; pass < 2; pass++) {
        int step = pass == 0 ? 2 : 1, iters = pass == 0 ? 1 : 2;
        for (int it = 0; it < iters; it++) {
            int from[2] = {bm[0], bm[1]};
            int ds[4][2] = {{0, -step}, {-step, 0}, {step, 0}, {0, step}};
            for (int k = 0; k < 4; k++) {
                int mx = from[0] + ds[k][0], my = from[1] + ds[k][1];
                if (iabs(mx - cx) > 3 || iabs(my - cy) > 3 || !mvb_contains(b, mx, my)) continue;
                luma_mc(p, 16, r->buf + pat(r, x0 + bx + (mx >> 2), y0 + by + (my >> 2)), r->stride, w, h, mx, my);
                uint32_t c = sad_wh(src + by * 16 + bx, 16, p, 16, w, h) + mvd_cost(e, lambda, mx - mvp[0]) +
                             mvd_cost(e, lambda, my - mvp[1]);
                if (c < best_cost) {
                    best_cost = c;
                    bm[0] = mx;
                    bm[1] = my;
                }
            }
            if (bm[0] == from[0] && bm[1] == from[1]) break;
        }
    }
    luma_mc(pred + by * 16 + bx, 16, r->buf + pat(r, x0 + bx + (bm[0] >> 2), y0 + by + (bm[1] >> 2)), r->stride, w, h,
            bm[0], bm[1]);
    out[0] = (int16_t)bm[0];
    out[1] = (int16_t)bm[1];
    return best_cost;
}

/* The 16x8, 8x16 and 8x8 shapes of one macroblock against the 16x16 result (cost16, its
 * prediction in py). On return py holds the winner's prediction, q its quadrant vectors (raster),
 * *cost its cost; returns the winning shape. The macroblock's motion grid holds the last shape
 * searched (the caller codes the winner over it). */
static int parts_analyse(ora_enc *e, int mb_x, int mb_y, const mbsrc_t *s, const mvb_t *b, uint32_t lambda,
                         const int F[2], const int16_t mv16[2], uint8_t *py, int q[4][2], uint32_t *cost) {
    int gx0 = mb_x * 4, gy0 = mb_y * 4, x0 = mb_x * 16, y0 = mb_y * 16, win = 0;
    for (int k = 0; k < 4; k++) {
        q[k][0] = mv16[0];
        q[k][1] = mv16[1];
    }
    for (int shape = 1; shape < 4; shape++) {
        uint8_t pred[256];
        int qs[4][2];
        uint32_t bits = 0, done = 0;
        for (int p = 0; p < PART_COUNT[shape]; p++) {
            int bx, by, w, h;
            part_rect(shape, p, &bx, &by, &w, &h);
            int gx = gx0 + bx / 4, gy = gy0 + by / 4;
            int16_t mvp[2], mv[2];
            pred_mv_shape(e, gx, gy, w / 4, shape, p, done, mvp);
            part_search(e, s->y, x0, y0, b, bx, by, w, h, mvp, mv16, F, lambda, mv, pred);
            fill_part(e, gx, gy, w / 4, h / 4, 0, mv[0], mv[1]);
            for (int yy = by / 4; yy < (by + h) / 4; yy++)
                for (int xx = bx / 4; xx < (bx + w) / 4; xx++) done |= 1u << (4 * yy + xx);
            bits += se_bits(mv[0] - mvp[0]) + se_bits(mv[1] - mvp[1]);
            for (int k = 0; k < 4; k++) {
                int kx = (k & 1) * 8, ky = (k >> 1) * 8;
                if (kx >= bx && kx < bx + w && ky >= by && ky < by + h) {
                    qs[k][0] = mv[0];
                    qs[k][1] = mv[1];
                }
            }
        }
        uint32_t c = satd16(s->y, 16, pred, 16) + lambda * (bits + PART_TYPE_BITS[shape]);
        if (c < *cost) {
            *cost = c;
            win = shape;
            memcpy(py, pred, 256);
            memcpy(q, qs, sizeof(qs));
        }
    }
    return win;
}

/* Chroma prediction of the four quadrants, each 4x4 block at its own vector (enc/inter.rs:229
 * predict_inter's chroma; the 1/8-sample bilinear filter). */
static void chroma_quads(const ora_enc *e, int mb_x, int mb_y, const int q[4][2], uint8_t *pu, uint8_t *pv) {
    for (int k = 0; k < 4; k++) {
        int bx = (k & 1) * 4, by = (k >> 1) * 4;
        int cx = mb_x * 8 + bx, cy = mb_y * 8 + by, fx = q[k][0] & 7, fy = q[k][1] & 7;
        int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
        for (int pl = 0; pl < 2; pl++) {
            const plane_t *cp = pl ? &e->refp.v : &e->refp.u;
            uint8_t *dst = (pl ? pv : pu) + by * 8 + bx;
            long off = pat(cp, cx + (q[k][0] >> 3), cy + (q[k][1] >> 3));
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) {
                    const uint8_t *s0 = cp->buf + off + (long)r * cp->stride + c, *s1 = s0 + cp->stride;
                    dst[r * 8 + c] = (uint8_t)((wa * s0[0] + wb * s0[1] + wc * s1[0] + wd * s1[1] + 32) >> 6);
                }
        }
    }
}

/* code_inter for a macroblock of `shape` (CAVLC syntax inline; CABAC from the records later):
 * mb_type ue(0..3), four sub_mb_type ue(0) for P_8x8, each partition's mvd against its 8.4.1.3
 * predictor from the grid as it fills, then the residual exactly as code_inter. */
static void code_inter_shape(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, int shape, const int q[4][2], int cbpl,
                             int cbpc, int16_t q4[16][16], int16_t q8[4][64], int use8, const chroma_levels_t *c,
                             uint8_t *py, uint8_t *pu, uint8_t *pv) {
    int idx = mb_y * e->mb_w + mb_x, gx0 = mb_x * 4, gy0 = mb_y * 4;
    int cbp = cbpl | (cbpc << 4);
    int qp_delta = cbp ? qp - e->cur_qp : 0, eff_qp = iclamp(e->cur_qp + qp_delta, 0, 51);
    int eff_qpc = chroma_qp(eff_qp);
    if (use8 && cbpl) e->mb_8x8[idx] = 1;
    if (w) {
        put_ue(w, (uint64_t)shape);
        if (shape == 3)
            for (int k = 0; k < 4; k++) put_ue(w, 0);
    }
    uint32_t done = 0;
    for (int p = 0; p < PART_COUNT[shape]; p++) {
        int bx, by, pw, ph;
        part_rect(shape, p, &bx, &by, &pw, &ph);
        int k = (bx >> 3) + (by >> 3) * 2;
        int16_t mvp[2];
        pred_mv_shape(e, gx0 + bx / 4, gy0 + by / 4, pw / 4, shape, p, done, mvp);
        if (w) {
            put_se(w, q[k][0] - mvp[0]);
            put_se(w, q[k][1] - mvp[1]);
        }
        fill_part(e, gx0 + bx / 4, gy0 + by / 4, pw / 4, ph / 4, 0, q[k][0], q[k][1]);
        for (int yy = by / 4; yy < (by + ph) / 4; yy++)
            for (int xx = bx / 4; xx < (bx + pw) / 4; xx++) done |= 1u << (4 * yy + xx);
    }
    if (e->cabac) {
        cabac_luma_nnz(e, mb_x, mb_y, cbpl, cbpc, q4, q8, use8);
    } else {
        if (w) put_ue(w, inter_cbp_golomb(cbp));
        if (cbp) {
            if (w) {
                if (e->t8x8) put32(w, 0, 1);
                put_se(w, qp_delta);
            }
            write_cavlc_residuals(e, w, mb_x, mb_y, cbpl, cbpc, q4, c);
        }
    }
    e->mb_cbp_luma[idx] = (uint8_t)cbpl;
    e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    e->cur_qp = eff_qp;
    e->mb_qps[idx] = (uint8_t)eff_qp;
    if (use8) {
        for (int i8 = 0; i8 < 4; i8++)
            if (cbpl & (1 << i8)) recon_8x8(py + (i8 >> 1) * 128 + (i8 & 1) * 8, q8[i8], eff_qp);
    } else {
        recon_luma_4x4(py, q4, cbpl, eff_qp);
    }
    recon_chroma(pu, pv, c, cbpc, eff_qpc);
    store_mb(e, mb_x, mb_y, py, pu, pv);
}

/* The record of a P-picture macroblock with partitions on: words 2-5 the quadrant vectors from
 * the motion grid (zero for intra), word 6 zero. */
static void parts_record(ora_enc *e, int mb_x, int mb_y) {
    ora_mb_decision *d = &e->dec[mb_y * e->mb_w + mb_x];
    int gs = e->mb_w * 4, q[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    int16_t zero[2] = {0, 0};
    if (d->mb_type <= 1 || d->mb_type >= 5)
        for (int k = 0; k < 4; k++) {
            long g = (long)(mb_y * 4 + (k >> 1) * 2) * gs + mb_x * 4 + (k & 1) * 2;
            q[k][0] = e->mv_grid[g][0];
            q[k][1] = e->mv_grid[g][1];
        }
    set_compact_record(d, q, zero);
}

/* enc.rs:3011 code_p_mb */
static void code_p_mb(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const mbsrc_t *s) {
    int mb_w = e->mb_w, gs = mb_w * 4, gx0 = mb_x * 4, gy0 = mb_y * 4, x0 = mb_x * 16, y0 = mb_y * 16;
    int idx = mb_y * mb_w + mb_x;
    ora_mb_decision *d = &e->dec[idx];
    int16_t *lv = e->levels + (long)idx * ORA_LEVELS_PER_MB;
    memset(lv, 0, ORA_LEVELS_PER_MB * sizeof(int16_t));
    int qpc = chroma_qp(qp);
    mvb_t b;
    mvb_init(&b, x0, y0, e->refp.y.w, e->refp.y.h);
    uint32_t lambda = T_LAMBDA[qp];
    int16_t mvp[2];
    pred_mv(e, gx0, gy0, mvp);
    int16_t pskip[2] = {mvp[0], mvp[1]};
    {
        int li = gy0 * gs + gx0 - 1, ti = (gy0 - 1) * gs + gx0;
        if (gx0 == 0 || gy0 == 0 || (e->ref_grid[li] == 0 && e->mv_grid[li][0] == 0 && e->mv_grid[li][1] == 0) ||
            (e->ref_grid[ti] == 0 && e->mv_grid[ti][0] == 0 && e->mv_grid[ti][1] == 0))
            pskip[0] = pskip[1] = 0;
    }
    d->mvp[0] = mvp[0];
    d->mvp[1] = mvp[1];
    d->qp = (uint8_t)qp;
    uint8_t py[256], pu[64], pv[64];
    if (e->sb.qp != qp) skip_bounds_init(&e->sb, qp, qpc, e->compact);
    /* Fast probes only next to a skipped macroblock (enc.rs:3050-3063). */
    int probe = !e->fast || (mb_x > 0 && e->mb_is_skip[idx - 1]) ||
                (mb_y > 0 && (e->mb_is_skip[idx - mb_w] || (mb_x > 0 && e->mb_is_skip[idx - mb_w - 1]) ||
                              (mb_x + 1 < mb_w && e->mb_is_skip[idx - mb_w + 1])));
    if (try_p_skip(e, mb_x, mb_y, qp, qpc, s, pskip, probe && mvb_contains(&b, pskip[0], pskip[1]), py, pu, pv)) {
        d->mb_type = 0;
        d->flags = 1;
        d->mv[0] = pskip[0];
        d->mv[1] = pskip[1];
        return;
    }
    int16_t cands[5][2];
    int n = 0;
    cands[n][0] = mvp[0];
    cands[n++][1] = mvp[1];
    cands[n][0] = 0;
    cands[n++][1] = 0;
    if (mb_x > 0) {
        cands[n][0] = e->mv_grid[gy0 * gs + gx0 - 1][0];
        cands[n++][1] = e->mv_grid[gy0 * gs + gx0 - 1][1];
    }
    if (mb_y > 0) {
        cands[n][0] = e->mv_grid[(gy0 - 1) * gs + gx0][0];
        cands[n++][1] = e->mv_grid[(gy0 - 1) * gs + gx0][1];
    }
    cands[n][0] = e->prev_mb_mv[idx][0];
    cands[n++][1] = e->prev_mb_mv[idx][1];
    ensure_mv_costs(e, qp);
    ms_t ms = {&e->refp.y, s->y, x0, y0, {mvp[0], mvp[1]}, lambda, e, b, d};
    int full[2];
    uint32_t full_cost, nb[4];
    ms_full_pel(&ms, cands, n, e->me_iters, full, &full_cost, nb);
    int16_t mv[2];
    uint32_t me_cost;
    if (full_cost - ms_mv_cost(&ms, full[0] * 4, full[1] * 4) <= e->sb.luma) {
        load_block(&e->refp.y, pat(&e->refp.y, x0 + full[0], y0 + full[1]), py, 16);
        mv[0] = (int16_t)(full[0] * 4);
        mv[1] = (int16_t)(full[1] * 4);
        me_cost = full_cost;
    } else if (e->fast) {
        ms_half_pel_toward(&ms, full, full_cost, nb, mv, &me_cost, py);
    } else {
        ms_sub_pel(&ms, &e->win, full, full_cost, 2, mv, &me_cost, py);
    }
    d->cost = me_cost;
    edges_t ed;
    get_edges(&e->cur.y, mb_x, mb_y, 16, &ed);
    /* Fast analyses intra only when the best of the V / H / DC SADs is within 1.75 times the inter
     * SAD (enc.rs:3119-3140, simd/h264_enc.rs:612 intra_sad_x3_16x16). */
    int try_intra = 1;
    if (e->fast) {
        uint32_t inter_sad = me_cost - ms_mv_cost(&ms, mv[0], mv[1]);
        int dc = i16_dc(&ed);
        uint32_t sv = 0, sh = 0, sd = 0;
        for (int i = 0; i < 256; i++) {
            int v = s->y[i];
            sv += (uint32_t)iabs(v - (ed.has_top ? ed.top[i % 16] : 0));
            sh += (uint32_t)iabs(v - (ed.has_left ? ed.left[i / 16] : 0));
            sd += (uint32_t)iabs(v - dc);
        }
        uint32_t best = sd;
        if (ed.has_top && sv < best) best = sv;
        if (ed.has_left && sh < best) best = sh;
        try_intra = 4 * (uint64_t)best < 7 * (uint64_t)inter_sad;
    }
    uint32_t inter_satd = 0;
    if (try_intra) {
        inter_satd = satd16(s->y, 16, py, 16);
        d->n_satd++;
    }
    uint32_t inter_cost = inter_satd + lambda * (se_bits(mv[0] - mvp[0]) + se_bits(mv[1] - mvp[1]) + 1);
    /* Partitions (ora_set_partitions): the 16x8, 8x16 and 8x8 shapes against the 16x16 cost. */
    int shape = 0, q[4][2];
    if (e->parts && inter_cost > PART_GATE * lambda) {
        shape = parts_analyse(e, mb_x, mb_y, s, &b, lambda, full, mv, py, q, &inter_cost);
    }
    if (inter_satd > 24 * lambda) {
        uint32_t i16_satd;
        int i16_mode;
        analyse_i16(&ed, s->y, &i16_satd, &i16_mode, &d->n_satd, !e->fast);
        if (i16_satd + lambda * (e->fast ? 16 : 8) < inter_cost) {
            begin_coded_mb(e, w);
            code_i16(e, w, mb_x, mb_y, qp, s, i16_mode);
            d->mv[0] = d->mv[1] = 0;
            return;
        }
    }
    if (shape)
        chroma_quads(e, mb_x, mb_y, q, pu, pv);
    else
        mc_chroma_mb(&e->refp, mb_x * 8, mb_y * 8, mv, pu, pv);
    int16_t q4[16][16], q8[4][64];
    int cbpl = quant_luma_4x4_inter(s->y, py, qp, q4, 0);
    int use8 = 0;
    if (e->t8x8 && e->cabac && cbpl) {
        int cbp8 = quant_luma_8x8_inter(s->y, py, qp, q8);
        int nnz4 = 0, nnz8 = 0;
        for (int i = 0; i < 256; i++) {
            nnz4 += ((int16_t *)q4)[i] != 0;
            nnz8 += ((int16_t *)q8)[i] != 0;
        }
        if (nnz8 <= nnz4) {
            use8 = 1;
            cbpl = cbp8;
        }
    }
    chroma_levels_t c;
    quant_chroma_inter(s, pu, pv, qpc, &c, 0);
    int cbpc = chroma_cbp(&c);
    d->mv[0] = mv[0];
    d->mv[1] = mv[1];
    if (shape == 0 && cbpl == 0 && cbpc == 0 && mv[0] == pskip[0] && mv[1] == pskip[1]) {
        code_skip(e, mb_x, mb_y, mv);
        store_mb(e, mb_x, mb_y, py, pu, pv);
        d->mb_type = 0;
        return;
    }
    begin_coded_mb(e, w);
    if (e->parts) {
        if (shape == 0)
            for (int k = 0; k < 4; k++) {
                q[k][0] = mv[0];
                q[k][1] = mv[1];
            }
        code_inter_shape(e, w, mb_x, mb_y, qp, shape, q, cbpl, cbpc, q4, q8, use8, &c, py, pu, pv);
    } else {
        code_inter(e, w, mb_x, mb_y, qp, mv, mvp, cbpl, cbpc, q4, q8, use8, &c, py, pu, pv);
    }
    d->mb_type = PART_MB_TYPE[shape];
    d->cbp_luma = (uint8_t)cbpl;
    d->cbp_chroma = (uint8_t)cbpc;
    d->transform_8x8 = (uint8_t)(use8 && cbpl);
    d->qp = e->mb_qps[idx];
    if (use8) {
        for (int i8 = 0; i8 < 4; i8++)
            for (int i = 0; i < 64; i++) lv[ORA_LV_LUMA + i8 * 64 + i] = q8[i8][T_ZIGZAG8[i]];
    } else {
        rec_luma4(lv, q4);
    }
    rec_chroma(lv, &c);
}

/* ------------------------------------------------------------------------------------------ */
/* Deblocking (enc.rs:3776-3942, simd/h264_deblock.rs).                                        */
/* ------------------------------------------------------------------------------------------ */
static void luma_deblock(uint8_t *buf, long q0, long xs, long ys, int alpha, int beta, const int tc0s[4], int strong) {
    for (int k = 0; k < 16; k++) { /* simd/h264_deblock.rs:69 */
        int tc0 = tc0s[k / 4];
        if (!strong && tc0 < 0) continue;
        long o = q0 + k * ys;
#define P(i) buf[o + (i)*xs]
        int p0 = P(-1), q0v = P(0), p1 = P(-2), q1 = P(1);
        if (iabs(p0 - q0v) >= alpha || iabs(p1 - p0) >= beta || iabs(q1 - q0v) >= beta) continue;
        int p2 = P(-3), q2 = P(2);
        int ap = iabs(p2 - p0) < beta, aq = iabs(q2 - q0v) < beta;
        if (!strong) {
            int tc = tc0 + ap + aq;
            int dd = iclamp((((q0v - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
            P(-1) = clip_u8(p0 + dd);
            P(0) = clip_u8(q0v - dd);
            if (ap) P(-2) = (uint8_t)(p1 + iclamp((p2 + ((p0 + q0v + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0));
            if (aq) P(1) = (uint8_t)(q1 + iclamp((q2 + ((p0 + q0v + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0));
        } else {
            int p3 = P(-4), q3 = P(3);
            int small = iabs(p0 - q0v) < ((alpha >> 2) + 2);
            if (ap && small) {
                P(-1) = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0v + q1 + 4) >> 3);
                P(-2) = (uint8_t)((p2 + p1 + p0 + q0v + 2) >> 2);
                P(-3) = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0v + 4) >> 3);
            } else {
                P(-1) = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
            }
            if (aq && small) {
                P(0) = (uint8_t)((p1 + 2 * p0 + 2 * q0v + 2 * q1 + q2 + 4) >> 3);
                P(1) = (uint8_t)((p0 + q0v + q1 + q2 + 2) >> 2);
                P(2) = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0v + p0 + 4) >> 3);
            } else {
                P(0) = (uint8_t)((2 * q1 + q0v + p1 + 2) >> 2);
            }
        }
#undef P
    }
}
static void chroma_deblock(uint8_t *buf, long q0, long xs, long ys, int alpha, int beta, const int tc0s[4], int strong) {
    for (int k = 0; k < 8; k++) { /* simd/h264_deblock.rs:206 */
        int tc0 = tc0s[k / 2];
        if (!strong && tc0 < 0) continue;
        long o = q0 + k * ys;
        int p1 = buf[o - 2 * xs], p0 = buf[o - xs], q0v = buf[o], q1 = buf[o + xs];
        if (iabs(p0 - q0v) >= alpha || iabs(p1 - p0) >= beta || iabs(q1 - q0v) >= beta) continue;
        if (strong) {
            buf[o - xs] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
            buf[o] = (uint8_t)((2 * q1 + q0v + p1 + 2) >> 2);
        } else {
            int tc = tc0 + 1;
            int dd = iclamp((((q0v - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
            buf[o - xs] = clip_u8(p0 + dd);
            buf[o] = clip_u8(q0v - dd);
        }
    }
}
/* enc.rs:3902 EdgeFilter::new + apply */
static void edge_filter(plane_t *p, int x, int y, int dir, int chroma, int qp, const int bs[4]) {
    int ia = iclamp(qp, 0, 51), alpha = T_DEBLOCK_ALPHA[ia], beta = T_DEBLOCK_BETA[ia];
    if (!alpha || !beta) return;
    int tc[5] = {-1, T_DEBLOCK_TC0[ia][0], T_DEBLOCK_TC0[ia][1], T_DEBLOCK_TC0[ia][2], 0};
    int tc0s[4];
    for (int k = 0; k < 4; k++) tc0s[k] = tc[imin(bs[k], 4)];
    int strong = bs[0] == 4;
    long q0 = pat(p, x, y), xs = dir == 0 ? 1 : p->stride, ys = dir == 0 ? p->stride : 1;
    if (chroma)
        chroma_deblock(p->buf, q0, xs, ys, alpha, beta, tc0s, strong);
    else
        luma_deblock(p->buf, q0, xs, ys, alpha, beta, tc0s, strong);
}
static uint32_t nz_row(const uint8_t *nnz, long g) {
    return (uint32_t)(nnz[g] != 0) | (uint32_t)(nnz[g + 1] != 0) << 1 | (uint32_t)(nnz[g + 2] != 0) << 2 |
           (uint32_t)(nnz[g + 3] != 0) << 3;
}
static uint32_t nz_col(const uint32_t rows[4], int c) {
    return ((rows[0] >> c) & 1) | (((rows[1] >> c) & 1) << 1) | (((rows[2] >> c) & 1) << 2) | (((rows[3] >> c) & 1) << 3);
}
/* enc.rs:3776 deblock */
static void deblock(ora_enc *e) {
    int mb_w = e->mb_w, mb_h = e->mb_h, gs = mb_w * 4;
#define QUIET(m) (!e->mb_is_intra[m] && e->mb_cbp_luma[m] == 0)
#define MVEQ(ga, gb) (e->mv_grid[ga][0] == e->mv_grid[gb][0] && e->mv_grid[ga][1] == e->mv_grid[gb][1])
    for (int mb_y = 0; mb_y < mb_h; mb_y++)
        for (int mb_x = 0; mb_x < mb_w; mb_x++) {
            int m = mb_y * mb_w + mb_x;
            long g0 = (long)mb_y * 4 * gs + mb_x * 4;
            if (QUIET(m) && (mb_x == 0 || (QUIET(m - 1) && MVEQ(g0 - 4, g0))) &&
                (mb_y == 0 || (QUIET(m - mb_w) && MVEQ(g0 - 4 * gs, g0))))
                continue;
            int intra = e->mb_is_intra[m], internal = intra || e->mb_cbp_luma[m] != 0, skip_odd = e->mb_8x8[m];
            uint32_t rows[4];
            for (int r = 0; r < 4; r++) rows[r] = nz_row(e->luma_nnz, g0 + r * gs);
            int qp_q = e->mb_qps[m], qpc_q = chroma_qp(qp_q);
            for (int dir = 0; dir < 2; dir++)
                for (int ed = 0; ed < 4; ed++) {
                    int pm;
                    if (ed == 0) {
                        if (dir == 0 && mb_x > 0)
                            pm = m - 1;
                        else if (dir == 1 && mb_y > 0)
                            pm = m - mb_w;
                        else
                            continue;
                    } else {
                        if (!internal || ((ed & 1) && skip_odd)) continue;
                        pm = m;
                    }
                    int bs[4];
                    if (intra || e->mb_is_intra[pm]) {
                        for (int k = 0; k < 4; k++) bs[k] = ed == 0 ? 4 : 3;
                    } else {
                        uint32_t coded;
                        if (dir == 0 && ed == 0) {
                            coded = nz_col(rows, 0);
                            for (int k = 0; k < 4; k++) coded |= (uint32_t)(e->luma_nnz[g0 + k * gs - 1] != 0) << k;
                        } else if (dir == 0)
                            coded = nz_col(rows, ed) | nz_col(rows, ed - 1);
                        else if (ed == 0)
                            coded = rows[0] | nz_row(e->luma_nnz, g0 - gs);
                        else
                            coded = rows[ed] | rows[ed - 1];
                        int motion = 0;
                        if (ed == 0) {
                            long gp = dir == 0 ? g0 - 1 : g0 - gs;
                            motion = e->ref_grid[g0] != e->ref_grid[gp] || iabs(e->mv_grid[g0][0] - e->mv_grid[gp][0]) >= 4 ||
                                     iabs(e->mv_grid[g0][1] - e->mv_grid[gp][1]) >= 4;
                        }
                        if (!coded && !motion) continue;
                        for (int k = 0; k < 4; k++) bs[k] = ((coded >> k) & 1) ? 2 : motion;
                    }
                    int qp, qpc;
                    if (pm == m) {
                        qp = qp_q;
                        qpc = qpc_q;
                    } else {
                        int qp_p = e->mb_qps[pm];
                        qp = (qp_q + qp_p + 1) >> 1;
                        qpc = (qpc_q + chroma_qp(qp_p) + 1) >> 1;
                    }
                    int x = dir == 0 ? mb_x * 16 + ed * 4 : mb_x * 16, y = dir == 0 ? mb_y * 16 : mb_y * 16 + ed * 4;
                    edge_filter(&e->cur.y, x, y, dir, 0, qp, bs);
                    if (!(ed & 1)) {
                        edge_filter(&e->cur.u, x / 2, y / 2, dir, 1, qpc, bs);
                        edge_filter(&e->cur.v, x / 2, y / 2, dir, 1, qpc, bs);
                    }
                }
        }
#undef QUIET
#undef MVEQ
}


/* ------------------------------------------------------------------------------------------ */
/* Compact P macroblock (enc/inter.rs:498-759, enc/compact.rs).                                 */
/* ------------------------------------------------------------------------------------------ */
/* enc.rs:73 LAMBDA_COMPACT: four fifths of LAMBDA, at least 1 */
static uint32_t lambda_compact(int qp) {
    uint32_t l = (uint32_t)T_LAMBDA[qp] * 4 / 5;
    return l > 1 ? l : 1;
}
/* enc.rs:282 mv_costs(qp, compact) */
static void ensure_mv_costs_c(ora_enc *e, int qp) {
    if (e->mv_costs_c_qp == qp) return;
    uint32_t lambda = lambda_compact(qp);
    for (int d = -MV_COST_RANGE; d <= MV_COST_RANGE; d++)
        e->mv_costs_c[d + MV_COST_RANGE] = (uint16_t)(lambda * se_bits(d));
    e->mv_costs_c_qp = qp;
}

/* compact.rs:24 HpelPlanes::build with simd/h264_enc.rs:905 hpel_planes_scalar: b, h and j of
 * the reference luma over [-pad + 2, w + pad - 11) x [-pad + 2, h + pad - 3), replicated
 * outward (compact.rs:55 extend_from) */
static void plane_extend_from(plane_t *p, int x0, int x1, int y0, int y1) {
    int s = p->stride, rows = p->h + 2 * p->pad;
    long bx0 = pat(p, x0, 0) % s, bx1 = pat(p, x1 - 1, 0) % s;
    long by0 = pat(p, 0, y0) / s, by1 = pat(p, 0, y1 - 1) / s;
    for (long y = by0; y <= by1; y++) {
        uint8_t *row = p->buf + y * s;
        uint8_t l = row[bx0], r = row[bx1];
        memset(row, l, (size_t)bx0);
        memset(row + bx1 + 1, r, (size_t)(s - bx1 - 1));
    }
    for (long y = 0; y < by0; y++) memcpy(p->buf + y * s, p->buf + by0 * s, (size_t)s);
    for (long y = by1 + 1; y < rows; y++) memcpy(p->buf + y * s, p->buf + by1 * s, (size_t)s);
}
static void hpel_planes_build(ora_enc *e) {
    const plane_t *f = &e->refp.y;
    int pad = f->pad, s = f->stride;
    int x_lo = -pad + 2, y_lo = -pad + 2, x_hi = f->w + pad - 11, y_hi = f->h + pad - 3;
    int *tv = (int *)malloc(sizeof(int) * (size_t)(x_hi - x_lo + 5));
    for (int y = y_lo; y < y_hi; y++) {
        for (int k = 0; k < x_hi - x_lo + 5; k++) {
            const uint8_t *r = f->buf + pat(f, x_lo + k - 2, y);
            tv[k] = tap6(r[-2 * s], r[-s], r[0], r[s], r[2 * s], r[3 * s]);
        }
        for (int x = x_lo; x < x_hi; x++) {
            const uint8_t *r = f->buf + pat(f, x, y);
            const int *t = tv + (x - x_lo);
            e->hh.buf[pat(&e->hh, x, y)] = clip_u8((tap6(r[-2], r[-1], r[0], r[1], r[2], r[3]) + 16) >> 5);
            e->hv.buf[pat(&e->hv, x, y)] = clip_u8((t[2] + 16) >> 5);
            e->hc.buf[pat(&e->hc, x, y)] = clip_u8((tap6(t[0], t[1], t[2], t[3], t[4], t[5]) + 512) >> 10);
        }
    }
    free(tv);
    plane_extend_from(&e->hh, x_lo, x_hi, y_lo, y_hi);
    plane_extend_from(&e->hv, x_lo, x_hi, y_lo, y_hi);
    plane_extend_from(&e->hc, x_lo, x_hi, y_lo, y_hi);
}

/* compact.rs:87 RefLuma::src: the block (or the two blocks whose rounded average is the
 * prediction) of the w x h block at (x0, y0) displaced by quarter-pel (mx, my) */
static void rl_blocks(const ora_enc *e, int x0, int y0, int mx, int my, const uint8_t **pa, const uint8_t **pb) {
    const plane_t *f = &e->refp.y;
    long o = pat(f, x0 + (mx >> 2), y0 + (my >> 2)), s = f->stride;
    const uint8_t *F = f->buf, *Hh = e->hh.buf, *V = e->hv.buf, *C = e->hc.buf;
    const uint8_t *a = NULL, *b = NULL;
    switch ((mx & 3) | ((my & 3) << 2)) {
    case 0: a = F + o; break;
    case 2: a = Hh + o; break;
    case 8: a = V + o; break;
    case 10: a = C + o; break;
    case 1: a = F + o; b = Hh + o; break;
    case 3: a = Hh + o; b = F + o + 1; break;
    case 4: a = F + o; b = V + o; break;
    case 12: a = V + o; b = F + o + s; break;
    case 5: a = Hh + o; b = V + o; break;
    case 7: a = Hh + o; b = V + o + 1; break;
    case 13: a = V + o; b = Hh + o + s; break;
    case 15: a = V + o + 1; b = Hh + o + s; break;
    case 6: a = Hh + o; b = C + o; break;
    case 14: a = C + o; b = Hh + o + s; break;
    case 9: a = V + o; b = C + o; break;
    default: a = C + o; b = V + o + 1; break; /* (3, 2) */
    }
    *pa = a;
    *pb = b;
}
/* compact.rs:130 RefLuma::pred into dst (stride ds) */
static void rl_pred(const ora_enc *e, int x0, int y0, int w, int h, int mx, int my, uint8_t *dst, int ds) {
    const uint8_t *a, *b;
    rl_blocks(e, x0, y0, mx, my, &a, &b);
    long s = e->refp.y.stride;
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++)
            dst[r * ds + c] = b ? (uint8_t)((a[r * s + c] + b[r * s + c] + 1) >> 1) : a[r * s + c];
}
/* simd/h264_me.rs:163 pixel_satd_wxh_scalar: 4x4 Hadamard sums over the block, halved at the end */
static uint32_t satd_wh(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h) {
    uint32_t sum = 0;
    for (int by = 0; by < h; by += 4)
        for (int bx = 0; bx < w; bx += 4) {
            int d[4][4];
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++) d[y][x] = a[(by + y) * sa + bx + x] - b[(by + y) * sb + bx + x];
            for (int y = 0; y < 4; y++) {
                int s01 = d[y][0] + d[y][1], d01 = d[y][0] - d[y][1];
                int s23 = d[y][2] + d[y][3], d23 = d[y][2] - d[y][3];
                d[y][0] = s01 + s23;
                d[y][1] = s01 - s23;
                d[y][2] = d01 + d23;
                d[y][3] = d01 - d23;
            }
            for (int c = 0; c < 4; c++) {
                int s01 = d[0][c] + d[1][c], d01 = d[0][c] - d[1][c];
                int s23 = d[2][c] + d[3][c], d23 = d[2][c] - d[3][c];
                sum += (uint32_t)(iabs(s01 + s23) + iabs(s01 - s23) + iabs(d01 + d23) + iabs(d01 - d23));
            }
        }
    return sum >> 1;
}

/* compact.rs:267 PartSearch + :453 FullPel */
typedef struct {
    ora_enc *e;
    const uint8_t *src; /* the partition's source samples, stride 16 */
    int x0, y0, w, h;   /* the partition in the picture */
    int mvp[2];
    uint32_t lambda;
    mvb_t b;            /* bounds of the macroblock's vectors */
    int lo[2], hi[2];   /* full-pel bounds */
    int best[2];
    uint32_t best_cost;
} ps_t;
typedef struct {
    int mv[2];
    uint32_t satd, mv_cost;
} me_t;
static uint32_t me_cost(const me_t *m) { return m->satd + m->mv_cost; }
static void ps_init(ps_t *p, ora_enc *e, const uint8_t *src, int x0, int y0, int w, int h, const int16_t mvp[2],
                    uint32_t lambda, const mvb_t *b) {
    p->e = e;
    p->src = src;
    p->x0 = x0;
    p->y0 = y0;
    p->w = w;
    p->h = h;
    p->mvp[0] = mvp[0];
    p->mvp[1] = mvp[1];
    p->lambda = lambda;
    p->b = *b;
    p->lo[0] = (b->min[0] + 3) >> 2;
    p->lo[1] = (b->min[1] + 3) >> 2;
    p->hi[0] = b->max[0] >> 2;
    p->hi[1] = b->max[1] >> 2;
    p->best[0] = p->best[1] = 0;
    p->best_cost = UINT32_MAX;
}
static uint32_t ps_mvd_cost(const ps_t *p, int d) { /* compact.rs:543 */
    if (d >= -MV_COST_RANGE && d <= MV_COST_RANGE) return p->e->mv_costs_c[d + MV_COST_RANGE];
    return p->lambda * se_bits(d);
}
static uint32_t ps_mv_cost(const ps_t *p, int mx, int my) { return ps_mvd_cost(p, mx - p->mvp[0]) + ps_mvd_cost(p, my - p->mvp[1]); }
/* compact.rs:557 cost: SAD or SATD against the prediction at quarter-pel (mx, my) */
static uint32_t ps_cost(const ps_t *p, int mx, int my, int satd) {
    uint8_t pred[256];
    rl_pred(p->e, p->x0, p->y0, p->w, p->h, mx, my, pred, 16);
    return satd ? satd_wh(p->src, 16, pred, 16, p->w, p->h) : sad_wh(p->src, 16, pred, 16, p->w, p->h);
}
/* compact.rs:492 FullPel::at */
static void fp_at(const ps_t *p, const int from[2], int dx, int dy, int out[2]) {
    int m0 = from[0] + dx, m1 = from[1] + dy;
    int ok = (uint32_t)(m0 - p->lo[0]) <= (uint32_t)(p->hi[0] - p->lo[0]) && (uint32_t)(m1 - p->lo[1]) <= (uint32_t)(p->hi[1] - p->lo[1]);
    out[0] = ok ? m0 : from[0];
    out[1] = ok ? m1 : from[1];
}
/* compact.rs:516 FullPel::take: four points in order, the first of equal costs winning;
 * returns the last that improved, or -1 */
static int fp_take(ps_t *p, int pts[4][2]) {
    const plane_t *f = &p->e->refp.y;
    int moved = -1;
    for (int k = 0; k < 4; k++) {
        long o = pat(f, p->x0 + pts[k][0], p->y0 + pts[k][1]);
        uint32_t c = sad_wh(p->src, 16, f->buf + o, f->stride, p->w, p->h) + ps_mv_cost(p, 4 * pts[k][0], 4 * pts[k][1]);
        if (c < p->best_cost) {
            p->best_cost = c;
            p->best[0] = pts[k][0];
            p->best[1] = pts[k][1];
            moved = k;
        }
    }
    return moved;
}
/* compact.rs:723 refine_qpel_k: quarter-pel diamonds by SATD, at most iters */
static me_t ps_refine_qpel(const ps_t *p, me_t m, int iters) {
    static const int D[4][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}};
    int best[2] = {m.mv[0], m.mv[1]}, prev[2] = {0, 0}, has_prev = 0;
    uint32_t best_satd = m.satd, best_cost = m.satd + ps_mv_cost(p, best[0], best[1]);
    for (int it = 0; it < iters; it++) {
        int from[2] = {best[0], best[1]};
        for (int k = 0; k < 4; k++) {
            int mx = from[0] + D[k][0], my = from[1] + D[k][1];
            if ((has_prev && mx == prev[0] && my == prev[1]) || !mvb_contains(&p->b, mx, my)) continue;
            uint32_t satd = ps_cost(p, mx, my, 1);
            uint32_t cost = satd + ps_mv_cost(p, mx, my);
            if (cost < best_cost) {
                best_cost = cost;
                best_satd = satd;
                best[0] = mx;
                best[1] = my;
            }
        }
        if (best[0] == from[0] && best[1] == from[1]) break;
        prev[0] = from[0];
        prev[1] = from[1];
        has_prev = 1;
    }
    me_t r = {{best[0], best[1]}, best_satd, ps_mv_cost(p, best[0], best[1])};
    return r;
}
/* compact.rs:651 refine: the predictor at its sub-pel position, a half-pel diamond by SAD,
 * then quarter-pel diamonds by SATD */
static me_t ps_refine(const ps_t *p, int sx, int sy, uint32_t start_cost, int qpel) {
    int best[2] = {sx, sy};
    uint32_t best_cost = start_cost;
    if ((p->mvp[0] != best[0] || p->mvp[1] != best[1]) && mvb_contains(&p->b, p->mvp[0], p->mvp[1])) {
        uint32_t c = ps_cost(p, p->mvp[0], p->mvp[1], 0) + ps_mv_cost(p, p->mvp[0], p->mvp[1]);
        if (c < best_cost) {
            best_cost = c;
            best[0] = p->mvp[0];
            best[1] = p->mvp[1];
        }
    }
    static const int DIRS[4][2] = {{0, -2}, {-2, 0}, {2, 0}, {0, 2}};
    int from[2] = {best[0], best[1]};
    for (int k = 0; k < 4; k++) {
        int mx = from[0] + DIRS[k][0], my = from[1] + DIRS[k][1];
        if (!mvb_contains(&p->b, mx, my)) continue;
        uint32_t c = ps_cost(p, mx, my, 0) + ps_mv_cost(p, mx, my);
        if (c < best_cost) {
            best_cost = c;
            best[0] = mx;
            best[1] = my;
        }
    }
    me_t m = {{best[0], best[1]}, ps_cost(p, best[0], best[1], 1), ps_mv_cost(p, best[0], best[1])};
    return ps_refine_qpel(p, m, qpel);
}
/* compact.rs:583 search_qpel: zero and the candidates (rounded to full-pel, clamped, distinct),
 * a radius-2 hexagon for up to range / 2 steps, a square of eight, then sub-pel refinement */
static me_t ps_search(ps_t *p, const int16_t (*cands)[2], int nc, int range, int qpel) {
    int pts[12][2] = {{0, 0}};
    int n = 1;
    for (int i = 0; i < nc; i++) {
        int m0 = iclamp((cands[i][0] + 2) >> 2, p->lo[0], p->hi[0]);
        int m1 = iclamp((cands[i][1] + 2) >> 2, p->lo[1], p->hi[1]);
        int dup = 0;
        for (int j = 0; j < n; j++)
            if (pts[j][0] == m0 && pts[j][1] == m1) dup = 1;
        if (n < 12 && !dup) {
            pts[n][0] = m0;
            pts[n][1] = m1;
            n++;
        }
    }
    for (int c = 0; c < n; c += 4) {
        int q[4][2];
        for (int k = 0; k < 4; k++) {
            int i = c + k < n ? c + k : 0;
            q[k][0] = pts[i][0];
            q[k][1] = pts[i][1];
        }
        fp_take(p, q);
    }
    static const int HEX[6][2] = {{-2, 0}, {-1, 2}, {1, 2}, {2, 0}, {1, -2}, {-1, -2}};
    static const int NEXT[6][3] = {{0, 1, 5}, {0, 1, 2}, {1, 2, 3}, {2, 3, 4}, {3, 4, 5}, {0, 4, 5}};
    int dir = -1;
    if (range >= 2) {
        int from[2] = {p->best[0], p->best[1]}, h[6][2];
        for (int k = 0; k < 6; k++) fp_at(p, from, HEX[k][0], HEX[k][1], h[k]);
        int qa[4][2] = {{h[0][0], h[0][1]}, {h[1][0], h[1][1]}, {h[2][0], h[2][1]}, {h[3][0], h[3][1]}};
        int qb[4][2] = {{h[4][0], h[4][1]}, {h[5][0], h[5][1]}, {from[0], from[1]}, {from[0], from[1]}};
        int a = fp_take(p, qa), b = fp_take(p, qb);
        dir = b >= 0 ? b + 4 : a;
    }
    for (int it = 1; it < range / 2 && dir >= 0; it++) {
        int from[2] = {p->best[0], p->best[1]}, q[4][2];
        for (int k = 0; k < 3; k++) fp_at(p, from, HEX[NEXT[dir][k]][0], HEX[NEXT[dir][k]][1], q[k]);
        q[3][0] = from[0];
        q[3][1] = from[1];
        int t = fp_take(p, q);
        dir = t >= 0 ? NEXT[dir][t] : -1;
    }
    {
        static const int SQ[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
        int from[2] = {p->best[0], p->best[1]}, qa[4][2], qb[4][2];
        for (int k = 0; k < 4; k++) {
            fp_at(p, from, SQ[k][0], SQ[k][1], qa[k]);
            fp_at(p, from, SQ[k + 4][0], SQ[k + 4][1], qb[k]);
        }
        fp_take(p, qa);
        fp_take(p, qb);
    }
    return ps_refine(p, p->best[0] * 4, p->best[1] * 4, p->best_cost, qpel);
}

#define COMPACT_ME_RANGE 16  /* compact.rs:758 ME_RANGE */
#define COMPACT_QPEL_REFINE 2 /* compact.rs:760 QPEL_REFINE */
#define P8X8_GATE 40         /* inter.rs:10 */

static void fill_part(ora_enc *e, int gx, int gy, int w4, int h4, int8_t r, int mx, int my) { /* h264.rs:3554 */
    int gs = e->mb_w * 4;
    for (int y = gy; y < gy + h4; y++)
        for (int x = gx; x < gx + w4; x++) {
            e->ref_grid[y * gs + x] = r;
            e->mv_grid[y * gs + x][0] = (int16_t)mx;
            e->mv_grid[y * gs + x][1] = (int16_t)my;
        }
}
/* The Compact decision record: words 2-5 hold the vectors of the four 8x8 quadrants (raster
 * order), word 6 the 16x16 search vector (mvr) the next picture takes as a candidate. */
static void set_compact_record(ora_mb_decision *d, const int q[4][2], const int16_t mvr[2]) {
    int16_t w[5][2];
    for (int k = 0; k < 4; k++) {
        w[k][0] = (int16_t)q[k][0];
        w[k][1] = (int16_t)q[k][1];
    }
    w[4][0] = mvr[0];
    w[4][1] = mvr[1];
    memcpy(d->mv, w[0], 4);
    memcpy(d->mvp, w[1], 4);
    memcpy(&d->cost, w[2], 4);
    memcpy(&d->n_sad, w[3], 4);
    memcpy(&d->n_satd, w[4], 4);
}

/* inter.rs:293 code_inter_p_body (CAVLC): P_L0_16x16 or P_8x8 (four P_L0_8x8) */
static void code_inter_p(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, int p8x8, const int q[4][2], int cbpl,
                         int cbpc, int16_t q4[16][16], const chroma_levels_t *c, uint8_t *py, uint8_t *pu, uint8_t *pv) {
    int idx = mb_y * e->mb_w + mb_x, gx0 = mb_x * 4, gy0 = mb_y * 4;
    int cbp = cbpl | (cbpc << 4);
    int qp_delta = cbp ? qp - e->cur_qp : 0, eff_qp = iclamp(e->cur_qp + qp_delta, 0, 51);
    int eff_qpc = chroma_qp(eff_qp);
    if (w) {
        put_ue(w, p8x8 ? 3 : 0);
        if (p8x8)
            for (int k = 0; k < 4; k++) put_ue(w, 0);
    }
    for (int part = 0; part < (p8x8 ? 4 : 1); part++) {
        int bx = p8x8 ? (part & 1) * 2 : 0, by = p8x8 ? (part >> 1) * 2 : 0, w4 = p8x8 ? 2 : 4;
        int16_t mvp[2];
        pred_mv_part(e, gx0 + bx, gy0 + by, w4, mvp);
        if (w) {
            put_se(w, q[part][0] - mvp[0]);
            put_se(w, q[part][1] - mvp[1]);
        }
        fill_part(e, gx0 + bx, gy0 + by, w4, w4, 0, q[part][0], q[part][1]);
    }
    if (w) put_ue(w, inter_cbp_golomb(cbp));
    if (cbp) {
        if (w) put_se(w, qp_delta);
        write_cavlc_residuals(e, w, mb_x, mb_y, cbpl, cbpc, q4, c);
    }
    e->mb_cbp_luma[idx] = (uint8_t)cbpl;
    e->mb_cbp_chroma[idx] = (uint8_t)cbpc;
    e->cur_qp = eff_qp;
    e->mb_qps[idx] = (uint8_t)eff_qp;
    recon_luma_4x4(py, q4, cbpl, eff_qp);
    recon_chroma(pu, pv, c, cbpc, eff_qpc);
    store_mb(e, mb_x, mb_y, py, pu, pv);
}

/* inter.rs:229 predict_inter (list 0 only), per 8x8 quadrant: luma from the reference and its
 * half-sample planes, chroma by the 1/8-sample bilinear filter. Motion compensation is per
 * sample, so quadrants give the prediction of any partitioning of equal motion. */
static void predict_quads(const ora_enc *e, int mb_x, int mb_y, const int q[4][2], uint8_t *py, uint8_t *pu, uint8_t *pv) {
    for (int k = 0; k < 4; k++) {
        int bx = (k & 1) * 8, by = (k >> 1) * 8;
        rl_pred(e, mb_x * 16 + bx, mb_y * 16 + by, 8, 8, q[k][0], q[k][1], py + by * 16 + bx, 16);
        int cx = mb_x * 8 + bx / 2, cy = mb_y * 8 + by / 2, fx = q[k][0] & 7, fy = q[k][1] & 7;
        int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
        for (int pl = 0; pl < 2; pl++) {
            const plane_t *cp = pl ? &e->refp.v : &e->refp.u;
            uint8_t *dst = (pl ? pv : pu) + (by / 2) * 8 + bx / 2;
            long off = pat(cp, cx + (q[k][0] >> 3), cy + (q[k][1] >> 3));
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) {
                    const uint8_t *s0 = cp->buf + off + (long)r * cp->stride + c, *s1 = s0 + cp->stride;
                    dst[r * 8 + c] = (uint8_t)((wa * s0[0] + wb * s0[1] + wc * s1[0] + wd * s1[1] + 32) >> 6);
                }
        }
    }
}

/* inter.rs:498 code_p_mb_body (P pictures, CAVLC) */
static void code_p_mb_compact(ora_enc *e, bits_t *w, int mb_x, int mb_y, int qp, const mbsrc_t *s) {
    int mb_w = e->mb_w, gs = mb_w * 4, gx0 = mb_x * 4, gy0 = mb_y * 4, x0 = mb_x * 16, y0 = mb_y * 16;
    int idx = mb_y * mb_w + mb_x;
    ora_mb_decision *d = &e->dec[idx];
    int16_t *lv = e->levels + (long)idx * ORA_LEVELS_PER_MB;
    memset(lv, 0, ORA_LEVELS_PER_MB * sizeof(int16_t));
    int qpc = chroma_qp(qp);
    mvb_t b;
    mvb_init(&b, x0, y0, e->refp.y.w, e->refp.y.h);
    uint32_t lambda = lambda_compact(qp);
    int16_t mvp[2];
    pred_mv_part(e, gx0, gy0, 4, mvp);
    int16_t pskip[2] = {mvp[0], mvp[1]};
    {
        int li = gy0 * gs + gx0 - 1, ti = (gy0 - 1) * gs + gx0;
        if (gx0 == 0 || gy0 == 0 || (e->ref_grid[li] == 0 && e->mv_grid[li][0] == 0 && e->mv_grid[li][1] == 0) ||
            (e->ref_grid[ti] == 0 && e->mv_grid[ti][0] == 0 && e->mv_grid[ti][1] == 0))
            pskip[0] = pskip[1] = 0;
    }
    d->qp = (uint8_t)qp;
    uint8_t py[256], pu[64], pv[64];
    if (e->sb.qp != qp) skip_bounds_init(&e->sb, qp, qpc, 1);
    /* The skip probe only next to a skipped macroblock (x264 fast P_Skip detection). */
    const uint8_t *sk = e->mb_is_skip;
    int probe = mvb_contains(&b, pskip[0], pskip[1]) &&
                ((mb_x > 0 && sk[idx - 1]) ||
                 (mb_y > 0 && (sk[idx - mb_w] || (mb_x > 0 && sk[idx - mb_w - 1]) || (mb_x + 1 < mb_w && sk[idx - mb_w + 1]))));
    const int skipq[4][2] = {{pskip[0], pskip[1]}, {pskip[0], pskip[1]}, {pskip[0], pskip[1]}, {pskip[0], pskip[1]}};
    if (try_p_skip(e, mb_x, mb_y, qp, qpc, s, pskip, probe, py, pu, pv)) {
        e->mvr[idx][0] = pskip[0];
        e->mvr[idx][1] = pskip[1];
        d->mb_type = 0;
        d->flags = 1;
        set_compact_record(d, skipq, pskip);
        return;
    }
    /* 1. 16x16 from the predictor, zero, the neighbours' and the previous picture's search vectors. */
    int16_t cands[11][2];
    int n = 0;
#define PUSH(v) do { cands[n][0] = (v)[0]; cands[n][1] = (v)[1]; n++; } while (0)
    PUSH(mvp);
    if (mb_x > 0) PUSH(e->mvr[idx - 1]);
    if (mb_y > 0) {
        PUSH(e->mvr[idx - mb_w]);
        if (mb_x > 0) PUSH(e->mvr[idx - mb_w - 1]);
        if (mb_x + 1 < mb_w) PUSH(e->mvr[idx - mb_w + 1]);
    }
    PUSH(e->prev_mvr[idx]);
    if (mb_x + 1 < mb_w) PUSH(e->prev_mvr[idx + 1]);
    if (mb_y + 1 < e->mb_h) PUSH(e->prev_mvr[idx + mb_w]);
#undef PUSH
    ensure_mv_costs_c(e, qp);
    ps_t ps;
    ps_init(&ps, e, s->y, x0, y0, 16, 16, mvp, lambda, &b);
    me_t m16 = ps_search(&ps, cands, n, COMPACT_ME_RANGE, 1);
    int16_t mvr16[2] = {(int16_t)m16.mv[0], (int16_t)m16.mv[1]};
    e->mvr[idx][0] = mvr16[0];
    e->mvr[idx][1] = mvr16[1];
    /* A 16x16 vector at the skip vector whose residual decimates away is P_Skip. */
    if (!probe && m16.mv[0] == pskip[0] && m16.mv[1] == pskip[1] && mvb_contains(&b, pskip[0], pskip[1]) &&
        try_p_skip(e, mb_x, mb_y, qp, qpc, s, pskip, 1, py, pu, pv)) {
        d->mb_type = 0;
        set_compact_record(d, skipq, mvr16);
        return;
    }
    int p8x8 = 0, q[4][2];
    for (int k = 0; k < 4; k++) {
        q[k][0] = m16.mv[0];
        q[k][1] = m16.mv[1];
    }
    uint32_t best_cost = me_cost(&m16);
    me_t best_me[4] = {m16, m16, m16, m16};
    /* 2. Four 8x8 searches seeded by the 16x16 vector and the partitions before them. */
    if (best_cost > lambda * P8X8_GATE) {
        me_t m8[4];
        int16_t cands8[5][2];
        cands8[0][0] = (int16_t)m16.mv[0];
        cands8[0][1] = (int16_t)m16.mv[1];
        for (int i = 0; i < 4; i++) {
            int bx = (i & 1) * 2, by = (i >> 1) * 2;
            int16_t mvp8[2];
            pred_mv_part(e, gx0 + bx, gy0 + by, 2, mvp8);
            ps_t p8;
            ps_init(&p8, e, s->y + by * 4 * 16 + bx * 4, x0 + bx * 4, y0 + by * 4, 8, 8, mvp8, lambda, &b);
            m8[i] = ps_search(&p8, cands8, i + 1, COMPACT_ME_RANGE, 1);
            cands8[i + 1][0] = (int16_t)m8[i].mv[0];
            cands8[i + 1][1] = (int16_t)m8[i].mv[1];
            fill_part(e, gx0 + bx, gy0 + by, 2, 2, 0, m8[i].mv[0], m8[i].mv[1]);
        }
        uint32_t cost8 = 0;
        for (int i = 0; i < 4; i++) cost8 += me_cost(&m8[i]);
        if (cost8 < best_cost) {
            p8x8 = 1;
            for (int i = 0; i < 4; i++) {
                q[i][0] = m8[i].mv[0];
                q[i][1] = m8[i].mv[1];
                best_me[i] = m8[i];
            }
        }
    }
    /* Quarter-pel refinement of the partitions chosen, each from the final vectors before it. */
    best_cost = 0;
    for (int part = 0; part < (p8x8 ? 4 : 1); part++) {
        int bx = p8x8 ? (part & 1) * 2 : 0, by = p8x8 ? (part >> 1) * 2 : 0, w4 = p8x8 ? 2 : 4;
        int16_t mvpp[2];
        pred_mv_part(e, gx0 + bx, gy0 + by, w4, mvpp);
        ps_t pr;
        ps_init(&pr, e, s->y + by * 4 * 16 + bx * 4, x0 + bx * 4, y0 + by * 4, w4 * 4, w4 * 4, mvpp, lambda, &b);
        me_t r = ps_refine_qpel(&pr, best_me[part], COMPACT_QPEL_REFINE);
        best_cost += me_cost(&r);
        fill_part(e, gx0 + bx, gy0 + by, w4, w4, 0, r.mv[0], r.mv[1]);
        for (int k = 0; k < 4; k++) {
            int kx = (k & 1) * 2, ky = (k >> 1) * 2;
            if (kx >= bx && kx < bx + w4 && ky >= by && ky < by + w4) {
                q[k][0] = r.mv[0];
                q[k][1] = r.mv[1];
            }
        }
    }
    /* 3. Intra 16x16 and 4x4 against the best inter cost. */
    edges_t ed;
    get_edges(&e->cur.y, mb_x, mb_y, 16, &ed);
    uint32_t i16_satd;
    int i16_mode;
    analyse_i16(&ed, s->y, &i16_satd, &i16_mode, NULL, 1);
    uint32_t i16_cost = i16_satd + lambda * 4;
    static const int zq[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (i16_cost < best_cost) {
        begin_coded_mb(e, w);
        code_i16(e, w, mb_x, mb_y, qp, s, i16_mode);
        set_compact_record(d, zq, mvr16);
        return;
    }
    uint32_t lo = i16_cost < best_cost ? i16_cost : best_cost;
    if (lo > lambda * 40 && (uint64_t)i16_cost * 2 < (uint64_t)best_cost * 3) {
        i4_choice_t ch;
        if (analyse_i4(e, mb_x, mb_y, qp, s, lo, &ch)) {
            begin_coded_mb(e, w);
            code_i4(e, w, mb_x, mb_y, qp, s, &ch);
            uint64_t packed;
            memcpy(&packed, d->mv, 8);
            set_compact_record(d, zq, mvr16);
            memcpy(d->mv, &packed, 8);
            return;
        }
    }
    /* 4. Prediction, residual and syntax. */
    predict_quads(e, mb_x, mb_y, (const int(*)[2])q, py, pu, pv);
    int16_t q4[16][16];
    int cbpl = quant_luma_4x4_inter(s->y, py, qp, q4, 1);
    chroma_levels_t c;
    quant_chroma_inter(s, pu, pv, qpc, &c, 1);
    int cbpc = chroma_cbp(&c);
    if (cbpl == 0 && cbpc == 0 && !p8x8 && q[0][0] == pskip[0] && q[0][1] == pskip[1]) {
        code_skip(e, mb_x, mb_y, pskip);
        store_mb(e, mb_x, mb_y, py, pu, pv);
        d->mb_type = 0;
        set_compact_record(d, skipq, mvr16);
        return;
    }
    begin_coded_mb(e, w);
    code_inter_p(e, w, mb_x, mb_y, qp, p8x8, (const int(*)[2])q, cbpl, cbpc, q4, &c, py, pu, pv);
    d->mb_type = p8x8 ? 5 : 1;
    d->cbp_luma = (uint8_t)cbpl;
    d->cbp_chroma = (uint8_t)cbpc;
    d->qp = e->mb_qps[idx];
    set_compact_record(d, (const int(*)[2])q, mvr16);
    rec_luma4(lv, q4);
    rec_chroma(lv, &c);
}

/* inter.rs:781 deblock_parts (P pictures): the boundary strength of each edge segment from the
 * coefficients and the motion of the two 4x4 blocks beside it */
static int uniform_inter(const ora_enc *e, long g0) { /* inter.rs:768 */
    int gs = e->mb_w * 4;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            long g = g0 + r * gs + c;
            if (e->luma_nnz[g] || e->ref_grid[g] != e->ref_grid[g0] || e->mv_grid[g][0] != e->mv_grid[g0][0] ||
                e->mv_grid[g][1] != e->mv_grid[g0][1])
                return 0;
        }
    return 1;
}
static void deblock_parts(ora_enc *e) {
    int mb_w = e->mb_w, mb_h = e->mb_h, gs = mb_w * 4;
    for (int mb_y = 0; mb_y < mb_h; mb_y++)
        for (int mb_x = 0; mb_x < mb_w; mb_x++) {
            int m = mb_y * mb_w + mb_x;
            long g0 = (long)mb_y * 4 * gs + mb_x * 4;
            int intra = e->mb_is_intra[m], t8 = e->mb_8x8[m];
            int qp_q = e->mb_qps[m], qpc_q = chroma_qp(qp_q);
            int edges = (!intra && uniform_inter(e, g0)) ? 1 : 4;
            for (int dir = 0; dir < 2; dir++)
                for (int ed = 0; ed < edges; ed++) {
                    int pm;
                    if (ed == 0) {
                        if (dir == 0 && mb_x > 0)
                            pm = m - 1;
                        else if (dir == 1 && mb_y > 0)
                            pm = m - mb_w;
                        else
                            continue;
                    } else {
                        if ((ed & 1) && t8) continue;
                        pm = m;
                    }
                    int bs[4], any = 0;
                    for (int k = 0; k < 4; k++) {
                        if (intra || e->mb_is_intra[pm]) {
                            bs[k] = ed == 0 ? 4 : 3;
                        } else {
                            long q = dir == 0 ? g0 + k * gs + ed : g0 + ed * gs + k;
                            long pp = dir == 0 ? q - 1 : q - gs;
                            if (e->luma_nnz[q] || e->luma_nnz[pp])
                                bs[k] = 2;
                            else
                                bs[k] = e->ref_grid[q] != e->ref_grid[pp] || iabs(e->mv_grid[q][0] - e->mv_grid[pp][0]) >= 4 ||
                                        iabs(e->mv_grid[q][1] - e->mv_grid[pp][1]) >= 4;
                        }
                        any |= bs[k];
                    }
                    if (!any) continue;
                    int qp, qpc;
                    if (pm == m) {
                        qp = qp_q;
                        qpc = qpc_q;
                    } else {
                        int qp_p = e->mb_qps[pm];
                        qp = (qp_q + qp_p + 1) >> 1;
                        qpc = (qpc_q + chroma_qp(qp_p) + 1) >> 1;
                    }
                    int x = dir == 0 ? mb_x * 16 + ed * 4 : mb_x * 16, y = dir == 0 ? mb_y * 16 : mb_y * 16 + ed * 4;
                    edge_filter(&e->cur.y, x, y, dir, 0, qp, bs);
                    if (!(ed & 1)) {
                        edge_filter(&e->cur.u, x / 2, y / 2, dir, 1, qpc, bs);
                        edge_filter(&e->cur.v, x / 2, y / 2, dir, 1, qpc, bs);
                    }
                }
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Headers and frame driver (h264.rs:239-1127, 3032-3165; enc.rs:2234-2514, 2643-2672).        */
/* ------------------------------------------------------------------------------------------ */
static void nal_append(bits_t *out, int ref_idc, int type, const uint8_t *rbsp, long n) { /* enc.rs:2643 */
    bits_push(out, 0);
    bits_push(out, 0);
    bits_push(out, 0);
    bits_push(out, 1);
    bits_push(out, (uint8_t)((ref_idc << 5) | type));
    int zeros = 0;
    for (long i = 0; i < n; i++) {
        uint8_t b = rbsp[i];
        if (zeros == 2 && b <= 3) {
            bits_push(out, 3);
            zeros = 0;
        }
        bits_push(out, b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
}
/* h264.rs:556 Sps::to_rbsp with new_profile's fields */
static void write_sps(const ora_enc *e, bits_t *w) {
    int profile_idc = e->profile == 0 ? 66 : (e->profile == 1 ? 77 : 100);
    int csf = e->profile == 0 ? 0xC0 : (e->profile == 1 ? 0x40 : 0x00);
    put32(w, (uint32_t)profile_idc, 8);
    put32(w, (uint32_t)csf, 8);
    put32(w, 40, 8);
    put_ue(w, 0);
    if (profile_idc == 100) {
        put_ue(w, 1);
        put_ue(w, 0);
        put_ue(w, 0);
        put32(w, 0, 1);
        put32(w, 0, 1);
    }
    put_ue(w, 0); /* log2_max_frame_num - 4 */
    put_ue(w, 0); /* poc type */
    put_ue(w, 2); /* log2_max_poc_lsb - 4 */
    put_ue(w, 2); /* max_num_ref_frames */
    put32(w, 0, 1);
    put_ue(w, (uint64_t)(e->mb_w - 1));
    put_ue(w, (uint64_t)(e->mb_h - 1));
    put32(w, 1, 1); /* frame_mbs_only */
    put32(w, 1, 1); /* direct_8x8_inference */
    int crop_r = (e->mb_w * 16 - e->width) / 2, crop_b = (e->mb_h * 16 - e->height) / 2;
    int has_crop = crop_r || crop_b;
    put32(w, (uint32_t)has_crop, 1);
    if (has_crop) {
        put_ue(w, 0);
        put_ue(w, (uint64_t)crop_r);
        put_ue(w, 0);
        put_ue(w, (uint64_t)crop_b);
    }
    put32(w, 1, 1); /* vui */
    put32(w, 1, 1); /* aspect_ratio_info_present */
    put32(w, 1, 8); /* SAR 1:1 */
    put32(w, 0, 1); /* overscan */
    put32(w, 1, 1); /* video_signal_type */
    put32(w, 5, 3);
    put32(w, 0, 1); /* limited range */
    put32(w, 1, 1); /* colour description: BT.709 */
    put32(w, 1, 8);
    put32(w, 1, 8);
    put32(w, 1, 8);
    put32(w, 0, 1); /* chroma_loc */
    put32(w, 1, 1); /* timing info */
    put_bits(w, (uint64_t)e->fps_den, 32);
    put_bits(w, (uint64_t)e->fps_num * 2, 32);
    put32(w, 1, 1);
    put32(w, 0, 1);
    put32(w, 0, 1);
    put32(w, 0, 1);
    put32(w, 0, 1); /* bitstream_restriction */
    bits_trailing(w);
}
/* h264.rs:806 Pps::to_rbsp */
static void write_pps(const ora_enc *e, bits_t *w) {
    put_ue(w, 0);
    put_ue(w, 0);
    put32(w, (uint32_t)e->cabac, 1);
    put32(w, 0, 1);
    put_ue(w, 0);
    put_ue(w, 0);
    put_ue(w, 0);
    put32(w, 0, 1);
    put32(w, 0, 2);
    put_se(w, e->qp - 26);
    put_se(w, 0);
    put_se(w, e->cqo); /* chroma_qp_index_offset (h264.rs:3087 sync_reorder) */
    put32(w, 1, 1);
    put32(w, 0, 1);
    put32(w, 0, 1);
    if (e->t8x8) {
        put32(w, 1, 1);
        put32(w, 0, 1);
        put_se(w, e->cqo);
    }
    bits_trailing(w);
}
long ora_headers(ora_enc *e, uint8_t *out, long cap) {
    bits_t o, s;
    bits_init(&o);
    bits_init(&s);
    write_sps(e, &s);
    nal_append(&o, 3, 7, s.buf, s.len);
    s.len = 0;
    s.n = 0;
    s.acc = 0;
    write_pps(e, &s);
    nal_append(&o, 3, 8, s.buf, s.len);
    long n = o.len;
    if (n <= cap) memcpy(out, o.buf, (size_t)n);
    bits_free(&o);
    bits_free(&s);
    return n <= cap ? n : -1;
}

static void copy_plane(const plane_t *p, uint8_t *dst) {
    for (int r = 0; r < p->h; r++) memcpy(dst + (long)r * p->w, p->buf + pat(p, 0, r), (size_t)p->w);
}

/* ------------------------------------------------------------------------------------------ */
/* CABAC slice data (Main/High): the syntax of code_skip / begin_coded_mb / code_inter /        */
/* code_i4 / code_i16 (enc.rs:2779-3306, :3441-3720) with encode_cabac_cbp_and_residuals and   */
/* encode_cabac_chroma_residuals (h264.rs:3327-3552), written after the frame's analysis from   */
/* its decision records and levels (entropy coding feeds nothing back into Balanced decisions). */
/* The arithmetic coder is ITU-T H.264 9.3.4.2 in its bit-serial form (PutBit with              */
/* bitsOutstanding); the reference (cabac.rs:504-611) writes the same bits byte by byte.        */
/* ------------------------------------------------------------------------------------------ */
#include "h264_cabac_tables.h"

typedef struct {
    bits_t *w;
    uint32_t low, range;
    int first, outstanding;
    uint8_t st[460]; /* pStateIdx << 1 | valMPS */
} cabac_t;

/* 9.3.1.1: context variables from (m, n) and SliceQPY; 9.3.4.1: encoder registers. */
static void cabac_init(cabac_t *c, bits_t *w, int slice_qp, int table) {
    c->w = w;
    c->low = 0;
    c->range = 510;
    c->first = 1;
    c->outstanding = 0;
    int q = iclamp(slice_qp, 0, 51);
    for (int i = 0; i < 460; i++) {
        int m = CABAC_INIT_MN[table][i][0], n = CABAC_INIT_MN[table][i][1];
        int pre = iclamp(((m * q) >> 4) + n, 1, 126);
        c->st[i] = (uint8_t)(pre <= 63 ? (63 - pre) << 1 : ((pre - 64) << 1) | 1);
    }
}
static void cabac_put_bit(cabac_t *c, int b) { /* 9.3.4.2 PutBit */
    if (c->first)
        c->first = 0;
    else
        put32(c->w, (uint32_t)b, 1);
    for (; c->outstanding > 0; c->outstanding--) put32(c->w, (uint32_t)(1 - b), 1);
}
static void cabac_renorm(cabac_t *c) { /* 9.3.4.2 RenormE */
    while (c->range < 256) {
        if (c->low < 256) {
            cabac_put_bit(c, 0);
        } else if (c->low >= 512) {
            c->low -= 512;
            cabac_put_bit(c, 1);
        } else {
            c->low -= 256;
            c->outstanding++;
        }
        c->range <<= 1;
        c->low <<= 1;
    }
}
static void cabac_bin(cabac_t *c, int ctx, int bin) { /* 9.3.4.2 EncodeDecision */
    int s = c->st[ctx] >> 1, mps = c->st[ctx] & 1;
    uint32_t rlps = CABAC_RANGE_LPS[s][(c->range >> 6) & 3];
    c->range -= rlps;
    if (bin != mps) {
        c->low += c->range;
        c->range = rlps;
        if (s == 0) mps = 1 - mps;
        s = CABAC_NEXT_LPS[s];
    } else if (s < 62) {
        s++;
    }
    c->st[ctx] = (uint8_t)((s << 1) | mps);
    cabac_renorm(c);
}
static void cabac_bypass(cabac_t *c, int bin) { /* 9.3.4.4 */
    c->low <<= 1;
    if (bin) c->low += c->range;
    if (c->low >= 1024) {
        cabac_put_bit(c, 1);
        c->low -= 1024;
    } else if (c->low < 512) {
        cabac_put_bit(c, 0);
    } else {
        c->low -= 512;
        c->outstanding++;
    }
}
static void cabac_terminate(cabac_t *c, int bin) { /* 9.3.4.5, EncodeFlush writes rbsp_stop_one_bit */
    c->range -= 2;
    if (bin) {
        c->low += c->range;
        c->range = 2;
        cabac_renorm(c);
        cabac_put_bit(c, (int)((c->low >> 9) & 1));
        put32(c->w, ((c->low >> 7) & 3) | 1, 2);
    } else {
        cabac_renorm(c);
    }
}

/* I-slice mb_type (cabac.rs:738 encode_intra_mb_type; Table 9-36 binarization, ctxIdx 3-10,
 * or 17-20 after the P-slice prefix): 0 I_NxN, 1..24 I_16x16. */
static void cabac_intra_type(cabac_t *c, int in_p, int inc, int t) {
    int b0 = in_p ? 17 : 3 + inc;
    cabac_bin(c, b0, t != 0);
    if (!t) return;
    cabac_terminate(c, 0);
    int k = t - 1, mode = k & 3, cc = (k >> 2) % 3, luma = k >= 12;
    cabac_bin(c, in_p ? 18 : 6, luma);
    cabac_bin(c, in_p ? 19 : 7, cc != 0);
    if (cc) cabac_bin(c, in_p ? 19 : 8, cc == 2);
    cabac_bin(c, in_p ? 20 : 9, mode >> 1);
    cabac_bin(c, in_p ? 20 : 10, mode & 1);
}
/* Exp-Golomb suffix of a UEGk binarization in bypass bins (9.3.2.3). */
static void cabac_eg(cabac_t *c, uint32_t v, int k) {
    while (v >= (1u << k)) {
        cabac_bypass(c, 1);
        v -= 1u << k;
        k++;
    }
    cabac_bypass(c, 0);
    while (k--) cabac_bypass(c, (int)((v >> k) & 1));
}
/* mvd_l0 component (cabac.rs:940; UEG3, uCoff 9, signed; ctxIdx 40-46 / 47-53). */
static void cabac_mvd(cabac_t *c, int comp, int amvd, int v) {
    int base = comp ? 47 : 40, a = iabs(v);
    cabac_bin(c, base + (amvd < 3 ? 0 : amvd <= 32 ? 1 : 2), a != 0);
    if (!a) return;
    for (int i = 1; i < 9 && i < a; i++) cabac_bin(c, base + 3 + imin(i - 1, 3), 1);
    if (a < 9)
        cabac_bin(c, base + 3 + imin(a - 1, 3), 0);
    else
        cabac_eg(c, (uint32_t)(a - 9), 3);
    cabac_bypass(c, v < 0);
}

/* One residual block (cabac.rs:981 encode_residual_block_cat, 9.3.2.3 / 9.3.3.1.3): coded_block_flag
 * (not for 8x8 blocks), the significance map, then levels from the last one down. */
static const int CB_CBF[5] = {85, 89, 93, 97, 101};
static const int CB_SIG[6] = {105, 120, 134, 149, 152, 402};
static const int CB_LAST[6] = {166, 181, 195, 210, 213, 417};
static const int CB_ABS[6] = {227, 237, 247, 257, 266, 426};
/* Table 9-43, frame-coded 8x8 blocks: ctxIdxInc of significant_coeff_flag and of
 * last_significant_coeff_flag by scan position. */
static const uint8_t SIG8_INC[63] = {0,  1,  2,  3,  4,  5,  5,  4,  4,  3,  3,  4,  4,  4,  5,  5,
                                     4,  4,  4,  4,  3,  3,  6,  7,  7,  7,  8,  9,  10, 9,  8,  7,
                                     7,  6,  11, 12, 13, 11, 6,  7,  8,  9,  14, 10, 9,  8,  6,  11,
                                     12, 13, 11, 6,  9,  14, 10, 9,  11, 12, 13, 11, 14, 10, 12};
static const uint8_t LAST8_INC[63] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2,
                                      2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 4, 4,
                                      4, 4, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 8, 8, 8};
static int cabac_block(cabac_t *c, int cat, int cbf_inc, const int16_t *sc, int n) {
    int last = n - 1;
    while (last >= 0 && sc[last] == 0) last--;
    if (cat < 5) {
        cabac_bin(c, CB_CBF[cat] + cbf_inc, last >= 0);
        if (last < 0) return 0;
    }
    for (int i = 0; i < n - 1 && i <= last; i++) {
        int si = cat == 5 ? SIG8_INC[i] : (cat == 3 ? imin(i, 2) : i);
        int li = cat == 5 ? LAST8_INC[i] : (cat == 3 ? imin(i, 2) : i);
        cabac_bin(c, CB_SIG[cat] + si, sc[i] != 0);
        if (sc[i]) cabac_bin(c, CB_LAST[cat] + li, i == last);
    }
    int eq1 = 0, gt1 = 0, gmax = cat == 3 ? 3 : 4, base = CB_ABS[cat];
    for (int i = last; i >= 0; i--) {
        if (!sc[i]) continue;
        int a = iabs(sc[i]) - 1;
        cabac_bin(c, base + (gt1 ? 0 : imin(4, 1 + eq1)), a != 0);
        if (a) {
            int c1 = base + 5 + imin(gt1, gmax);
            for (int k = 1; k < 14 && k < a; k++) cabac_bin(c, c1, 1);
            if (a < 14)
                cabac_bin(c, c1, 0);
            else
                cabac_eg(c, (uint32_t)(a - 14), 0);
            gt1++;
        } else {
            eq1++;
        }
        cabac_bypass(c, sc[i] < 0);
    }
    return 1;
}

/* Neighbour state of the picture as coded so far, for the ctxIdxInc derivations of 9.3.3.1.1. */
typedef struct {
    int mb_w, mb_h;
    uint8_t *skip, *intra, *i16, *cbpl, *cbpc, *cmode, *t8, *dc_cbf, *cdc_cbf[2];
    uint8_t *cbf4;        /* luma 4x4 (8x8 blocks: all four entries) coded_block_flag, 4 mb_w x 4 mb_h */
    uint8_t *ccbf[2];     /* chroma AC coded_block_flag, 2 mb_w x 2 mb_h per plane */
    uint8_t *modes;       /* intra 4x4 modes (DC for other macroblocks) */
    int16_t (*mvd)[2];    /* mvd of each 4x4 block (0 for skip and intra) */
} cabac_nb_t;

static void cabac_nb_init(cabac_nb_t *s, int mb_w, int mb_h) {
    int n = mb_w * mb_h;
    s->mb_w = mb_w;
    s->mb_h = mb_h;
    uint8_t **f[] = {&s->skip, &s->intra, &s->i16, &s->cbpl, &s->cbpc, &s->cmode, &s->t8, &s->dc_cbf, &s->cdc_cbf[0],
                     &s->cdc_cbf[1]};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) *f[i] = (uint8_t *)calloc((size_t)n, 1);
    s->cbf4 = (uint8_t *)calloc((size_t)n * 16, 1);
    s->ccbf[0] = (uint8_t *)calloc((size_t)n * 4, 1);
    s->ccbf[1] = (uint8_t *)calloc((size_t)n * 4, 1);
    s->modes = (uint8_t *)malloc((size_t)n * 16);
    memset(s->modes, 2, (size_t)n * 16);
    s->mvd = calloc((size_t)n * 16, sizeof(*s->mvd));
}
static void cabac_nb_free(cabac_nb_t *s) {
    uint8_t *f[] = {s->skip, s->intra, s->i16, s->cbpl, s->cbpc, s->cmode, s->t8, s->dc_cbf, s->cdc_cbf[0],
                    s->cdc_cbf[1], s->cbf4, s->ccbf[0], s->ccbf[1], s->modes};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) free(f[i]);
    free(s->mvd);
}

/* coded_block_flag ctxIdxInc of a luma 4x4 block at grid (gx, gy) (9.3.3.1.1.9 for
 * ctxBlockCat 1 and 2): a neighbour block counts its flag when its macroblock is available,
 * not skipped, and its 8x8 quadrant is coded (an 8x8-transform quadrant's flag is 1); a
 * neighbour outside the picture counts 1 for an intra macroblock, 0 for an inter one. */
static int cbf4_cond(const cabac_nb_t *s, int gx, int gy, int intra) {
    if (gx < 0 || gy < 0) return intra;
    int m = (gy >> 2) * s->mb_w + (gx >> 2);
    if (s->skip[m]) return 0;
    int q = ((gy & 2) ? 2 : 0) + ((gx & 2) ? 1 : 0);
    if (!((s->cbpl[m] >> q) & 1)) return 0;
    return s->cbf4[gy * s->mb_w * 4 + gx];
}
/* Chroma AC (ctxBlockCat 4): the neighbour's flag when it codes chroma AC. */
static int ccbf_cond(const cabac_nb_t *s, int p, int cx, int cy, int intra) {
    if (cx < 0 || cy < 0) return intra;
    int m = (cy >> 1) * s->mb_w + (cx >> 1);
    if (s->skip[m] || s->cbpc[m] != 2) return 0;
    return s->ccbf[p][cy * s->mb_w * 2 + cx];
}

/* coded_block_pattern (9.3.2.6 / 9.3.3.1.1.4), transform_size_8x8_flag for inter macroblocks,
 * mb_qp_delta and the residual blocks: the part of a macroblock after its prediction syntax. */
static void cabac_cbp_residual(cabac_t *c, cabac_nb_t *s, int mb_x, int mb_y, int intra, int t8_flag_here,
                               int cbpl, int cbpc, int use8, const int16_t *lv, int *prev_dqp) {
    int mw = s->mb_w, m = mb_y * mw + mb_x;
    for (int q = 0; q < 4; q++) {
        int bx = q & 1, by = q >> 1, a, b;
        if (bx) a = !((cbpl >> (q - 1)) & 1);
        else a = mb_x > 0 && (s->skip[m - 1] || !((s->cbpl[m - 1] >> (by * 2 + 1)) & 1));
        if (by) b = !((cbpl >> (q - 2)) & 1);
        else b = mb_y > 0 && (s->skip[m - mw] || !((s->cbpl[m - mw] >> (2 + bx)) & 1));
        cabac_bin(c, 73 + a + 2 * b, (cbpl >> q) & 1);
    }
    int la = mb_x > 0 && !s->skip[m - 1] ? s->cbpc[m - 1] : 0;
    int ta = mb_y > 0 && !s->skip[m - mw] ? s->cbpc[m - mw] : 0;
    cabac_bin(c, 77 + (la != 0) + 2 * (ta != 0), cbpc != 0);
    if (cbpc) cabac_bin(c, 81 + (la == 2) + 2 * (ta == 2), cbpc == 2);
    s->cbpl[m] = (uint8_t)cbpl;
    s->cbpc[m] = (uint8_t)cbpc;
    if (!cbpl && !cbpc) {
        *prev_dqp = 0;
        return;
    }
    if (t8_flag_here && cbpl) {
        int inc = (mb_x > 0 && s->t8[m - 1]) + (mb_y > 0 && s->t8[m - mw]);
        cabac_bin(c, 399 + inc, use8);
    }
    cabac_bin(c, 60 + (*prev_dqp != 0), 0); /* mb_qp_delta 0: fixed QP */
    *prev_dqp = 0;
    int gs = mw * 4;
    if (use8) {
        for (int q = 0; q < 4; q++) {
            int gx = mb_x * 4 + (q & 1) * 2, gy = mb_y * 4 + (q >> 1) * 2, on = (cbpl >> q) & 1;
            if (on) cabac_block(c, 5, 0, lv + ORA_LV_LUMA + q * 64, 64);
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) s->cbf4[(gy + dy) * gs + gx + dx] = (uint8_t)on;
        }
    } else {
        for (int b = 0; b < 16; b++) {
            int bx, by;
            blk_xy(b, &bx, &by);
            int gx = mb_x * 4 + bx, gy = mb_y * 4 + by, f = 0;
            if ((cbpl >> (b >> 2)) & 1) {
                int inc = cbf4_cond(s, gx - 1, gy, intra) + 2 * cbf4_cond(s, gx, gy - 1, intra);
                f = cabac_block(c, 2, inc, lv + ORA_LV_LUMA + b * 16, 16);
            }
            s->cbf4[gy * gs + gx] = (uint8_t)f;
        }
    }
    if (cbpc >= 1)
        for (int p = 0; p < 2; p++) {
            int a = mb_x > 0 ? (!s->skip[m - 1] && s->cdc_cbf[p][m - 1]) : intra;
            int b = mb_y > 0 ? (!s->skip[m - mw] && s->cdc_cbf[p][m - mw]) : intra;
            s->cdc_cbf[p][m] = (uint8_t)cabac_block(c, 3, a + 2 * b, lv + ORA_LV_CHROMA_DC + p * 4, 4);
        }
    if (cbpc >= 2)
        for (int p = 0; p < 2; p++)
            for (int b = 0; b < 4; b++) {
                int cx = mb_x * 2 + (b & 1), cy = mb_y * 2 + (b >> 1);
                int inc = ccbf_cond(s, p, cx - 1, cy, intra) + 2 * ccbf_cond(s, p, cx, cy - 1, intra);
                s->ccbf[p][cy * mw * 2 + cx] =
                    (uint8_t)cabac_block(c, 4, inc, lv + ORA_LV_CHROMA_AC + (p * 4 + b) * 16 + 1, 15);
            }
}

static void cabac_chroma_mode(cabac_t *c, const cabac_nb_t *s, int mb_x, int mb_y, int mode) {
    int mw = s->mb_w, m = mb_y * mw + mb_x;
    int a = mb_x > 0 && s->intra[m - 1] && s->cmode[m - 1] != 0;
    int b = mb_y > 0 && s->intra[m - mw] && s->cmode[m - mw] != 0;
    cabac_bin(c, 64 + a + b, mode != 0);
    if (mode) {
        cabac_bin(c, 67, mode != 1);
        if (mode != 1) cabac_bin(c, 67, mode == 3);
    }
}

/* The slice data of the picture just analysed: every macroblock's syntax, then end_of_slice_flag
 * (terminate) after each (enc.rs:2382-2434). */
static void cabac_slice(ora_enc *e, bits_t *w, int is_idr, int slice_qp) {
    cabac_t c;
    cabac_init(&c, w, slice_qp, is_idr ? 0 : 1);
    cabac_nb_t s;
    int mw = e->mb_w, mh = e->mb_h, gs = mw * 4, prev_dqp = 0;
    cabac_nb_init(&s, mw, mh);
    for (int mb_y = 0; mb_y < mh; mb_y++)
        for (int mb_x = 0; mb_x < mw; mb_x++) {
            int m = mb_y * mw + mb_x;
            const ora_mb_decision *d = &e->dec[m];
            const int16_t *lv = e->levels + (size_t)m * ORA_LEVELS_PER_MB;
            int sa = mb_x > 0 && !s.skip[m - 1], sb = mb_y > 0 && !s.skip[m - mw];
            if (!is_idr) cabac_bin(&c, 11 + sa + sb, d->mb_type == 0);
            if (d->mb_type == 0) {
                s.skip[m] = 1;
                prev_dqp = 0;
            } else if (d->mb_type == 1 && !e->parts) { /* P_L0_16x16 (enc.rs:3196 code_inter) */
                cabac_bin(&c, 14, 0);
                cabac_bin(&c, 15, 0);
                cabac_bin(&c, 16, 0);
                int gx = mb_x * 4, gy = mb_y * 4, mvd[2] = {d->mv[0] - d->mvp[0], d->mv[1] - d->mvp[1]};
                for (int k = 0; k < 2; k++) {
                    int amvd = (gx > 0 ? iabs(s.mvd[gy * gs + gx - 1][k]) : 0) + (gy > 0 ? iabs(s.mvd[(gy - 1) * gs + gx][k]) : 0);
                    cabac_mvd(&c, k, amvd, mvd[k]);
                }
                int use8 = d->transform_8x8 && d->cbp_luma;
                s.t8[m] = (uint8_t)use8;
                cabac_cbp_residual(&c, &s, mb_x, mb_y, 0, e->t8x8, d->cbp_luma, d->cbp_chroma, use8, lv, &prev_dqp);
                for (int by = 0; by < 4; by++)
                    for (int bx = 0; bx < 4; bx++) {
                        s.mvd[(gy + by) * gs + gx + bx][0] = (int16_t)mvd[0];
                        s.mvd[(gy + by) * gs + gx + bx][1] = (int16_t)mvd[1];
                    }
            } else if (d->mb_type == 1 || d->mb_type >= 5) {
                /* Partitions on: P_L0_16x16, P_L0_L0_16x8, P_L0_L0_8x16 or P_8x8 (four P_L0_8x8);
                 * mb_type bins of Table 9-37 (ctxIdx 14-17), sub_mb_type "1" (ctxIdx 21), then
                 * each partition's mvd against its 8.4.1.3 predictor over the final motion grid. */
                int shape = d->mb_type == 1 ? 0 : (d->mb_type == 6 ? 1 : (d->mb_type == 7 ? 2 : 3));
                cabac_bin(&c, 14, 0);
                cabac_bin(&c, 15, shape == 1 || shape == 2);
                if (shape == 1 || shape == 2)
                    cabac_bin(&c, 17, shape == 1);
                else
                    cabac_bin(&c, 16, shape == 3);
                if (shape == 3)
                    for (int k = 0; k < 4; k++) cabac_bin(&c, 21, 1);
                uint32_t done = 0;
                for (int p = 0; p < PART_COUNT[shape]; p++) {
                    int bx, by, pw, ph;
                    part_rect(shape, p, &bx, &by, &pw, &ph);
                    int gx = mb_x * 4 + bx / 4, gy = mb_y * 4 + by / 4;
                    int16_t mvp[2];
                    pred_mv_shape(e, gx, gy, pw / 4, shape, p, done, mvp);
                    int mvd[2] = {e->mv_grid[gy * gs + gx][0] - mvp[0], e->mv_grid[gy * gs + gx][1] - mvp[1]};
                    for (int k = 0; k < 2; k++) {
                        int amvd = (gx > 0 ? iabs(s.mvd[gy * gs + gx - 1][k]) : 0) + (gy > 0 ? iabs(s.mvd[(gy - 1) * gs + gx][k]) : 0);
                        cabac_mvd(&c, k, amvd, mvd[k]);
                    }
                    for (int yy = 0; yy < ph / 4; yy++)
                        for (int xx = 0; xx < pw / 4; xx++) {
                            s.mvd[(gy + yy) * gs + gx + xx][0] = (int16_t)mvd[0];
                            s.mvd[(gy + yy) * gs + gx + xx][1] = (int16_t)mvd[1];
                            done |= 1u << (4 * (by / 4 + yy) + bx / 4 + xx);
                        }
                }
                int use8 = d->transform_8x8 && d->cbp_luma;
                s.t8[m] = (uint8_t)use8;
                cabac_cbp_residual(&c, &s, mb_x, mb_y, 0, e->t8x8, d->cbp_luma, d->cbp_chroma, use8, lv, &prev_dqp);
            } else if (d->mb_type == 2) { /* I_16x16 (enc.rs:3587 code_i16) */
                int t = 1 + d->i16_mode + 4 * d->cbp_chroma + (d->cbp_luma ? 12 : 0);
                if (is_idr) {
                    cabac_intra_type(&c, 0, (mb_x > 0 && s.i16[m - 1]) + (mb_y > 0 && s.i16[m - mw]), t);
                } else {
                    cabac_bin(&c, 14, 1);
                    cabac_intra_type(&c, 1, 0, t);
                }
                s.intra[m] = s.i16[m] = 1;
                cabac_chroma_mode(&c, &s, mb_x, mb_y, d->chroma_mode);
                s.cmode[m] = d->chroma_mode;
                cabac_bin(&c, 60 + (prev_dqp != 0), 0); /* mb_qp_delta 0 */
                prev_dqp = 0;
                int a = mb_x > 0 ? s.dc_cbf[m - 1] : 1, b = mb_y > 0 ? s.dc_cbf[m - mw] : 1;
                s.dc_cbf[m] = (uint8_t)cabac_block(&c, 0, a + 2 * b, lv + ORA_LV_LUMA_DC, 16);
                s.cbpl[m] = d->cbp_luma;
                s.cbpc[m] = d->cbp_chroma;
                if (d->cbp_luma)
                    for (int k = 0; k < 16; k++) {
                        int bx, by;
                        blk_xy(k, &bx, &by);
                        int gx = mb_x * 4 + bx, gy = mb_y * 4 + by;
                        int inc = cbf4_cond(&s, gx - 1, gy, 1) + 2 * cbf4_cond(&s, gx, gy - 1, 1);
                        s.cbf4[gy * gs + gx] = (uint8_t)cabac_block(&c, 1, inc, lv + ORA_LV_LUMA + k * 16 + 1, 15);
                    }
                int cbpc = d->cbp_chroma;
                if (cbpc >= 1)
                    for (int p = 0; p < 2; p++) {
                        int pa = mb_x > 0 ? (!s.skip[m - 1] && s.cdc_cbf[p][m - 1]) : 1;
                        int pb = mb_y > 0 ? (!s.skip[m - mw] && s.cdc_cbf[p][m - mw]) : 1;
                        s.cdc_cbf[p][m] = (uint8_t)cabac_block(&c, 3, pa + 2 * pb, lv + ORA_LV_CHROMA_DC + p * 4, 4);
                    }
                if (cbpc >= 2)
                    for (int p = 0; p < 2; p++)
                        for (int k = 0; k < 4; k++) {
                            int cx = mb_x * 2 + (k & 1), cy = mb_y * 2 + (k >> 1);
                            int inc = ccbf_cond(&s, p, cx - 1, cy, 1) + 2 * ccbf_cond(&s, p, cx, cy - 1, 1);
                            s.ccbf[p][cy * mw * 2 + cx] =
                                (uint8_t)cabac_block(&c, 4, inc, lv + ORA_LV_CHROMA_AC + (p * 4 + k) * 16 + 1, 15);
                        }
            } else { /* I_4x4 (enc.rs:3441 code_i4) */
                if (is_idr) {
                    cabac_intra_type(&c, 0, (mb_x > 0 && s.i16[m - 1]) + (mb_y > 0 && s.i16[m - mw]), 0);
                } else {
                    cabac_bin(&c, 14, 1);
                    cabac_intra_type(&c, 1, 0, 0);
                }
                if (e->t8x8) cabac_bin(&c, 399 + (mb_x > 0 && s.t8[m - 1]) + (mb_y > 0 && s.t8[m - mw]), 0);
                uint64_t packed;
                memcpy(&packed, d->mv, sizeof(packed));
                for (int k = 0; k < 16; k++) {
                    int bx, by;
                    blk_xy(k, &bx, &by);
                    int gx = mb_x * 4 + bx, gy = mb_y * 4 + by;
                    int mode = (int)((packed >> (4 * k)) & 15), pm = 2;
                    if (gx > 0 && gy > 0) pm = imin(s.modes[gy * gs + gx - 1], s.modes[(gy - 1) * gs + gx]);
                    cabac_bin(&c, 68, mode == pm);
                    if (mode != pm) {
                        int rem = mode < pm ? mode : mode - 1;
                        for (int i = 0; i < 3; i++) cabac_bin(&c, 69, (rem >> i) & 1);
                    }
                    s.modes[gy * gs + gx] = (uint8_t)mode;
                }
                s.intra[m] = 1;
                cabac_chroma_mode(&c, &s, mb_x, mb_y, d->chroma_mode);
                s.cmode[m] = d->chroma_mode;
                cabac_cbp_residual(&c, &s, mb_x, mb_y, 1, 0, d->cbp_luma, d->cbp_chroma, 0, lv, &prev_dqp);
            }
            cabac_terminate(&c, m + 1 == mw * mh);
        }
    cabac_nb_free(&s);
}

/* enc.rs:223 EncState::prepare */
static void prepare(ora_enc *e, int slice_qp) {
    int n = e->mb_w * e->mb_h;
    memset(e->luma_nnz, 0, (size_t)n * 16);
    memset(e->chroma_nnz[0], 0, (size_t)n * 4);
    memset(e->chroma_nnz[1], 0, (size_t)n * 4);
    memset(e->intra4x4_modes, 2, (size_t)n * 16);
    memset(e->mb_cbp_luma, 0, (size_t)n);
    memset(e->mb_cbp_chroma, 0, (size_t)n);
    memset(e->mb_qps, slice_qp, (size_t)n);
    memset(e->mb_is_intra, 0, (size_t)n);
    memset(e->mb_is_i16, 0, (size_t)n);
    memset(e->mb_is_skip, 0, (size_t)n);
    memset(e->mb_8x8, 0, (size_t)n);
    memset(e->dec, 0, (size_t)n * sizeof(ora_mb_decision));
    memset(e->levels, 0, (size_t)n * ORA_LEVELS_PER_MB * sizeof(int16_t));
}

/* enc.rs:2234 encode_frame_body for FixedQp, Balanced, no B pictures */
long ora_encode_frame(ora_enc *e, const uint8_t *y, const uint8_t *u, const uint8_t *v, int sy, int su, int sv,
                      uint8_t *out, long cap) {
    int mb_w = e->mb_w, mb_h = e->mb_h;
    unsigned gop_idx = e->frame_count % (unsigned)e->keyint;
    int is_idr = (e->frame_count % (unsigned)e->keyint) == 0 || !e->has_ref;
    int slice_qp = e->qp;
    frame_t f = {y, u, v, sy, su, sv, e->width, e->height};
    bits_t w;
    bits_init(&w);
    bits_t *wp = e->cabac ? NULL : &w;
    if (wp) { /* h264.rs:1071 SliceHeader::write */
        put_ue(wp, 0);
        put_ue(wp, is_idr ? 7 : 5);
        put_ue(wp, 0);
        put32(wp, gop_idx & 15, 4);
        if (is_idr) put_ue(wp, e->frame_count / (unsigned)e->keyint);
        put32(wp, (gop_idx * 2) & 63, 6);
        if (!is_idr) {
            put32(wp, 0, 1);
            put32(wp, 0, 1);
        }
        if (is_idr) {
            put32(wp, 0, 1);
            put32(wp, 0, 1);
        } else {
            put32(wp, 0, 1);
        }
        put_se(wp, slice_qp - e->qp);
        int dis = slice_qp == 0 || e->fast ? 1 : 0; /* enc.rs:2334 */
        put_ue(wp, (uint64_t)dis);
        if (dis != 1) {
            put_se(wp, 0);
            put_se(wp, 0);
        }
    }
    prepare(e, slice_qp);
    e->is_idr = is_idr;
    e->skip_run = 0;
    e->cur_qp = slice_qp;
    skip_bounds_init(&e->sb, slice_qp, chroma_qp(slice_qp), e->compact);
    mbsrc_t s;
    for (int mb_y = 0; mb_y < mb_h; mb_y++)
        for (int mb_x = 0; mb_x < mb_w; mb_x++) {
            load_mb(&f, mb_x, mb_y, &s);
            if (is_idr)
                code_intra(e, wp, mb_x, mb_y, slice_qp, &s);
            else if (e->compact)
                code_p_mb_compact(e, wp, mb_x, mb_y, slice_qp, &s);
            else {
                code_p_mb(e, wp, mb_x, mb_y, slice_qp, &s);
                if (e->parts) parts_record(e, mb_x, mb_y);
            }
            uint16_t nnz = 0;
            for (int by = 0; by < 4; by++)
                for (int bx = 0; bx < 4; bx++)
                    if (e->luma_nnz[(mb_y * 4 + by) * mb_w * 4 + mb_x * 4 + bx]) nnz |= (uint16_t)(1u << (by * 4 + bx));
            e->dec[mb_y * mb_w + mb_x].nnz = nnz;
        }
    if (wp) {
        if (!is_idr && e->skip_run > 0) put_ue(wp, (uint64_t)e->skip_run);
        bits_trailing(wp);
    } else if (e->cabac) {
        /* h264.rs:1071 SliceHeader::write with cabac_init_idc 0 for P slices, then
         * cabac_alignment_one_bit up to the byte boundary and the arithmetic-coded slice data
         * (enc.rs:2346-2440). */
        wp = &w;
        put_ue(wp, 0);
        put_ue(wp, is_idr ? 7 : 5);
        put_ue(wp, 0);
        put32(wp, gop_idx & 15, 4);
        if (is_idr) put_ue(wp, e->frame_count / (unsigned)e->keyint);
        put32(wp, (gop_idx * 2) & 63, 6);
        if (is_idr) {
            put32(wp, 0, 1);
            put32(wp, 0, 1);
        } else {
            put32(wp, 0, 1);
            put32(wp, 0, 1);
            put32(wp, 0, 1);
            put_ue(wp, 0); /* cabac_init_idc */
        }
        put_se(wp, slice_qp - e->qp);
        int dis = slice_qp == 0 || e->fast ? 1 : 0;
        put_ue(wp, (uint64_t)dis);
        if (dis != 1) {
            put_se(wp, 0);
            put_se(wp, 0);
        }
        while (wp->n & 7) put32(wp, 1, 1);
        cabac_slice(e, wp, is_idr, slice_qp);
        while (wp->n & 7) put32(wp, 0, 1);
    }
    copy_plane(&e->cur.y, e->pre_y);
    copy_plane(&e->cur.u, e->pre_u);
    copy_plane(&e->cur.v, e->pre_v);
    if (slice_qp != 0 && !e->fast) {
        if (e->compact || e->parts)
            deblock_parts(e);
        else
            deblock(e);
    }
    plane_extend(&e->cur.y);
    plane_extend(&e->cur.u);
    plane_extend(&e->cur.v);
    /* Keep the undeblocked picture for inspection: the deblocked one becomes the reference. */
    pic_t t = e->cur;
    e->cur = e->refp;
    e->refp = t;
    e->has_ref = 1;
    if (e->compact) {
        hpel_planes_build(e);
        int16_t(*t)[2] = e->prev_mvr;
        e->prev_mvr = e->mvr;
        e->mvr = t;
    }
    int gs = mb_w * 4;
    for (int my = 0; my < mb_h; my++)
        for (int mx = 0; mx < mb_w; mx++) {
            e->prev_mb_mv[my * mb_w + mx][0] = e->mv_grid[(my * 4) * gs + mx * 4][0];
            e->prev_mb_mv[my * mb_w + mx][1] = e->mv_grid[(my * 4) * gs + mx * 4][1];
        }
    e->last_idr = is_idr;
    long n = 0;
    bits_t o;
    bits_init(&o);
    if (wp) {
        if (is_idr) {
            uint8_t hdr[256];
            long hn = ora_headers(e, hdr, sizeof(hdr));
            for (long i = 0; i < hn; i++) bits_push(&o, hdr[i]);
        }
        nal_append(&o, 3, is_idr ? 5 : 1, w.buf, w.len);
        n = o.len;
    }
    long ret = n;
    if (n > cap)
        ret = -1;
    else if (n)
        memcpy(out, o.buf, (size_t)n);
    bits_free(&o);
    bits_free(&w);
    e->frame_count++;
    return ret;
}

void ora_get_decisions(const ora_enc *e, ora_mb_decision *out) {
    memcpy(out, e->dec, sizeof(ora_mb_decision) * (size_t)(e->mb_w * e->mb_h));
}
void ora_get_levels(const ora_enc *e, int16_t *out) {
    memcpy(out, e->levels, sizeof(int16_t) * (size_t)(e->mb_w * e->mb_h) * ORA_LEVELS_PER_MB);
}
void ora_get_recon(const ora_enc *e, int which, uint8_t *y, uint8_t *u, uint8_t *v) {
    int W = e->mb_w * 16, H = e->mb_h * 16;
    if (which == 0) {
        memcpy(y, e->pre_y, (size_t)W * H);
        memcpy(u, e->pre_u, (size_t)W * H / 4);
        memcpy(v, e->pre_v, (size_t)W * H / 4);
        return;
    }
    copy_plane(&e->refp.y, y);
    copy_plane(&e->refp.u, u);
    copy_plane(&e->refp.v, v);
}

/* ------------------------------------------------------------------------------------------ */
/* Kernel-level exports.                                                                       */
/* ------------------------------------------------------------------------------------------ */
uint32_t ora_sad16(const uint8_t *a, int sa, const uint8_t *b, int sb) { return sad_wh(a, sa, b, sb, 16, 16); }
uint32_t ora_satd16(const uint8_t *a, int sa, const uint8_t *b, int sb) { return satd16(a, sa, b, sb); }
uint32_t ora_satd4(const uint8_t *a, int sa, const uint8_t *b, int sb) { return satd4(a, sa, b, sb); }
void ora_fdct4(const int16_t *in, int16_t *out) { fdct4(in, out); }
void ora_idct4(const int16_t *in, int16_t *out) { idct4(in, out); }
void ora_fdct8(const int16_t *in, int16_t *out) { fdct8(in, out); }
void ora_luma_mc(uint8_t *dst, int ds, const uint8_t *src, int ss, int w, int h, int mx, int my) {
    luma_mc(dst, ds, src, ss, w, h, mx, my);
}
void ora_chroma_mc8(uint8_t *dst, const uint8_t *src, int ss, int fx, int fy) { chroma_mc8(dst, src, ss, fx, fy); }
void ora_window_pred(uint8_t *dst, const uint8_t *src, int ss, int dx, int dy) {
    /* A plane view whose sample (0,0) is src: the window is anchored at full-pel (0,0). */
    plane_t p = {(uint8_t *)src, ss, 0, 0, 0};
    static hpel_win_t w;
    ms_t m;
    memset(&m, 0, sizeof(m));
    m.r = &p;
    hpel_window(src - ss - 1, ss, &w);
    hpel_window_centre(&w);
    winsrc_t s = ms_window_src(&m, &w, 0, 0, dx, dy);
    winsrc_pred(&s, dst);
}
uint32_t ora_decimate_score(const int16_t *scanned, int n) { return decimate_score(scanned, n); }
void ora_dequant8(const int16_t *levels, int qp, int32_t *out) {
    init_tables();
    dequant8(levels, qp, out);
}
void ora_idct8(const int32_t *in, int32_t *out) {
    /* The residual idct8_add adds: rows, then columns, then (x + 32) >> 6. */
    int32_t g[64], col[8], o[8];
    for (int i = 0; i < 8; i++) idct8_1d(in + i * 8, g + i * 8);
    for (int j = 0; j < 8; j++) {
        for (int i = 0; i < 8; i++) col[i] = g[i * 8 + j];
        idct8_1d(col, o);
        for (int i = 0; i < 8; i++) out[i * 8 + j] = (int32_t)((uint32_t)o[i] + 32u) >> 6;
    }
}
int ora_quant8_mf(int rem, int pos) {
    init_tables();
    return QUANT8_MF16[rem][pos];
}
int ora_dequant8_norm(int rem, int pos) {
    init_tables();
    return DEQUANT8_POS[rem][pos];
}
int ora_cavlc_block(const int16_t *scanned, int n, int nc, uint8_t *out, int cap) {
    bits_t w;
    bits_init(&w);
    cavlc_block(&w, scanned, n, nc, n == 4);
    int nb = (int)bits_written(&w);
    bits_align_zero(&w);
    if (w.len <= cap) memcpy(out, w.buf, (size_t)w.len);
    bits_free(&w);
    return nb;
}
/* Test hook: the Compact prediction (RefLuma::pred over the half-sample planes) of the bw x bh
 * block at (x0, y0) of a w x h picture displaced by quarter-pel (mx, my). */
int ora_compact_pred(const uint8_t *pic, int w, int h, int x0, int y0, int bw, int bh, int mx, int my, uint8_t *dst) {
    ora_enc *e = ora_create(w, h, 30, 1, 26, 0, 30);
    if (!e || e->mb_w * 16 != w || e->mb_h * 16 != h) {
        ora_destroy(e);
        return -1;
    }
    ora_set_compact(e, 1);
    for (int r = 0; r < h; r++) memcpy(e->refp.y.buf + pat(&e->refp.y, 0, r), pic + (long)r * w, (size_t)w);
    plane_extend(&e->refp.y);
    hpel_planes_build(e);
    rl_pred(e, x0, y0, bw, bh, mx, my, dst, bw);
    ora_destroy(e);
    return 0;
}
void ora_predict_intra4x4(int mode, const uint8_t *top8, const uint8_t *left4, int tl, int has_top, int has_left,
                          uint8_t *dst) {
    pred4(mode, has_top ? top8 : NULL, has_left ? left4 : NULL, tl, dst, 4);
}
