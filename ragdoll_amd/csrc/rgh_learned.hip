// Learned motion proposals: feature extraction, bfloat16 training and FP8 inference of the
// proposal network on MFMA (gfx950), behind the C ABI of include/ragdoll_learned.h.
//
// The network (192 -> 128 -> 128 -> 176, ReLU) is three GEMMs. Training keeps fp32 master
// parameters and Adam state; every GEMM takes bfloat16 operands on v_mfma_f32_16x16x32_bf16
// with fp32 accumulation, and activations are stored in bfloat16 between layers. Inference
// keeps the whole quantized network (62 KB of E4M3 weights) in LDS, runs persistent
// workgroups of four waves over 64-macroblock tiles on v_mfma_f32_16x16x32_fp8_fp8, and
// reduces each macroblock's 169 displacement logits to its four best cells in the same kernel.
//
// No counterpart in the reference (its motion search, crates/ragdoll-codec/src/enc.rs:1182 and
// :3081-3095, has no learned component).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ragdoll_learned.h"

void rgh_note_error(const char *what, const char *detail, const char *file, int line);

namespace {

constexpr int F_IN = RGH_PNET_FEATURES, HID = RGH_PNET_HIDDEN, OUT = RGH_PNET_OUTPUTS;
constexpr int NCLS = RGH_PNET_CELLS * RGH_PNET_CELLS;  // 169 displacement classes
constexpr int MODE0 = RGH_PNET_MODE_OFFSET;
// Offsets into the flat parameter vector.
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + F_IN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * OUT, NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == RGH_PNET_PARAMS, "parameter layout");
// Quantized weights: [out][in] per layer.
constexpr int Q_W1 = 0, Q_W2 = Q_W1 + HID * F_IN, Q_W3 = Q_W2 + HID * HID, Q_BYTES = Q_W3 + OUT * HID;
static_assert(Q_BYTES == RGH_PNET_W8_BYTES, "quantized layout");
constexpr float FP8_MAX = 448.0f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (no NaN reaches here)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// E4M3 code of x, |x| <= 448, round to nearest even (v_cvt_pk_fp8_f32).
__device__ __forceinline__ uint32_t fp8_code(float x) {
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xFFu;
}

// ---------------------------------------------------------------------------------------
// bfloat16 GEMM (training): 64x64 tile per workgroup of four waves, each wave 32x32 as 2x2
// MFMA 16x16x32 blocks, k in steps of 32 through LDS. Operands are read with arbitrary element
// strides (the backward products read transposed activations) and zero-filled past the edges.
struct GemmArgs {
    const uint16_t *A;
    long sam, sak;
    const uint16_t *B;
    long sbk, sbn;
    float *C;
    uint16_t *Cb;
    long ldc;
    const float *bias;
    const uint16_t *mask;
    int M, N, K, relu, accumulate;
};

constexpr int GT = 64, GK = 32, GLD = GK + 8;  // LDS row: 32 operands + 8 pad (80 B, 16-B aligned)

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    __shared__ __align__(16) uint16_t As[GT * GLD];  // [m][k]
    __shared__ __align__(16) uint16_t Bs[GT * GLD];  // [n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
#pragma unroll
        for (int e = 0; e < GT * GK / 256; e++) {
            const int i = tid + 256 * e, r = i >> 5, c = i & 31, k = k0 + c;
            const int m = m0 + r, n = n0 + r;
            As[r * GLD + c] = (m < g.M && k < g.K) ? g.A[(long)m * g.sam + (long)k * g.sak] : (uint16_t)0;
            Bs[r * GLD + c] = (n < g.N && k < g.K) ? g.B[(long)k * g.sbk + (long)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
        // Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds row (A) / column (B)
        // l % 16 at k = 8 * (l / 16) + [0, 8).
        const int kq = 8 * (lane >> 4);
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = __builtin_bit_cast(bf16x8, *(const uint4 *)&As[(wm + i * 16 + (lane & 15)) * GLD + kq]);
            b[i] = __builtin_bit_cast(bf16x8, *(const uint4 *)&Bs[(wn + i * 16 + (lane & 15)) * GLD + kq]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // Result layout: lane l holds column l % 16, rows 4 * (l / 16) + [0, 4).
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q, n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const long o = (long)m * g.ldc + n;
                float v = acc[i][j][q];
                if (g.accumulate) v += g.C[o];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = fmaxf(v, 0.0f);
                if (g.mask && g.mask[o] == 0) v = 0.0f;
                if (g.C) g.C[o] = v;
                if (g.Cb) g.Cb[o] = f2bf(v);
            }
}

// Softmax cross-entropy of both heads; dZ = (softmax - onehot) / batch (fp32 and bfloat16).
__global__ void k_xent(const float *Z, const rgh_pnet_label *lab, int B, float inv_b, float *dZ, uint16_t *dZb,
                       float *loss) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= B) return;
    const float *z = Z + (long)r * OUT;
    float *gz = dZ + (long)r * OUT;
    uint16_t *gb = dZb + (long)r * OUT;
    const int d = lab[r].disp, md = lab[r].mode;
    float ld = 0.0f, lm = 0.0f;
    if (d >= 0 && d < NCLS) {
        float mx = z[0];
        for (int i = 1; i < NCLS; i++) mx = fmaxf(mx, z[i]);
        float s = 0.0f;
        for (int i = 0; i < NCLS; i++) s += expf(z[i] - mx);
        const float inv = 1.0f / s;
        ld = mx + logf(s) - z[d];
        for (int i = 0; i < NCLS; i++) gz[i] = (expf(z[i] - mx) * inv - (i == d ? 1.0f : 0.0f)) * inv_b;
    } else {
        for (int i = 0; i < NCLS; i++) gz[i] = 0.0f;
    }
    {
        const float *zm = z + MODE0;
        const float mx = fmaxf(zm[0], fmaxf(zm[1], zm[2]));
        const float e0 = expf(zm[0] - mx), e1 = expf(zm[1] - mx), e2 = expf(zm[2] - mx), s = e0 + e1 + e2;
        lm = mx + logf(s) - zm[md];
        const float e[3] = {e0, e1, e2};
        for (int i = 0; i < 3; i++) gz[MODE0 + i] = (e[i] / s - (i == md ? 1.0f : 0.0f)) * inv_b;
    }
    for (int i = MODE0 + 3; i < OUT; i++) gz[i] = 0.0f;
    for (int i = 0; i < OUT; i++) gb[i] = f2bf(gz[i]);
    atomicAdd(&loss[0], ld);
    atomicAdd(&loss[1], lm);
}

// Column sums in a fixed order: partial sums over 64-row chunks, then over the chunks.
__global__ void k_colsum_part(const float *G, int M, int N, float *part) {
    const int n = threadIdx.x, c = blockIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    const int r1 = min(M, (c + 1) * 64);
    for (int r = c * 64; r < r1; r++) s += G[(long)r * N + n];
    part[(long)c * N + n] = s;
}
__global__ void k_colsum_final(const float *part, int chunks, int N, float *out) {
    const int n = threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += part[(long)c * N + n];
    out[n] = s;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, int n, float lr, float c1,
                       float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float *x, uint16_t *y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------------------
// Features: one wave per macroblock.
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 4x4-mean images: cell (cx, cy) of a picture, -6 <= cx < 4 * mb_w + 6 and likewise cy, is
// the rounded mean of the samples (4cx + i, 4cy + j), coordinates clamped to the picture. One
// image per distinct plane of a features batch; every macroblock's 64x64 window and its own 4x4
// cells are then read from it instead of from the samples.
constexpr int MPAD = 6;
struct MeanPlane {
    const uint8_t *y;
    uint8_t *m;
    int stride, W, H, mw, mh;  // mw = 4 * mb_w + 12, mh likewise
};
__global__ __launch_bounds__(256) void k_means(const MeanPlane *planes) {
    const MeanPlane &P = planes[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P.mw * P.mh) return;
    const int cx = i % P.mw - MPAD, cy = i / P.mw - MPAD;
    int s = 0;
    for (int j = 0; j < 4; j++) {
        const uint8_t *row = P.y + (long)min(max(4 * cy + j, 0), P.H - 1) * P.stride;
        for (int k = 0; k < 4; k++) s += row[min(max(4 * cx + k, 0), P.W - 1)];
    }
    P.m[i] = (uint8_t)((s + 8) >> 4);
}

struct FeatPic {
    const uint8_t *cur;           // luma samples (mean absolute deviation)
    const uint8_t *mcur, *mprev;  // mean images of the picture and the previous picture
    uint16_t *features;
    int cs, W, H, mw, qp;
};

// Features: one wave per macroblock, four waves per workgroup, over the batch's (picture,
// macroblock) items (item = picture * max_mbs + macroblock). The loop runs the same number of
// rounds in every wave of a workgroup (its barriers are workgroup-wide); a wave whose item is
// past the end or past its picture's macroblocks only joins the barriers.
// Sum over the wave: DPP within rows of 16, then the four row sums by readlane.
__device__ __forceinline__ int wave_total(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
           __builtin_amdgcn_readlane(v, 48);
}

__global__ __launch_bounds__(256) void k_features(const FeatPic *pics, int max_mbs, long items) {
    __shared__ uint32_t rc_all[4][16 * 5];  // window cells, 16 bytes per row + a pad dword
    __shared__ uint32_t sc_all[4][4];       // the macroblock's cells, one row of 4 per dword
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *rc = rc_all[wave], *sc = sc_all[wave];
    for (long base = (long)blockIdx.x * 4; base < items; base += (long)gridDim.x * 4) {
        const long it = base + wave;
        const FeatPic *P = it < items ? &pics[it / max_mbs] : nullptr;
        const int mb = (int)(it % max_mbs), mb_w = P ? (P->W + 15) >> 4 : 0;
        const bool valid = P && mb < mb_w * ((P->H + 15) >> 4);
        if (valid) {
            const int bx = mb % mb_w, by = mb / mb_w, x0 = bx * 16, y0 = by * 16;
            // The previous picture's window cells (4 bx - 6 + c, 4 by - 6 + r) are cells
            // (4 bx + c, 4 by + r) of the offset mean image (rows of 4 * mb_w + 12 cells, so
            // every dword below is aligned): lane = row * 4 + dword.
            rc[(lane >> 2) * 5 + (lane & 3)] =
                *(const uint32_t *)(P->mprev + (long)(4 * by + (lane >> 2)) * P->mw + 4 * bx + 4 * (lane & 3));
            if (lane < 16) rc[lane * 5 + 4] = 0;
            if (lane < 4) {  // column 4 bx + 6 is not dword-aligned: bytes
                const uint8_t *q = P->mcur + (long)(4 * by + MPAD + lane) * P->mw + 4 * bx + MPAD;
                sc[lane] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
            }
            // The macroblock's samples: four per lane (row lane / 4, columns 4 * (lane % 4) + [0, 4)).
            const uint8_t *row = P->cur + (long)clampi(y0 + (lane >> 2), 0, P->H - 1) * P->cs;
            const int sx = x0 + (lane & 3) * 4;
            int px[4];
            if (sx + 3 < P->W && (((uintptr_t)(row + sx)) & 3) == 0) {
                const uint32_t w = *(const uint32_t *)(row + sx);
                for (int i = 0; i < 4; i++) px[i] = (int)((w >> (8 * i)) & 0xFF);
            } else {
                for (int i = 0; i < 4; i++) px[i] = row[min(sx + i, P->W - 1)];
            }
            const int mean = (wave_total(px[0] + px[1] + px[2] + px[3]) + 128) >> 8;
            const int mad = (wave_total(abs(px[0] - mean) + abs(px[1] - mean) + abs(px[2] - mean) + abs(px[3] - mean)) +
                             128) >> 8;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            uint16_t *f = P->features + (long)mb * F_IN;
            const uint32_t s0 = sc[0], s1 = sc[1], s2 = sc[2], s3 = sc[3];
            for (int e = lane; e < NCLS; e += 64) {
                const int dy = e / 13, dx = e % 13;  // cells displaced by (dx - 6, dy - 6)
                const int d = dx >> 2, sh = 8 * (dx & 3);
                const uint32_t *rw = rc + dy * 5 + d;
                const uint32_t b = (uint32_t)sh >> 3;  // byte offset of the window in its dword pair
                uint32_t acc = 0;
                acc = __builtin_amdgcn_sad_u8(s0, __builtin_amdgcn_alignbyte(rw[1], rw[0], b), acc);
                acc = __builtin_amdgcn_sad_u8(s1, __builtin_amdgcn_alignbyte(rw[6], rw[5], b), acc);
                acc = __builtin_amdgcn_sad_u8(s2, __builtin_amdgcn_alignbyte(rw[11], rw[10], b), acc);
                acc = __builtin_amdgcn_sad_u8(s3, __builtin_amdgcn_alignbyte(rw[16], rw[15], b), acc);
                f[e] = f2bf((float)min((int)acc, 2047) * (1.0f / 256.0f));
            }
            if (lane < 16) f[NCLS + lane] = f2bf((float)((sc[lane >> 2] >> (8 * (lane & 3))) & 0xFF) * (1.0f / 64.0f));
            if (lane == 16) f[NCLS + 16] = f2bf((float)mad * (1.0f / 16.0f));
            if (lane == 17) f[NCLS + 17] = f2bf((float)P->qp * (1.0f / 32.0f));
            if (lane >= 18 && lane < F_IN - NCLS) f[NCLS + lane] = 0;
        }
        // The wave reuses its LDS rows next round: every lane's reads end before the next writes.
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

__global__ void k_labels(const rgh_mb_decision *d, int n, rgh_pnet_label *l) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = d[i].mb_type;
    rgh_pnet_label o;
    if (t == 2 || t == 3 || t == 4) {
        o.disp = -1;
        o.mode = 2;
    } else {
        const int cx = clampi((d[i].mv[0] + 8) >> 4, -6, 6), cy = clampi((d[i].mv[1] + 8) >> 4, -6, 6);
        o.disp = (int16_t)((cy + 6) * 13 + cx + 6);
        o.mode = t == 0 ? 0 : 1;
    }
    l[i] = o;
}

__global__ void k_gather(const uint16_t *F, const rgh_pnet_label *L, const int *idx, int n, uint16_t *Fo,
                         rgh_pnet_label *Lo) {
    const int i = blockIdx.x, t = threadIdx.x;
    if (i >= n) return;
    const long s = idx[i];
    for (int k = t; k < F_IN; k += blockDim.x) Fo[(long)i * F_IN + k] = F[s * F_IN + k];
    if (t == 0) Lo[i] = L[s];
}

__global__ void k_fp8_encode(const float *x, int n, uint8_t *o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (uint8_t)fp8_code(x[i]);
}

// ---------------------------------------------------------------------------------------
// FP8 inference. A workgroup of four waves takes 64 macroblocks at a time; every layer is one
// 64 x N product in which wave w owns a slice of the output columns for all 64 rows, so each
// k-step issues 4 x (columns / 16) independent MFMAs from shared A fragments. Weights stay in
// LDS for the workgroup's lifetime (rows padded by 8 bytes against bank conflicts); the
// activations of the tile pass between layers through LDS as E4M3, the logits as fp32.
constexpr int IT = 64;
constexpr int XS = F_IN + 8, HS = HID + 8, ZS = OUT + 1;
constexpr int QL_W1 = 0, QL_W2 = QL_W1 + HID * XS, QL_W3 = QL_W2 + HID * HS, QL_BYTES = QL_W3 + OUT * HS;
struct InferArgs {
    const uint16_t *X;
    int M;
    const uint8_t *w8;
    const float *qs;    // [3][176] column scales, [3] input multipliers
    const float *bias;  // flat parameter vector (biases read at their offsets)
    rgh_proposal *out;
    float *logits;
};
struct InferLds {
    uint8_t w[QL_BYTES];
    uint8_t x[IT * XS];
    uint8_t h1[IT * HS];
    uint8_t h2[IT * HS];
    float z[IT * ZS];
    float scale[3][OUT];
    float bias[3][OUT];
    float mul[3];
    float tv[IT][4][4];
    int ti[IT][4][4];
};
static_assert(sizeof(InferLds) <= 160 * 1024, "LDS budget");
static_assert(XS % 8 == 0 && HS % 8 == 0 && QL_W2 % 8 == 0 && QL_W3 % 8 == 0, "8-byte operand alignment");

// Columns [16 nb0, 16 (nb0 + NB)) of act(in[64][K] . w[N][K]^T) for the tile's 64 rows: scale,
// bias and ReLU, then E4M3 into out8 (row stride OS) or fp32 into outf (row stride ZS).
template <int K, int IS, int NB, int OS>
__device__ __forceinline__ void fp8_layer(const uint8_t *in, const uint8_t *w, int nb0, int nbn, const float *scale,
                                          const float *bias, float mul_next, uint8_t *out8, float *outf, int lane) {
    const int kq = 8 * (lane >> 4), rl = lane & 15;
    f32x4 acc[4][NB];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        long a[4], b[NB];
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = *(const long *)&in[(r * 16 + rl) * IS + k0 + kq];
#pragma unroll
        for (int j = 0; j < NB; j++) b[j] = j < nbn ? *(const long *)&w[((nb0 + j) * 16 + rl) * IS + k0 + kq] : 0L;
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[r], b[j], acc[r][j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j >= nbn) continue;
        const int n = (nb0 + j) * 16 + rl;
        const float sc = scale[n], bi = bias[n];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = r * 16 + 4 * (lane >> 4) + q;
                const float v = acc[r][j][q] * sc + bi;
                if (out8)
                    out8[m * OS + n] = (uint8_t)fp8_code(fminf(fmaxf(v, 0.0f) * mul_next, FP8_MAX));
                else
                    outf[m * ZS + n] = v;
            }
    }
}

__global__ __launch_bounds__(256) void k_infer(InferArgs g) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    InferLds &S = *(InferLds *)lds_raw;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // Weights into padded rows: 8-byte chunks.
    {
        const int rows[3] = {HID, HID, OUT}, kin[3] = {F_IN, HID, HID}, src[3] = {Q_W1, Q_W2, Q_W3},
                  dst[3] = {QL_W1, QL_W2, QL_W3}, str[3] = {XS, HS, HS};
        for (int l = 0; l < 3; l++) {
            const int per = kin[l] / 8, total = rows[l] * per;
            for (int i = tid; i < total; i += 256) {
                const int r = i / per, c = (i % per) * 8;
                *(uint2 *)&S.w[dst[l] + r * str[l] + c] = *(const uint2 *)&g.w8[src[l] + r * kin[l] + c];
            }
        }
    }
    for (int i = tid; i < 3 * OUT; i += 256) {
        const int l = i / OUT, n = i % OUT;
        S.scale[l][n] = g.qs[i];
        const int boff = l == 0 ? OFF_B1 : (l == 1 ? OFF_B2 : OFF_B3);
        S.bias[l][n] = n < (l == 2 ? OUT : HID) ? g.bias[boff + n] : 0.0f;
    }
    if (tid < 3) S.mul[tid] = g.qs[3 * OUT + tid];
    __syncthreads();
    const int ntiles = (g.M + IT - 1) / IT;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r0 = t * IT;
        // Inputs: 64 rows x 192 bfloat16 -> E4M3, 8 per item.
        const float mul0 = S.mul[0];
        for (int e = tid; e < IT * (F_IN / 8); e += 256) {
            const int rr = e / (F_IN / 8), c8 = (e % (F_IN / 8)) * 8, row = r0 + rr;
            uint32_t lo = 0, hi = 0;
            if (row < g.M) {
                const uint4 v = *(const uint4 *)&g.X[(long)row * F_IN + c8];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float a = fminf(fmaxf(bf2f((uint16_t)(w[j] & 0xFFFF)) * mul0, -FP8_MAX), FP8_MAX);
                    const float b = fminf(fmaxf(bf2f((uint16_t)(w[j] >> 16)) * mul0, -FP8_MAX), FP8_MAX);
                    const uint32_t p = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xFFFFu;
                    if (j < 2)
                        lo |= p << (16 * j);
                    else
                        hi |= p << (16 * (j - 2));
                }
            }
            *(uint2 *)&S.x[rr * XS + c8] = make_uint2(lo, hi);
        }
        __syncthreads();
        fp8_layer<F_IN, XS, 2, HS>(S.x, S.w + QL_W1, 2 * wave, 2, S.scale[0], S.bias[0], S.mul[1], S.h1, nullptr, lane);
        __syncthreads();
        fp8_layer<HID, HS, 2, HS>(S.h1, S.w + QL_W2, 2 * wave, 2, S.scale[1], S.bias[1], S.mul[2], S.h2, nullptr, lane);
        __syncthreads();
        // 11 column blocks: waves 0-2 take three, wave 3 two.
        fp8_layer<HID, HS, 3, HS>(S.h2, S.w + QL_W3, 3 * wave, wave == 3 ? 2 : 3, S.scale[2], S.bias[2], 0.0f, nullptr,
                                  S.z, lane);
        __syncthreads();
        if (g.logits)
            for (int e = tid; e < IT * OUT; e += 256) {
                const int rr = e / OUT, row = r0 + rr;
                if (row < g.M) g.logits[(long)row * OUT + e % OUT] = S.z[rr * ZS + e % OUT];
            }
        // The four largest displacement logits of each row, earliest index first among equals:
        // four threads per row scan a quarter each (an insertion keeps v0 >= v1 >= v2 >= v3 and
        // a later equal logit never moves up), then the first merges the quarters in order.
        {
            const int rr = tid >> 2, q = tid & 3;
            const float *zz = S.z + rr * ZS;
            float v0 = -INFINITY, v1 = -INFINITY, v2 = -INFINITY, v3 = -INFINITY;
            int i0 = -1, i1 = -1, i2 = -1, i3 = -1;
            auto insert = [&](float v, int i) {
                if (!(v > v3)) return;
                v3 = v;
                i3 = i;
                if (v3 > v2) {
                    v3 = v2; i3 = i2; v2 = v; i2 = i;
                    if (v2 > v1) {
                        v2 = v1; i2 = i1; v1 = v; i1 = i;
                        if (v1 > v0) {
                            v1 = v0; i1 = i0; v0 = v; i0 = i;
                        }
                    }
                }
            };
            const int lo = q * 43, hi = min(NCLS, lo + 43);
            for (int i = lo; i < hi; i++) insert(zz[i], i);
            S.tv[rr][q][0] = v0; S.tv[rr][q][1] = v1; S.tv[rr][q][2] = v2; S.tv[rr][q][3] = v3;
            S.ti[rr][q][0] = i0; S.ti[rr][q][1] = i1; S.ti[rr][q][2] = i2; S.ti[rr][q][3] = i3;
            __syncthreads();
            if (q == 0 && r0 + rr < g.M) {
                v0 = v1 = v2 = v3 = -INFINITY;
                i0 = i1 = i2 = i3 = -1;
                for (int qq = 0; qq < 4; qq++)
                    for (int k = 0; k < 4; k++)
                        if (S.ti[rr][qq][k] >= 0) insert(S.tv[rr][qq][k], S.ti[rr][qq][k]);
                const int c0 = i0 < 0 ? 84 : i0, c1 = i1 < 0 ? 84 : i1, c2 = i2 < 0 ? 84 : i2, c3 = i3 < 0 ? 84 : i3;
                rgh_proposal o;
                o.cell[0][0] = (int8_t)(c0 % 13 - 6);
                o.cell[0][1] = (int8_t)(c0 / 13 - 6);
                o.cell[1][0] = (int8_t)(c1 % 13 - 6);
                o.cell[1][1] = (int8_t)(c1 / 13 - 6);
                o.cell[2][0] = (int8_t)(c2 % 13 - 6);
                o.cell[2][1] = (int8_t)(c2 / 13 - 6);
                o.cell[3][0] = (int8_t)(c3 % 13 - 6);
                o.cell[3][1] = (int8_t)(c3 / 13 - 6);
                const float m0 = zz[MODE0], m1 = zz[MODE0 + 1], m2 = zz[MODE0 + 2], mx = fmaxf(m0, fmaxf(m1, m2));
                const float e0 = expf(m0 - mx), e1 = expf(m1 - mx), e2 = expf(m2 - mx), inv = 255.0f / (e0 + e1 + e2);
                o.mode_prob[0] = (uint8_t)(e0 * inv + 0.5f);
                o.mode_prob[1] = (uint8_t)(e1 * inv + 0.5f);
                o.mode_prob[2] = (uint8_t)(e2 * inv + 0.5f);
                for (int k = 0; k < 5; k++) o.reserved[k] = 0;
                g.out[r0 + rr] = o;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Features and FP8 inference in one pass (rgh_pnet_propose_batch). A workgroup of four waves
// takes 64 macroblocks of one picture at a time: each wave computes the features of 16 of them,
// four lanes a macroblock, straight into the tile's E4M3 input rows (the values k_features
// stores, bfloat16-rounded, then scaled and encoded as k_infer encodes them: a table of the 2048
// possible cost-volume codes), then the three layers and the top-4 reduction run as in k_infer. The weights live in VGPRs (each wave holds the B fragments of its own output
// columns: 64 registers), so LDS holds only the tile's activations (~60 KB with the logits
// aliasing the input rows) and two workgroups share a CU, one computing features while the other
// multiplies. The features never leave the chip.
struct ProposePic {
    const uint8_t *cur;           // luma samples (mean absolute deviation)
    const uint8_t *mcur, *mprev;  // mean images of the picture and the previous picture
    int cs, W, H, mw, qp;
    int row0;                     // first output row of the picture
};
struct ProposeArgs {
    const ProposePic *pics;
    const int *tile0;  // tile0[i]: first tile of picture i; tile0[n]: the tile count
    int n;
    const uint8_t *w8;
    const float *qs;
    const float *bias;
    rgh_proposal *out;
    float *logits;
};
struct ProposeLds {
    union {
        struct {
            uint8_t x[IT * XS];
            uint8_t h1[IT * HS];
        };
        float z[IT * ZS];
    };
    union {
        uint8_t h2[IT * HS];
        struct {
            float tv[IT][4][4];
            int ti[IT][4][4];
        };
    };
    float scale[3][OUT];
    float bias[3][OUT];
    float mul[3];
    uint8_t code[2048];  // E4M3 input code of a cost-volume value (clamped SAD 0..2047)
};
static_assert(sizeof(ProposeLds) <= 64 * 1024, "two workgroups per CU");

// fp8_layer with the B fragments in registers: bw[ks][j] holds the 8 weights (k 32 ks + 8 (lane /
// 16) .. + 7) of column 16 (nb0 + j) + lane % 16.
template <int K, int IS, int NB, int OS>
__device__ __forceinline__ void fp8_layer_r(const uint8_t *in, const long (&bw)[K / 32][NB], int nb0, int nbn,
                                            const float *scale, const float *bias, float mul_next, uint8_t *out8,
                                            float *outf, int lane) {
    const int kq = 8 * (lane >> 4), rl = lane & 15;
    f32x4 acc[4][NB];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[r][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K / 32; ks++) {
        long a[4];
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = *(const long *)&in[(r * 16 + rl) * IS + 32 * ks + kq];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < NB; j++)
                acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[r], bw[ks][j], acc[r][j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j >= nbn) continue;
        const int n = (nb0 + j) * 16 + rl;
        const float sc = scale[n], bi = bias[n];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = r * 16 + 4 * (lane >> 4) + q;
                const float v = acc[r][j][q] * sc + bi;
                if (out8)
                    out8[m * OS + n] = (uint8_t)fp8_code(fminf(fmaxf(v, 0.0f) * mul_next, FP8_MAX));
                else
                    outf[m * ZS + n] = v;
            }
    }
}

// The B fragments of one layer's columns [16 nb0, 16 (nb0 + nbn)) from the quantized weights
// (global, [out][K] rows).
template <int K, int NB>
__device__ __forceinline__ void load_frags(const uint8_t *w, int nb0, int nbn, long (&bw)[K / 32][NB], int lane) {
    const int kq = 8 * (lane >> 4), rl = lane & 15;
#pragma unroll
    for (int ks = 0; ks < K / 32; ks++)
#pragma unroll
        for (int j = 0; j < NB; j++)
            bw[ks][j] = j < nbn ? *(const long *)&w[((nb0 + j) * 16 + rl) * K + 32 * ks + kq] : 0L;
}

// E4M3 input code of one feature value: bfloat16-rounded (as k_features stores it), then scaled
// and saturated as k_infer encodes its input.
__device__ __forceinline__ uint8_t feature_code(float f, float mul0) {
    const float a = fminf(fmaxf(bf2f(f2bf(f)) * mul0, -FP8_MAX), FP8_MAX);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(a, 0.0f, 0, false) & 0xFF);
}

// The inputs of one tile's features (see k_propose): the picture's metadata and, for lane 4 m + q
// of wave w (row 16 w + m), window rows dy0 .. dy0 + 6 of the previous picture's mean image
// (dy0 = 0, 4, 7, 10 for q = 0..3), the macroblock's own 4x4 cells and its source rows 4 q ..
// 4 q + 3 (loads only).
struct TileIn {
    uint32_t win[7][4], own[4], src[4][4];
    int count, qp;
    long orow0;  // output row of the tile's first macroblock
};
__device__ __forceinline__ void tile_inputs(const ProposeArgs &g, int t, int wave, int lane, TileIn &in) {
    // The picture of tile t (tile0 ascending).
    int lo = 0, hi = g.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g.tile0[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ProposePic P = g.pics[lo];
    const int mb_w = (P.W + 15) >> 4, n_mbs = mb_w * ((P.H + 15) >> 4);
    const int mb0 = (t - g.tile0[lo]) * IT, count = min(IT, n_mbs - mb0);
    in.count = count;
    in.qp = P.qp;
    in.orow0 = (long)P.row0 + mb0;
    const int m = lane >> 2, q = lane & 3, rr = 16 * wave + m;
    const bool valid = rr < count;
    const int mb = mb0 + rr, bx = valid ? mb % mb_w : 0, by = valid ? mb / mb_w : 0;
    const int dy0 = q == 0 ? 0 : 1 + 3 * q;
    const uint8_t *wb = P.mprev + (long)(4 * by + dy0) * P.mw + 4 * bx;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const uint32_t *wr = (const uint32_t *)(wb + (long)min(j, 15 - dy0) * P.mw);
#pragma unroll
        for (int c = 0; c < 4; c++) in.win[j][c] = valid ? wr[c] : 0u;
    }
    const uint8_t *ob = P.mcur + (long)(4 * by + MPAD) * P.mw + 4 * bx + MPAD;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint8_t *o = ob + (long)r * P.mw;
        in.own[r] = valid ? (uint32_t)o[0] | ((uint32_t)o[1] << 8) | ((uint32_t)o[2] << 16) | ((uint32_t)o[3] << 24) : 0u;
    }
    const int x0 = bx * 16, y0 = by * 16;
    const bool fast = x0 + 15 < P.W && ((((uintptr_t)P.cur + x0) | (uintptr_t)P.cs) & 3) == 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint8_t *row = P.cur + (long)clampi(y0 + 4 * q + r, 0, P.H - 1) * P.cs;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            uint32_t v = 0;
            if (valid) {
                if (fast) {
                    v = *(const uint32_t *)(row + x0 + 4 * c);
                } else {
                    for (int i = 0; i < 4; i++) v |= (uint32_t)row[min(x0 + 4 * c + i, P.W - 1)] << (8 * i);
                }
            }
            in.src[r][c] = v;
        }
    }
}

__global__ __launch_bounds__(256, 2) void k_propose(ProposeArgs g) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    ProposeLds &S = *(ProposeLds *)lds_raw;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    long b1[F_IN / 32][2], b2[HID / 32][2], b3[HID / 32][3];
    load_frags<F_IN, 2>(g.w8 + Q_W1, 2 * wave, 2, b1, lane);
    load_frags<HID, 2>(g.w8 + Q_W2, 2 * wave, 2, b2, lane);
    load_frags<HID, 3>(g.w8 + Q_W3, 3 * wave, wave == 3 ? 2 : 3, b3, lane);
    for (int i = tid; i < 3 * OUT; i += 256) {
        const int l = i / OUT, n = i % OUT;
        S.scale[l][n] = g.qs[i];
        const int boff = l == 0 ? OFF_B1 : (l == 1 ? OFF_B2 : OFF_B3);
        S.bias[l][n] = n < (l == 2 ? OUT : HID) ? g.bias[boff + n] : 0.0f;
    }
    if (tid < 3) S.mul[tid] = g.qs[3 * OUT + tid];
    __syncthreads();
    const float mul0 = S.mul[0];
    for (int v = tid; v < 2048; v += 256) S.code[v] = feature_code((float)v * (1.0f / 256.0f), mul0);
    __syncthreads();
    const int ntiles = g.tile0[g.n];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        TileIn in;
        tile_inputs(g, t, wave, lane, in);
        // ---- Features of rows 16 wave .. 16 wave + 15, four lanes a macroblock: lane 4 m + q of
        // wave w computes row 16 w + m's cost-volume entries of displacement rows dy in
        // [dy0(q), dy0(q + 1)) (dy0 = 0, 4, 7, 10, 13) from the previous picture's window cells
        // (4 bx - 6 + c, 4 by - 6 + r) = cells (4 bx + c, 4 by + r) of the offset mean image, held
        // in registers (window rows dy0 .. dy0 + 6), the macroblock's own cells (4 bx + c, 4 by +
        // r) of the current mean image, and a quarter of the mean absolute deviation (source rows
        // 4 q .. 4 q + 3, reduced over the quad).
        const int count = in.count;
        const long orow0 = in.orow0;
        const int m = lane >> 2, q = lane & 3, rr = 16 * wave + m;
        const bool valid = rr < count;
        const int dy0 = q == 0 ? 0 : 1 + 3 * q;
        const uint32_t(&win)[7][4] = in.win;
        const uint32_t(&own)[4] = in.own;
        const uint32_t(&src)[4][4] = in.src;
        uint8_t *xr = S.x + rr * XS;
        // Mean and mean absolute deviation of the macroblock (as k_features: rounded / 256).
        uint32_t sum = 0;
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) sum = __builtin_amdgcn_sad_u8(src[r][c], 0u, sum);
        sum += __builtin_amdgcn_update_dpp(0, (int)sum, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        sum += __builtin_amdgcn_update_dpp(0, (int)sum, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        const uint32_t mean = (sum + 128) >> 8, mean4 = mean * 0x01010101u;
        uint32_t dev = 0;
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) dev = __builtin_amdgcn_sad_u8(src[r][c], mean4, dev);
        dev += __builtin_amdgcn_update_dpp(0, (int)dev, 0xB1, 0xF, 0xF, false);
        dev += __builtin_amdgcn_update_dpp(0, (int)dev, 0x4E, 0xF, 0xF, false);
        const uint32_t mad = (dev + 128) >> 8;
        // Cost volume: entry (dy0 + t) * 13 + dx for t < 4 (lane q 0) or 3, dx 0..12.
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t == 3 && q != 0) break;
#pragma unroll
            for (int dx = 0; dx < 13; dx++) {
                const int d = dx >> 2, bsh = dx & 3;
                uint32_t acc = 0;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t w = d + 1 < 4 ? __builtin_amdgcn_alignbyte(win[t + r][d + 1], win[t + r][d], (uint32_t)bsh)
                                                 : (bsh ? __builtin_amdgcn_alignbyte(0u, win[t + r][d], (uint32_t)bsh)
                                                        : win[t + r][d]);
                    acc = __builtin_amdgcn_sad_u8(own[r], w, acc);
                }
                if (valid) xr[(dy0 + t) * 13 + dx] = S.code[min(acc, 2047u)];
            }
        }
        if (valid) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t oq = q == 0 ? own[0] : q == 1 ? own[1] : q == 2 ? own[2] : own[3];
                xr[NCLS + 4 * q + i] = feature_code((float)((oq >> (8 * i)) & 0xFF) * (1.0f / 64.0f), mul0);
            }
            if (q == 0) {
                xr[NCLS + 16] = feature_code((float)mad * (1.0f / 16.0f), mul0);
                xr[NCLS + 17] = feature_code((float)in.qp * (1.0f / 32.0f), mul0);
            } else if (q == 1) {
#pragma unroll
                for (int i = 18; i < F_IN - NCLS; i++) xr[NCLS + i] = 0;
            }
        } else {
            // Rows past the picture: zero inputs, outputs discarded.
#pragma unroll
            for (int i = 0; i < 12; i++) *(uint32_t *)&xr[48 * q + 4 * i] = 0;
        }
        __syncthreads();
        fp8_layer_r<F_IN, XS, 2, HS>(S.x, b1, 2 * wave, 2, S.scale[0], S.bias[0], S.mul[1], S.h1, nullptr, lane);
        __syncthreads();
        fp8_layer_r<HID, HS, 2, HS>(S.h1, b2, 2 * wave, 2, S.scale[1], S.bias[1], S.mul[2], S.h2, nullptr, lane);
        __syncthreads();
        fp8_layer_r<HID, HS, 3, HS>(S.h2, b3, 3 * wave, wave == 3 ? 2 : 3, S.scale[2], S.bias[2], 0.0f, nullptr, S.z,
                                    lane);
        __syncthreads();
        if (g.logits)
            for (int e = tid; e < count * OUT; e += 256) g.logits[(orow0 + e / OUT) * OUT + e % OUT] = S.z[(e / OUT) * ZS + e % OUT];
        // The four largest displacement logits of each row (as k_infer).
        {
            const int rr = tid >> 2, q = tid & 3;
            const float *zz = S.z + rr * ZS;
            float v0 = -INFINITY, v1 = -INFINITY, v2 = -INFINITY, v3 = -INFINITY;
            int i0 = -1, i1 = -1, i2 = -1, i3 = -1;
            auto insert = [&](float v, int i) {
                if (!(v > v3)) return;
                v3 = v;
                i3 = i;
                if (v3 > v2) {
                    v3 = v2; i3 = i2; v2 = v; i2 = i;
                    if (v2 > v1) {
                        v2 = v1; i2 = i1; v1 = v; i1 = i;
                        if (v1 > v0) {
                            v1 = v0; i1 = i0; v0 = v; i0 = i;
                        }
                    }
                }
            };
            // A quarter of the row (43 logits, 40 in the last quarter), eight loads at a time.
            const int lo4 = q * 43, hi4 = min(NCLS, lo4 + 43);
#pragma unroll
            for (int b0 = 0; b0 < 48; b0 += 8) {
                float v[8];
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = lo4 + b0 + k < hi4 ? zz[lo4 + b0 + k] : -INFINITY;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if (lo4 + b0 + k < hi4) insert(v[k], lo4 + b0 + k);
            }
            S.tv[rr][q][0] = v0; S.tv[rr][q][1] = v1; S.tv[rr][q][2] = v2; S.tv[rr][q][3] = v3;
            S.ti[rr][q][0] = i0; S.ti[rr][q][1] = i1; S.ti[rr][q][2] = i2; S.ti[rr][q][3] = i3;
            __syncthreads();
            if (q == 0 && rr < count) {
                v0 = v1 = v2 = v3 = -INFINITY;
                i0 = i1 = i2 = i3 = -1;
                for (int qq = 0; qq < 4; qq++)
                    for (int k = 0; k < 4; k++)
                        if (S.ti[rr][qq][k] >= 0) insert(S.tv[rr][qq][k], S.ti[rr][qq][k]);
                const int c0 = i0 < 0 ? 84 : i0, c1 = i1 < 0 ? 84 : i1, c2 = i2 < 0 ? 84 : i2, c3 = i3 < 0 ? 84 : i3;
                rgh_proposal o;
                o.cell[0][0] = (int8_t)(c0 % 13 - 6);
                o.cell[0][1] = (int8_t)(c0 / 13 - 6);
                o.cell[1][0] = (int8_t)(c1 % 13 - 6);
                o.cell[1][1] = (int8_t)(c1 / 13 - 6);
                o.cell[2][0] = (int8_t)(c2 % 13 - 6);
                o.cell[2][1] = (int8_t)(c2 / 13 - 6);
                o.cell[3][0] = (int8_t)(c3 % 13 - 6);
                o.cell[3][1] = (int8_t)(c3 / 13 - 6);
                const float m0 = zz[MODE0], m1 = zz[MODE0 + 1], m2 = zz[MODE0 + 2], mx = fmaxf(m0, fmaxf(m1, m2));
                const float e0 = expf(m0 - mx), e1 = expf(m1 - mx), e2 = expf(m2 - mx), inv = 255.0f / (e0 + e1 + e2);
                o.mode_prob[0] = (uint8_t)(e0 * inv + 0.5f);
                o.mode_prob[1] = (uint8_t)(e1 * inv + 0.5f);
                o.mode_prob[2] = (uint8_t)(e2 * inv + 0.5f);
                for (int k = 0; k < 5; k++) o.reserved[k] = 0;
                g.out[orow0 + rr] = o;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Host side.
rgh_status hip_status(hipError_t e) {
    if (e == hipSuccess) return RGH_OK;
    if (e == hipErrorOutOfMemory) return RGH_ERR_OOM;
    return RGH_ERR_DEVICE;
}
#define TRY(expr)                                                             \
    do {                                                                      \
        hipError_t e__ = (expr);                                              \
        if (e__ != hipSuccess) {                                              \
            rgh_note_error(#expr, hipGetErrorString(e__), "rgh_learned.hip", __LINE__); \
            return hip_status(e__);                                           \
        }                                                                     \
    } while (0)
#define FAIL(code, what)                                                      \
    do {                                                                      \
        rgh_note_error(__func__, what, "rgh_learned.hip", __LINE__);          \
        return code;                                                          \
    } while (0)

// E4M3 (OCP, no infinities) code of x, round to nearest even; |x| >= 448 saturates to 448.
uint8_t host_fp8(float x) {
    const uint8_t sign = std::signbit(x) ? 0x80 : 0;
    const float a = std::fabs(x);
    if (a >= FP8_MAX) return sign | 0x7E;
    if (a < 0.015625f) return sign | (uint8_t)std::nearbyint(a * 512.0f);  // subnormal (8 -> 2^-6, code 8)
    int e = std::ilogb(a);
    int q = (int)std::nearbyint((std::ldexp(a, -e) - 1.0f) * 8.0f);
    if (q == 8) {
        e++;
        q = 0;
    }
    const int code = ((e + 7) << 3) | q;
    return sign | (uint8_t)std::min(code, 0x7E);
}

struct Pnet {
    int device = 0, cap = 0, step = 0, quantized = 0;
    int launch_limit = 0;  // most workgroups a features or inference launch uses (0: no limit)
    // Features batch state: device copies of the descriptors and the mean images.
    void *fpics = nullptr, *mplanes = nullptr;
    uint8_t *means = nullptr;
    size_t fpics_cap = 0, mplanes_cap = 0, means_cap = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr, *qs = nullptr, *loss = nullptr;
    uint16_t *pb = nullptr;
    uint8_t *w8 = nullptr;
    uint16_t *h1b = nullptr, *h2b = nullptr, *dzb = nullptr, *dh2b = nullptr, *dh1b = nullptr;
    float *zf = nullptr, *dzf = nullptr, *dh2f = nullptr, *dh1f = nullptr, *part = nullptr;
    std::vector<void *> owned;
};

template <typename T>
hipError_t dalloc(Pnet *p, T **out, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    p->owned.push_back(q);
    *out = (T *)q;
    return hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), p->stream);
}

void free_net(Pnet *p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (void *q : p->owned) (void)hipFree(q);
    if (p->fpics) (void)hipFree(p->fpics);
    if (p->mplanes) (void)hipFree(p->mplanes);
    if (p->means) (void)hipFree(p->means);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

hipError_t gemm(Pnet *p, const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, float *C,
                uint16_t *Cb, long ldc, const float *bias, const uint16_t *mask, int M, int N, int K, int relu,
                int accumulate) {
    GemmArgs g{A, sam, sak, B, sbk, sbn, C, Cb, ldc, bias, mask, M, N, K, relu, accumulate};
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, p->stream, g);
    return hipGetLastError();
}

hipError_t colsum(Pnet *p, const float *G, int M, int N, float *out) {
    const int chunks = (M + 63) / 64;
    hipLaunchKernelGGL(k_colsum_part, dim3(chunks), dim3(256), 0, p->stream, G, M, N, p->part);
    hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(256), 0, p->stream, (const float *)p->part, chunks, N, out);
    return hipGetLastError();
}

// Forward pass over n <= cap rows: H1b, H2b (bfloat16) and logits Z (fp32).
hipError_t forward(Pnet *p, const uint16_t *X, int n, float *Z) {
    hipError_t e;
    if ((e = gemm(p, X, F_IN, 1, p->pb + OFF_W1, HID, 1, nullptr, p->h1b, HID, p->params + OFF_B1, nullptr, n, HID,
                  F_IN, 1, 0)) != hipSuccess)
        return e;
    if ((e = gemm(p, p->h1b, HID, 1, p->pb + OFF_W2, HID, 1, nullptr, p->h2b, HID, p->params + OFF_B2, nullptr, n,
                  HID, HID, 1, 0)) != hipSuccess)
        return e;
    return gemm(p, p->h2b, HID, 1, p->pb + OFF_W3, OUT, 1, Z, nullptr, OUT, p->params + OFF_B3, nullptr, n, OUT, HID,
                0, 0);
}

// Loss gradients of a batch into p->grads (and the summed losses into p->loss).
hipError_t backward(Pnet *p, const uint16_t *X, const rgh_pnet_label *L, int B) {
    hipError_t e;
    if ((e = hipMemsetAsync(p->loss, 0, 2 * sizeof(float), p->stream)) != hipSuccess) return e;
    if ((e = forward(p, X, B, p->zf)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_xent, dim3((B + 127) / 128), dim3(128), 0, p->stream, (const float *)p->zf, L, B,
                       1.0f / (float)B, p->dzf, p->dzb, p->loss);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // Layer 3: dW3 = H2^T dZ, db3 = sum dZ, dH2 = (dZ W3^T) * [H2 > 0].
    if ((e = gemm(p, p->h2b, 1, HID, p->dzb, OUT, 1, p->grads + OFF_W3, nullptr, OUT, nullptr, nullptr, HID, OUT, B, 0,
                  0)) != hipSuccess)
        return e;
    if ((e = colsum(p, p->dzf, B, OUT, p->grads + OFF_B3)) != hipSuccess) return e;
    if ((e = gemm(p, p->dzb, OUT, 1, p->pb + OFF_W3, 1, OUT, p->dh2f, p->dh2b, HID, nullptr, p->h2b, B, HID, OUT, 0,
                  0)) != hipSuccess)
        return e;
    // Layer 2.
    if ((e = gemm(p, p->h1b, 1, HID, p->dh2b, HID, 1, p->grads + OFF_W2, nullptr, HID, nullptr, nullptr, HID, HID, B, 0,
                  0)) != hipSuccess)
        return e;
    if ((e = colsum(p, p->dh2f, B, HID, p->grads + OFF_B2)) != hipSuccess) return e;
    if ((e = gemm(p, p->dh2b, HID, 1, p->pb + OFF_W2, 1, HID, p->dh1f, p->dh1b, HID, nullptr, p->h1b, B, HID, HID, 0,
                  0)) != hipSuccess)
        return e;
    // Layer 1.
    if ((e = gemm(p, X, 1, F_IN, p->dh1b, HID, 1, p->grads + OFF_W1, nullptr, HID, nullptr, nullptr, F_IN, HID, B, 0,
                  0)) != hipSuccess)
        return e;
    return colsum(p, p->dh1f, B, HID, p->grads + OFF_B1);
}

// Grows a device buffer to at least `bytes`.
hipError_t grow(void **buf, size_t *cap, size_t bytes) {
    if (*cap >= bytes) return hipSuccess;
    if (*buf) {
        hipError_t e = hipFree(*buf);
        if (e != hipSuccess) return e;
    }
    *buf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(buf, bytes);
    if (e == hipSuccess) *cap = bytes;
    return e;
}

// Checks a batch of pictures, then computes one 4x4-mean image per distinct (plane, stride, size)
// into p->means (k_means, enqueued on p->stream); mean_of[i] receives the (cur, prev) mean
// planes of picture i.
rgh_status prepare_means(Pnet *p, const rgh_pnet_picture *pics, int32_t n, bool need_features,
                         std::vector<std::pair<MeanPlane, MeanPlane>> &mean_of, int &max_mbs) {
    max_mbs = 0;
    for (int i = 0; i < n; i++) {
        const rgh_pnet_picture &q = pics[i];
        if (!q.cur_y || !q.prev_y || (need_features && !q.features)) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer");
        if (q.width < 1 || q.height < 1 || q.width > 16384 || q.height > 16384 || q.cur_stride < q.width ||
            q.prev_stride < q.width)
            FAIL(RGH_ERR_INVALID_DIMENSIONS, "bad picture size or stride");
        if (q.qp < 0 || q.qp > 51) FAIL(RGH_ERR_INVALID_ARGUMENT, "qp outside 0..51");
        max_mbs = std::max(max_mbs, ((q.width + 15) / 16) * ((q.height + 15) / 16));
    }
    if (n == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    std::vector<MeanPlane> planes;
    std::vector<size_t> offs;
    size_t total = 0;
    int max_cells = 0;
    auto plane_of = [&](const uint8_t *y, int stride, int W, int H) {
        for (size_t k = 0; k < planes.size(); k++)
            if (planes[k].y == y && planes[k].stride == stride && planes[k].W == W && planes[k].H == H) return (int)k;
        // Cells -6 .. 4 * mb + 6 in each direction: every macroblock's window and cells.
        MeanPlane m{y, nullptr, stride, W, H, 4 * ((W + 15) / 16) + 2 * MPAD, 4 * ((H + 15) / 16) + 2 * MPAD};
        planes.push_back(m);
        offs.push_back(total);
        total += (size_t)m.mw * m.mh;
        max_cells = std::max(max_cells, m.mw * m.mh);
        return (int)planes.size() - 1;
    };
    std::vector<std::pair<int, int>> pair_of(n);
    for (int i = 0; i < n; i++)
        pair_of[i] = {plane_of(pics[i].cur_y, pics[i].cur_stride, pics[i].width, pics[i].height),
                      plane_of(pics[i].prev_y, pics[i].prev_stride, pics[i].width, pics[i].height)};
    if (planes.size() > 65535) FAIL(RGH_ERR_INVALID_ARGUMENT, "too many distinct planes");
    TRY(grow((void **)&p->means, &p->means_cap, total));
    for (size_t k = 0; k < planes.size(); k++) planes[k].m = p->means + offs[k];
    mean_of.resize((size_t)n);
    for (int i = 0; i < n; i++) mean_of[(size_t)i] = {planes[(size_t)pair_of[i].first], planes[(size_t)pair_of[i].second]};
    TRY(grow(&p->mplanes, &p->mplanes_cap, sizeof(MeanPlane) * planes.size()));
    TRY(hipMemcpyAsync(p->mplanes, planes.data(), sizeof(MeanPlane) * planes.size(), hipMemcpyHostToDevice,
                       p->stream));
    hipLaunchKernelGGL(k_means, dim3((max_cells + 255) / 256, (unsigned)planes.size()), dim3(256), 0, p->stream,
                       (const MeanPlane *)p->mplanes);
    TRY(hipGetLastError());
    return RGH_OK;
}

}  // namespace

extern "C" {

rgh_status rgh_pnet_create(int32_t device, int32_t max_batch, void **net) {
    if (!net || max_batch < 1 || max_batch > (1 << 20)) FAIL(RGH_ERR_INVALID_ARGUMENT, "bad net pointer or max_batch");
    *net = nullptr;
    int count = 0;
    TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) FAIL(RGH_ERR_INVALID_ARGUMENT, "no such device");
    TRY(hipSetDevice(device));
    Pnet *p = new Pnet;
    p->device = device;
    p->cap = max_batch;
    hipError_t e = hipStreamCreate(&p->stream);
    const size_t B = (size_t)max_batch;
    if (e == hipSuccess) e = dalloc(p, &p->params, NPARAM);
    if (e == hipSuccess) e = dalloc(p, &p->grads, NPARAM);
    if (e == hipSuccess) e = dalloc(p, &p->m, NPARAM);
    if (e == hipSuccess) e = dalloc(p, &p->v, NPARAM);
    if (e == hipSuccess) e = dalloc(p, &p->pb, NPARAM);
    if (e == hipSuccess) e = dalloc(p, &p->w8, Q_BYTES);
    if (e == hipSuccess) e = dalloc(p, &p->qs, RGH_PNET_SCALES);
    if (e == hipSuccess) e = dalloc(p, &p->loss, 2);
    if (e == hipSuccess) e = dalloc(p, &p->h1b, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->h2b, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->dh2b, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->dh1b, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->dzb, B * OUT);
    if (e == hipSuccess) e = dalloc(p, &p->zf, B * OUT);
    if (e == hipSuccess) e = dalloc(p, &p->dzf, B * OUT);
    if (e == hipSuccess) e = dalloc(p, &p->dh2f, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->dh1f, B * HID);
    if (e == hipSuccess) e = dalloc(p, &p->part, ((B + 63) / 64) * OUT);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        rgh_note_error("rgh_pnet_create", hipGetErrorString(e), "rgh_learned.hip", __LINE__);
        free_net(p);
        return hip_status(e);
    }
    *net = p;
    return RGH_OK;
}

rgh_status rgh_pnet_destroy(void *net) {
    free_net((Pnet *)net);
    return RGH_OK;
}

rgh_status rgh_pnet_set_launch_limit(void *net, int32_t max_workgroups) {
    Pnet *p = (Pnet *)net;
    if (!p || max_workgroups < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null net or negative limit");
    p->launch_limit = max_workgroups;
    return RGH_OK;
}

rgh_status rgh_pnet_set_params(void *net, const float *params, int64_t n) {
    Pnet *p = (Pnet *)net;
    if (!p || !params || n != NPARAM) FAIL(RGH_ERR_INVALID_ARGUMENT, "expected RGH_PNET_PARAMS parameters");
    TRY(hipSetDevice(p->device));
    TRY(hipMemcpyAsync(p->params, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, p->stream));
    TRY(hipMemsetAsync(p->m, 0, NPARAM * sizeof(float), p->stream));
    TRY(hipMemsetAsync(p->v, 0, NPARAM * sizeof(float), p->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, p->stream, (const float *)p->params, p->pb,
                       NPARAM);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    p->step = 0;
    p->quantized = 0;
    return RGH_OK;
}

rgh_status rgh_pnet_get_params(void *net, float *params, int64_t n) {
    Pnet *p = (Pnet *)net;
    if (!p || !params || n != NPARAM) FAIL(RGH_ERR_INVALID_ARGUMENT, "expected RGH_PNET_PARAMS parameters");
    TRY(hipSetDevice(p->device));
    TRY(hipMemcpyAsync(params, p->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_features_batch(void *net, const rgh_pnet_picture *pics, int32_t n) {
    Pnet *p = (Pnet *)net;
    if (!p || !pics || n < 0 || n > 65535) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or bad picture count");
    std::vector<std::pair<MeanPlane, MeanPlane>> mean_of;
    int max_mbs = 0;
    rgh_status st = prepare_means(p, pics, n, true, mean_of, max_mbs);
    if (st != RGH_OK || n == 0) return st;
    std::vector<FeatPic> fp(n);
    for (int i = 0; i < n; i++) {
        const MeanPlane &c = mean_of[(size_t)i].first, &q = mean_of[(size_t)i].second;
        fp[i] = FeatPic{pics[i].cur_y, c.m, q.m, pics[i].features, pics[i].cur_stride, pics[i].width, pics[i].height,
                        c.mw, pics[i].qp};
    }
    TRY(grow(&p->fpics, &p->fpics_cap, sizeof(FeatPic) * (size_t)n));
    TRY(hipMemcpyAsync(p->fpics, fp.data(), sizeof(FeatPic) * (size_t)n, hipMemcpyHostToDevice, p->stream));
    const long items = (long)n * max_mbs;
    const long groups = (items + 3) / 4;
    hipLaunchKernelGGL(k_features, dim3((unsigned)std::min<long>(groups, 1L << 30)), dim3(256), 0, p->stream,
                       (const FeatPic *)p->fpics, max_mbs, items);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_propose_batch(void *net, const rgh_pnet_picture *pics, int32_t n, rgh_proposal *proposals,
                                  float *logits) {
    Pnet *p = (Pnet *)net;
    if (!p || !pics || !proposals || n < 0 || n > 65535) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or bad picture count");
    if (!p->quantized) FAIL(RGH_ERR_INVALID_ARGUMENT, "rgh_pnet_quantize has not run since the parameters changed");
    std::vector<std::pair<MeanPlane, MeanPlane>> mean_of;
    int max_mbs = 0;
    rgh_status st = prepare_means(p, pics, n, false, mean_of, max_mbs);
    if (st != RGH_OK || n == 0) return st;
    // Descriptors, then the tile table (first tile of each picture, and the count) behind them.
    std::vector<ProposePic> pp((size_t)n);
    std::vector<int> tile0((size_t)n + 1);
    int rows = 0, tiles = 0;
    for (int i = 0; i < n; i++) {
        const MeanPlane &c = mean_of[(size_t)i].first, &q = mean_of[(size_t)i].second;
        const int mbs = ((pics[i].width + 15) / 16) * ((pics[i].height + 15) / 16);
        pp[(size_t)i] = ProposePic{pics[i].cur_y, c.m, q.m, pics[i].cur_stride, pics[i].width, pics[i].height, c.mw,
                                   pics[i].qp, rows};
        tile0[(size_t)i] = tiles;
        rows += mbs;
        tiles += (mbs + IT - 1) / IT;
    }
    tile0[(size_t)n] = tiles;
    const size_t dbytes = sizeof(ProposePic) * (size_t)n, tbytes = sizeof(int) * ((size_t)n + 1);
    TRY(grow(&p->fpics, &p->fpics_cap, dbytes + tbytes));
    TRY(hipMemcpyAsync(p->fpics, pp.data(), dbytes, hipMemcpyHostToDevice, p->stream));
    TRY(hipMemcpyAsync((uint8_t *)p->fpics + dbytes, tile0.data(), tbytes, hipMemcpyHostToDevice, p->stream));
    int cus = 0;
    TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    int grid = std::max(1, std::min(tiles, 2 * cus));
    if (p->launch_limit > 0) grid = std::min(grid, p->launch_limit);
    ProposeArgs g{(const ProposePic *)p->fpics, (const int *)((uint8_t *)p->fpics + dbytes), n, p->w8, p->qs,
                  p->params, proposals, logits};
    TRY(hipFuncSetAttribute((const void *)k_propose, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sizeof(ProposeLds)));
    hipLaunchKernelGGL(k_propose, dim3(grid), dim3(256), sizeof(ProposeLds), p->stream, g);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_features(void *net, const uint8_t *cur, int32_t cur_stride, const uint8_t *prev,
                             int32_t prev_stride, int32_t width, int32_t height, int32_t qp, uint16_t *features) {
    const rgh_pnet_picture q{cur, prev, features, cur_stride, prev_stride, width, height, qp, 0};
    return rgh_pnet_features_batch(net, &q, 1);
}

rgh_status rgh_pnet_labels(void *net, const rgh_mb_decision *dec, int32_t n, rgh_pnet_label *labels) {
    Pnet *p = (Pnet *)net;
    if (!p || !dec || !labels || n < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    if (n == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    hipLaunchKernelGGL(k_labels, dim3((n + 255) / 256), dim3(256), 0, p->stream, dec, (int)n, labels);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_gather(void *net, const uint16_t *F, const rgh_pnet_label *L, const int32_t *idx, int32_t n,
                           uint16_t *Fo, rgh_pnet_label *Lo) {
    Pnet *p = (Pnet *)net;
    if (!p || !F || !L || !idx || !Fo || !Lo || n < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    if (n == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    hipLaunchKernelGGL(k_gather, dim3(n), dim3(64), 0, p->stream, F, L, (const int *)idx, (int)n, Fo, Lo);
    TRY(hipGetLastError());
    return RGH_OK;
}

rgh_status rgh_pnet_train_step(void *net, const uint16_t *X, const rgh_pnet_label *L, int32_t batch, float lr,
                               float *loss) {
    Pnet *p = (Pnet *)net;
    if (!p || !X || !L || batch < 1 || batch > p->cap || !(lr > 0.0f)) FAIL(RGH_ERR_INVALID_ARGUMENT, "bad batch or rate");
    TRY(hipSetDevice(p->device));
    TRY(backward(p, X, L, batch));
    p->step++;
    const float c1 = 1.0f - std::pow(0.9f, (float)p->step), c2 = 1.0f - std::pow(0.999f, (float)p->step);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, p->stream, p->params, (const float *)p->grads,
                       p->m, p->v, p->pb, NPARAM, lr, c1, c2);
    TRY(hipGetLastError());
    p->quantized = 0;
    if (loss) {
        TRY(hipMemcpyAsync(loss, p->loss, 2 * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        TRY(hipStreamSynchronize(p->stream));
    }
    return RGH_OK;
}

rgh_status rgh_pnet_gradients(void *net, const uint16_t *X, const rgh_pnet_label *L, int32_t batch, float *grads,
                              float *logits) {
    Pnet *p = (Pnet *)net;
    if (!p || !X || !L || batch < 1 || batch > p->cap) FAIL(RGH_ERR_INVALID_ARGUMENT, "bad batch");
    TRY(hipSetDevice(p->device));
    TRY(backward(p, X, L, batch));
    if (grads) TRY(hipMemcpyAsync(grads, p->grads, NPARAM * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    if (logits)
        TRY(hipMemcpyAsync(logits, p->zf, (size_t)batch * OUT * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_forward(void *net, const uint16_t *X, int32_t n, float *logits) {
    Pnet *p = (Pnet *)net;
    if (!p || !X || !logits || n < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    TRY(hipSetDevice(p->device));
    for (int r = 0; r < n; r += p->cap) {
        const int k = std::min(p->cap, n - r);
        TRY(forward(p, X + (size_t)r * F_IN, k, logits + (size_t)r * OUT));
    }
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

// Quantizes the current parameters with the given hidden-activation maxima (see the header).
static rgh_status quantize_ranges(Pnet *p, float h1max, float h2max) {
    std::vector<float> P(NPARAM);
    TRY(hipMemcpyAsync(P.data(), p->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    // Layer inputs: features as they are (all below 8); hidden activations scaled so that their
    // maximum maps to 448.
    const float mul[3] = {1.0f, h1max > 0.0f ? FP8_MAX / h1max : 1.0f, h2max > 0.0f ? FP8_MAX / h2max : 1.0f};
    std::vector<uint8_t> w8(Q_BYTES);
    std::vector<float> qs(RGH_PNET_SCALES, 0.0f);
    const int offs[3] = {OFF_W1, OFF_W2, OFF_W3}, qoff[3] = {Q_W1, Q_W2, Q_W3}, kin[3] = {F_IN, HID, HID},
              nout[3] = {HID, HID, OUT};
    for (int l = 0; l < 3; l++) {
        for (int c = 0; c < nout[l]; c++) {
            float wmax = 0.0f;
            for (int r = 0; r < kin[l]; r++) wmax = std::max(wmax, std::fabs(P[offs[l] + r * nout[l] + c]));
            const float ws = wmax > 0.0f ? wmax / FP8_MAX : 1.0f, inv = 1.0f / ws;
            for (int r = 0; r < kin[l]; r++) w8[qoff[l] + c * kin[l] + r] = host_fp8(P[offs[l] + r * nout[l] + c] * inv);
            qs[l * OUT + c] = ws / mul[l];
        }
        qs[3 * OUT + l] = mul[l];
    }
    TRY(hipMemcpyAsync(p->w8, w8.data(), Q_BYTES, hipMemcpyHostToDevice, p->stream));
    TRY(hipMemcpyAsync(p->qs, qs.data(), qs.size() * sizeof(float), hipMemcpyHostToDevice, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    p->quantized = 1;
    return RGH_OK;
}

rgh_status rgh_pnet_quantize(void *net, const uint16_t *calib, int32_t n) {
    Pnet *p = (Pnet *)net;
    if (!p || !calib || n < 1) FAIL(RGH_ERR_INVALID_ARGUMENT, "calibration rows required");
    TRY(hipSetDevice(p->device));
    const int k = std::min<int>(n, p->cap);
    TRY(forward(p, calib, k, p->zf));
    std::vector<uint16_t> h1((size_t)k * HID), h2((size_t)k * HID);
    TRY(hipMemcpyAsync(h1.data(), p->h1b, h1.size() * 2, hipMemcpyDeviceToHost, p->stream));
    TRY(hipMemcpyAsync(h2.data(), p->h2b, h2.size() * 2, hipMemcpyDeviceToHost, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    auto bf = [](uint16_t h) {
        uint32_t u = (uint32_t)h << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    float m1 = 0.0f, m2 = 0.0f;
    for (uint16_t h : h1) m1 = std::max(m1, bf(h));
    for (uint16_t h : h2) m2 = std::max(m2, bf(h));
    return quantize_ranges(p, m1, m2);
}

rgh_status rgh_pnet_quantize_ranges(void *net, float h1_max, float h2_max) {
    Pnet *p = (Pnet *)net;
    if (!p || !(h1_max >= 0.0f) || !(h2_max >= 0.0f)) FAIL(RGH_ERR_INVALID_ARGUMENT, "bad activation maxima");
    TRY(hipSetDevice(p->device));
    return quantize_ranges(p, h1_max, h2_max);
}

rgh_status rgh_pnet_infer(void *net, const uint16_t *X, int32_t n, rgh_proposal *out, float *logits) {
    Pnet *p = (Pnet *)net;
    if (!p || !X || !out || n < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    if (!p->quantized) FAIL(RGH_ERR_INVALID_ARGUMENT, "rgh_pnet_quantize has not run since the parameters changed");
    if (n == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    int cus = 0;
    TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    int grid = std::max(1, std::min((n + IT - 1) / IT, cus));
    if (p->launch_limit > 0) grid = std::min(grid, p->launch_limit);
    InferArgs g{X, (int)n, p->w8, p->qs, p->params, out, logits};
    TRY(hipFuncSetAttribute((const void *)k_infer, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(InferLds)));
    hipLaunchKernelGGL(k_infer, dim3(grid), dim3(256), sizeof(InferLds), p->stream, g);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_get_quantized(void *net, uint8_t *w8, float *scales) {
    Pnet *p = (Pnet *)net;
    if (!p) FAIL(RGH_ERR_INVALID_ARGUMENT, "null net");
    if (!p->quantized) FAIL(RGH_ERR_INVALID_ARGUMENT, "rgh_pnet_quantize has not run since the parameters changed");
    TRY(hipSetDevice(p->device));
    if (w8) TRY(hipMemcpyAsync(w8, p->w8, Q_BYTES, hipMemcpyDeviceToHost, p->stream));
    if (scales) TRY(hipMemcpyAsync(scales, p->qs, RGH_PNET_SCALES * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_fp8_encode(void *net, const float *x, int32_t n, uint8_t *out) {
    Pnet *p = (Pnet *)net;
    if (!p || !x || !out || n < 0) FAIL(RGH_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    if (n == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    hipLaunchKernelGGL(k_fp8_encode, dim3((n + 255) / 256), dim3(256), 0, p->stream, x, (int)n, out);
    TRY(hipGetLastError());
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

rgh_status rgh_pnet_gemm_bf16(void *net, const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk,
                              int64_t sbn, float *C, uint16_t *Cb, int64_t ldc, const float *bias, const uint16_t *mask,
                              int32_t M, int32_t N, int32_t K, int32_t relu, int32_t accumulate) {
    Pnet *p = (Pnet *)net;
    if (!p || !A || !B || (!C && !Cb) || M < 0 || N < 0 || K < 0 || ldc < N || (accumulate && !C))
        FAIL(RGH_ERR_INVALID_ARGUMENT, "bad GEMM arguments");
    if (M == 0 || N == 0) return RGH_OK;
    TRY(hipSetDevice(p->device));
    TRY(gemm(p, A, sam, sak, B, sbk, sbn, C, Cb, ldc, bias, mask, M, N, K, relu, accumulate));
    TRY(hipStreamSynchronize(p->stream));
    return RGH_OK;
}

}  // extern "C"
