// In-loop deblocking (ITU-T H.264 8.7) of an analysed P frame, one wave64 per macroblock in
// the same anti-diagonal wavefront order as the analysis, and border extension of the deblocked
// picture, which becomes the next reference. Device code only: the frame-step kernel that runs
// these items lives with the analysis (rgh_analyse.hip).
//
// Restates ragdoll's `EncState::deblock` (crates/ragdoll-codec/src/h264/enc.rs:3776-3881), its
// `EdgeFilter` (:3902-3942), the edge kernels of crates/ragdoll-simd/src/h264_deblock.rs
// (:69 luma, :206 chroma scalar references) and `Plane::extend_borders` (enc.rs:164-178).
//
// Wavefront safety: macroblock (x, y) filters its left and top edges, writing up to 3 samples
// into its left and top neighbours and reading 4; every macroblock whose filtering it must see
// ((x-1, y), (x, y-1), (x+1, y-1)) lies on an earlier step t = x + 2y, every macroblock that
// must see its result lies on a later one, and macroblocks of one step are 32 or more samples
// apart. Each macroblock works on a 20x20 luma (10x12 chroma) copy in LDS and writes it back.
//
// Border extension is split by macroblock: the padding samples whose nearest interior sample
// lies in border macroblock (x, y) are written once that macroblock is final, i.e. after the
// deblock items (x + 1, y) and (x, y + 1) that still modify it.
#pragma once
#include "rgh_device.h"
#include "rgh_tables.h"

namespace {

struct DbLds {
    uint8_t y[20 * 20];   // rows y0-4 .., columns x0-4 ..
    uint8_t u[10 * 12];   // rows cy0-2 .., columns cx0-4 ..
    uint8_t v[10 * 12];
};

struct Edge {
    int on;       // filter this edge
    int alpha, beta, strong;
    int tc0[4];   // per 4-line segment; -1 = segment not filtered
};

// enc.rs:3911 EdgeFilter::new
__device__ Edge make_edge(int qp, const int bs[4]) {
    Edge E;
    int ia = iclamp(qp, 0, 51);
    E.alpha = c_alpha[ia];
    E.beta = c_beta[ia];
    E.on = E.alpha != 0 && E.beta != 0;
    E.strong = bs[0] == 4;
#pragma unroll
    for (int k = 0; k < 4; k++) E.tc0[k] = bs[k] == 0 ? -1 : (bs[k] >= 4 ? 0 : (int)c_tc0[ia][bs[k] - 1]);
    return E;
}

// tc0 of 4-line segment g, selected without indexing the array (a dynamic index would put
// the edge in scratch memory).
__device__ __forceinline__ int seg_tc0(const Edge &E, int g) {
    return g == 0 ? E.tc0[0] : (g == 1 ? E.tc0[1] : (g == 2 ? E.tc0[2] : E.tc0[3]));
}

// simd/h264_deblock.rs:69 h264_luma_deblock_scalar, one line (p3..q3 at q0 - 4s .. q0 + 3s).
__device__ void luma_line(uint8_t *b, int q0, int s, const Edge &E, int k) {
    int tc0 = seg_tc0(E, k >> 2);
    if (!E.strong && tc0 < 0) return;
    int p0 = b[q0 - s], q0v = b[q0], p1 = b[q0 - 2 * s], q1 = b[q0 + s];
    if (iabs(p0 - q0v) >= E.alpha || iabs(p1 - p0) >= E.beta || iabs(q1 - q0v) >= E.beta) return;
    int p2 = b[q0 - 3 * s], q2 = b[q0 + 2 * s];
    int ap = iabs(p2 - p0) < E.beta, aq = iabs(q2 - q0v) < E.beta;
    if (!E.strong) {
        int tc = tc0 + ap + aq;
        int d = iclamp((((q0v - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        b[q0 - s] = (uint8_t)clip255(p0 + d);
        b[q0] = (uint8_t)clip255(q0v - d);
        if (ap) b[q0 - 2 * s] = (uint8_t)(p1 + iclamp((p2 + ((p0 + q0v + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0));
        if (aq) b[q0 + s] = (uint8_t)(q1 + iclamp((q2 + ((p0 + q0v + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0));
    } else {
        int p3 = b[q0 - 4 * s], q3 = b[q0 + 3 * s];
        int small = iabs(p0 - q0v) < ((E.alpha >> 2) + 2);
        if (ap && small) {
            b[q0 - s] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0v + q1 + 4) >> 3);
            b[q0 - 2 * s] = (uint8_t)((p2 + p1 + p0 + q0v + 2) >> 2);
            b[q0 - 3 * s] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0v + 4) >> 3);
        } else {
            b[q0 - s] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        }
        if (aq && small) {
            b[q0] = (uint8_t)((p1 + 2 * p0 + 2 * q0v + 2 * q1 + q2 + 4) >> 3);
            b[q0 + s] = (uint8_t)((p0 + q0v + q1 + q2 + 2) >> 2);
            b[q0 + 2 * s] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0v + p0 + 4) >> 3);
        } else {
            b[q0] = (uint8_t)((2 * q1 + q0v + p1 + 2) >> 2);
        }
    }
}

// simd/h264_deblock.rs:206 h264_chroma_deblock_scalar, one line (p1..q1).
__device__ void chroma_line(uint8_t *b, int q0, int s, const Edge &E, int k) {
    int tc0 = seg_tc0(E, k >> 1);
    if (!E.strong && tc0 < 0) return;
    int p1 = b[q0 - 2 * s], p0 = b[q0 - s], q0v = b[q0], q1 = b[q0 + s];
    if (iabs(p0 - q0v) >= E.alpha || iabs(p1 - p0) >= E.beta || iabs(q1 - q0v) >= E.beta) return;
    if (E.strong) {
        b[q0 - s] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        b[q0] = (uint8_t)((2 * q1 + q0v + p1 + 2) >> 2);
    } else {
        int tc = tc0 + 1;
        int d = iclamp((((q0v - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        b[q0 - s] = (uint8_t)clip255(p0 + d);
        b[q0] = (uint8_t)clip255(q0v - d);
    }
}

__device__ __forceinline__ int is_intra(const rgh_mb_decision &d) {
    return d.mb_type == RGH_MB_I16X16 || d.mb_type == RGH_MB_I4X4 || d.mb_type == RGH_MB_I_PCM;
}
__device__ __forceinline__ int quiet(const rgh_mb_decision &d) { return !is_intra(d) && d.cbp_luma == 0; }
__device__ __forceinline__ uint32_t nz_col(uint32_t nnz, int c) {
    return ((nnz >> c) & 1) | (((nnz >> (4 + c)) & 1) << 1) | (((nnz >> (8 + c)) & 1) << 2) | (((nnz >> (12 + c)) & 1) << 3);
}

// Vector of quadrant k (raster) of a record in the quadrant layout (words 2-5); zero for intra.
__device__ __forceinline__ uint32_t quad_word(const rgh_mb_decision &d, int k) {
    const uint32_t w0 = (uint32_t)(uint16_t)d.mv[0] | ((uint32_t)(uint16_t)d.mv[1] << 16);
    const uint32_t w1 = (uint32_t)(uint16_t)d.mvp[0] | ((uint32_t)(uint16_t)d.mvp[1] << 16);
    const uint32_t w2 = d.cost, w3 = (uint32_t)d.n_sad | ((uint32_t)d.n_subpel << 16);
    return is_intra(d) ? 0u : (k == 0 ? w0 : (k == 1 ? w1 : (k == 2 ? w2 : w3)));
}
// Boundary strength 0 / 1 of two inter 4x4 blocks without coefficients: their vectors (packed
// x, y) four or more quarter samples apart in a component (8.7.2.1, one reference).
__device__ __forceinline__ int motion_bs(uint32_t a, uint32_t b) {
    return iabs((int)(int16_t)(a & 0xFFFF) - (int)(int16_t)(b & 0xFFFF)) >= 4 ||
           iabs((int)(int16_t)(a >> 16) - (int)(int16_t)(b >> 16)) >= 4;
}

__device__ __forceinline__ void deblock_mb_inline(const RghJobD &J, int mb_x, int mb_y, DbLds &S) {
    const int lane = threadIdx.x, mb_w = J.mb_w, m = mb_y * mb_w + mb_x;
    const rgh_mb_decision Q = load_dec(&J.dec[m]);
    rgh_mb_decision L{}, T{};
    if (mb_x > 0) L = load_dec(&J.dec[m - 1]);
    if (mb_y > 0) T = load_dec(&J.dec[m - mb_w]);
    auto mveq = [](const rgh_mb_decision &a, const rgh_mb_decision &b) { return a.mv[0] == b.mv[0] && a.mv[1] == b.mv[1]; };
    // Compact pictures and pictures with P partitions: every 4x4 block carries its quadrant's
    // vector (enc/inter.rs:781 deblock_parts); the internal edges of an inter macroblock without coefficients and with
    // one vector throughout are not filtered.
    const bool parts = J.compact != 0 || J.parts != 0;
    const bool uniform = !is_intra(Q) && Q.nnz == 0 && quad_word(Q, 0) == quad_word(Q, 1) &&
                         quad_word(Q, 0) == quad_word(Q, 2) && quad_word(Q, 0) == quad_word(Q, 3);
    if (!parts && quiet(Q) && (mb_x == 0 || (quiet(L) && mveq(L, Q))) && (mb_y == 0 || (quiet(T) && mveq(T, Q)))) return;

    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8;
    // Stage luma rows y0-4 .. y0+15, columns x0-4 .. x0+15 (5 aligned dwords a row) and chroma
    // rows cy0-2 .. cy0+7, columns cx0-4 .. cx0+7 (3 aligned dwords a row) of both planes; all
    // loads are issued before the LDS stores.
    uint32_t ly[2], lc;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int i = lane + 64 * k;
        ly[k] = i < 100 ? *(const RGH_G uint32_t *)&J.cur_y[(long)(y0 - 4 + i / 5) * J.sy + x0 - 4 + (i % 5) * 4] : 0u;
    }
    {
        const int p = lane / 30, i = lane % 30;
        const RGH_G uint8_t *pl = p ? J.cur_v : J.cur_u;
        lc = lane < 60 ? *(const RGH_G uint32_t *)&pl[(long)(cy0 - 2 + i / 3) * J.sc + cx0 - 4 + (i % 3) * 4] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int i = lane + 64 * k;
        if (i < 100) *(uint32_t *)&S.y[(i / 5) * 20 + (i % 5) * 4] = ly[k];
    }
    if (lane < 60) {
        const int p = lane / 30, i = lane % 30;
        *(uint32_t *)&(p ? S.v : S.u)[(i / 3) * 12 + (i % 3) * 4] = lc;
    }
    __syncthreads();

    const int intra = is_intra(Q), internal = parts ? !uniform : (intra || Q.cbp_luma != 0), skip_odd = Q.transform_8x8;
    const int qp_q = Q.qp, qpc_q = chroma_qp_of(qp_q, J.cqo);
    for (int dir = 0; dir < 2; dir++) {
        for (int e = 0; e < 4; e++) {
            // The macroblock on the p side of the edge: the left or top neighbour for edge 0, else
            // this one (selected field by field: a pointer to a local record would live in
            // scratch memory).
            const bool ext = e == 0;
            if (ext) {
                if (!(dir == 0 ? mb_x > 0 : mb_y > 0)) continue;
            } else if (!internal || ((e & 1) && skip_odd)) {
                continue;
            }
            const rgh_mb_decision &N = dir == 0 ? L : T;
            const int p_intra = ext ? is_intra(N) : intra;
            const uint32_t p_nnz = ext ? N.nnz : Q.nnz;
            const int p_mvx = ext ? N.mv[0] : Q.mv[0], p_mvy = ext ? N.mv[1] : Q.mv[1];
            const int qp_p = ext ? N.qp : Q.qp;
            int bs[4];
            if (intra || p_intra) {
#pragma unroll
                for (int k = 0; k < 4; k++) bs[k] = ext ? 4 : 3;
            } else if (parts) {
                // Segment k: q block (e, k) (vertical edges) or (k, e) of this macroblock, p
                // block one to the left / above, in the neighbour for edge 0.
                int any = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int qx = dir == 0 ? e : k, qy = dir == 0 ? k : e;
                    const int px = dir == 0 ? (e ? e - 1 : 3) : k, py = dir == 0 ? k : (e ? e - 1 : 3);
                    const int qn = (Q.nnz >> (4 * qy + qx)) & 1, pn = (p_nnz >> (4 * py + px)) & 1;
                    const uint32_t qv = quad_word(Q, (qx >> 1) + (qy & 2));
                    const int pk = (px >> 1) + (py & 2);
                    const uint32_t pv = ext ? (dir == 0 ? quad_word(L, pk) : quad_word(T, pk)) : quad_word(Q, pk);
                    bs[k] = (qn | pn) ? 2 : motion_bs(qv, pv);
                    any |= bs[k];
                }
                if (!any) continue;
            } else {
                uint32_t coded;
                if (dir == 0 && ext)
                    coded = nz_col(Q.nnz, 0) | nz_col(p_nnz, 3);
                else if (dir == 0)
                    coded = nz_col(Q.nnz, e) | nz_col(Q.nnz, e - 1);
                else if (ext)
                    coded = (Q.nnz & 0xF) | (p_nnz >> 12);
                else
                    coded = ((Q.nnz >> (4 * e)) & 0xF) | ((Q.nnz >> (4 * (e - 1))) & 0xF);
                int motion = ext && (iabs(Q.mv[0] - p_mvx) >= 4 || iabs(Q.mv[1] - p_mvy) >= 4);
                if (!coded && !motion) continue;
#pragma unroll
                for (int k = 0; k < 4; k++) bs[k] = ((coded >> k) & 1) ? 2 : motion;
            }
            int qp = qp_q, qpc = qpc_q;
            if (ext) {
                qp = (qp_q + qp_p + 1) >> 1;
                qpc = (qpc_q + chroma_qp_of(qp_p, J.cqo) + 1) >> 1;
            }
            Edge EL = make_edge(qp, bs);
            if (EL.on && lane < 16) {
                if (dir == 0)
                    luma_line(S.y, (4 + lane) * 20 + 4 + e * 4, 1, EL, lane);
                else
                    luma_line(S.y, (4 + e * 4) * 20 + 4 + lane, 20, EL, lane);
            }
            if (!(e & 1)) {
                Edge EC = make_edge(qpc, bs);
                if (EC.on && lane >= 16 && lane < 32) {
                    int k = lane & 7;
                    uint8_t *pl = lane < 24 ? S.u : S.v;
                    if (dir == 0)
                        chroma_line(pl, (2 + k) * 12 + 4 + e * 2, 1, EC, k);
                    else
                        chroma_line(pl, (2 + e * 2) * 12 + 4 + k, 12, EC, k);
                }
            }
            __syncthreads();
        }
    }
    // Write back all but the untouched top-left corner, as aligned dwords.
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int i = lane + 64 * k, r = i / 5, c = i % 5;
        if (i < 100 && !(r < 4 && c == 0))
            *(RGH_G uint32_t *)&J.cur_y[(long)(y0 - 4 + r) * J.sy + x0 - 4 + c * 4] = *(const uint32_t *)&S.y[r * 20 + c * 4];
    }
    if (lane < 60) {
        const int p = lane / 30, i = lane % 30, r = i / 3, c = i % 3;
        RGH_G uint8_t *pl = p ? J.cur_v : J.cur_u;
        if (!(r < 2 && c == 0))
            *(RGH_G uint32_t *)&pl[(long)(cy0 - 2 + r) * J.sc + cx0 - 4 + c * 4] = *(const uint32_t *)&(p ? S.v : S.u)[r * 12 + c * 4];
    }
}

// enc.rs:164 extend_borders, restricted to the padding samples whose nearest interior sample
// lies in border macroblock (X, Y) of the current picture, all three planes. Rows and columns
// are dword aligned (strides are multiples of 32, paddings of 16).
__device__ __forceinline__ void pad_mb_inline(const RghJobD &J, int X, int Y) {
    const int mb_w = J.mb_w, mb_h = J.mb_h;
    if (X != 0 && X != mb_w - 1 && Y != 0 && Y != mb_h - 1) return;
    const int lane = threadIdx.x;
#pragma unroll 1
    for (int p = 0; p < 3; p++) {
        const int B = p ? 8 : 16, pad = p ? RGH_PAD_C : RGH_PAD_Y, s = p ? J.sc : J.sy;
        const int w = mb_w * B, h = mb_h * B;
        RGH_G uint8_t *o = p == 0 ? J.cur_y : (p == 1 ? J.cur_u : J.cur_v);
        const int r0 = Y == 0 ? -pad : Y * B, r1 = Y == mb_h - 1 ? h + pad : (Y + 1) * B;
        const int c0 = X == 0 ? -pad : X * B, c1 = X == mb_w - 1 ? s - pad : (X + 1) * B;
        const int nd = (c1 - c0) >> 2, n = nd * (r1 - r0);
        for (int i = lane; i < n; i += 64) {
            const int r = r0 + i / nd, c = c0 + 4 * (i % nd);
            const bool cin = c >= 0 && c < w;
            if (cin && r >= 0 && r < h) continue;
            const RGH_G uint8_t *srow = o + (long)iclamp(r, 0, h - 1) * s;
            const uint32_t v = cin ? *(const RGH_G uint32_t *)(srow + c) : 0x01010101u * srow[c < 0 ? 0 : w - 1];
            *(RGH_G uint32_t *)(o + (long)r * s + c) = v;
        }
    }
}

}  // namespace

// Edge replication of the padding of job blockIdx.z's current picture (enc.rs:164
// extend_borders): plane blockIdx.y (0 Y, 1 U, 2 V), padded row blockIdx.x.
extern "C" __global__ __launch_bounds__(256) void rgh_k_extend(const RghJob *__restrict__ jobs) {
    const RghJobD J = device_view(jobs[blockIdx.z]);
    const int plane = blockIdx.y;
    const int pad = plane ? RGH_PAD_C : RGH_PAD_Y;
    const int w = plane ? J.mb_w * 8 : J.mb_w * 16, h = plane ? J.mb_h * 8 : J.mb_h * 16;
    const int s = plane ? J.sc : J.sy;
    const int prow = blockIdx.x;
    if (prow >= h + 2 * pad) return;
    RGH_G uint8_t *base = plane == 0 ? J.cur_y : (plane == 1 ? J.cur_u : J.cur_v);  // sample (0, 0)
    const int y = prow - pad;
    const int sy = iclamp(y, 0, h - 1);
    const bool interior_row = y >= 0 && y < h;
    for (int c = threadIdx.x; c < s; c += blockDim.x) {
        int x = c - pad;
        if (interior_row && x >= 0 && x < w) continue;
        base[(long)y * s + x] = base[(long)sy * s + iclamp(x, 0, w - 1)];
    }
}

// Out-of-line forms, for kernels whose inliner budget is better spent elsewhere.
__device__ void deblock_mb(const RghJobD &J, int mb_x, int mb_y, DbLds &S) { deblock_mb_inline(J, mb_x, mb_y, S); }
__device__ void pad_mb(const RghJobD &J, int X, int Y) { pad_mb_inline(J, X, Y); }
