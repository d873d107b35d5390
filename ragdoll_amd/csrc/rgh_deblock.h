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

// The deblocking tables indexed by QP, entry q in lane q (lanes 0-51): alpha, beta, tc0 for bS
// 1-3 and the chroma QP of luma QP q. A macroblock loads them with its records, in one round
// trip, so that no edge waits for a table read.
struct DbTables {
    uint32_t alpha, beta, tc1, tc2, tc3, cqp;
};
__device__ __forceinline__ DbTables load_db_tables(int cqo) {
    const int q = min((int)threadIdx.x, 51);
    return DbTables{c_alpha[q], c_beta[q], c_tc0[q][0], c_tc0[q][1], c_tc0[q][2],
                    c_chroma_qp[iclamp(q + cqo, 0, 51)]};
}
// Entry q (uniform, 0..51) of a table register.
__device__ __forceinline__ int db_at(uint32_t reg, int q) { return (int)__builtin_amdgcn_readlane(reg, q); }

// enc.rs:3911 EdgeFilter::new
__device__ __forceinline__ Edge make_edge(int qp, const int bs[4], const DbTables &T) {
    Edge E;
    const int ia = iclamp(qp, 0, 51);
    E.alpha = db_at(T.alpha, ia);
    E.beta = db_at(T.beta, ia);
    E.on = E.alpha != 0 && E.beta != 0;
    E.strong = bs[0] == 4;
    const int t1 = db_at(T.tc1, ia), t2 = db_at(T.tc2, ia), t3 = db_at(T.tc3, ia);
#pragma unroll
    for (int k = 0; k < 4; k++)
        E.tc0[k] = bs[k] == 0 ? -1 : (bs[k] >= 4 ? 0 : (bs[k] == 1 ? t1 : (bs[k] == 2 ? t2 : t3)));
    return E;
}

// tc0 of 4-line segment g, selected without indexing the array (a dynamic index would put
// the edge in scratch memory).
__device__ __forceinline__ int seg_tc0(const Edge &E, int g) {
    return g == 0 ? E.tc0[0] : (g == 1 ? E.tc0[1] : (g == 2 ? E.tc0[2] : E.tc0[3]));
}

// simd/h264_deblock.rs:69 h264_luma_deblock_scalar on one line of samples in registers: v[0..7]
// hold p3 p2 p1 p0 q0 q1 q2 q3; tc0 is the line's segment's (-1: not filtered unless bS 4).
// Every output is selected, so lanes whose line is not filtered keep their samples; a sample is
// stored as its low byte (the reference's u8 cast).
__device__ __forceinline__ void luma_px(int *v, const Edge &E, int tc0) {
    const int p3 = v[0], p2 = v[1], p1 = v[2], p0 = v[3], q0 = v[4], q1 = v[5], q2 = v[6], q3 = v[7];
    const bool f = (E.strong || tc0 >= 0) && iabs(p0 - q0) < E.alpha && iabs(p1 - p0) < E.beta &&
                   iabs(q1 - q0) < E.beta;
    const bool ap = iabs(p2 - p0) < E.beta, aq = iabs(q2 - q0) < E.beta;
    if (!E.strong) {
        const int tc = tc0 + (int)ap + (int)aq;
        const int d = iclamp((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        const int avg = (p0 + q0 + 1) >> 1;
        v[3] = f ? clip255(p0 + d) : p0;
        v[4] = f ? clip255(q0 - d) : q0;
        v[2] = f && ap ? p1 + iclamp((p2 + avg - (p1 << 1)) >> 1, -tc0, tc0) : p1;
        v[5] = f && aq ? q1 + iclamp((q2 + avg - (q1 << 1)) >> 1, -tc0, tc0) : q1;
    } else {
        const bool small = iabs(p0 - q0) < ((E.alpha >> 2) + 2);
        const bool sp = f && ap && small, sq = f && aq && small;
        v[3] = sp ? (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3 : (f ? (2 * p1 + p0 + q1 + 2) >> 2 : p0);
        v[2] = sp ? (p2 + p1 + p0 + q0 + 2) >> 2 : p1;
        v[1] = sp ? (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3 : p2;
        v[4] = sq ? (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3 : (f ? (2 * q1 + q0 + p1 + 2) >> 2 : q0);
        v[5] = sq ? (p0 + q0 + q1 + q2 + 2) >> 2 : q1;
        v[6] = sq ? (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3 : q2;
    }
}

// simd/h264_deblock.rs:206 h264_chroma_deblock_scalar on one line in registers: v[0..3] hold
// p1 p0 q0 q1.
__device__ __forceinline__ void chroma_px(int *v, const Edge &E, int tc0) {
    const int p1 = v[0], p0 = v[1], q0 = v[2], q1 = v[3];
    const bool f = (E.strong || tc0 >= 0) && iabs(p0 - q0) < E.alpha && iabs(p1 - p0) < E.beta &&
                   iabs(q1 - q0) < E.beta;
    if (E.strong) {
        v[1] = f ? (2 * p1 + p0 + q1 + 2) >> 2 : p0;
        v[2] = f ? (2 * q1 + q0 + p1 + 2) >> 2 : q0;
    } else {
        const int tc = tc0 + 1;
        const int d = iclamp((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        v[1] = f ? clip255(p0 + d) : p0;
        v[2] = f ? clip255(q0 - d) : q0;
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
    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8;
    // One round trip: the three records, the tables, and the samples (luma rows y0-4 .. y0+15,
    // columns x0-4 .. x0+15, 5 aligned dwords a row; chroma rows cy0-2 .. cy0+7, columns
    // cx0-4 .. cx0+7, 3 aligned dwords a row, of both planes), the samples whether or not the
    // records leave any edge to filter.
    const rgh_mb_decision Q = load_dec(&J.dec[m]);
    rgh_mb_decision L{}, T{};
    if (mb_x > 0) L = load_dec(&J.dec[m - 1]);
    if (mb_y > 0) T = load_dec(&J.dec[m - mb_w]);
    const DbTables tab = load_db_tables(J.cqo);
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
    auto mveq = [](const rgh_mb_decision &a, const rgh_mb_decision &b) { return a.mv[0] == b.mv[0] && a.mv[1] == b.mv[1]; };
    // Compact pictures and pictures with P partitions: every 4x4 block carries its quadrant's
    // vector (enc/inter.rs:781 deblock_parts); the internal edges of an inter macroblock without coefficients and with
    // one vector throughout are not filtered.
    const bool parts = J.compact != 0 || J.parts != 0;
    const bool uniform = !is_intra(Q) && Q.nnz == 0 && quad_word(Q, 0) == quad_word(Q, 1) &&
                         quad_word(Q, 0) == quad_word(Q, 2) && quad_word(Q, 0) == quad_word(Q, 3);
    if (!parts && quiet(Q) && (mb_x == 0 || (quiet(L) && mveq(L, Q))) && (mb_y == 0 || (quiet(T) && mveq(T, Q)))) return;

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
    const int qp_q = Q.qp, qpc_q = db_at(tab.cqp, qp_q);
    // The filter parameters of edge e (0: the macroblock's left or top edge) in direction dir
    // (0 vertical edges, 1 horizontal edges), luma and chroma; false when the edge is not
    // filtered. The macroblock on the p side of the edge is the left or top neighbour for edge 0,
    // else this one (selected field by field: a pointer to a local record would live in scratch
    // memory).
    auto edge_params = [&](int dir, int e, Edge &EL, Edge &EC) -> bool {
        const bool ext = e == 0;
        if (ext) {
            if (!(dir == 0 ? mb_x > 0 : mb_y > 0)) return false;
        } else if (!internal || ((e & 1) && skip_odd)) {
            return false;
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
            if (!any) return false;
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
            if (!coded && !motion) return false;
#pragma unroll
            for (int k = 0; k < 4; k++) bs[k] = ((coded >> k) & 1) ? 2 : motion;
        }
        int qp = qp_q, qpc = qpc_q;
        if (ext) {
            qp = (qp_q + qp_p + 1) >> 1;
            qpc = (qpc_q + db_at(tab.cqp, qp_p) + 1) >> 1;
        }
        EL = make_edge(qp, bs, tab);
        EC = make_edge(qpc, bs, tab);
        return true;
    };
    // Each line crosses the four edges of a direction in one lane, in edge order, with its
    // samples in registers: lanes 0-15 a luma line (row or column k = lane of the macroblock,
    // the 4 samples before it included), lanes 16-23 a U line and 24-31 a V line (k = lane & 7).
    // Lines of one direction are independent, so the only barrier is between the directions.
    const bool lum = lane < 16, chr = lane >= 16 && lane < 32;
    const int k = lum ? lane : (lane & 7);
    uint8_t *cpl = lane < 24 ? S.u : S.v;
    // Vertical edges: luma row 4 + k (columns 0-19 of the copy), chroma row 2 + k (columns 0-11).
    {
        int b[20];
        uint32_t w[5] = {0, 0, 0, 0, 0};
        if (lum) {
#pragma unroll
            for (int i = 0; i < 5; i++) w[i] = *(const uint32_t *)&S.y[(4 + k) * 20 + 4 * i];
        } else if (chr) {
#pragma unroll
            for (int i = 0; i < 3; i++) w[i] = *(const uint32_t *)&cpl[(2 + k) * 12 + 4 * i];
        }
#pragma unroll
        for (int i = 0; i < 20; i++) b[i] = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFFu);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            Edge EL, EC;
            if (!edge_params(0, e, EL, EC)) continue;
            if (lum && EL.on) luma_px(b + 4 * e, EL, seg_tc0(EL, k >> 2));
            if (!(e & 1) && chr && EC.on) chroma_px(b + 2 + 2 * e, EC, seg_tc0(EC, k >> 1));
        }
#pragma unroll
        for (int i = 0; i < 5; i++)
            w[i] = (uint32_t)(b[4 * i] & 0xFF) | (uint32_t)(b[4 * i + 1] & 0xFF) << 8 |
                   (uint32_t)(b[4 * i + 2] & 0xFF) << 16 | (uint32_t)(b[4 * i + 3] & 0xFF) << 24;
        if (lum) {
#pragma unroll
            for (int i = 0; i < 5; i++) *(uint32_t *)&S.y[(4 + k) * 20 + 4 * i] = w[i];
        } else if (chr) {
#pragma unroll
            for (int i = 0; i < 3; i++) *(uint32_t *)&cpl[(2 + k) * 12 + 4 * i] = w[i];
        }
    }
    __syncthreads();
    // Horizontal edges: luma column 4 + k (rows 0-19), chroma column 4 + k (rows 0-9).
    {
        int b[20];
        if (lum) {
#pragma unroll
            for (int i = 0; i < 20; i++) b[i] = S.y[i * 20 + 4 + k];
        } else {
#pragma unroll
            for (int i = 0; i < 20; i++) b[i] = chr && i < 10 ? cpl[i * 12 + 4 + k] : 0;
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            Edge EL, EC;
            if (!edge_params(1, e, EL, EC)) continue;
            if (lum && EL.on) luma_px(b + 4 * e, EL, seg_tc0(EL, k >> 2));
            if (!(e & 1) && chr && EC.on) chroma_px(b + 2 * e, EC, seg_tc0(EC, k >> 1));
        }
        // Rows 1-18 (luma) and 1-8 (chroma) are the ones an edge can modify.
        if (lum) {
#pragma unroll
            for (int i = 1; i < 19; i++) S.y[i * 20 + 4 + k] = (uint8_t)b[i];
        } else if (chr) {
#pragma unroll
            for (int i = 1; i < 9; i++) cpl[i * 12 + 4 + k] = (uint8_t)b[i];
        }
    }
    __syncthreads();
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
