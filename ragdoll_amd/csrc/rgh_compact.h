// Compact-preset P analysis (enc/inter.rs:498-759 code_p_mb_body, enc/compact.rs PartSearch):
// one wave64 per macroblock, in the frame-step schedule of the Balanced analysis. Included by
// rgh_analyse.hip inside its anonymous namespace, after the shared helpers.
//
// The reference searches over whole-picture half-sample planes built after deblocking
// (compact.rs:24 HpelPlanes). Here every sub-sample prediction is computed where it is needed,
// from integer reference samples staged in LDS, by the standard's formulas (8.4.2.2.1); the
// planes' values are those formulas over the region the vectors can reach (28 samples beyond
// the picture, inside the planes' computed area), so the predictions are identical.
#pragma once

// LDS of a Compact macroblock: the shared analysis state and the staged reference regions: a
// search region around the best full-sample candidate and a window around the predictor, for
// the 16x16 partition and for each 8x8 partition.
constexpr int CR16_S = 64, CR16_N = 58;  // 16x16 search region: stride, rows (58 x 58 used)
constexpr int CW16_S = 28, CW16_N = 25;  // 16x16 predictor window (25 x 25 used)
constexpr int CR8_S = 56, CR8_N = 50;    // 8x8 search region (50 x 50 used)
constexpr int CW8_S = 20, CW8_N = 17;    // 8x8 predictor window (17 x 17 used)
constexpr int CP_R16 = 0, CP_W16 = CP_R16 + CR16_S * CR16_N, CP_R8 = CP_W16 + 704, CP_W8 = CP_R8 + 4 * CR8_S * CR8_N;
constexpr int CP_BYTES = CP_W8 + 4 * 352;
static_assert(CW16_S * CW16_N <= 704 && CW8_S * CW8_N <= 352, "window sizes");

struct CompactLds {
    MbLds a;
    alignas(16) uint8_t c[CP_BYTES];
    // Partition state, uniform across the wave: results of the 8x8 searches (slots 0-3) and the
    // refined quadrant vectors (slots 4-7), 8 ints each (x, y, satd, mv cost, region base,
    // stride, x, y); the 16x16 candidates (x at ST_C16X, y at ST_C16Y) and the 8x8 ones.
    alignas(16) int st[128];
};
constexpr int ST_C16X = 64, ST_C16Y = 80, ST_C8X = 96, ST_C8Y = 104;

// A staged region: byte offset in CompactLds::c, stride, picture coordinates of its first
// sample.
struct Reg {
    int base, st, ox, oy;
};

__device__ __forceinline__ int rg(const CompactLds &S, const Reg &R, int X, int Y) {
    return S.c[R.base + (Y - R.oy) * R.st + (X - R.ox)];
}
// Vertical 6-tap sum at column X between rows Y and Y + 1.
__device__ __forceinline__ int rg_tv(const CompactLds &S, const Reg &R, int X, int Y) {
    return tap6(rg(S, R, X, Y - 2), rg(S, R, X, Y - 1), rg(S, R, X, Y), rg(S, R, X, Y + 1), rg(S, R, X, Y + 2),
                rg(S, R, X, Y + 3));
}
__device__ __forceinline__ int rg_b(const CompactLds &S, const Reg &R, int X, int Y) {  // (X + 1/2, Y)
    return clip255((tap6(rg(S, R, X - 2, Y), rg(S, R, X - 1, Y), rg(S, R, X, Y), rg(S, R, X + 1, Y), rg(S, R, X + 2, Y),
                         rg(S, R, X + 3, Y)) + 16) >> 5);
}
__device__ __forceinline__ int rg_h(const CompactLds &S, const Reg &R, int X, int Y) {  // (X, Y + 1/2)
    return clip255((rg_tv(S, R, X, Y) + 16) >> 5);
}
__device__ __forceinline__ int rg_j(const CompactLds &S, const Reg &R, int X, int Y) {  // (X + 1/2, Y + 1/2)
    return clip255((tap6(rg_tv(S, R, X - 2, Y), rg_tv(S, R, X - 1, Y), rg_tv(S, R, X, Y), rg_tv(S, R, X + 1, Y),
                         rg_tv(S, R, X + 2, Y), rg_tv(S, R, X + 3, Y)) + 512) >> 10);
}
// The luma prediction sample at integer position (X, Y) plus quarter-sample phase f = fx | fy << 2:
// the pairs averaged are those of compact.rs:87 RefLuma::src (8.4.2.2.1).
__device__ int rg_pel(const CompactLds &S, const Reg &R, int X, int Y, int f) {
    auto avg = [](int a, int b) { return (a + b + 1) >> 1; };
    switch (f) {
    case 0: return rg(S, R, X, Y);
    case 2: return rg_b(S, R, X, Y);
    case 8: return rg_h(S, R, X, Y);
    case 10: return rg_j(S, R, X, Y);
    case 1: return avg(rg(S, R, X, Y), rg_b(S, R, X, Y));
    case 3: return avg(rg_b(S, R, X, Y), rg(S, R, X + 1, Y));
    case 4: return avg(rg(S, R, X, Y), rg_h(S, R, X, Y));
    case 12: return avg(rg_h(S, R, X, Y), rg(S, R, X, Y + 1));
    case 5: return avg(rg_b(S, R, X, Y), rg_h(S, R, X, Y));
    case 7: return avg(rg_b(S, R, X, Y), rg_h(S, R, X + 1, Y));
    case 13: return avg(rg_h(S, R, X, Y), rg_b(S, R, X, Y + 1));
    case 15: return avg(rg_h(S, R, X + 1, Y), rg_b(S, R, X, Y + 1));
    case 6: return avg(rg_b(S, R, X, Y), rg_j(S, R, X, Y));
    case 14: return avg(rg_j(S, R, X, Y), rg_b(S, R, X, Y + 1));
    case 9: return avg(rg_h(S, R, X, Y), rg_j(S, R, X, Y));
    default: return avg(rg_j(S, R, X, Y), rg_h(S, R, X + 1, Y));  // (3, 2)
    }
}

// Whether every reference macroblock covering picture box [x0, x1] x [y0, y1] (luma, inclusive)
// is final for launch seq (its deblock/pad item stamped in an earlier launch), for this lane's
// share of them; the caller ballots.
__device__ __forceinline__ int box_waits(const RghJobD &J, int x0, int x1, int y0, int y1, uint32_t seq) {
    const int mb_w = J.mb_w, lane = threadIdx.x;
    const int ax0 = iclamp(x0 >> 4, 0, mb_w - 1), ax1 = iclamp(x1 >> 4, 0, mb_w - 1);
    const int ay0 = iclamp(y0 >> 4, 0, J.mb_h - 1), ay1 = iclamp(y1 >> 4, 0, J.mb_h - 1);
    const int nx = ax1 - ax0 + 1, n = nx * (ay1 - ay0 + 1);
    int wait = 0;
    for (int i = lane; i < n; i += 64) {
        const int a = ax0 + i % nx, b = ay0 + i / nx;
        wait |= !stamp_final(J.ref_fin[(b + 1) * (mb_w + 1) + a + 1], J.ref_serial, seq);
    }
    return wait;
}

// Stages reference luma columns [X, X + cols) x rows [Y, Y + rows) (from the aligned dword at or
// below X) into S.c at base with stride st; returns the region. All loads are issued before the
// LDS stores. PER bounds the dwords per lane.
template <int PER>
__device__ Reg stage_ref(const RghJobD &J, CompactLds &S, int base, int st, int X, int Y, int cols, int rows) {
    const int lane = threadIdx.x, x4 = X & ~3, ndw = (X - x4 + cols + 3) >> 2, n = ndw * rows;
    const RGH_G uint8_t *src = J.ref_y + (long)Y * J.sy + x4;
    uint32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lane + 64 * k, r = i / ndw, j = i - r * ndw;
        v[k] = i < n ? *(const RGH_G uint32_t *)(src + (long)r * J.sy + 4 * j) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lane + 64 * k, r = i / ndw, j = i - r * ndw;
        if (i < n) *(uint32_t *)&S.c[base + r * st + 4 * j] = v[k];
    }
    __syncthreads();
    return Reg{base, st, x4, Y};
}

// SAD (satd false) or SATD (compact.rs:557 cost; simd/h264_me.rs:163) of the W x W partition at
// picture (bx, by), source at (px, py) of the macroblock, against its prediction at quarter-pel
// (mx, my) from region R. Lane layout: 4x4 block lane >> 2 (raster within the partition), row
// lane & 3; W = 8 uses lanes 0-15. Uniform arguments; every lane must call.
template <int W>
__device__ uint32_t part_cost(const CompactLds &S, const Reg &R, int bx, int by, int px, int py, int mx, int my,
                              bool satd) {
    constexpr int NB = W / 4;
    const int lane = threadIdx.x, blk = lane >> 2, j = lane & 3, bc = blk % NB, br = blk / NB;
    uint32_t v = 0;
    if (W == 16 || lane < 16) {
        const int X = bx + 4 * bc + (mx >> 2), Y = by + 4 * br + j + (my >> 2), f = (mx & 3) | ((my & 3) << 2);
        const uint32_t sw = *(const uint32_t *)&S.a.src[(py + 4 * br + j) * 16 + px + 4 * bc];
        int d[4];
        uint32_t pw = 0;
#pragma unroll
        for (int x = 0; x < 4; x++) {
            const int p = rg_pel(S, R, X + x, Y, f);
            pw |= (uint32_t)p << (8 * x);
            d[x] = (int)byte_of(sw, x) - p;
        }
        if (!satd) {
            v = __builtin_amdgcn_sad_u8(sw, pw, 0);
        } else {
            const int s01 = d[0] + d[1], d01 = d[0] - d[1], s23 = d[2] + d[3], d23 = d[2] - d[3];
            const int h[4] = {s01 + s23, s01 - s23, d01 + d23, d01 - d23};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int q1 = __builtin_amdgcn_update_dpp(0, h[c], 0xB1, 0xF, 0xF, false);  // row j ^ 1
                const int a = (j & 1) ? q1 - h[c] : h[c] + q1;
                const int q2 = __builtin_amdgcn_update_dpp(0, a, 0x4E, 0xF, 0xF, false);  // row j ^ 2
                v += (uint32_t)iabs((j & 2) ? q2 - a : a + q2);
            }
        }
    }
    const uint32_t t = wave_sum(v);
    return satd ? t >> 1 : t;
}

// Full-sample SADs of the W x W partition at four full-sample displacements pts (from region R):
// lanes 16p..16p+15 measure point p.
template <int W>
__device__ void take4_sads(const CompactLds &S, const Reg &R, int bx, int by, int px, int py, const int *ptx,
                           const int *pty, uint32_t out[4]) {
    const int lane = threadIdx.x, p = lane >> 4;
    const int dx = p == 0 ? ptx[0] : (p == 1 ? ptx[1] : (p == 2 ? ptx[2] : ptx[3]));
    const int dy = p == 0 ? pty[0] : (p == 1 ? pty[1] : (p == 2 ? pty[2] : pty[3]));
    uint32_t v = 0;
    if (W == 16) {
        const int r = lane & 15;
        const uint8_t *row = &S.c[R.base + (by + dy + r - R.oy) * R.st];
        const int x = bx + dx - R.ox;
#pragma unroll
        for (int k = 0; k < 4; k++)
            v += __builtin_amdgcn_sad_u8(*(const uint32_t *)&S.a.src[(py + r) * 16 + px + 4 * k], lds4u(row, x + 4 * k), 0);
    } else {
        const int r = (lane >> 1) & 7, hh = lane & 1;
        const uint8_t *row = &S.c[R.base + (by + dy + r - R.oy) * R.st];
        v = __builtin_amdgcn_sad_u8(*(const uint32_t *)&S.a.src[(py + r) * 16 + px + 4 * hh], lds4u(row, bx + dx - R.ox + 4 * hh),
                                     0);
    }
    // Row sums (16 lanes each) by DPP, read out per point.
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = (uint32_t)__builtin_amdgcn_readlane((int)v, 16 * k);
}

// Motion vector predictor neighbour: reference index (-2 unavailable, -1 intra) and vector.
struct Nb {
    int r, x, y;
};
// enc/inter.rs:143 mvp_from, reference 0, rule 0.
__device__ __forceinline__ void mvp_from(const Nb &a, const Nb &b, const Nb &c, int &mx, int &my) {
    auto med = [](int p, int q, int s) {
        const int hi = p > q ? p : q, lo = p < q ? p : q, t = lo > s ? lo : s;
        return hi < t ? hi : t;
    };
    const int match = (a.r == 0) + (b.r == 0) + (c.r == 0);
    if (match == 1) {
        mx = a.r == 0 ? a.x : (b.r == 0 ? b.x : c.x);
        my = a.r == 0 ? a.y : (b.r == 0 ? b.y : c.y);
    } else if (match == 0 && b.r == -2 && c.r == -2 && a.r != -2) {
        mx = a.x;
        my = a.y;
    } else {
        mx = med(a.x, b.x, c.x);
        my = med(a.y, b.y, c.y);
    }
}

// One analysed partition: quarter-pel vector, SATD, vector cost, and the staged region its
// neighbourhood lies in.
struct MeR {
    int x, y;
    uint32_t satd, mvc;
    Reg reg;
};

// Per-macroblock constants of the searches.
struct CpMb {
    int x0, y0;
    int bmin0, bmin1, bmax0, bmax1;  // quarter-pel vector bounds of the macroblock (enc.rs:1104)
    uint32_t lambda;
};
__device__ __forceinline__ bool cp_inb(const CpMb &M, int x, int y) {
    return x >= M.bmin0 && x <= M.bmax0 && y >= M.bmin1 && y <= M.bmax1;
}
__device__ __forceinline__ uint32_t cp_mvc(const CpMb &M, int mvpx, int mvpy, int x, int y) {
    return M.lambda * (se_bits(x - mvpx) + se_bits(y - mvpy));
}

// compact.rs:723 refine_qpel_k: quarter-pel diamonds by SATD around m within its region.
template <int W>
__device__ MeR refine_qpel(const CompactLds &S, const CpMb &M, int bx, int by, int px, int py, int mvpx, int mvpy, MeR m,
                           int iters) {
    int bxv = m.x, byv = m.y, prevx = 0, prevy = 0, has_prev = 0;
    uint32_t best_satd = m.satd, best_cost = m.satd + cp_mvc(M, mvpx, mvpy, bxv, byv);
    for (int it = 0; it < iters; it++) {
        const int fx = bxv, fy = byv;
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int mx = fx + (k == 2 ? -1 : (k == 3 ? 1 : 0)), my = fy + (k == 0 ? -1 : (k == 1 ? 1 : 0));
            if ((has_prev && mx == prevx && my == prevy) || !cp_inb(M, mx, my)) continue;
            const uint32_t satd = part_cost<W>(S, m.reg, bx, by, px, py, mx, my, true);
            const uint32_t cost = satd + cp_mvc(M, mvpx, mvpy, mx, my);
            if (cost < best_cost) {
                best_cost = cost;
                best_satd = satd;
                bxv = mx;
                byv = my;
            }
        }
        if (bxv == fx && byv == fy) break;
        prevx = fx;
        prevy = fy;
        has_prev = 1;
    }
    m.x = bxv;
    m.y = byv;
    m.satd = best_satd;
    m.mvc = cp_mvc(M, mvpx, mvpy, bxv, byv);
    return m;
}

// compact.rs:583 search_qpel (thorough, one quarter-pel diamond) of the W x W partition at
// (px, py) of the macroblock: zero and the candidates (quarter-pel, rounded to full samples,
// clamped, distinct, at most 12 points) by SAD, a radius-2 hexagon for up to 7 further steps, a
// square of eight, the predictor at its own sub-sample position, a half-sample diamond by SAD,
// then a quarter-sample diamond by SATD. cx / cy: nc candidate vectors. Regions: search region
// at rbase (stride rst), predictor window at wbase (stride wst). Returns false when a reference
// macroblock it reads is not final yet.
// cxo / cyo: offsets of the nc candidate vectors in S.st. outer (when has_outer): an already staged
// and checked region with picture box [ox0, ox1] x [oy0, oy1]; whatever lies inside it is read
// from it instead of being staged again.
template <int W, int RPER, int WPER>
__device__ bool part_search(const RghJobD &J, CompactLds &S, const CpMb &M, int px, int py, int mvpx, int mvpy, int cxo,
                            int cyo, int nc, int rbase, int rst, int wbase, int wst, bool has_outer, const Reg &outer,
                            int ox0, int ox1, int oy0, int oy1, uint32_t seq, MeR &out, Reg &sreg, int &sx0, int &sy0) {
    auto inside = [&](int x0_, int x1_, int y0_, int y1_) {
        return has_outer && x0_ >= ox0 && x1_ <= ox1 && y0_ >= oy0 && y1_ <= oy1;
    };
    const int lane = threadIdx.x, bx = M.x0 + px, by = M.y0 + py;
    const int lo0 = (M.bmin0 + 3) >> 2, lo1 = (M.bmin1 + 3) >> 2, hi0 = M.bmax0 >> 2, hi1 = M.bmax1 >> 2;
    int ptx[12], pty[12], n = 1;
    ptx[0] = pty[0] = 0;
#pragma unroll
    for (int i = 1; i < 12; i++) ptx[i] = pty[i] = 0;
#pragma unroll
    for (int i = 0; i < 11; i++) {
        if (i < nc) {
            const int cvx = __builtin_amdgcn_readfirstlane(S.st[cxo + i]), cvy = __builtin_amdgcn_readfirstlane(S.st[cyo + i]);
            const int m0 = iclamp((cvx + 2) >> 2, lo0, hi0), m1 = iclamp((cvy + 2) >> 2, lo1, hi1);
            bool dup = false;
#pragma unroll
            for (int k = 0; k < 12; k++)
                if (k < n && ptx[k] == m0 && pty[k] == m1) dup = true;
            if (!dup) {
                // n <= i + 1 < 12: write slot n without indexing by it.
#pragma unroll
                for (int k = 1; k < 12; k++)
                    if (k == n) {
                        ptx[k] = m0;
                        pty[k] = m1;
                    }
                n++;
            }
        }
    }
    // Round trip: the candidate blocks, and the window around the predictor (both read from the
    // outer region instead when they lie inside it).
    const bool mvp_in = cp_inb(M, mvpx, mvpy);
    const int wx = bx + (mvpx >> 2) - 4, wy = by + (mvpy >> 2) - 4;
    bool cand_in = true;
#pragma unroll
    for (int k = 0; k < 12; k++)
        if (k < n) cand_in = cand_in && inside(bx + ptx[k], bx + ptx[k] + W - 1, by + pty[k], by + pty[k] + W - 1);
    const bool win_in = !mvp_in || inside(wx, wx + W + 8, wy, wy + W + 8);
    int best0 = 0, best1 = 0;
    uint32_t best_cost = 0xFFFFFFFFu;
    Reg wreg = outer;
    if (cand_in && win_in) {
        // Zero and the candidates, four at a time (padding repeats the zero vector, which cannot
        // win again).
#pragma unroll
        for (int c = 0; c < 12; c += 4) {
            if (c >= n) continue;
            int qx[4], qy[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                qx[k] = c + k < n ? ptx[c + k] : 0;
                qy[k] = c + k < n ? pty[c + k] : 0;
            }
            uint32_t sads[4];
            take4_sads<W>(S, outer, bx, by, px, py, qx, qy, sads);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t cst = sads[k] + cp_mvc(M, mvpx, mvpy, 4 * qx[k], 4 * qy[k]);
                if (cst < best_cost) {
                    best_cost = cst;
                    best0 = qx[k];
                    best1 = qy[k];
                }
            }
        }
    } else {
        constexpr int NB = W / 4;
        const int blk = lane >> 2, j = lane & 3, bc = blk % NB, br = blk / NB;
        const bool on = W == 16 || lane < 16;
        uint32_t cw[12];
#pragma unroll
        for (int k = 0; k < 12; k++)
            cw[k] = (k < n && on) ? ld4u(J.ref_y + (long)(by + pty[k] + 4 * br + j) * J.sy + bx + ptx[k] + 4 * bc) : 0u;
        int wait = 0;
        // Candidate blocks span at most two macroblocks each way: their corners name them all.
        if (lane < 4 * n) {
            const int k = lane >> 2, kx = lane & 1, ky = (lane >> 1) & 1;
            int qx = 0, qy = 0;
#pragma unroll
            for (int i = 0; i < 12; i++)
                if (i == k) {
                    qx = ptx[i];
                    qy = pty[i];
                }
            const int X = bx + qx + kx * (W - 1), Y = by + qy + ky * (W - 1);
            const int a = iclamp(X >> 4, 0, J.mb_w - 1), b = iclamp(Y >> 4, 0, J.mb_h - 1);
            wait |= !stamp_final(J.ref_fin[(b + 1) * (J.mb_w + 1) + a + 1], J.ref_serial, seq);
        }
        if (mvp_in) {
            wait |= box_waits(J, wx, wx + W + 8, wy, wy + W + 8, seq);
            wreg = stage_ref<WPER>(J, S, wbase, wst, wx, wy, W + 9, W + 9);
        }
        if (__ballot(wait)) return false;
        // Candidates in order: the first strictly cheapest wins.
#pragma unroll
        for (int k = 0; k < 12; k++) {
            if (k >= n) continue;
            const uint32_t sd = wave_sum(on ? __builtin_amdgcn_sad_u8(*(const uint32_t *)&S.a.src[(py + 4 * br + j) * 16 + px + 4 * bc],
                                                                      cw[k], 0)
                                            : 0u);
            const uint32_t c = sd + cp_mvc(M, mvpx, mvpy, 4 * ptx[k], 4 * pty[k]);
            if (c < best_cost) {
                best_cost = c;
                best0 = ptx[k];
                best1 = pty[k];
            }
        }
    }
    // Round trip: the search region, 21 samples around the best candidate's block (the hexagon
    // and square move at most 17, the refinements 5 quarter samples and their filter taps 3).
    Reg rreg;
    {
        // Kept inside the padded picture: reads never reach more than 32 samples beyond the
        // picture (vectors keep blocks within 28, filter taps add 3, quarter positions 1), so a
        // region moved inward still covers them.
        const int X = iclamp(bx + best0 - 21, -(RGH_PAD_Y - 4), J.mb_w * 16 + RGH_PAD_Y - (W + 42));
        const int Y = iclamp(by + best1 - 21, -(RGH_PAD_Y - 4), J.mb_h * 16 + RGH_PAD_Y - (W + 42));
        sx0 = X;
        sy0 = Y;
        if (inside(X, X + W + 41, Y, Y + W + 41)) {
            rreg = outer;
        } else {
            const int w2 = box_waits(J, X, X + W + 41, Y, Y + W + 41, seq);
            rreg = stage_ref<RPER>(J, S, rbase, rst, X, Y, W + 42, W + 42);
            if (__ballot(w2)) return false;
        }
    }
    // Takes of four points in order (compact.rs:516): returns the last that improved, or -1.
    auto take = [&](const int *qx, const int *qy) {
        uint32_t sads[4];
        take4_sads<W>(S, rreg, bx, by, px, py, qx, qy, sads);
        int moved = -1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t c = sads[k] + cp_mvc(M, mvpx, mvpy, 4 * qx[k], 4 * qy[k]);
            if (c < best_cost) {
                best_cost = c;
                best0 = qx[k];
                best1 = qy[k];
                moved = k;
            }
        }
        return moved;
    };
    // compact.rs:492 at: from + d, or from when that leaves the vectors allowed.
    auto at = [&](int fx, int fy, int dx, int dy, int &ox, int &oy) {
        const int m0 = fx + dx, m1 = fy + dy;
        const bool ok = (uint32_t)(m0 - lo0) <= (uint32_t)(hi0 - lo0) && (uint32_t)(m1 - lo1) <= (uint32_t)(hi1 - lo1);
        ox = ok ? m0 : fx;
        oy = ok ? m1 : fy;
    };
    // Hexagon of radius 2 in circular order; after a step toward point d only the three points
    // around it are new (x264 hex2).
    // Point k of the hexagon: (-2,0) (-1,2) (1,2) (2,0) (1,-2) (-1,-2).
    auto hx = [](int k) { return k == 0 ? -2 : (k == 3 ? 2 : (k == 1 || k == 5 ? -1 : 1)); };
    auto hy = [](int k) { return k == 0 || k == 3 ? 0 : (k < 3 ? 2 : -2); };
    int dir;
    {
        const int fx = best0, fy = best1;
        int qx[8], qy[8];
#pragma unroll
        for (int k = 0; k < 6; k++) at(fx, fy, hx(k), hy(k), qx[k], qy[k]);
        qx[6] = qx[7] = fx;
        qy[6] = qy[7] = fy;
        const int a = take(qx, qy), b = take(qx + 4, qy + 4);
        dir = b >= 0 ? b + 4 : a;
    }
    for (int it = 1; it < 8 && dir >= 0; it++) {
        const int fx = best0, fy = best1;
        // The three points around direction d, in compact.rs NEXT order: {0,1,5} {0,1,2} {1,2,3}
        // {2,3,4} {3,4,5} {0,4,5}.
        const int k0 = dir == 0 ? 0 : (dir == 5 ? 0 : dir - 1), k1 = dir == 0 ? 1 : (dir == 5 ? 4 : dir);
        const int k2 = dir == 0 ? 5 : (dir == 5 ? 5 : dir + 1);
        int qx[4], qy[4];
        at(fx, fy, hx(k0), hy(k0), qx[0], qy[0]);
        at(fx, fy, hx(k1), hy(k1), qx[1], qy[1]);
        at(fx, fy, hx(k2), hy(k2), qx[2], qy[2]);
        qx[3] = fx;
        qy[3] = fy;
        const int t = take(qx, qy);
        dir = t == 0 ? k0 : (t == 1 ? k1 : (t == 2 ? k2 : -1));
    }
    {
        const int fx = best0, fy = best1;
        const int SX[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, SY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
        int qx[8], qy[8];
#pragma unroll
        for (int k = 0; k < 8; k++) at(fx, fy, SX[k], SY[k], qx[k], qy[k]);
        take(qx, qy);
        take(qx + 4, qy + 4);
    }
    // compact.rs:651 refine: the predictor at its own sub-sample position, then the half-sample
    // diamond by SAD, then the quarter-sample diamond by SATD.
    MeR m;
    m.x = 4 * best0;
    m.y = 4 * best1;
    m.reg = rreg;
    uint32_t cost = best_cost;
    if (mvp_in && (mvpx != m.x || mvpy != m.y)) {
        const uint32_t c = part_cost<W>(S, wreg, bx, by, px, py, mvpx, mvpy, false) + cp_mvc(M, mvpx, mvpy, mvpx, mvpy);
        if (c < cost) {
            cost = c;
            m.x = mvpx;
            m.y = mvpy;
            m.reg = wreg;
        }
    }
    {
        const int fx = m.x, fy = m.y;
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int mx = fx + (k == 1 ? -2 : (k == 2 ? 2 : 0)), my = fy + (k == 0 ? -2 : (k == 3 ? 2 : 0));
            if (!cp_inb(M, mx, my)) continue;
            const uint32_t c = part_cost<W>(S, m.reg, bx, by, px, py, mx, my, false) + cp_mvc(M, mvpx, mvpy, mx, my);
            if (c < cost) {
                cost = c;
                m.x = mx;
                m.y = my;
            }
        }
    }
    m.satd = part_cost<W>(S, m.reg, bx, by, px, py, m.x, m.y, true);
    m.mvc = cp_mvc(M, mvpx, mvpy, m.x, m.y);
    out = refine_qpel<W>(S, M, bx, by, px, py, mvpx, mvpy, m, 1);
    sreg = rreg;
    return true;
}

// Partition state in LDS (written by every lane with the same values, read back as scalars).
__device__ __forceinline__ void put_mer(CompactLds &S, int slot, const MeR &m) {
    int *p = &S.st[8 * slot];
    p[0] = m.x;
    p[1] = m.y;
    p[2] = (int)m.satd;
    p[3] = (int)m.mvc;
    p[4] = m.reg.base;
    p[5] = m.reg.st;
    p[6] = m.reg.ox;
    p[7] = m.reg.oy;
}
__device__ __forceinline__ MeR get_mer(const CompactLds &S, int slot) {
    const int *p = &S.st[8 * slot];
    MeR m;
    m.x = __builtin_amdgcn_readfirstlane(p[0]);
    m.y = __builtin_amdgcn_readfirstlane(p[1]);
    m.satd = (uint32_t)__builtin_amdgcn_readfirstlane(p[2]);
    m.mvc = (uint32_t)__builtin_amdgcn_readfirstlane(p[3]);
    m.reg.base = __builtin_amdgcn_readfirstlane(p[4]);
    m.reg.st = __builtin_amdgcn_readfirstlane(p[5]);
    m.reg.ox = __builtin_amdgcn_readfirstlane(p[6]);
    m.reg.oy = __builtin_amdgcn_readfirstlane(p[7]);
    return m;
}

// The lane's four luma prediction samples of the 16x16 macroblock (row lane >> 2, columns
// 4 (lane & 3)..+3): from one vector and region, or (quads) from the refined quadrant vectors and
// regions in LDS slots 4-7.
__device__ uint32_t cp_predict(CompactLds &S, const CpMb &M, const Reg &R0, int mx0, int my0, bool quads) {
    const int lane = threadIdx.x, r = lane >> 2, c0 = (lane & 3) * 4, k = (c0 >> 3) + 2 * (r >> 3);
    Reg R = R0;
    int mx = mx0, my = my0;
    if (quads) {
        const int *p = &S.st[8 * (4 + k)];
        mx = p[0];
        my = p[1];
        R = Reg{p[4], p[5], p[6], p[7]};
    }
    const int X = M.x0 + c0 + (mx >> 2), Y = M.y0 + r + (my >> 2), f = (mx & 3) | ((my & 3) << 2);
    uint32_t w = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) w |= (uint32_t)rg_pel(S, R, X + x, Y, f) << (8 * x);
    return w;
}

// enc.rs:2906 try_p_skip from the prediction at the skip vector (luma predw of the lane, row
// lane >> 2; chroma pu / pv of the lane's chroma sample): the residual certainly vanishes by the
// SAD / SSD bounds, or quantizes (Compact: intra-style rounding) to decimated nothing.
__device__ bool cp_skip_check(const RghJobD &J, CompactLds &S, uint32_t srcw, uint32_t predw, uint32_t src_u,
                              uint32_t src_v, uint32_t pu, uint32_t pv, int qp, int qpc) {
    const int lane = threadIdx.x;
    const int fine = J.compact ? 1 : 0;
    const uint32_t sad = wave_sum(__builtin_amdgcn_sad_u8(srcw, predw, 0));
    if (sad >= J.sb_probe) return false;
    int luma_zero = sad <= J.sb_luma || wave_sum(ssd4(srcw, predw)) <= J.sb_luma_ssd;
    if (!luma_zero) {
        // enc.rs:2163 probe_skip_luma: no level above 1 and a total score below 6.
        *(uint32_t *)&S.a.pred[lane * 4] = predw;
        __syncthreads();
        int big = 0;
        uint32_t score = 0;
        if (lane < 16) {
            int c[16];
            sub_fdct4(S.a.src, S.a.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, c);
            quant4(c, qp, fine);
            uint32_t m = zz_mask(c, &big);
            score = decimate_mask(m, big);
        }
        luma_zero = __ballot(big) == 0 && wave_sum(score) < 6;
    }
    if (!luma_zero) return false;
    const uint32_t du = (uint32_t)iabs((int)src_u - (int)pu), dv = (uint32_t)iabs((int)src_v - (int)pv);
    int chroma_zero = wave_sum(du) <= J.sb_chroma || wave_sum(du * du) <= J.sb_chroma_ssd;
    if (chroma_zero) chroma_zero = wave_sum(dv) <= J.sb_chroma || wave_sum(dv * dv) <= J.sb_chroma_ssd;
    if (!chroma_zero) {
        // enc.rs:2186 probe_skip_chroma: no DC level and AC scores below 7.
        S.a.predu[lane] = (uint8_t)pu;
        S.a.predv[lane] = (uint8_t)pv;
        __syncthreads();
        quant_chroma(S.a, qpc, fine);
        int dcnz = 0;
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int i = 0; i < 4; i++) dcnz |= S.a.cdc[p][i];
        const int su = S.a.cscore[0] + S.a.cscore[1] + S.a.cscore[2] + S.a.cscore[3];
        const int sv = S.a.cscore[4] + S.a.cscore[5] + S.a.cscore[6] + S.a.cscore[7];
        chroma_zero = !dcnz && su < 7 && sv < 7;
    }
    return chroma_zero != 0;
}

// One macroblock of a Compact P picture (enc/inter.rs:498 code_p_mb_body for CAVLC): decision
// record (quadrant vectors and the 16x16 search vector, see ragdoll_hip.h), levels and
// reconstruction. Returns false, having written nothing, when reference samples it reads are
// not final yet.
__device__ bool analyse_compact_mb(const RghJobD &J, int mb_x, int mb_y, CompactLds &S, uint32_t seq) {
    const int lane = threadIdx.x, rr = lane >> 2, c0 = (lane & 3) * 4;
    const int mb_w = J.mb_w, mb_h = J.mb_h, idx = mb_y * mb_w + mb_x;
    const int x0 = mb_x * 16, y0 = mb_y * 16, cx0 = mb_x * 8, cy0 = mb_y * 8, W = mb_w * 16;
    const int qp = J.qp, qpc = chroma_qp_of(qp, J.cqo);
    const uint32_t lam4 = c_lambda[qp] * 4u / 5u, lambda = lam4 > 1 ? lam4 : 1u;  // enc.rs:73 LAMBDA_COMPACT
    const bool has_l = mb_x > 0, has_t = mb_y > 0, has_tr = has_t && mb_x + 1 < mb_w, has_tl = has_t && has_l;

    // ---- Round trip 1: source, the neighbours' and the previous picture's records, edges. ----
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
    // Records: 0 left, 1 top, 2 top-left, 3 top-right (this picture); 4 co-located, 5 right,
    // 6 below (previous picture). Unavailable ones read a stand-in and are ignored.
    uint32_t recw = 0;
    if (lane < 56) {
        const int rec = lane >> 3, w = lane & 7;
        int ri = idx;
        bool cur = rec < 4;
        if (rec == 0 && has_l) ri = idx - 1;
        if (rec == 1 && has_t) ri = idx - mb_w;
        if (rec == 2 && has_tl) ri = idx - mb_w - 1;
        if (rec == 3 && has_tr) ri = idx - mb_w + 1;
        if (rec == 5 && mb_x + 1 < mb_w) ri = idx + 1;
        if (rec == 6 && mb_y + 1 < mb_h) ri = idx + mb_w;
        if (cur && ri == idx) cur = false;  // stand-in from the previous picture
        const RGH_G uint32_t *dp = (const RGH_G uint32_t *)(cur ? &J.dec[ri] : &J.prev[ri]);
        recw = dp[w];
    }
    // Records 4-6 are final once the previous picture's deblock/pad item (x + 1, y + 1) is
    // stamped: it runs after every analysis of that picture up to 6 wavefront steps past this
    // macroblock's, which include (x, y), (x + 1, y) and (x, y + 1). The candidates' own boxes do
    // not imply it when every candidate points away from this macroblock.
    int own_pending = 0;
    if (lane == 56 && J.ref_fin)
        own_pending = !stamp_final(J.ref_fin[(mb_y + 1) * (mb_w + 1) + mb_x + 1], J.ref_serial, seq);
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
    {
        const uint32_t eb = (ew >> eshift) & 0xFF;
        if (lane < 16) {
            S.a.eleft[lane] = has_l ? (uint8_t)eb : 0;
        } else if (lane < 32) {
            S.a.cleft[(lane >> 3) & 1][lane & 7] = has_l ? (uint8_t)eb : 0;
        } else if (lane < 36) {
            *(uint32_t *)&S.a.etop[4 * (lane - 32)] = has_t ? ew : 0u;
        } else if (lane < 37) {
        } else if (lane < 41) {
            *(uint32_t *)&S.a.ctop[(lane - 37) >> 1][4 * ((lane - 37) & 1)] = has_t ? ew : 0u;
        } else if (lane == 41) {
            S.a.etl = has_l && has_t ? (int)eb : 128;
        } else if (lane < 44) {
            S.a.ctl[lane - 42] = has_l && has_t ? (int)eb : 128;
        }
    }
    *(uint32_t *)&S.a.src[lane * 4] = srcw;
    S.a.srcu[lane] = (uint8_t)src_u;
    S.a.srcv[lane] = (uint8_t)src_v;
    const uint32_t tr_word = has_tr ? (uint32_t)__builtin_amdgcn_readlane((int)ew, 36) : 0u;
    auto word = [&](int rec, int w) { return (uint32_t)__builtin_amdgcn_readlane((int)recw, rec * 8 + w); };
    // Neighbour motion at 4x4 granularity: reference index (-2 unavailable, -1 intra) and the
    // vector of the quadrant holding the 4x4 block.
    auto nb = [&](int rec, bool avail, int quad) {
        Nb n{-2, 0, 0};
        if (avail) {
            const uint32_t t = word(rec, 0) & 0xFF;
            if (t == RGH_MB_I16X16 || t == RGH_MB_I4X4) {
                n.r = -1;
            } else {
                const uint32_t v = word(rec, 2 + quad);
                n.r = 0;
                n.x = (int16_t)(v & 0xFFFF);
                n.y = (int16_t)(v >> 16);
            }
        }
        return n;
    };
    auto mvr_of = [&](int rec, int &x, int &y) {
        const uint32_t v = word(rec, 6);
        x = (int16_t)(v & 0xFFFF);
        y = (int16_t)(v >> 16);
    };
    const Nb Lq1 = nb(0, has_l, 1), Lq3 = nb(0, has_l, 3), Tq2 = nb(1, has_t, 2), Tq3 = nb(1, has_t, 3);
    const Nb TLq3 = nb(2, has_tl, 3), TRq2 = nb(3, has_tr, 2);
    __syncthreads();

    CpMb M;
    M.x0 = x0;
    M.y0 = y0;
    M.bmin0 = -4 * (x0 + RGH_MV_MARGIN);
    M.bmin1 = max(-4 * (y0 + RGH_MV_MARGIN), -2048);
    M.bmax0 = 4 * (W - 16 - x0 + RGH_MV_MARGIN);
    M.bmax1 = min(4 * (mb_h * 16 - 16 - y0 + RGH_MV_MARGIN), 2044);
    M.lambda = lambda;

    // The 16x16 predictor (A: left block, B: top block, C: top-right else D: top-left) and the
    // skip vector (enc/inter.rs:517-529).
    int mvpx, mvpy;
    mvp_from(Lq1, Tq2, has_tr ? TRq2 : TLq3, mvpx, mvpy);
    int pskx = mvpx, psky = mvpy;
    if (!has_l || !has_t || (Lq1.r == 0 && Lq1.x == 0 && Lq1.y == 0) || (Tq2.r == 0 && Tq2.x == 0 && Tq2.y == 0))
        pskx = psky = 0;
    auto is_skip = [&](int rec) { return (word(rec, 0) & 0xFF) == RGH_MB_P_SKIP; };
    const bool probe = cp_inb(M, pskx, psky) &&
                       ((has_l && is_skip(0)) || (has_t && (is_skip(1) || (has_tl && is_skip(2)) || (has_tr && is_skip(3)))));

    if (__ballot(own_pending)) return false;

    rgh_mb_decision D;
    memset(&D, 0, sizeof(D));
    D.qp = (uint8_t)qp;
    int mb_type = -1, mvrx = pskx, mvry = psky;
    uint32_t predw = 0, pu = 0, pv = 0;
    int cbpl = 0, cbpc = 0, i16_mode = 0, cmode = 0;
    uint64_t modes = 0;

    // ---- The skip probe next to a skipped macroblock (x264 fast P_Skip detection). ----
    if (probe) {
        const int X = x0 + (pskx >> 2) - 4, Y = y0 + (psky >> 2) - 4;
        const int w = box_waits(J, X, X + 24, Y, Y + 24, seq);
        const Reg R = stage_ref<3>(J, S, CP_W16, CW16_S, X, Y, 25, 25);
        if (__ballot(w)) return false;
        predw = cp_predict(S, M, R, pskx, psky, false);
        const ChromaTaps ct = chroma_taps(J, mb_x, mb_y, pskx, psky);
        chroma_mc_combine(ct, pskx, psky, &pu, &pv);
        if (cp_skip_check(J, S, srcw, predw, src_u, src_v, pu, pv, qp, qpc)) {
            mb_type = 0;
            D.flags = 1;
        }
    }

    if (mb_type < 0) {
        // ---- 16x16: the predictor, the neighbours' and the previous picture's search vectors. ----
        int nc = 0;
        auto push = [&](int x, int y) {
            S.st[ST_C16X + nc] = x;
            S.st[ST_C16Y + nc] = y;
            nc++;
        };
        int vx, vy;
        push(mvpx, mvpy);
        if (has_l) {
            mvr_of(0, vx, vy);
            push(vx, vy);
        }
        if (has_t) {
            mvr_of(1, vx, vy);
            push(vx, vy);
            if (has_l) {
                mvr_of(2, vx, vy);
                push(vx, vy);
            }
            if (mb_x + 1 < mb_w) {
                mvr_of(3, vx, vy);
                push(vx, vy);
            }
        }
        mvr_of(4, vx, vy);
        push(vx, vy);
        if (mb_x + 1 < mb_w) {
            mvr_of(5, vx, vy);
            push(vx, vy);
        }
        if (mb_y + 1 < mb_h) {
            mvr_of(6, vx, vy);
            push(vx, vy);
        }
        MeR m16;
        Reg r16;
        int r16x = 0, r16y = 0;
        if (!part_search<16, 15, 3>(J, S, M, 0, 0, mvpx, mvpy, ST_C16X, ST_C16Y, nc, CP_R16, CR16_S, CP_W16, CW16_S, false,
                                    Reg{0, 0, 0, 0}, 0, 0, 0, 0, seq, m16, r16, r16x, r16y))
            return false;
        mvrx = m16.x;
        mvry = m16.y;
        // A 16x16 vector at the skip vector whose residual decimates away is P_Skip.
        if (!probe && m16.x == pskx && m16.y == psky && cp_inb(M, pskx, psky)) {
            predw = cp_predict(S, M, m16.reg, pskx, psky, false);
            const ChromaTaps ct = chroma_taps(J, mb_x, mb_y, pskx, psky);
            chroma_mc_combine(ct, pskx, psky, &pu, &pv);
            if (cp_skip_check(J, S, srcw, predw, src_u, src_v, pu, pv, qp, qpc)) mb_type = 0;
        }
        if (mb_type < 0) {
            bool p8 = false;
            uint32_t best_cost = m16.satd + m16.mvc;
            // ---- Four 8x8 searches seeded by the 16x16 vector and the partitions before them. ----
            // Predictor of partition i (8.4.1.3, enc/inter.rs:104): A left, B top, C top-right
            // else D top-left; inside the macroblock from the partitions before it (LDS slots
            // base .. base + 2).
            auto part_mvp = [&](int i, int base, int &ox, int &oy) {
                auto own = [&](int k) {
                    const MeR t = get_mer(S, base + k);
                    return Nb{0, t.x, t.y};
                };
                Nb A, B, C;
                if (i == 0) {
                    A = Lq1;
                    B = Tq2;
                    C = has_t ? Tq3 : Nb{-2, 0, 0};
                } else if (i == 1) {
                    A = own(0);
                    B = Tq3;
                    C = has_tr ? TRq2 : (has_t ? Tq2 : Nb{-2, 0, 0});
                } else if (i == 2) {
                    A = Lq3;
                    B = own(0);
                    C = own(1);
                } else {
                    A = own(2);
                    B = own(1);
                    C = own(0);
                }
                mvp_from(A, B, C, ox, oy);
            };
            if (best_cost > lambda * 40u) {
                S.st[ST_C8X] = m16.x;
                S.st[ST_C8Y] = m16.y;
                uint32_t cost8 = 0;
#pragma unroll 1
                for (int i = 0; i < 4; i++) {
                    int p8x, p8y;
                    part_mvp(i, 0, p8x, p8y);
                    const int px = (i & 1) * 8, py = (i >> 1) * 8;
                    MeR m8;
                    Reg dummy;
                    int dx, dy;
                    if (!part_search<8, 11, 2>(J, S, M, px, py, p8x, p8y, ST_C8X, ST_C8Y, i + 1, CP_R8 + i * CR8_S * CR8_N, CR8_S,
                                               CP_W8 + i * 352, CW8_S, true, r16, r16x, r16x + 57, r16y, r16y + 57, seq, m8, dummy,
                                               dx, dy))
                        return false;
                    put_mer(S, i, m8);
                    S.st[ST_C8X + i + 1] = m8.x;
                    S.st[ST_C8Y + i + 1] = m8.y;
                    cost8 += m8.satd + m8.mvc;
                }
                p8 = cost8 < best_cost;
            }
            // ---- Quarter-pel refinement of the partitions chosen (enc/inter.rs:655-678), into
            // slots 4-7. ----
            best_cost = 0;
            if (!p8) {
                const MeR r = refine_qpel<16>(S, M, x0, y0, 0, 0, mvpx, mvpy, m16, 2);
                best_cost = r.satd + r.mvc;
                for (int k = 0; k < 4; k++) put_mer(S, 4 + k, r);
            } else {
#pragma unroll 1
                for (int i = 0; i < 4; i++) {
                    int p8x, p8y;
                    part_mvp(i, 4, p8x, p8y);
                    const MeR r = refine_qpel<8>(S, M, x0 + (i & 1) * 8, y0 + (i >> 1) * 8, (i & 1) * 8, (i >> 1) * 8, p8x,
                                                 p8y, get_mer(S, i), 2);
                    best_cost += r.satd + r.mvc;
                    put_mer(S, 4 + i, r);
                }
            }
            // ---- Intra 16x16 and 4x4 against the best inter cost (enc/inter.rs:681-700). ----
            uint16_t n_satd = 0;
            const I16Analysis ia = analyse_i16(S.a, mb_x, mb_y, n_satd);
            const uint32_t i16_cost = ia.cost + lambda * 4;
            if (i16_cost < best_cost) {
                mb_type = 2;
                i16_mode = ia.mode;
                const I16Code ic = code_i16(S.a, mb_x, mb_y, qp, qpc, ia, src_u, src_v);
                cbpl = ic.cbpl;
                cbpc = ic.cbpc;
                cmode = ic.cmode;
            } else {
                const uint32_t lo = i16_cost < best_cost ? i16_cost : best_cost;
                bool i4 = false;
                if (lo > lambda * 40u && (uint64_t)i16_cost * 2 < (uint64_t)best_cost * 3) {
                    __syncthreads();
                    if (lane < 16)
                        S.a.cv[(1 + lane) * CV_W + 3] = S.a.eleft[lane];
                    else if (lane < 20)
                        *(uint32_t *)&S.a.cv[4 + 4 * (lane - 16)] = *(const uint32_t *)&S.a.etop[4 * (lane - 16)];
                    else if (lane == 20)
                        *(uint32_t *)&S.a.cv[20] = tr_word;
                    else if (lane == 21)
                        S.a.cv[3] = (uint8_t)S.a.etl;
                    __syncthreads();
                    // The neighbours' intra 4x4 modes (DC for anything but I_4x4).
                    const uint64_t all_dc = 0x2222222222222222ull;
                    const uint64_t lmodes = has_l && (word(0, 0) & 0xFF) == RGH_MB_I4X4
                                                ? ((uint64_t)word(0, 3) << 32) | word(0, 2) : all_dc;
                    const uint64_t tmodes = has_t && (word(1, 0) & 0xFF) == RGH_MB_I4X4
                                                ? ((uint64_t)word(1, 3) << 32) | word(1, 2) : all_dc;
                    i4 = analyse_i4(S.a, mb_x, mb_y, mb_w, qp, c_lambda[qp], lo, lmodes, tmodes, modes, cbpl);
                }
                if (i4) {
                    // enc.rs:3441 code_i4: the canvas holds the luma reconstruction.
                    mb_type = 3;
                    *(uint32_t *)&S.a.pred[lane * 4] = *(const uint32_t *)&S.a.cv[(1 + rr) * CV_W + 4 + c0];
                    const ChromaIntra ci = chroma_intra(S.a, mb_x, mb_y, src_u, src_v);
                    cmode = ci.mode;
                    S.a.predu[lane] = (uint8_t)ci.pu;
                    S.a.predv[lane] = (uint8_t)ci.pv;
                    __syncthreads();
                    quant_chroma(S.a, qpc, 1);
                    cbpc = chroma_cbp(S.a);
                    recon_chroma(S.a, cbpc, qpc);
                } else {
                    cbpl = 0;
                }
            }
            if (mb_type < 0) {
                // ---- Prediction, residual (intra-style rounding) and reconstruction. ----
                predw = cp_predict(S, M, Reg{0, 0, 0, 0}, 0, 0, true);
                *(uint32_t *)&S.a.pred[lane * 4] = predw;
                {
                    const int k = ((lane & 7) >> 2) + 2 * ((lane >> 3) >> 2);
                    const int mx = S.st[8 * (4 + k)], my = S.st[8 * (4 + k) + 1];
                    const ChromaTaps ct = chroma_taps(J, mb_x, mb_y, mx, my);
                    chroma_mc_combine(ct, mx, my, &pu, &pv);
                }
                S.a.predu[lane] = (uint8_t)pu;
                S.a.predv[lane] = (uint8_t)pv;
                __syncthreads();
                // enc.rs:1810 quant_luma_4x4_inter: decimated 8x8 quadrants, then the whole MB.
                if (lane < 16) {
                    int c[16], big;
                    sub_fdct4(S.a.src, S.a.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, c);
                    quant4(c, qp, 1);
                    const uint32_t m = zz_mask(c, &big);
#pragma unroll
                    for (int i = 0; i < 16; i++) S.a.coef[lane][i] = (int16_t)c[i];
                    S.a.bnz[lane] = m != 0;
                    S.a.bbig[lane] = big;
                    S.a.bscore[lane] = (int32_t)decimate_mask(m, big);
                }
                __syncthreads();
                int keep = 0;
                {
                    uint32_t total = 0;
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        const int nz = S.a.bnz[4 * g] | S.a.bnz[4 * g + 1] | S.a.bnz[4 * g + 2] | S.a.bnz[4 * g + 3];
                        const int big = S.a.bbig[4 * g] | S.a.bbig[4 * g + 1] | S.a.bbig[4 * g + 2] | S.a.bbig[4 * g + 3];
                        if (!nz) continue;
                        if (big) {
                            keep |= 1 << g;
                            total += 9;
                            continue;
                        }
                        const uint32_t sc =
                            (uint32_t)(S.a.bscore[4 * g] + S.a.bscore[4 * g + 1] + S.a.bscore[4 * g + 2] + S.a.bscore[4 * g + 3]);
                        if (sc >= 4) keep |= 1 << g;
                        total += sc;
                    }
                    if (total < 6) keep = 0;
                }
                cbpl = keep;
                if (lane < 16 && !((keep >> (lane >> 2)) & 1))
#pragma unroll
                    for (int i = 0; i < 16; i++) S.a.coef[lane][i] = 0;
                __syncthreads();
                // enc.rs:2084 quant_chroma_inter: a plane's AC goes when it scores below 7.
                quant_chroma(S.a, qpc, 1);
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    const int any = S.a.cmask[4 * p] | S.a.cmask[4 * p + 1] | S.a.cmask[4 * p + 2] | S.a.cmask[4 * p + 3];
                    const int sc = S.a.cscore[4 * p] + S.a.cscore[4 * p + 1] + S.a.cscore[4 * p + 2] + S.a.cscore[4 * p + 3];
                    if (any && sc < 7) {
                        __syncthreads();
                        if (lane < 4) {
                            S.a.cmask[4 * p + lane] = 0;
#pragma unroll
                            for (int i = 0; i < 16; i++) S.a.craster[p][lane][i] = 0;
                        }
                        __syncthreads();
                    }
                }
                cbpc = chroma_cbp(S.a);
                if (cbpl == 0 && cbpc == 0 && !p8 && S.st[32] == pskx && S.st[33] == psky) {
                    mb_type = 0;
                } else {
                    mb_type = p8 ? RGH_MB_P_8X8 : 1;
                    if (lane < 16 && ((cbpl >> (lane >> 2)) & 1)) {
                        int blk[16];
#pragma unroll
                        for (int i = 0; i < 16; i++) blk[i] = S.a.coef[lane][i];
                        dequant4(blk, qp);
                        add_idct4(S.a.pred, 16, blk_x(lane) * 4, blk_y(lane) * 4, blk);
                    }
                    __syncthreads();
                    recon_chroma(S.a, cbpc, qpc);
                }
            }
        }
    }

    // ---- Outputs: reconstruction, decision record and levels. ----
    uint32_t qw[4];
    if (mb_type == 0) {
        // P_Skip: the prediction at the skip vector is the reconstruction.
        *(uint32_t *)&S.a.pred[lane * 4] = predw;
        S.a.predu[lane] = (uint8_t)pu;
        S.a.predv[lane] = (uint8_t)pv;
        cbpl = cbpc = 0;
        for (int k = 0; k < 4; k++) qw[k] = (uint32_t)(uint16_t)pskx | ((uint32_t)(uint16_t)psky << 16);
    } else if (mb_type == 2 || mb_type == 3) {
        for (int k = 0; k < 4; k++) qw[k] = 0;
    } else {
        for (int k = 0; k < 4; k++)
            qw[k] = (uint32_t)(uint16_t)S.st[8 * (4 + k)] | ((uint32_t)(uint16_t)S.st[8 * (4 + k) + 1] << 16);
    }
    __syncthreads();
    *(RGH_G uint32_t *)&J.cur_y[(long)(y0 + rr) * J.sy + x0 + c0] = *(const uint32_t *)&S.a.pred[lane * 4];
    J.cur_u[(long)(cy0 + (lane >> 3)) * J.sc + cx0 + (lane & 7)] = S.a.predu[lane];
    J.cur_v[(long)(cy0 + (lane >> 3)) * J.sc + cx0 + (lane & 7)] = S.a.predv[lane];
    if (J.pre_y) {
        *(RGH_G uint32_t *)&J.pre_y[(long)(y0 + rr) * W + x0 + c0] = *(const uint32_t *)&S.a.pred[lane * 4];
        J.pre_u[(long)(cy0 + (lane >> 3)) * (W / 2) + cx0 + (lane & 7)] = S.a.predu[lane];
        J.pre_v[(long)(cy0 + (lane >> 3)) * (W / 2) + cx0 + (lane & 7)] = S.a.predv[lane];
    }
    RGH_G int16_t *lv = J.levels + (long)idx * RGH_LEVELS_PER_MB;
    uint32_t nnz = 0;
    if (mb_type != 0) {
        int nzb = 0;
        if (mb_type == 3) {
            for (int i = lane; i < 256; i += 64) lv[i] = S.a.q4i[i >> 4][zz4(i & 15)];
            if (lane < 16)
#pragma unroll
                for (int i = 0; i < 16; i++) nzb |= S.a.q4i[lane][i] != 0;
        } else {
            for (int i = lane; i < 256; i += 64) lv[i] = S.a.coef[i >> 4][zz4(i & 15)];
            if (lane < 16 && !(mb_type == 2 && !cbpl))
#pragma unroll
                for (int i = 0; i < 16; i++) nzb |= S.a.coef[lane][i] != 0;
        }
        const uint64_t bl = __ballot(nzb);
#pragma unroll
        for (int b = 0; b < 16; b++)
            if ((bl >> b) & 1) nnz |= 1u << (blk_y(b) * 4 + blk_x(b));
        if (lane < 16) lv[256 + lane] = mb_type == 2 ? S.a.dcq[zz4(lane)] : 0;
        if (lane < 8) lv[272 + lane] = S.a.cdc[lane >> 2][lane & 3];
        for (int i = lane; i < 128; i += 64) {
            const int p = i >> 6, b = (i >> 4) & 3, k = i & 15;
            lv[280 + i] = (k == 0 || cbpc < 2) ? 0 : S.a.craster[p][b][zz4(k)];
        }
    } else {
        if (lane < RGH_LEVELS_PER_MB / 8) {  // 816 bytes, 16 a lane
            ((RGH_G uint64_t *)lv)[2 * lane] = 0;
            ((RGH_G uint64_t *)lv)[2 * lane + 1] = 0;
        }
    }
    if (lane == 0) {
        D.mb_type = (uint8_t)mb_type;
        D.i16_mode = (uint8_t)(mb_type == 2 ? i16_mode : 0);
        D.chroma_mode = (uint8_t)(mb_type == 2 || mb_type == 3 ? cmode : 0);
        D.cbp_luma = (uint8_t)cbpl;
        D.cbp_chroma = (uint8_t)cbpc;
        D.nnz = (uint16_t)nnz;
        const uint32_t w[5] = {qw[0], qw[1], qw[2], qw[3], ((uint32_t)(uint16_t)mvrx) | ((uint32_t)(uint16_t)mvry << 16)};
        __builtin_memcpy((uint8_t *)&D + 8, w, 20);
        if (mb_type == 3) __builtin_memcpy(D.mv, &modes, sizeof(modes));
        store_dec(&J.dec[idx], D);
    }
    return true;
}
