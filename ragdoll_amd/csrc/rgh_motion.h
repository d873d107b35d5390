// Motion vector prediction of the entropy coders: the luma vector predictor of ITU-T H.264
// 8.4.1.3 over the picture's 4x4 blocks as coded so far, one reference picture. Host code.
#pragma once

#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/ragdoll_hip.h"

namespace rgh_motion {

// Partition shapes of a P macroblock, with the rectangle (in samples, within the macroblock) of
// partition p, the partition count and the mb_type of each (Table 7-13).
enum Shape { S16x16 = 0, S16x8 = 1, S8x16 = 2, S8x8 = 3 };
inline void part_rect(int shape, int p, int &bx, int &by, int &w, int &h) {
    bx = shape == S8x16 ? 8 * p : (shape == S8x8 ? 8 * (p & 1) : 0);
    by = shape == S16x8 ? 8 * p : (shape == S8x8 ? 8 * (p >> 1) : 0);
    w = shape <= S16x8 ? 16 : 8;
    h = (shape == S16x16 || shape == S8x16) ? 16 : 8;
}
inline int part_count(int shape) { return shape == S16x16 ? 1 : (shape == S8x8 ? 4 : 2); }
// The shape of an inter record's mb_type, or -1.
inline int shape_of(int mb_type) {
    return mb_type == RGH_MB_P_L0_16X16 ? S16x16
           : mb_type == RGH_MB_P_L0_L0_16X8 ? S16x8
           : mb_type == RGH_MB_P_L0_L0_8X16 ? S8x16
           : mb_type == RGH_MB_P_8X8 ? S8x8 : -1;
}

// The quadrant vectors (raster order) of a record in the quadrant layout (words 2-5).
inline void quads_of(const rgh_mb_decision &d, int q[4][2]) {
    const uint8_t *p = (const uint8_t *)&d + 8;
    for (int k = 0; k < 4; k++) {
        int16_t v[2];
        memcpy(v, p + 4 * k, 4);
        q[k][0] = v[0];
        q[k][1] = v[1];
    }
}

// Reference index (-2 not coded yet, -1 intra, 0 list 0) and vector of every 4x4 block.
struct MotionGrid {
    int gs, gh;
    std::vector<int8_t> ref;
    std::vector<int16_t> mv;  // x, y per block
    MotionGrid(int w4, int h4) : gs(w4), gh(h4), ref((size_t)w4 * h4, -2), mv((size_t)w4 * h4 * 2, 0) {}
    void fill(int gx, int gy, int w4, int h4, int r, int mx, int my) {
        for (int y = gy; y < gy + h4; y++)
            for (int x = gx; x < gx + w4; x++) {
                const size_t i = (size_t)y * gs + x;
                ref[i] = (int8_t)r;
                mv[2 * i] = (int16_t)mx;
                mv[2 * i + 1] = (int16_t)my;
            }
    }
    static int median3(int a, int b, int c) {
        const int mx = a > b ? a : b, mn = a < b ? a : b, t = mn > c ? mn : c;
        return mx < t ? mx : t;
    }
    // 8.4.1.3 for partition `part` of `shape` whose top-left 4x4 block is (gx, gy), pw blocks
    // wide. Neighbours A (left), B (above), C (above right, else D above left) by 6.4.11.7: a
    // block is available when it lies in the picture and is coded (a block of a later partition
    // of this macroblock, or right of it in its rows, is not: the grid marks it -2); B and C take
    // A's values when neither is available but A is (8.4.1.3.1); then the directional rule of
    // 16x8 (B for the upper, A for the lower partition) and 8x16 (A for the left, C for the right
    // partition) when that neighbour uses reference 0, else the one neighbour that does, else the
    // component-wise median. Unavailable and intra neighbours count as reference -1, vector 0.
    void pred(int gx, int gy, int pw, int shape, int part, int out[2]) const {
        struct N {
            bool av;
            int r, x, y;
        };
        auto at = [&](int x, int y) -> N {
            if (x < 0 || y < 0 || x >= gs || y >= gh) return {false, -1, 0, 0};
            const size_t i = (size_t)y * gs + x;
            if (ref[i] == -2) return {false, -1, 0, 0};
            if (ref[i] != 0) return {true, -1, 0, 0};
            return {true, 0, mv[2 * i], mv[2 * i + 1]};
        };
        N a = at(gx - 1, gy), b = at(gx, gy - 1), c = at(gx + pw, gy - 1);
        if (!c.av) c = at(gx - 1, gy - 1);
        if (!b.av && !c.av && a.av) b = c = a;
        const N *dir = nullptr;
        if (shape == S16x8) dir = part == 0 ? &b : &a;
        if (shape == S8x16) dir = part == 0 ? &a : &c;
        if (dir && dir->r == 0) {
            out[0] = dir->x;
            out[1] = dir->y;
            return;
        }
        const int hits = (a.r == 0) + (b.r == 0) + (c.r == 0);
        if (hits == 1) {
            const N &n = a.r == 0 ? a : (b.r == 0 ? b : c);
            out[0] = n.x;
            out[1] = n.y;
            return;
        }
        out[0] = median3(a.x, b.x, c.x);
        out[1] = median3(a.y, b.y, c.y);
    }
};

}  // namespace rgh_motion
