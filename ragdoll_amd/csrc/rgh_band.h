// Placement of a wavefront step's items on the XCDs (device code and its host-side test).
#pragma once

#ifdef __HIPCC__
#define RGH_BAND_FN __host__ __device__ __forceinline__
#else
#define RGH_BAND_FN static inline
#endif

// Items of a wavefront step on XCDs by macroblock row. A device deals the workgroups of a launch
// round-robin over its 8 XCDs in workgroup order, so workgroup column bx of a grid row whose
// width is a multiple of 8 runs on XCD (bx + d) mod 8 for a d fixed within the launch. The items
// of one frame's step lie on consecutive rows y0, y0 + 1, ...; rows are grouped in bands of
// RGH_BAND_ROWS, and column bx takes the row whose band k satisfies (k + h) mod 8 = bx mod 8, so
// that a band's items run on one XCD from launch to launch while the dispatcher starts where it
// did: the next macroblock of a row reuses most of the search region its left neighbour brought
// into that XCD's L2 cache, and the rows below a band's first reuse the rows above. Any 8 x
// RGH_BAND_ROWS consecutive rows hold each band residue RGH_BAND_ROWS times, so every aligned
// group of 8 x RGH_BAND_ROWS columns maps one to one onto the group's rows, whatever XCD each
// workgroup actually runs on (the mapping only places work; it never decides what runs).
#ifndef RGH_BAND_ROWS
#define RGH_BAND_ROWS 8
#endif
#define RGH_ITEM_GROUP (8 * RGH_BAND_ROWS)
RGH_BAND_FN int band_item_row(int bx, int y0, int h) {
    constexpr int R = RGH_BAND_ROWS;
    const int q0 = y0 + (bx & ~(RGH_ITEM_GROUP - 1));  // first row of bx's group
    const int c = bx & 7, jj = (bx >> 3) & (R - 1);   // band residue wanted, and which of its R rows
    const int k0 = q0 / R, o = q0 & (R - 1);
    const int k = k0 + ((c - h - k0) & 7);             // first band at or after k0 with (k + h) % 8 == c
    // Band k lies wholly in the group, or it is the group's partial first band (then its rows
    // q0 .. (k0 + 1) R - 1 come first and the group's partial last band, k0 + 8, supplies the rest).
    return k * R >= q0 ? k * R + jj : q0 + jj + (jj >= R - o ? 7 * R : 0);
}
