"""The placement of a wavefront step's items on XCDs (ragdoll_amd/csrc/rgh_band.h).

band_item_row decides which macroblock row each grid column of a launch analyses. It must be a
one-to-one map of every aligned group of 8 x RGH_BAND_ROWS columns onto that many consecutive
rows, whatever the first row of the step and the picture's band phase, or an item would run
twice and another never; and it places row y on column residue ((y / RGH_BAND_ROWS) + h) mod 8.
Checked exhaustively on the host for the shipped band size and the other powers of two.
"""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "rgh_band.h"
int main(void) {
    const int R = RGH_BAND_ROWS, G = RGH_ITEM_GROUP;
    static int seen[4 * 128];
    for (int y0 = 0; y0 < 300; y0++)
        for (int h = 0; h < 8; h++)
            for (int groups = 1; groups <= 4; groups++) {
                memset(seen, 0, sizeof(seen));
                for (int bx = 0; bx < groups * G; bx++) {
                    const int y = band_item_row(bx, y0, h);
                    if (y < y0 || y >= y0 + groups * G || seen[y - y0]++) {
                        printf("not one to one: R %d y0 %d h %d bx %d -> %d\n", R, y0, h, bx, y);
                        return 1;
                    }
                    if (((y / R) + h) % 8 != bx % 8) {
                        printf("wrong residue: R %d y0 %d h %d bx %d -> %d\n", R, y0, h, bx, y);
                        return 1;
                    }
                }
            }
    printf("ok %d\n", R);
    return 0;
}
"""


@pytest.mark.parametrize("rows", [None, 2, 4, 8, 16])
def test_band_item_row_is_one_to_one(rows):
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "band.cpp")
        exe = os.path.join(tmp, "band")
        with open(src, "w") as f:
            f.write(PROGRAM)
        cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "ragdoll_amd", "csrc"), src, "-o", exe]
        if rows:
            cmd.insert(1, "-DRGH_BAND_ROWS=%d" % rows)
        subprocess.run(cmd, check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0, out.stdout
        assert out.stdout.startswith("ok")
