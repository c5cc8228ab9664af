"""Batch tables of the tile kernels (csrc/tt_tables.cpp), checked on the CPU.

The kernels add each batch's 16 element results into the nodes at the same time, so a batch must not hold
two elements that share a node; every element of a tile's box that touches one of the tile's nodes must be
listed exactly once; padding slots own no corner.  A small program builds the tables of several grids
with the library's own builder and checks these properties, the owned-corner masks against a direct
count, and the batch counts against ceil(elements / 16) for the interior tiles (26 for 8x8x4 tiles,
16 for 8x8x2).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "topopt_in_petsc_transient_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

PROG = r"""
#include "tt_tables.cpp"
#include <array>
#include <cstdio>
#include <set>
using namespace tt;
template <int TZ>
static int check(const Grid& g, bool report) {
    std::vector<int4> codes;
    std::vector<int2> at(64);
    build_tab<TZ>(g, codes, at.data());
    int bad = 0;
    for (int c = 0; c < 64; c++) {
        const bool first[3] = {(c & 1) != 0, (c & 4) != 0, (c & 16) != 0};
        const bool last[3] = {(c & 2) != 0, (c & 8) != 0, (c & 32) != 0};
        const int full[3] = {8, 8, TZ};
        const int lext[3] = {g.nx - 8 * ((g.nx + 7) / 8 - 1), g.ny - 8 * ((g.ny + 7) / 8 - 1),
                             g.nz - TZ * ((g.nz + TZ - 1) / TZ - 1)};
        int o[3];
        for (int d = 0; d < 3; d++) o[d] = last[d] ? lext[d] : full[d];
        // the elements the table must hold, with their owned-corner masks
        std::set<std::pair<int, int>> want, got;
        const int E[3] = {9, 9, TZ + 1};
        for (int ez = 0; ez < E[2]; ez++)
            for (int ey = 0; ey < E[1]; ey++)
                for (int ex = 0; ex < E[0]; ex++) {
                    const int e[3] = {ex, ey, ez};
                    bool in = true;
                    for (int d = 0; d < 3; d++)
                        in = in && e[d] >= (first[d] ? 1 : 0) && e[d] <= (last[d] ? o[d] - 1 : E[d] - 1);
                    if (!in) continue;
                    int own = 0;
                    for (int a = 0; a < 8; a++) {
                        const int h[3] = {ex + cx_of(a), ey + cy_of(a), ez + cz_of(a)};
                        if (h[0] >= 1 && h[0] <= o[0] && h[1] >= 1 && h[1] <= o[1] && h[2] >= 1 && h[2] <= o[2])
                            own |= 1 << a;
                    }
                    if (own) want.insert({(ez * 10 + ey) * 10 + ex, own});
                }
        for (int b = 0; b < at[c].y; b++) {
            std::vector<int> box;
            for (int k = 0; k < 16; k++) {
                const int4 w = codes[at[c].x + b * 16 + k];
                const int own = w.x >> 22, bi = w.x & 4095;
                if (!own) continue;
                if (!got.insert({bi, own}).second) bad++;  // listed twice
                for (int q : box)
                    if (std::abs(q % 10 - bi % 10) <= 1 && std::abs(q / 10 % 10 - bi / 10 % 10) <= 1 &&
                        std::abs(q / 100 - bi / 100) <= 1)
                        bad++;  // two elements of one batch share a node
                box.push_back(bi);
            }
        }
        if (got != want) bad++;
        if (c == 0 && !first[0] && !last[0] && g.nx > 17 && g.ny > 17 && g.nz > 2 * TZ + 1 &&
            at[0].y != (int)((want.size() + 15) / 16))
            bad++;
        if (report && c == 0) std::printf("TZ %d interior: %zu elements, %d batches\n", TZ, want.size(), at[0].y);
    }
    return bad;
}
int main() {
    int bad = 0;
    const std::array<int, 3> grids[] = {{128, 64, 64}, {64, 32, 32}, {32, 16, 16}, {16, 8, 8}, {40, 24, 12}};
    for (const auto& d : grids) {
        const Grid g = make_grid(d[0], d[1], d[2], 2.0, 1.0, 1.0);
        bad += check<4>(g, d[0] == 128) + check<2>(g, d[0] == 128);
    }
    std::printf("%d problems\n", bad);
    return bad != 0;
}
"""


def test_batch_tables(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc absent")
    src = tmp_path / "tables.cpp"
    src.write_text(PROG)
    exe = tmp_path / "tables"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", CSRC, "-o", str(exe), str(src)], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TZ 4 interior: 405 elements, 26 batches" in r.stdout
    assert "TZ 2 interior: 243 elements, 16 batches" in r.stdout
