"""Index helpers shared by the kernels (csrc/tt_internal.h), checked on the CPU.

* grid_ijk: (i, j, k) of a node or element index from two 32-bit divisions, used by the restriction,
  prolongation and PDE-filter kernels; compared with 64-bit division and remainder over every index of
  small odd grids and over samples of grids with up to 2^31 - 1 points.
* sens_dof: the sensitivity kernel's element dof order.  It must be a permutation of the 24 dofs, and
  the six dofs of lane group q must lie at consecutive memory offsets 3 n + t (t = 0..5) from the node n
  of the low-x corner of the group's edge (0 for q = 0, + nx for q = 1, + nx ny for q = 2, both for
  q = 3), with the corners placed as in the reference's element numbering (corner_id).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "topopt_in_petsc_transient_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

PROG = r"""
#include "tt_internal.h"
#include <cstdint>
#include <cstdio>
using namespace tt;
static int check_grid(long nx, long ny, long nz, long stride) {
    int bad = 0;
    const long n = nx * ny * nz;
    for (long m = 0; m < n; m += stride) {
        int i, j, k;
        grid_ijk((unsigned)m, (unsigned)nx, (unsigned)ny, i, j, k);
        if (i != m % nx || j != (m / nx) % ny || k != m / (nx * ny)) bad++;
    }
    int i, j, k;  // the last index
    grid_ijk((unsigned)(n - 1), (unsigned)nx, (unsigned)ny, i, j, k);
    if (i != nx - 1 || j != ny - 1 || k != nz - 1) bad++;
    return bad;
}
int main() {
    int bad = 0;
    bad += check_grid(13, 9, 9, 1);
    bad += check_grid(129, 65, 65, 1);
    bad += check_grid(7, 1, 5, 1);
    bad += check_grid(1, 1, 1, 1);
    bad += check_grid(1291, 1291, 1288, 7919);  // 2,146,617,208 points, just under 2^31
    bad += check_grid(65536, 32767, 1, 4099);    // 2,147,418,112 points
    // sens_dof: a permutation, and contiguous per lane group
    int seen[24] = {0};
    for (int k = 0; k < 24; k++) seen[sens_dof(k)]++;
    for (int d = 0; d < 24; d++)
        if (seen[d] != 1) bad++;
    const long nx = 10, ny = 7;
    for (int q = 0; q < 4; q++) {
        const long x0 = 3 * (nx * (q & 1) + nx * ny * (q >> 1));
        for (int t = 0; t < 6; t++) {
            const int dof = sens_dof(4 * t + q), a = dof / 3, c = dof % 3;
            long node = -1;
            for (int z = 0; z < 2; z++)
                for (int y = 0; y < 2; y++)
                    for (int x = 0; x < 2; x++)
                        if (corner_id(x, y, z) == a) node = x + nx * y + nx * ny * z;
            if (3 * node + c != x0 + t) bad++;
        }
    }
    std::printf("%d problems\n", bad);
    return bad != 0;
}
"""


def test_index_helpers(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc absent")
    src = tmp_path / "idx.cpp"
    src.write_text(PROG)
    exe = tmp_path / "idx"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", CSRC, "-o", str(exe), str(src)], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 problems" in r.stdout
