"""The main.cc-order design loop through the drop-in classes (topopt_gpu driver, GPU).

Runs the 32x16x16 cantilever design loop (dt 0.01, 100 steps, 3-level multigrid, PDE
filter rmin 0.08 with the reference's default projection, volume fraction and initial design
0.12 (-volfrac 0.12 -V0 0.12), the half-sine y-line pulse (-tt_loadloc 2 -tt_loadsignal 1
-tt_pulse_steps 20), 5 MMA iterations) and compares its per-iteration history and final design with the oracle's
run of the same loop (tests/golden/design_loop_32x16x16.npz), both with solves at rtol
1e-12.  The design loop amplifies solver round-off a little each iteration, so the bar is
1e-8 relative on the objective history and 1e-7 on the final densities (FP64), 1e-5 and
1e-4 in the mixed-precision mode.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.gpu_helpers import rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "topopt_in_petsc_transient_amd", "topopt_gpu")
GOLD = os.path.join(ROOT, "tests", "golden", "design_loop_32x16x16.npz")

ARGS = ["-nx", "33", "-ny", "17", "-nz", "17", "-xcmax", "2", "-ycmax", "1", "-zcmax", "1", "-nlvls", "3",
        "-boundaries", "1", "-tt_loadloc", "2", "-tt_loadsignal", "1", "-tt_pulse_steps", "20", "-optimizationRegion", "2",
        "-objective", "3", "-Emax", "1", "-Ef", "1e-9", "-Rmax", "1", "-Rf", "1e-3", "-penalK", "3", "-penalM", "1",
        "-sInt", "1", "-massInterpolation", "0", "-dalph", "1e-2", "-dbeta", "1e-4", "-nu", "0.3", "-ti", "1",
        "-t1", "1.0", "-ncp", "1", "-tsc", "100", "-filter", "2", "-rmin", "0.08", "-projectionFilter", "1",
        "-beta", "0.1", "-eta", "0.15", "-volfrac", "0.12", "-V0", "0.12", "-movlim", "0.2", "-maxItr", "5",
        "-ksp_rtol", "1e-12", "-ksp_max_it", "500", "-filter_rtol", "1e-13"]


@pytest.mark.parametrize("precision,tol_f,tol_x", [(0, 1e-8, 1e-7), (1, 1e-5, 1e-4)])
def test_drop_in_design_loop_matches_oracle(tmp_path, precision, tol_f, tol_x):
    assert os.path.exists(DRIVER), "build the driver first (__graft_entry__.build())"
    hist_file = tmp_path / "hist.jsonl"
    cmd = [DRIVER, *ARGS, "-precision", str(precision), "-history", str(hist_file)]
    print("running", " ".join(cmd), flush=True)
    # (run in the temporary directory: the driver writes the reference's restart files into ./)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=tmp_path)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    lines = [json.loads(s) for s in hist_file.read_text().splitlines() if s.strip()]
    its = [d for d in lines if "itr" in d]
    xphys = np.array([d for d in lines if "xphys" in d][0]["xphys"])
    g = np.load(GOLD, allow_pickle=False)
    hist = g["hist"]  # columns: fx, scaled fx, gx, ch, mnd
    assert len(its) == len(hist)
    for k, d in enumerate(its):
        assert abs(d["fx"] - hist[k, 0]) <= tol_f * abs(hist[k, 0]), k
        assert abs(d["gx"] - hist[k, 2]) <= tol_f * 1e-2 + tol_f * abs(hist[k, 2]), k
        assert abs(d["ch"] - hist[k, 3]) <= tol_x, k
        assert abs(d["mnd"] - hist[k, 4]) <= tol_f * abs(hist[k, 4]), k
    assert rel(xphys, g["xphys"]) < tol_x
