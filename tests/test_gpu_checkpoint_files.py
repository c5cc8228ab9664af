"""Checkpoint states given by the caller and the reference's checkpoint files cp%04d.dat (GPU).

* tt_set_checkpoint: block start states taken from a forward sweep replace it; the gradient recomputed
  block by block from them equals the gradient of the sweep (solves at rtol 1e-12: 1e-10), with every
  state kept and with binomial checkpoint slots.
* The drop-in driver writes cp0000.dat (the initial state, LinearElasticity.cc:1128-1137) and, with -ncp 2
  (main.cc:85-87 calls GenerateCheckPoints only then), cp0001.dat and cp0002.dat (U, DU, DDU at steps 50
  and 100, :1261-1276) in PETSc binary form, and reads them back for the gradient (:1651-1672): the first
  design iteration's objective equals the -ncp 1 run's, whose gradient uses the device checkpoints.
* Files written from the oracle's states (as the reference would write them) drive the library's gradient:
  fx and dfdx against the oracle.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import problem, random_density, rel
from tests.test_gpu_design_loop import ARGS, DRIVER
from topopt_in_petsc_transient_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slots", [0, 5])
def test_gradient_from_given_block_states(slots):
    po, pg = problem(nsteps=30, slots=slots)
    x = random_density(O.nel(po), seed=31)
    c = capi.Context(pg)
    c.set_density(x)
    c.complete_initial_state()
    c.forward()
    blocks = [0, 10, 20]
    saved = {t: c.get_state(t) for t in blocks}
    fx_ref, df_ref = c.objective_gradient()
    c.complete_initial_state()  # a fresh initial state, no forward sweep
    for t in reversed(blocks):
        c.set_checkpoint(t, *saved[t])
    fx, df = c.objective_gradient()
    print(f"slots {slots}: fx {fx:.15e} vs {fx_ref:.15e}, dfdx rel diff {rel(df, df_ref):.2e}", flush=True)
    assert abs(fx - fx_ref) <= 1e-10 * abs(fx_ref) and rel(df, df_ref) < 1e-10
    # the oracle agrees as well
    ref = O.forward(po, x)
    fo, dfo = O.objective_gradient(po, x, ref)
    assert abs(fx - fo) <= 1e-8 * abs(fo) and rel(df, dfo) < 1e-8
    c.close()


def _driver(tmp_path, name, extra):
    wd = tmp_path / name
    wd.mkdir()
    hist = wd / "hist.jsonl"
    cmd = [DRIVER, *ARGS, "-precision", "0", "-maxItr", "1", "-workdir", str(wd), "-history", str(hist),
           "-outputPhysics0", "0", "-outputPhysics1", "0", "-outputDesign", "0", *extra]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=wd)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = [json.loads(s) for s in hist.read_text().splitlines() if s.strip()]
    return wd, [d for d in lines if "itr" in d]


def test_driver_writes_and_reads_checkpoint_files(tmp_path):
    wd1, one = _driver(tmp_path, "ncp1", ["-ncp", "1", "-tsc", "100"])
    wd2, two = _driver(tmp_path, "ncp2", ["-ncp", "2", "-tsc", "50"])
    assert (wd1 / "cp0000.dat").exists() and not (wd1 / "cp0001.dat").exists()
    for f in ("cp0000.dat", "cp0001.dat", "cp0002.dat"):
        assert (wd2 / f).exists(), f
    assert abs(two[0]["fx"] - one[0]["fx"]) <= 1e-9 * abs(one[0]["fx"])
    # the files hold U, DU, DDU of steps 0, 50 and 100 of the first design
    nd = 3 * 33 * 17 * 17
    u0, v0, _ = capi.petsc_read_vecs(wd2 / "cp0000.dat", [nd] * 3)
    assert not u0.any() and not v0.any()  # the initial state: u0 = v0 = 0
    assert np.array_equal(capi.petsc_read_vecs(wd1 / "cp0000.dat", [nd] * 3)[2],
                          capi.petsc_read_vecs(wd2 / "cp0000.dat", [nd] * 3)[2])
    u50, v50, a50 = capi.petsc_read_vecs(wd2 / "cp0001.dat", [nd] * 3)
    u100, v100, a100 = capi.petsc_read_vecs(wd2 / "cp0002.dat", [nd] * 3)
    assert min(np.linalg.norm(q) for q in (u50, v50, a50, u100, v100, a100)) > 0


def test_gradient_from_oracle_written_checkpoint_files(tmp_path):
    """Checkpoint files written from the oracle's states (as the reference writes them) drive the library's
    block-wise gradient: fx and dfdx against the oracle's adjoint."""
    po, pg = problem(nsteps=24, slots=4)
    x = random_density(O.nel(po), seed=33)
    ref = O.forward(po, x)
    fo, dfo = O.objective_gradient(po, x, ref)
    nd = ref.shape[2]
    for c, t in enumerate((0, 8, 16)):
        capi.petsc_write_vecs(tmp_path / f"cp{c:04d}.dat", [ref[t, 0], ref[t, 1], ref[t, 2]])
    c = capi.Context(pg)
    c.set_density(x)
    c.complete_initial_state()
    for k, t in reversed(list(enumerate((0, 8, 16)))):
        u, v, a = capi.petsc_read_vecs(tmp_path / f"cp{k:04d}.dat", [nd] * 3)
        c.set_checkpoint(t, u, v, a)
    fx, df = c.objective_gradient()
    print(f"fx {fx:.15e} (oracle {fo:.15e}), dfdx rel err {rel(df, dfo):.2e}", flush=True)
    assert abs(fx - fo) <= 1e-8 * abs(fo) and rel(df, dfo) < 1e-8
    c.close()
