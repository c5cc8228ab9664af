"""Field output of the drop-in driver (MPIIO.cc, main.cc:39-49, 102-105, 167-176) on the GPU: the three
files main.cc writes by default, read back with the layout bin2vtu.py reads (restated below), and
their fields checked against the library run on the same designs.
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import problem, rel
from tests.test_gpu_design_loop import ARGS, DRIVER
from topopt_in_petsc_transient_amd import capi

pytestmark = pytest.mark.gpu

NX, NY, NZ = 33, 17, 17


def read_output(path):
    """The layout bin2vtu.py reads: info string, header, names (each string ends in 0x01), float32 points,
    unsigned-long connectivity, offsets and types, then datasets (step, point fields, cell fields)."""
    b = path.read_bytes()
    pos = 0

    def string():
        nonlocal pos
        end = b.index(b"\x01", pos)
        s = b[pos:end].decode()
        pos = end + 1
        return s

    def arr(dtype, n):
        nonlocal pos
        a = np.frombuffer(b, dtype=dtype, count=n, offset=pos)
        pos += a.nbytes
        return a

    info = string()
    ndom = int(arr("<u8", 1)[0])
    npts, ncells, npf, ncf = (int(v) for v in arr("<u8", 4 * ndom))
    nodes_per_el = int(arr("<u8", 1)[0])
    pnames, cnames = string(), string()
    pts = arr("<f4", 3 * npts).reshape(npts, 3)
    conn = arr("<u8", nodes_per_el * ncells).reshape(ncells, nodes_per_el)
    offs, types = arr("<u8", ncells), arr("<u8", ncells)
    data = []
    while pos < len(b):
        step = int(arr("<u8", 1)[0])
        pf = [arr("<f4", npts) for _ in range(npf)]
        cf = [arr("<f4", ncells) for _ in range(ncf)]
        data.append((step, pf, cf))
    return dict(info=info, ndom=ndom, npts=npts, ncells=ncells, npf=npf, ncf=ncf, npe=nodes_per_el, pnames=pnames,
                cnames=cnames, pts=pts, conn=conn, offs=offs, types=types, data=data)


def _state_u(x, steps):
    """Displacements of the library's FP64 run of the design-loop problem (the driver's arguments) on
    density x."""
    _, pg = problem(nelx=32, nely=16, nelz=16, nsteps=100, dt=0.01, pulse_steps=20, nlvls=3, rtol=1e-12)
    c = capi.Context(pg)
    c.set_density(np.asarray(x, dtype=np.float64))
    c.complete_initial_state()
    c.forward()
    out = {t: c.get_state(t)[0] for t in steps}
    c.close()
    return out


def test_driver_writes_the_reference_outputs(tmp_path):
    assert os.path.exists(DRIVER), "build the driver first (__graft_entry__.build())"
    hist = tmp_path / "h.jsonl"
    cmd = [DRIVER, *ARGS, "-precision", "0", "-history", str(hist)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    lines = [json.loads(s) for s in hist.read_text().splitlines() if s.strip()]
    its = [d for d in lines if "itr" in d]
    xfinal = np.array([d for d in lines if "xphys" in d][0]["xphys"])

    design = read_output(tmp_path / "outputDesign_00000.dat")
    phys0 = read_output(tmp_path / "outputPhysics0_00000.dat")
    phys1 = read_output(tmp_path / "outputPhysics1_00000.dat")
    p = O.cantilever_params(32, 16, 16)
    for f, npf, ncf, names in ((design, 0, 3, ("", "x, xTilde, xPhys")), (phys0, 3, 1, ("ux, uy, uz", "xPhys")),
                               (phys1, 3, 1, ("ux, uy, uz", "xPhys"))):
        assert f["info"] == "TopOpt result version 1.1\n"
        assert (f["ndom"], f["npts"], f["ncells"], f["npf"], f["ncf"], f["npe"]) == (1, NX * NY * NZ, 8192, npf, ncf, 8)
        assert (f["pnames"], f["cnames"]) == names
        n = np.arange(NX * NY * NZ)
        i, j, k = n % NX, (n // NX) % NY, n // (NX * NY)
        assert np.array_equal(f["pts"], np.stack([2.0 * i / 32, 1.0 * j / 16, 1.0 * k / 16], 1).astype(np.float32))
        for e in (0, 31, 4095, 8191):
            assert list(f["conn"][e]) == list(O.element_nodes(p, e)), e
        assert np.array_equal(f["offs"], 8 * np.arange(1, 8193)) and np.all(f["types"] == 12)

    # design output: one dataset per iteration (main.cc:102-105), densities in [0, 1]
    assert [d[0] for d in design["data"]] == [d["itr"] for d in its]
    for _, pf, cf in design["data"]:
        assert pf == [] and all(np.all((c >= 0) & (c <= 1)) for c in cf)
    # physics of the first design (steps 0..100) and of the final one (main.cc:167-176)
    for f, xphys in ((phys0, design["data"][0][2][2]), (phys1, xfinal)):
        assert [d[0] for d in f["data"]] == list(range(101))
        assert all(not np.any(u) for u in f["data"][0][1])  # u0 = 0
        assert rel(f["data"][50][2][0], xphys) < 1e-7
        ref = _state_u(xphys, (30, 100))
        for t in (30, 100):
            u = np.stack(f["data"][t][1], 1).ravel()  # ux, uy, uz per node
            assert rel(u, ref[t]) < 1e-5, t
    assert rel(phys1["data"][1][2][0], xfinal) < 1e-7


def test_reader_restatement_on_a_hand_built_file(tmp_path):
    """The reader above against bytes packed from the layout of MPIIO.cc:409-586, 724-905."""
    head = b"TopOpt result version 1.1\n\x01" + struct.pack("<6Q", 1, 2, 1, 1, 1, 8) + b"u\x01x\x01"
    body = struct.pack("<6f", 0, 0, 0, 1, 0, 0) + struct.pack("<8Q", *range(8)) + struct.pack("<2Q", 8, 12)
    data = struct.pack("<Q", 7) + struct.pack("<2f", 0.5, -0.5) + struct.pack("<f", 0.25)
    f = tmp_path / "o.dat"
    f.write_bytes(head + body + data)
    o = read_output(f)
    assert o["npts"] == 2 and o["ncells"] == 1 and o["data"][0][0] == 7
    assert list(o["data"][0][1][0]) == [0.5, -0.5] and list(o["data"][0][2][0]) == [0.25]
