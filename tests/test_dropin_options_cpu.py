"""Option semantics of the drop-in TopOpt / LinearElasticity (CPU, no GPU work).

A small program parses a command line through the drop-in's PetscInitialize, builds TopOpt
(TopOpt.cc:71-518) and the tt_params the LinearElasticity constructor would create its context
with (LinearElasticityParams), and prints them.  The checks follow the reference's behaviour:

* PETSc matches option names without regard to case: the job scripts' -Rmin reaches the
  rmin that TopOpt.cc:363 reads;
* V0 is set to the default volume fraction (0.3) before -volfrac is read (TopOpt.cc:113-114 vs
  :373), so only -V0 changes the initial design;
* the reference never reads -loadsignal (Load, LinearElasticity.cc:593-615, always applies the
  Gaussian-cosine signal), so the drop-in must not either;
* dt = t1 / (Ts - 1) whatever -t0 says (LinearElasticity.cc:96-99);
* the outer solver defaults to the reference's rtol 1e-10 and 200 iterations (:2750-2754).

When the reference tree is present, every JobScripts/Job_fig*.sh command line is parsed as well.
"""
import json
import os
import re
import shlex
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "topopt_in_petsc_transient_amd")
HIPCC = "/opt/rocm/bin/hipcc"
JOBS = "/root/reference/JobScripts"

PROG = r"""
#include <cstdio>
#include <string>
#include "topopt_dropin.h"
int main(int argc, char** argv) {
    PetscInitialize(&argc, &argv, PETSC_NULL, "");
    TopOpt* opt = new TopOpt();
    // -test_output f: one design dataset through the field writer (MPIIO), x = 0.25, xTilde = 0.5, xPhys = 0.75
    char out[PETSC_MAX_PATH_LEN];
    PetscBool has_out = PETSC_FALSE;
    PetscOptionsGetString(NULL, NULL, "-test_output", out, sizeof(out), &has_out);
    if (has_out) {
        VecSet(opt->x, 0.25);
        VecSet(opt->xTilde, 0.5);
        VecSet(opt->xPhys, 0.75);
        MPIIO* o = new MPIIO(opt->da_nodes, 0, "", 3, "x, xTilde, xPhys", out);
        if (o->WriteVTK(opt->da_nodes, opt->x, opt->xTilde, opt->xPhys, 7) != 0) return 3;
        delete o;
    }
    tt_params p;
    const int rc = LinearElasticityParams(opt->da_nodes, opt->nu, opt->nlvls, opt->ncp, opt->tsc, opt->t1, opt->ti,
                                          opt->objective, opt->sInt, opt->mInt, opt->t0, opt->boundaries,
                                          opt->loadloc, opt->reduction, opt->basis, opt->dalpha, opt->dbeta,
                                          opt->omega0, opt->optimizationRegion, &p);
    std::printf("{\"rc\": %d, \"rmin\": %.17g, \"V0\": %.17g, \"volfrac\": %.17g, \"loadloc\": %d, "
                "\"boundaries\": %d, \"projection\": %d, \"filter\": %d, \"t1\": %.17g, \"Ts\": %d, "
                "\"p_loadloc\": %d, \"p_loadsignal\": %d, \"p_pulse_steps\": %d, \"p_t1\": %.17g, "
                "\"p_nsteps\": %d, \"p_rtol\": %.17g, \"p_maxit\": %d, \"p_nelx\": %d, \"p_nely\": %d, "
                "\"p_nelz\": %d, \"p_dalpha\": %.17g, \"p_dbeta\": %.17g, \"p_objective\": %d, "
                "\"p_cheb\": %d, \"p_coarse_pre\": %d, \"p_coarse_post\": %d}\n",
                rc, opt->rmin, opt->V0, opt->volfrac, opt->loadloc, opt->boundaries, (int)opt->projectionFilter,
                opt->filter, opt->t1, opt->Ts, p.loadloc, p.loadsignal, p.pulse_steps, p.t1, p.nsteps, p.rtol,
                p.maxit, p.nelx, p.nely, p.nelz, p.dalpha, p.dbeta, p.objective, p.cheb_degree, p.mg_coarse_pre,
                p.mg_coarse_post);
    delete opt;
    return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc absent")
    lib = os.path.join(PKG, "libtopopt_gpu.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    d = tmp_path_factory.mktemp("opts")
    src = d / "opts.cpp"
    src.write_text(PROG)
    out = d / "opts"
    # under the sanitizer test run (tests/test_sanitizers_cpu.py) the program's host code is built with the
    # sanitizers too and linked against the host-sanitized library
    san = os.environ.get("TT_SANITIZE_CXXFLAGS", "").split()
    libdir = os.path.dirname(os.environ["TT_LIB"]) if san and os.environ.get("TT_LIB") else PKG
    link_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
                "-shared-libsan"] if san else []
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-xhip", *san, *link_san, "-I",
           os.path.join(ROOT, "include"), str(src), os.path.join(PKG, "csrc", "topopt_classes.cpp"), "-L", libdir,
           "-ltopopt_gpu", f"-Wl,-rpath,{libdir}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


def parse(prog, args):
    r = subprocess.run([prog, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [s for s in r.stdout.splitlines() if s.startswith("{")][-1]
    return json.loads(line)


def test_defaults_are_the_reference_defaults(prog):
    d = parse(prog, [])
    assert d["rc"] == 0
    assert (d["rmin"], d["V0"], d["volfrac"], d["boundaries"], d["loadloc"]) == (1.5, 0.3, 0.3, 0, 0)
    assert (d["t1"], d["Ts"], d["p_nsteps"], d["p_t1"]) == (50.0, 51, 50, 50.0)
    assert (d["p_loadsignal"], d["p_rtol"], d["p_maxit"]) == (0, 1e-10, 200)


def test_option_names_match_without_case(prog):
    d = parse(prog, ["-Rmin", "2.4", "-VOLFRAC", "0.2", "-NX", "33"])
    assert d["rmin"] == 2.4 and d["volfrac"] == 0.2 and d["p_nelx"] == 32


def test_v0_is_the_default_volfrac_unless_given(prog):
    d = parse(prog, ["-volfrac", "0.12"])
    assert d["volfrac"] == 0.12 and d["V0"] == 0.3
    d = parse(prog, ["-volfrac", "0.12", "-V0", "0.12"])
    assert d["V0"] == 0.12


def test_loadsignal_is_not_an_option_of_the_reference(prog):
    for v in ("0", "1"):
        d = parse(prog, ["-boundaries", "1", "-loadsignal", v])
        assert d["p_loadsignal"] == 0 and d["p_loadloc"] == 1


def test_own_load_preset_has_its_own_names(prog):
    d = parse(prog, ["-tt_loadloc", "2", "-tt_loadsignal", "1", "-tt_pulse_steps", "20"])
    assert (d["p_loadloc"], d["p_loadsignal"], d["p_pulse_steps"]) == (2, 1, 20)


def test_load_location_without_preset_is_rejected(prog):
    assert parse(prog, ["-loadloc", "2"])["rc"] != 0


def test_time_grid_ignores_t0(prog):
    d = parse(prog, ["-t0", "5", "-t1", "2.0", "-ncp", "2", "-tsc", "10"])
    assert (d["p_t1"], d["p_nsteps"]) == (2.0, 20)


def test_solver_options(prog):
    d = parse(prog, ["-ksp_rtol", "1e-8", "-ksp_max_it", "50"])
    assert (d["p_rtol"], d["p_maxit"]) == (1e-8, 50)
    assert (d["p_cheb"], d["p_coarse_pre"], d["p_coarse_post"]) == (2, 1, 0)
    d = parse(prog, ["-cheb_degree", "3", "-mg_coarse_pre", "2", "-MG_Coarse_Post", "1"])
    assert (d["p_cheb"], d["p_coarse_pre"], d["p_coarse_post"]) == (3, 2, 1)


def test_field_writer_layout(prog, tmp_path):
    """The single-rank MPIIO writer (MPIIO.cc:32-586): header, names, points, cells and one design dataset,
    read back with the layout bin2vtu.py reads."""
    import numpy as np

    from tests.test_gpu_output import read_output
    f = tmp_path / "design.dat"
    parse(prog, ["-nx", "9", "-ny", "5", "-nz", "5", "-nlvls", "2", "-test_output", str(f)])
    d = read_output(f)
    assert (d["ndom"], d["npts"], d["ncells"], d["npf"], d["ncf"], d["npe"]) == (1, 225, 128, 0, 3, 8)
    assert d["cnames"].replace(" ", "") == "x,xTilde,xPhys"
    assert np.array_equal(d["types"], np.full(128, 12)) and np.array_equal(d["offs"], 8 * np.arange(1, 129))
    # the default MBB box: 60 x 20 x 20, node (i, j, k) at (7.5 i, 5 j, 5 k)
    assert np.allclose(d["pts"][[0, 1, 9, 224]], [[0, 0, 0], [7.5, 0, 0], [0, 5, 0], [60, 20, 20]])
    assert len(d["data"]) == 1 and d["data"][0][0] == 7
    for field, v in zip(d["data"][0][2], (0.25, 0.5, 0.75)):
        assert np.all(field == np.float32(v))


def test_boolean_values_without_case(prog):
    assert parse(prog, ["-projectionFilter", "FALSE"])["projection"] == 0
    assert parse(prog, ["-projectionFilter", "true"])["projection"] == 1


def _job_lines():
    if not os.path.isdir(JOBS):
        return []
    out = []
    for f in sorted(os.listdir(JOBS)):
        if not re.match(r"Job_fig.*\.sh$", f):
            continue
        txt = open(os.path.join(JOBS, f)).read().replace("\\\n", " ")
        for line in txt.splitlines():
            if "./topopt" in line and not line.lstrip().startswith("#"):
                out.append((f, shlex.split(line.split("./topopt", 1)[1].replace("\\", " "))))
    return out


@pytest.mark.skipif(not os.path.isdir(JOBS), reason="reference job scripts absent")
def test_reference_job_scripts_parse_as_the_reference_reads_them(prog):
    jobs = _job_lines()
    assert len(jobs) == 7
    for name, args in jobs:
        low = [a.lower() for a in args]

        def val(opt, default):
            return float(args[low.index(opt) + 1]) if opt in low else default

        d = parse(prog, args)
        assert d["rc"] == 0, name
        assert d["rmin"] == val("-rmin", 1.5), name          # -Rmin 2.4 in every script
        assert d["V0"] == val("-v0", 0.3), name              # -V0 1.0 in every script
        assert d["volfrac"] == val("-volfrac", 0.3), name
        assert d["p_loadsignal"] == 0, name                  # -loadsignal is never read by the reference
        assert d["p_loadloc"] == int(val("-loadloc", d["loadloc"])), name
        assert d["p_t1"] == val("-t1", 50.0), name
        assert d["p_nsteps"] == int(val("-ncp", 2)) * int(val("-tsc", 25)), name
        assert d["p_rtol"] == 1e-10 and d["p_maxit"] == 200, name
