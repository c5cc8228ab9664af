"""Golden fixtures (tests/golden/cantilever_8x4x4.npz, written by make_golden.py).

CPU: the oracle still reproduces the committed vectors (guards the oracle itself).
GPU: the HIP path reproduces them through the C-ABI, FP64 within 1e-8 and the mixed
path within 1e-5 relative.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import rel

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "cantilever_8x4x4.npz")


def _load():
    return dict(np.load(GOLD, allow_pickle=False))


def test_oracle_reproduces_golden_vectors():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = _load()
    fresh = mg.cantilever_case()
    for k in ("xphys", "states", "dfdx", "xtilde", "mma_xnew", "ke", "me"):
        assert rel(fresh[k], g[k]) < 1e-12, k
    assert abs(fresh["fx"] - g["fx"]) <= 1e-12 * abs(g["fx"])


def _gpu_params(g, precision):
    from topopt_in_petsc_transient_amd import capi
    nelx, nely, nelz = (int(v) for v in g["nel"])
    nsteps, dt = int(g["nsteps"]), float(g["dt"])
    return capi.default_params(nelx=nelx, nely=nely, nelz=nelz, nsteps=nsteps, t1=dt * nsteps,
                               pulse_steps=int(g["pulse_steps"]), nlvls=2, precision=precision, rtol=1e-12,
                               maxit=500, rmin=float(g["filter_rmin"]), filter_rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_gpu_path_reproduces_golden_vectors(mixed):
    from topopt_in_petsc_transient_amd import capi
    g = _load()
    tol = 1e-5 if mixed else 1e-8
    c = capi.Context(_gpu_params(g, capi.TT_PREC_MIXED if mixed else capi.TT_PREC_FP64))
    c.set_density(g["xphys"])
    c.complete_initial_state()
    c.forward()
    for t in range(1, int(g["nsteps"]) + 1):
        u, v, a = c.get_state(t)
        for q, arr in enumerate((u, v, a)):
            assert rel(arr, g["states"][t, q]) < tol, (t, q)
    fx, df = c.objective_gradient()
    assert abs(fx - g["fx"]) <= tol * abs(g["fx"])
    assert rel(df, g["dfdx"]) < tol
    assert rel(c.pde_filter(g["xphys"]), g["xtilde"]) < 1e-10
    m = capi.MMA(c)
    x = g["xphys"]
    lo, hi = m.outer_movelimit(0.0, 1.0, 0.2, x)
    xnew = m.update(x, g["dfdx"] / abs(float(g["fx"])), float(g["mma_gx"][0]), np.full(x.size, 1.0 / x.size), lo, hi)
    assert rel(xnew, g["mma_xnew"]) < 1e-10
