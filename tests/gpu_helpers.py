"""Shared helpers of the GPU parity tests: matched problem set-ups for the HIP
library (through its C-ABI) and the FP64 CPU oracle."""
import numpy as np

from oracle import oracle as O
from topopt_in_petsc_transient_amd import capi


def problem(nelx=16, nely=8, nelz=8, nsteps=30, dt=0.01, pulse_steps=10, objective=3, nlvls=2,
            precision=capi.TT_PREC_FP64, rtol=1e-12, slots=0, rmin=0.08, lx=2.0, ly=1.0, lz=1.0, **kw):
    """Cantilever pulse problem (x=0 face clamped, y-line load at x=lx, z=0, half-sine pulse)."""
    common = dict(nelx=nelx, nely=nely, nelz=nelz, nsteps=nsteps, objective=objective)
    po = O.cantilever_params(dt=dt, pulse_steps=pulse_steps, lx=lx, ly=ly, lz=lz, rtol=1e-13, **common)
    pg = capi.default_params(t1=dt * nsteps, pulse_steps=pulse_steps, xc=(0.0, lx, 0.0, ly, 0.0, lz), nlvls=nlvls,
                             precision=precision, rtol=rtol, checkpoint_slots=slots, rmin=rmin,
                             filter_rtol=1e-13, maxit=500, **common)
    for k, v in kw.items():
        setattr(po, k, v)
        setattr(pg, k, v)
    return po, pg


def random_density(n, seed=0, lo=0.05, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (d if d > 0 else 1.0)
