"""Writes the golden fixtures of the parity tests from the FP64 CPU oracle.

The reference ships no fixtures for this path (and cannot be built here), so the
vectors come from the oracle, which the pinning tests check against closed forms,
conservation laws, direct solves and finite differences.  Run from the repository
root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

SEED = 20240601


def cantilever_case():
    """8x4x4 cantilever, 12 Newmark steps (dt 0.01, 5-step pulse), objective 3, solver rtol 1e-13."""
    p = O.cantilever_params(8, 4, 4, nsteps=12, dt=0.01, pulse_steps=5, rtol=1e-13)
    rng = np.random.default_rng(SEED)
    x = rng.uniform(0.05, 1.0, O.nel(p))
    st = O.forward(p, x)
    fx, df = O.objective_gradient(p, x, st)
    xt = O.pde_filter(p, 0.5, x, rtol=1e-14)
    # one MMA step (a = 0, c = 1000, d = 0; asymptotes 0.5/0.7/1.2, outer move limit 0.2)
    mma = O.MMA(len(x), 1, x)
    g = np.array([x.mean() - 0.3])
    dg = np.full((1, len(x)), 1.0 / len(x))
    xmin, xmax = O.MMA.outer_movelimit(0.0, 1.0, 0.2, x)
    xnew = mma.update(x, df / abs(fx), g, dg, xmin, xmax)
    return dict(nel=np.array([8, 4, 4]), nsteps=np.array(12), dt=np.array(0.01), pulse_steps=np.array(5),
                seed=np.array(SEED), rtol=np.array(1e-13), xphys=x, states=st, fx=np.array(fx), dfdx=df,
                filter_rmin=np.array(0.5), xtilde=xt, mma_gx=g, mma_xnew=xnew,
                ke=O.hex8_ke(0.25, 0.25, 0.25, 0.3), me=O.hex8_me(0.25, 0.25, 0.25))


def design_loop_case():
    """The 32x16x16 cantilever design loop (dt 0.01, 100 steps, 20-step pulse, objective 3, PDE filter
    rmin 0.08 with the default projection beta 0.1 / eta 0.15, volume fraction and start 0.12, move
    limit 0.2, 5 MMA iterations), solves at rtol 1e-12.  About 90 s on 8 cores."""
    p = O.cantilever_params(32, 16, 16, nsteps=100, dt=0.01, pulse_steps=20, rtol=1e-12)
    x, xp, hist = O.design_loop(p, 0.08, 0.12, 0.12, 5, projection=True, beta=0.1, eta=0.15, movlim=0.2)
    return dict(hist=hist, x=x, xphys=xp, rtol=np.array(1e-12))


if __name__ == "__main__":
    out = os.path.join(HERE, "cantilever_8x4x4.npz")
    np.savez_compressed(out, **cantilever_case())
    print("wrote", out, os.path.getsize(out), "bytes")
    if "--design-loop" in sys.argv:
        out = os.path.join(HERE, "design_loop_32x16x16.npz")
        np.savez_compressed(out, **design_loop_case())
        print("wrote", out, os.path.getsize(out), "bytes")
