"""Adjoint sensitivities of the HIP path (GPU).

* gradcheck: the C-ABI objective/gradient pair wrapped (here, in the test harness
  only) in a torch.autograd.Function; torch.autograd.gradcheck compares the
  adjoint gradient with central finite differences (h = 1e-6, FP64) along 20
  seeded random density directions on a 16x8x8 grid with 30 Newmark steps.
* the adjoint sensitivities on the 32x16x16 cantilever mesh (100 steps, dt 0.01)
  against the oracle's restatement of the reference's sensitivity loops
  (LinearElasticity.cc:1625-2204): FP64 within 1e-8, mixed within 1e-5 relative.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.gpu_helpers import problem, rel
from topopt_in_petsc_transient_amd import capi

pytestmark = pytest.mark.gpu


class TransientObjective(torch.autograd.Function):
    """fx(xPhys) through tt_forward + tt_objective_gradient; backward returns dfdx."""

    @staticmethod
    def forward(ctx, x, solver):
        c = solver
        c.set_density(x.detach().cpu().numpy())
        c.complete_initial_state()
        c.forward()
        fx, df = c.objective_gradient()
        ctx.save_for_backward(torch.from_numpy(df).to(x.dtype))
        return torch.tensor(fx, dtype=x.dtype)

    @staticmethod
    def backward(ctx, g):
        (df,) = ctx.saved_tensors
        return g * df, None


def test_gradcheck_along_random_directions():
    """Central differences at h = 1e-6 carry round-off and solver noise of about
    |fx| * 1e-14 / h (solves at rtol 1e-13), so gradcheck runs with atol = 3e-8 |fx|
    and rtol = 1e-5; a second, tighter check at h = 1e-4 (truncation ~1e-8 relative,
    noise ~1e-10 |fx|) requires 1e-6 relative agreement."""
    _, pg = problem(nelx=16, nely=8, nelz=8, nsteps=30, dt=0.01, pulse_steps=10, rtol=1e-13)
    c = capi.Context(pg)
    n = pg.nelx * pg.nely * pg.nelz
    rng = np.random.default_rng(2024)
    x0 = torch.from_numpy(rng.uniform(0.2, 0.9, n))
    fx0 = float(TransientObjective.apply(x0, c))
    for k in range(20):
        d = torch.from_numpy(rng.standard_normal(n))
        d /= torch.linalg.norm(d)
        t = torch.zeros(1, dtype=torch.float64, requires_grad=True)
        f = lambda s: TransientObjective.apply(x0 + s * d, c)  # noqa: E731
        assert torch.autograd.gradcheck(f, (t,), eps=1e-6, atol=3e-8 * abs(fx0), rtol=1e-5, nondet_tol=0.0), k
        y = f(t)
        (g,) = torch.autograd.grad(y, t)
        h = 1e-4
        with torch.no_grad():
            fd = (float(f(t + h)) - float(f(t - h))) / (2 * h)
        assert abs(float(g) - fd) <= 1e-6 * abs(fd) + 1e-10 * abs(fx0), (k, float(g), fd)


@pytest.mark.parametrize("precision,tol", [(capi.TT_PREC_FP64, 1e-8), (capi.TT_PREC_MIXED, 1e-5)])
def test_sensitivities_on_the_cantilever_mesh(precision, tol):
    po, pg = problem(nelx=32, nely=16, nelz=16, nsteps=100, dt=0.01, pulse_steps=20, nlvls=3,
                     precision=precision, rtol=1e-12, slots=10)
    x = np.full(O.nel(po), 0.12)
    x += np.random.default_rng(12).uniform(-0.05, 0.05, x.size)
    st = O.forward(po, x)
    fx_ref, df_ref = O.objective_gradient(po, x, st)
    c = capi.Context(pg)
    c.set_density(x)
    c.complete_initial_state()
    c.forward()
    fx, df = c.objective_gradient()
    assert abs(fx - fx_ref) <= tol * abs(fx_ref)
    assert rel(df, df_ref) < tol
