"""HIP path against the FP64 CPU oracle, through the C-ABI (small grids).

Bars: FP64 path within 1e-8 relative of the oracle, mixed-precision path
within 1e-5 relative; integer indexing (BC mask, load location) exact.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import problem, random_density, rel
from topopt_in_petsc_transient_amd import capi

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-8
MIXED_TOL = 1e-5


def _ctx(pg, x):
    c = capi.Context(pg)
    c.set_density(x)
    return c


@pytest.mark.parametrize("which,cK,cM", [(capi.TT_OP_K, 1.0, 0.0), (capi.TT_OP_M, 0.0, 1.0), (capi.TT_OP_A, None, None)])
def test_operator_apply_fp64_and_fp32(which, cK, cM):
    po, pg = problem(nelx=16, nely=8, nelz=8)
    x = random_density(O.nel(po), seed=1)
    c = _ctx(pg, x)
    if cK is None:
        a = O.newmark_consts(1, po.t1 / po.nsteps)
        cK, cM = 1.0 + a[2] * po.dbeta, a[5] + a[2] * po.dalpha
    N = O.bc_mask(po)
    u = np.random.default_rng(2).standard_normal(c.nd) * N
    ref = O.apply_bc(po, x, cK, cM, u, N)
    y = np.zeros(c.nd)
    c.op_apply(which, u, y, prec=0)
    assert rel(y, ref) < 1e-13
    y32 = np.zeros(c.nd, dtype=np.float32)
    c.op_apply(which, u.astype(np.float32), y32, prec=1)
    assert rel(y32, ref) < 1e-5


def test_forward_states_fp64():
    po, pg = problem()
    x = random_density(O.nel(po), seed=3)
    states = O.forward(po, x)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    for t in (0, 1, 5, 10, 29, 30):
        u, v, a = c.get_state(t)
        for q, arr in enumerate((u, v, a)):
            if np.linalg.norm(states[t, q]) == 0:
                assert np.linalg.norm(arr) == 0
            else:
                assert rel(arr, states[t, q]) < FP64_TOL, (t, q)


@pytest.mark.parametrize("precision", [capi.TT_PREC_FP64, capi.TT_PREC_MIXED])
def test_newmark_solves_meet_rtol_on_the_true_residual(precision):
    """At rtol >= 1e-9 the solver stops on the recursive residual without recomputing b - A x; the true
    residual of every step (the equation-of-motion residual M a + C v + K u - f, which equals
    A u_t - b_t for the Newmark corrector) must still meet rtol ||b_t||."""
    rtol = 1e-8
    po, pg = problem(nsteps=20, rtol=rtol, precision=precision)
    x = random_density(O.nel(po), seed=12)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    K = O.assemble(po, x, 1.0, 0.0)
    M = O.assemble(po, x, 0.0, 1.0)
    C = po.dalpha * M + po.dbeta * K
    a = O.newmark_consts(1, po.t1 / po.nsteps)
    A = (1.0 + a[2] * po.dbeta) * K + (a[5] + a[2] * po.dalpha) * M
    N = O.bc_mask(po)
    F = O.load_location(po)
    for t in range(1, po.nsteps + 1):
        u, v, acc = c.get_state(t)
        r = (M @ acc + C @ v + K @ u - F * O.load_intensity(po, t)) * N
        b = (A @ u) * N - r
        assert np.linalg.norm(r) <= rtol * np.linalg.norm(b) * (1.0 + 1e-6), t


@pytest.mark.parametrize("precision", [capi.TT_PREC_FP64, capi.TT_PREC_MIXED])
def test_graph_replay_matches_eager_bitwise(precision):
    """Krylov iterations replayed as captured HIP graphs (graphs = 1) give bit-identical states,
    objective and gradient to eager launches, over many replays with eager work and host
    synchronisation between them (forward and adjoint sweeps; the revolve test covers replays
    inside recomputation)."""
    out = []
    for graphs in (0, 1):
        _, pg = problem(nsteps=30, slots=0, precision=precision, rtol=1e-10)
        pg.graphs = graphs
        x = random_density(pg.nelx * pg.nely * pg.nelz, seed=13)
        c = _ctx(pg, x)
        c.complete_initial_state()
        c.forward()
        states = [c.get_state(t) for t in (1, 7, 30)]
        fx, df = c.objective_gradient()
        out.append((states, fx, df))
        c.close()
    (s0, f0, d0), (s1, f1, d1) = out
    for a, b in zip(s0, s1):
        for q in range(3):
            assert np.array_equal(a[q], b[q])
    assert f0 == f1
    assert np.array_equal(d0, d1)


def test_objective_gradient_fp64():
    po, pg = problem()
    x = random_density(O.nel(po), seed=4)
    states = O.forward(po, x)
    fx_ref, df_ref = O.objective_gradient(po, x, states)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    fx, df = c.objective_gradient()
    assert abs(fx - fx_ref) / abs(fx_ref) < FP64_TOL
    assert rel(df, df_ref) < FP64_TOL


@pytest.mark.parametrize("objective", [0, 1, 2, 4, 5, 6])
def test_other_objectives_fp64(objective):
    po, pg = problem(nelx=8, nely=4, nelz=4, nsteps=12, dt=0.02, pulse_steps=5, objective=objective, nlvls=2)
    x = random_density(O.nel(po), seed=5)
    states = O.forward(po, x)
    fx_ref, df_ref = O.objective_gradient(po, x, states)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    fx, df = c.objective_gradient()
    assert abs(fx - fx_ref) / abs(fx_ref) < FP64_TOL
    assert rel(df, df_ref) < FP64_TOL


# Newmark steps of forward sweep + gradient for 30 steps: the 30 steps of the sweep plus the binomial
# recomputation (the adjoint step at t reuses state t from the step at t + 1, so it adds none).
# 29 slots hold every split point of the first descent: nothing is recomputed.
REVOLVE_FORWARD_STEPS = {2: 127, 4: 70, 29: 30}


@pytest.mark.parametrize("slots", sorted(REVOLVE_FORWARD_STEPS))
def test_revolve_matches_store_all_bitwise(slots):
    _, pg = problem(nsteps=30)
    x = random_density(pg.nelx * pg.nely * pg.nelz, seed=6)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    fx0, df0 = c.objective_gradient()
    _, pg2 = problem(nsteps=30, slots=slots)
    c2 = _ctx(pg2, x)
    c2.complete_initial_state()
    c2.forward()
    fx1, df1 = c2.objective_gradient()
    s = c2.stats()
    assert fx1 == fx0
    assert np.array_equal(df1, df0)
    assert s.adjoint_steps == 30 and s.forward_steps == REVOLVE_FORWARD_STEPS[slots]
    # a second gradient after the same forward sweep starts again from the kept state T
    fx2, df2 = c2.objective_gradient()
    assert fx2 == fx0
    assert np.array_equal(df2, df0)


def test_mixed_precision_against_oracle():
    po, pg = problem(precision=capi.TT_PREC_MIXED, rtol=1e-10)
    x = random_density(O.nel(po), seed=7)
    states = O.forward(po, x)
    fx_ref, df_ref = O.objective_gradient(po, x, states)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    u, _, _ = c.get_state(30)
    assert rel(u, states[30, 0]) < MIXED_TOL
    fx, df = c.objective_gradient()
    assert abs(fx - fx_ref) / abs(fx_ref) < MIXED_TOL
    assert rel(df, df_ref) < MIXED_TOL


_SCHED_REF = {}


def _sched_reference():
    """Oracle states, objective and gradient of the 32x16x16, 12-step problem (computed once)."""
    if not _SCHED_REF:
        po, _ = problem(nelx=32, nely=16, nelz=16, nsteps=12, nlvls=3)
        x = random_density(O.nel(po), seed=21)
        states = O.forward(po, x)
        fx, df = O.objective_gradient(po, x, states)
        _SCHED_REF.update(x=x, states=states, fx=fx, df=df)
    return _SCHED_REF


@pytest.mark.parametrize("pre,post", [(1, 0), (1, 1), (2, 2), (3, 1)])
@pytest.mark.parametrize("precision", [capi.TT_PREC_FP64, capi.TT_PREC_MIXED])
def test_intermediate_level_smoothing_schedules(precision, pre, post):
    """Three multigrid levels, so one intermediate level: each smoothing schedule there (mg_coarse_pre steps
    before the coarse-grid correction, mg_coarse_post after; (1, 0) is the default, a single damped-Jacobi
    step formed together with the residual) converges to the oracle's states, objective and gradient."""
    ref = _sched_reference()
    _, pg = problem(nelx=32, nely=16, nelz=16, nsteps=12, nlvls=3, precision=precision,
                    rtol=1e-12 if precision == capi.TT_PREC_FP64 else 1e-10)
    pg.mg_coarse_pre, pg.mg_coarse_post = pre, post
    tol = FP64_TOL if precision == capi.TT_PREC_FP64 else MIXED_TOL
    c = _ctx(pg, ref["x"])
    c.complete_initial_state()
    c.forward()
    for t in (1, 6, 12):
        u, v, a = c.get_state(t)
        assert rel(u, ref["states"][t, 0]) < tol and rel(a, ref["states"][t, 2]) < tol, t
    fx, df = c.objective_gradient()
    assert abs(fx - ref["fx"]) / abs(ref["fx"]) < tol
    assert rel(df, ref["df"]) < tol
    assert c.stats().max_relres <= pg.rtol * (1.0 + 1e-6)


def test_intermediate_level_schedule_bounds_are_checked():
    _, pg = problem(nelx=16, nely=8, nelz=8, nsteps=2, nlvls=3)
    pg.mg_coarse_post = 17
    with pytest.raises(capi.TTError):
        capi.Context(pg)


def test_pde_filter_and_projection():
    po, pg = problem()
    x = random_density(O.nel(po), seed=8, lo=0.0, hi=1.0)
    c = capi.Context(pg)
    ref = O.pde_filter(po, 0.08, x)
    out = c.pde_filter(x)
    assert rel(out, ref) < 1e-10
    for proj in (False, True):
        xt_r, xp_r = O.filter_project(po, 0.08, x, projection=proj, beta=2.0, eta=0.3)
        xt, xp = c.filter_project(x, projection=proj, beta=2.0, eta=0.3)
        assert rel(xt, xt_r) < 1e-10 and rel(xp, xp_r) < 1e-10
        df = np.random.default_rng(9).standard_normal(x.size)
        dg = np.full(x.size, 1.0 / x.size)
        a_r, b_r = O.filter_gradients(po, 0.08, xt_r, df, dg, projection=proj, beta=2.0, eta=0.3)
        a, b = c.filter_gradients(xt, df, dg, projection=proj, beta=2.0, eta=0.3)
        assert rel(a, a_r) < 1e-10 and rel(b, b_r) < 1e-10
    assert abs(c.mnd(x) - np.mean(4 * x * (1 - x))) < 1e-14


def test_pde_filter_large_radius_device_loop():
    """The filter's PCG runs with its scalars on the device and few host looks: at the job scripts' ratio
    R/h ~ 3.3 (rmin 0.714 on h = 1/16, hundreds of iterations) it matches the oracle; a second call (whose
    first look follows the first call's count) gives the same answer, and a zero input gives zero."""
    rmin = 0.714
    po, pg = problem(nelx=32, nely=16, nelz=16, rmin=rmin)
    x = random_density(O.nel(po), seed=21, lo=0.0, hi=1.0)
    c = capi.Context(pg)
    ref = O.pde_filter(po, rmin, x)
    out = c.pde_filter(x)
    assert rel(out, ref) < 1e-10
    assert np.array_equal(c.pde_filter(x), out)
    assert not c.pde_filter(np.zeros_like(x)).any()
    assert rel(c.pde_filter(x), ref) < 1e-10


def test_mma_update_sequence_matches_oracle():
    po, pg = problem(nelx=8, nely=4, nelz=4)
    n = O.nel(po)
    c = capi.Context(pg)
    rng = np.random.default_rng(10)
    x = np.full(n, 0.3)
    m_ref = O.MMA(n, 1, x)
    m_ref.set_asymptotes(0.1, 0.6, 1.05)
    m = capi.MMA(c)
    m.set_asymptotes(0.1, 0.6, 1.05)
    xr = x.copy()
    xg = x.copy()
    xold_r = x.copy()
    xold_g = x.copy()
    for it in range(5):
        df = rng.standard_normal(n) - 0.5
        dg = np.full(n, 1.0 / n)
        gx = np.mean(xr) - 0.25
        lo_r, hi_r = O.MMA.outer_movelimit(0.0, 1.0, 0.2, xr)
        lo, hi = m.outer_movelimit(0.0, 1.0, 0.2, xg)
        lo_s, hi_s = O.MMA.outer_movelimit(0.0, 1.0, 0.2, xg)  # same input: element-wise ops are exact
        assert np.array_equal(lo, lo_s) and np.array_equal(hi, hi_s)
        xr = m_ref.update(xr, df, gx, dg, lo_r, hi_r)
        xg = m.update(xg, df, gx, dg, lo, hi)
        assert rel(xg, xr) < 1e-10, it
        ch_r, xold_r = O.MMA.design_change(xr, xold_r)
        ch, xold_g = m.design_change(xg, xold_g)
        assert abs(ch - ch_r) < 1e-10
        k_r = m_ref.kkt_residual(xr, df, gx, dg, lo_r, hi_r)
        k = m.kkt_residual(xg, df, gx, dg, lo, hi)
        assert abs(k[0] - k_r[0]) <= 1e-8 * max(1.0, abs(k_r[0]))


def test_nonconvergence_is_an_error():
    po, pg = problem(nelx=8, nely=4, nelz=4, nsteps=3, maxit=1, rtol=1e-14)
    x = random_density(O.nel(po), seed=11)
    c = _ctx(pg, x)
    with pytest.raises(capi.TTError) as e:
        c.complete_initial_state()
        c.forward()
    assert e.value.code == capi.TT_ERR_NOT_CONVERGED


def test_nonfinite_density_is_an_error():
    _, pg = problem(nelx=8, nely=4, nelz=4, nsteps=3)
    c = capi.Context(pg)
    x = np.full(pg.nelx * pg.nely * pg.nelz, 0.5)
    x[3] = np.nan
    with pytest.raises(capi.TTError) as e:
        c.set_density(x)
    assert e.value.code == capi.TT_ERR_NONFINITE


# The reference's other presets (LinearElasticity.cc:222-335 regions, :337-591 boundaries, :593-675
# loads, :185-219 time integration, :1580-1623 interpolations) on a 12 x 4 x 4 box of 12 x 4 x 8
# elements, over a time window that contains the Gaussian-cosine signal's peak (t0 = 8).
@pytest.mark.parametrize("boundaries,loadloc,loadsignal,ti,sInt,mInt,region,objective", [
    (0, 0, 0, 1, 1, 0, 0, 3),
    (1, 1, 0, 1, 1, 0, 1, 3),
    (1, 2, 1, 0, 2, 1, 3, 3),
    (0, 1, 1, 1, 1, 1, 2, 6),
])
def test_reference_presets_fp64(boundaries, loadloc, loadsignal, ti, sInt, mInt, region, objective):
    kw = dict(boundaries=boundaries, loadloc=loadloc, loadsignal=loadsignal, ti=ti, sInt=sInt, mInt=mInt,
              region=region)
    po, pg = problem(nelx=12, nely=4, nelz=8, nsteps=24, dt=0.5, pulse_steps=6, objective=objective, nlvls=2,
                     lx=12.0, ly=4.0, lz=4.0, **kw)
    if sInt == 2:
        po.penalK = pg.penalK = 4.0
    x = random_density(O.nel(po), seed=17)
    assert np.array_equal(O.bc_mask(po) > 0, _gpu_free_mask(pg))
    states = O.forward(po, x)
    fx_ref, df_ref = O.objective_gradient(po, x, states)
    c = _ctx(pg, x)
    c.complete_initial_state()
    c.forward()
    for t in (1, 12, 24):
        u, v, a = c.get_state(t)
        assert rel(u, states[t, 0]) < FP64_TOL and rel(a, states[t, 2]) < FP64_TOL, t
    fx, df = c.objective_gradient()
    assert fx_ref != 0.0
    assert abs(fx - fx_ref) / abs(fx_ref) < FP64_TOL
    assert rel(df, df_ref) < FP64_TOL


def _gpu_free_mask(pg):
    """Free dofs seen by the GPU path: K applied to the unit vector field is the identity on clamped dofs."""
    c = capi.Context(pg)
    n = pg.nelx * pg.nely * pg.nelz
    c.set_density(np.full(n, 0.5))
    nd = c.nd
    e = np.zeros(nd)
    y = np.zeros(nd)
    free = np.ones(nd, dtype=bool)
    for comp in range(3):  # probe one component at a time: clamped rows return the input unchanged
        e[:] = 0.0
        e[comp::3] = 1e-3
        c.op_apply(capi.TT_OP_M, e, y, prec=0)
        clamped = np.isclose(y[comp::3], 1e-3, rtol=0, atol=1e-15)
        free[comp::3] = ~clamped
    c.close()
    return free
