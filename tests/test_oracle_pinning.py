"""Pins the FP64 CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, golden vectors or fixtures for this path, and its
arithmetic sits on PETSc, which is absent here, so the oracle cannot be compared
with a reference build.  Each check below is chosen so that a plausible mistake
in the restatement (a dropped term, a wrong sign, index or coefficient, a
transposed operand) fails it: closed forms, invariants, rigid-body and patch
tests, conservation laws, independent direct solves and finite differences.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import oracle as O

NU = 0.3


def _corner(a):
    """Reference coordinates (0/1) of hex8 corner a in the order of LinearElasticity.cc:3067-3074."""
    return ((1 if a in (1, 2, 5, 6) else 0), (1 if a in (2, 3, 6, 7) else 0), (1 if a >= 4 else 0))


def _node_xyz(dx, dy, dz):
    return np.array([[c[0] * dx, c[1] * dy, c[2] * dz] for c in map(_corner, range(8))])


def _lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


# ---------------------------------------------------------------- A1 index maps
def test_element_connectivity_matches_dmda_definition():
    p = O.cantilever_params(3, 2, 2, nsteps=1)
    nx, ny = 4, 3
    count = np.zeros(O.nn(p), dtype=int)
    for k, j, i in itertools.product(range(2), range(2), range(3)):
        e = i + 3 * (j + 2 * k)
        nodes = O.element_nodes(p, e)
        for a, n in enumerate(nodes):
            cx, cy, cz = _corner(a)
            assert n == (i + cx) + nx * ((j + cy) + ny * (k + cz))
            count[n] += 1
    # every node lies in between 1 and 8 elements; the single interior node column in 8
    assert count.min() == 1 and count.max() == 8
    assert count.sum() == 8 * O.nel(p)


# ---------------------------------------------------------------- A2 stiffness
@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0), (0.5, 0.25, 0.125)])
def test_ke_symmetry_rank_and_rigid_modes(dims):
    ke = O.hex8_ke(*dims, NU)
    assert np.allclose(ke, ke.T, atol=1e-14 * np.abs(ke).max())
    w = np.linalg.eigvalsh(ke)
    tol = 1e-10 * w.max()
    assert np.sum(np.abs(w) < tol) == 6 and w.min() > -tol  # PSD, rank 18
    xyz = _node_xyz(*dims)
    modes = []
    for c in range(3):  # translations
        u = np.zeros((8, 3))
        u[:, c] = 1.0
        modes.append(u.ravel())
    for axis in np.eye(3):  # rotations u = w x X
        modes.append(np.cross(axis, xyz).ravel())
    for u in modes:
        assert np.linalg.norm(ke @ u) < 1e-12 * np.abs(ke).max() * np.linalg.norm(u)


@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0), (0.5, 0.25, 0.125)])
def test_ke_patch_test_uniform_strain_energy(dims):
    """Trilinear hex8 reproduces linear fields exactly: u^T K u = V (lambda tr(eps)^2 + 2 mu eps:eps)."""
    ke = O.hex8_ke(*dims, NU)
    xyz = _node_xyz(*dims)
    V = dims[0] * dims[1] * dims[2]
    lam, mu = _lame(1.0, NU)
    rng = np.random.default_rng(1)
    for _ in range(5):
        G = rng.standard_normal((3, 3))  # displacement gradient (strain + rotation)
        eps = 0.5 * (G + G.T)
        u = (xyz @ G.T).ravel()
        energy = u @ ke @ u
        exact = V * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert energy == pytest.approx(exact, rel=1e-12)


def test_ke_scaling_and_diagonal_closed_form():
    k1 = O.hex8_ke(1.0, 1.0, 1.0, NU)
    k2 = O.hex8_ke(2.0, 2.0, 2.0, NU)
    assert np.allclose(k2, 2.0 * k1, rtol=1e-13, atol=1e-15)
    # K_(a,x),(a,x) = int (lam + 2 mu) N_a,x^2 + mu (N_a,y^2 + N_a,z^2) dV with
    # int N_a,x^2 dV = dy dz / (9 dx) etc. for the trilinear shape functions of a dx x dy x dz box
    lam, mu = _lame(1.0, NU)
    dx, dy, dz = 0.5, 0.25, 0.125
    ke = O.hex8_ke(dx, dy, dz, NU)
    ix, iy, iz = dy * dz / (9 * dx), dx * dz / (9 * dy), dx * dy / (9 * dz)
    expect = [(lam + 2 * mu) * ix + mu * (iy + iz), (lam + 2 * mu) * iy + mu * (ix + iz),
              (lam + 2 * mu) * iz + mu * (ix + iy)]
    assert np.allclose(np.diag(ke), np.tile(expect, 8), rtol=1e-13)


# ---------------------------------------------------------------- A3 mass
@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0), (0.5, 0.25, 0.125)])
def test_me_closed_form(dims):
    """Consistent trilinear mass: M_ab = V/216 * 2^(3 - #differing coordinates), times I3."""
    me = O.hex8_me(*dims)
    V = dims[0] * dims[1] * dims[2]
    for a, b in itertools.product(range(8), range(8)):
        hd = sum(x != y for x, y in zip(_corner(a), _corner(b)))
        blk = me[3 * a:3 * a + 3, 3 * b:3 * b + 3]
        assert np.allclose(blk, V / 216.0 * 2 ** (3 - hd) * np.eye(3), rtol=1e-13, atol=1e-16)
    u = np.tile([1.0, 0.0, 0.0], 8)
    assert u @ me @ u == pytest.approx(V, rel=1e-13)  # total mass, density 1


# ---------------------------------------------------------------- A22 PDE-filter element matrix
def test_pdefilter_element_matrix_properties():
    dims = (0.5, 0.25, 0.125)
    V = dims[0] * dims[1] * dims[2]
    m8 = O.hex8_me(*dims)[0::3, 0::3]
    assert np.allclose(O.pdefilter_ke(*dims, 0.0), m8, rtol=1e-13, atol=1e-16)  # R = 0: scalar mass
    R = 0.03
    kf = O.pdefilter_ke(*dims, R)
    assert np.allclose(kf, kf.T, atol=1e-16)
    lap = (kf - m8) / R ** 2
    assert np.allclose(lap.sum(axis=1), 0.0, atol=1e-12 * np.abs(lap).max())  # constants in the null space
    w = np.linalg.eigvalsh(lap)
    assert np.sum(np.abs(w) < 1e-10 * w.max()) == 1 and w.min() > -1e-10 * w.max()
    # Dirichlet energy of the linear field phi = x: int |grad phi|^2 = V
    phi = _node_xyz(*dims)[:, 0]
    assert phi @ lap @ phi == pytest.approx(V, rel=1e-12)


# ---------------------------------------------------------------- A5 interpolation
@pytest.mark.parametrize("scheme,penal", [(0, 1.0), (1, 3.0), (2, 4.0)])
def test_interpolation_limits_and_derivative(scheme, penal):
    vmin, vmax = 1e-3, 1.7
    assert O.interp_c(scheme, vmin, vmax, penal, 0.0)[0] == pytest.approx(vmin, rel=1e-15)
    assert O.interp_c(scheme, vmin, vmax, penal, 1.0)[0] == pytest.approx(vmax, rel=1e-15)
    for x in (0.2, 0.5, 0.9):
        h = 1e-6
        fd = (O.interp_c(scheme, vmin, vmax, penal, x + h)[0] - O.interp_c(scheme, vmin, vmax, penal, x - h)[0]) / (2 * h)
        assert O.interp_c(scheme, vmin, vmax, penal, x)[1] == pytest.approx(fd, rel=1e-8)
    if scheme == 1:  # SIMP closed form
        assert O.interp_c(1, vmin, vmax, 3.0, 0.5)[0] == pytest.approx(vmin + 0.125 * (vmax - vmin), rel=1e-15)


# ---------------------------------------------------------------- A8 time integration
def test_newmark_constants_reproduce_the_schemes():
    # the reference's own numbers for beta = 1/4, gamma = 1/2, dt = 0.01
    assert np.allclose(O.newmark_consts(1, 0.01), [-1.0, 0.0, 200.0, 400.0, 1.0, 40000.0], rtol=1e-14, atol=1e-14)
    rng = np.random.default_rng(3)
    dt = 0.037
    u0, v0, a0, u1 = rng.standard_normal(4)
    for ti in (0, 1):
        a = O.newmark_consts(ti, dt)
        v1 = a[0] * v0 + a[1] * a0 + a[2] * (u1 - u0)
        acc1 = -a[3] * v0 - a[4] * a0 + a[5] * (u1 - u0)
        if ti == 1:  # Newmark average acceleration, beta = 1/4, gamma = 1/2
            assert u1 == pytest.approx(u0 + dt * v0 + dt * dt * (0.25 * a0 + 0.25 * acc1), rel=1e-12)
            assert v1 == pytest.approx(v0 + dt * (0.5 * a0 + 0.5 * acc1), rel=1e-12)
        else:  # backward Euler
            assert v1 == pytest.approx((u1 - u0) / dt, rel=1e-12)
            assert acc1 == pytest.approx((v1 - v0) / dt, rel=1e-12)


# ---------------------------------------------------------------- A7 / A9 boundary and load
def test_clamp_and_line_load_locations():
    p = O.cantilever_params(4, 3, 2, nsteps=10, pulse_steps=4)
    nx, ny, nz = 5, 4, 3
    N = O.bc_mask(p).reshape(nz, ny, nx, 3)
    assert np.all(N[:, :, 0, :] == 0.0) and np.all(N[:, :, 1:, :] == 1.0)  # x = 0 face clamped
    F = O.load_location(p).reshape(nz, ny, nx, 3)
    expect = np.zeros_like(F)
    expect[0, :, nx - 1, 1] = 1.0  # y-dof on the x = xmax, z = 0 edge
    expect[0, 0, nx - 1, 1] = expect[0, ny - 1, nx - 1, 1] = 0.5  # half weight at the line ends
    assert np.array_equal(F, expect)
    for n in range(12):  # half-sine pulse, divided by nely
        s = np.sin(np.pi * n / 4) if n <= 4 else 0.0
        assert O.load_intensity(p, n) == pytest.approx(s / 3, abs=1e-15)


# ---------------------------------------------------------------- A6 / A11 operator and solve
def _small(nsteps=10, **kw):
    p = O.cantilever_params(4, 2, 2, nsteps=nsteps, pulse_steps=4, **kw)
    x = np.random.default_rng(5).uniform(0.2, 1.0, O.nel(p))
    return p, x


def test_matrix_free_apply_equals_assembled_matrix():
    p, x = _small()
    u = np.random.default_rng(6).standard_normal(3 * O.nn(p))
    for cK, cM in ((1.0, 0.0), (0.0, 1.0), (1.3, 4000.0)):
        A = O.assemble(p, x, cK, cM)
        assert np.allclose(O.apply(p, x, cK, cM, u), A @ u, rtol=1e-13, atol=1e-13 * np.abs(A @ u).max())


def test_pcg_matches_direct_solve():
    p, x = _small()
    N = O.bc_mask(p)
    cK, cM = 1.01, 4e4
    A = O.assemble(p, x, cK, cM).tolil()
    free = N > 0
    b = np.random.default_rng(8).standard_normal(len(N)) * N
    Af = A.tocsr()[free][:, free]
    xd = np.zeros_like(b)
    xd[free] = spla.spsolve(Af.tocsc(), b[free])
    xs, it, relres = O.pcg(p, x, cK, cM, b, rtol=1e-13)
    assert relres <= 1e-13
    assert np.linalg.norm(xs - xd) <= 1e-11 * np.linalg.norm(xd)


def test_undamped_newmark_conserves_energy_after_the_pulse():
    """Trapezoidal Newmark conserves 1/2 v^T M v + 1/2 u^T K u exactly when f = 0 and C = 0."""
    p, x = _small(nsteps=16, dalpha=0.0, dbeta=0.0, rtol=1e-14)
    st = O.forward(p, x)
    K = O.assemble(p, x, 1.0, 0.0)
    M = O.assemble(p, x, 0.0, 1.0)
    E = [0.5 * s[1] @ (M @ s[1]) + 0.5 * s[0] @ (K @ s[0]) for s in st]
    E = np.array(E)
    assert E[4] > 0
    assert np.allclose(E[5:], E[5], rtol=1e-9)  # load is zero from step 5 on
    assert not np.allclose(E[1:5], E[5], rtol=1e-3)  # while loaded it changes


def test_newmark_step_satisfies_the_equation_of_motion():
    """M a_t + C v_t + K u_t = N f_t on free dofs (C = alpha M + beta K) at every step."""
    p, x = _small(nsteps=8, rtol=1e-14)
    st = O.forward(p, x)
    K = O.assemble(p, x, 1.0, 0.0)
    M = O.assemble(p, x, 0.0, 1.0)
    C = p.dalpha * M + p.dbeta * K
    N = O.bc_mask(p)
    F = O.load_location(p)
    for t in range(1, p.nsteps + 1):
        u, v, a = st[t]
        r = (M @ a + C @ v + K @ u - F * O.load_intensity(p, t)) * N
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(K @ u)


def test_nonzero_initial_displacement_is_masked_and_starts_the_motion():
    """-solutionAsU0: u0 = N.u0_in, v0 = 0, M a0 = N.(f(0) - K u0) on free dofs, and every later step
    satisfies the equation of motion."""
    p, x = _small(nsteps=6, rtol=1e-14)
    nd = 3 * O.nn(p)
    u0 = np.random.default_rng(7).standard_normal(nd) * 1e-2
    st = O.forward(p, x, u0=u0)
    N = O.bc_mask(p)
    assert np.array_equal(st[0, 0], N * u0) and not st[0, 1].any()
    K = O.assemble(p, x, 1.0, 0.0)
    M = O.assemble(p, x, 0.0, 1.0)
    C = p.dalpha * M + p.dbeta * K
    F = O.load_location(p)
    for t in range(0, p.nsteps + 1):
        u, v, a = st[t]
        r = (M @ a + C @ v + K @ u - F * O.load_intensity(p, t)) * N
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(K @ u), t
    assert np.linalg.norm(st[3, 0] - O.forward(p, x)[3, 0]) > 1e-3 * np.linalg.norm(st[3, 0])


def test_adjoint_gradient_from_a_nonzero_initial_displacement():
    p, x = _small(nsteps=6, rtol=1e-14)
    u0 = np.random.default_rng(8).standard_normal(3 * O.nn(p)) * 1e-2
    fx, df = O.objective_gradient(p, x, O.forward(p, x, u0=u0))
    rng = np.random.default_rng(12)
    h = 1e-6
    for _ in range(2):
        d = rng.standard_normal(len(x))
        d /= np.linalg.norm(d)
        fp = O.objective(p, x + h * d, O.forward(p, x + h * d, u0=u0))
        fm = O.objective(p, x - h * d, O.forward(p, x - h * d, u0=u0))
        assert df @ d == pytest.approx((fp - fm) / (2 * h), rel=2e-6, abs=1e-9 * abs(fx))


# ---------------------------------------------------------------- A13-A18 adjoint sensitivities
@pytest.mark.parametrize("objective", [3, 0, 1, 2, 4, 5, 6])
def test_adjoint_gradient_matches_central_differences(objective):
    p, x = _small(nsteps=8, rtol=1e-14, objective=objective)
    st = O.forward(p, x)
    fx, df = O.objective_gradient(p, x, st)
    assert np.isfinite(fx) and fx != 0.0
    rng = np.random.default_rng(11)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(len(x))
        d /= np.linalg.norm(d)
        fp = O.objective(p, x + h * d, O.forward(p, x + h * d))
        fm = O.objective(p, x - h * d, O.forward(p, x - h * d))
        fd = (fp - fm) / (2 * h)
        assert df @ d == pytest.approx(fd, rel=2e-6, abs=1e-9 * abs(fx))


# ---------------------------------------------------------------- A22 / A23 filter and projection
def test_pde_filter_preserves_constants_and_is_symmetric():
    p = O.cantilever_params(6, 4, 3, nsteps=1)
    n = O.nel(p)
    assert np.allclose(O.pde_filter(p, 0.3, np.full(n, 0.37)), 0.37, rtol=1e-10)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    assert O.pde_filter(p, 0.3, a) @ b == pytest.approx(a @ O.pde_filter(p, 0.3, b), rel=1e-10)
    # smoothing: a checkerboard is damped
    ijk = np.indices((3, 4, 6)).sum(axis=0).ravel()
    cb = np.where(ijk % 2 == 0, 1.0, -1.0)
    assert np.linalg.norm(O.pde_filter(p, 0.3, cb)) < 0.5 * np.linalg.norm(cb)


def test_smooth_projection_closed_form_and_chain_rule():
    beta, eta = 8.0, 0.3
    den = np.tanh(beta * eta) + np.tanh(beta * (1 - eta))
    for x in (0.0, 0.1, 0.3, 0.77, 1.0):
        ref = (np.tanh(beta * eta) + np.tanh(beta * (x - eta))) / den
        assert O.smooth_projection(x, beta, eta) == pytest.approx(ref, rel=1e-14, abs=1e-15)
        h = 1e-6
        fd = (O.smooth_projection(x + h, beta, eta) - O.smooth_projection(x - h, beta, eta)) / (2 * h)
        assert O.chainrule_projection(x, beta, eta) == pytest.approx(fd, rel=1e-7)


# ---------------------------------------------------------------- A24 MMA
def test_mma_solves_a_separable_problem_with_a_volume_constraint():
    """min sum (x - t)^2 s.t. mean(x) <= V, 0 <= x <= 1: x* = clip(t - mu, 0, 1), mean(x*) = V."""
    n, V = 40, 0.3
    t = np.random.default_rng(4).uniform(0.0, 1.2, n)
    lo, hi = 0.0, 2.0  # bisection for mu
    for _ in range(200):
        mu = 0.5 * (lo + hi)
        lo, hi = (mu, hi) if np.clip(t - mu, 0, 1).mean() > V else (lo, mu)
    xstar = np.clip(t - mu, 0, 1)
    x = np.full(n, 0.5)
    mma = O.MMA(n, 1, x)
    for _ in range(60):
        df = 2 * (x - t)
        g = np.array([x.mean() - V])
        dg = np.full((1, n), 1.0 / n)
        xmin, xmax = O.MMA.outer_movelimit(0.0, 1.0, 0.2, x)
        x = mma.update(x, df, g, dg, xmin, xmax)
    assert np.abs(x - xstar).max() < 1e-4
    assert x.mean() <= V + 1e-6
