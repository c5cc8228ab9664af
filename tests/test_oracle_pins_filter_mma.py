"""Oracle pins of the PDE-filter radius and of the MMA update step (CPU).

Each check is chosen so that a plausible restatement mistake fails it, and each
test also asserts that a wrong variant (a radius off by sqrt(3), swapped or
other asymptote factors) fails the same check.

* PDE filter (PDEFilter.cc:32, R = rmin / (2 sqrt 3); PDEFilter.cc:193-216): the
  filter solves -R^2 lap u + u = x with natural boundary conditions, so a
  cosine mode cos(k x) with k = m pi / Lx (zero slope at both x faces) is an
  eigenfunction damped by 1 / (1 + R^2 k^2) (Lazarov & Sigmund 2011).  On an
  x-elongated mesh with k h <= 0.3 the discrete filter matches that to 2 %.
* MMA (MMA.cc:510-934): after each update x must be the minimiser of the convex
  separable subproblem that Svanberg's method defines at the current point.  With
  m = 1 its dual is one-dimensional and monotone, so the test solves it by
  bisection, independently of the oracle's primal-dual interior-point solver.
  A 2-variable, 3-iteration case checks the moving asymptotes and the move
  limits against literals computed by hand from the published asymptote rule
  and the clamps of MMA.cc:545-560.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- PDE filter radius
def _elongated(nelx=128, nyz=2, lx=2.0):
    h = lx / nelx
    p = O.cantilever_params(nelx, nyz, nyz, nsteps=1, lx=lx, ly=nyz * h, lz=nyz * h)
    xc = (np.arange(nelx) + 0.5) * h
    return p, xc, h


def _cosine_damping_error(p, xc, nyz, rmin_used, R_model, k):
    x = np.tile(np.cos(k * xc), nyz * nyz)
    out = O.pde_filter(p, rmin_used, x)
    expect = x / (1.0 + (R_model * k) ** 2)
    return np.linalg.norm(out - expect) / np.linalg.norm(expect)


@pytest.mark.parametrize("R", [0.05, 0.1])
@pytest.mark.parametrize("m", [4, 8, 12])
def test_pde_filter_damps_cosine_modes_by_the_helmholtz_factor(R, m):
    """R = rmin / (2 sqrt 3) (PDEFilter.cc:32): mode k is scaled by 1 / (1 + R^2 k^2)."""
    lx, nyz = 2.0, 2
    p, xc, h = _elongated(128, nyz, lx)
    k = m * math.pi / lx
    assert k * h <= 0.3
    rmin = R * 2.0 * math.sqrt(3.0)
    assert _cosine_damping_error(p, xc, nyz, rmin, R, k) < 0.02


@pytest.mark.parametrize("wrong", [math.sqrt(3.0), 2.0 * math.sqrt(3.0), 0.5])
def test_pde_filter_radius_variants_fail_the_damping_check(wrong):
    """A radius conversion off by sqrt(3) (R = rmin / 2), by 2 sqrt(3) (R = rmin) or by 1/2 fails."""
    lx, nyz = 2.0, 2
    p, xc, h = _elongated(128, nyz, lx)
    R = 0.1
    k = 8 * math.pi / lx
    # the check as above, with the filter run at a radius that differs by the factor `wrong`
    assert _cosine_damping_error(p, xc, nyz, R * 2.0 * math.sqrt(3.0) * wrong, R, k) > 0.2


# ---------------------------------------------------------------- MMA: the subproblem by dual bisection
FEPS = 1.0e-6  # MMA.cc:581


def _subproblem(x, df, g, dg, xmin, xmax, L, U):
    """Svanberg's convex separable approximation at x (m = 1, the reference's regularisation:
    0.001 |df| + feps / (2 (U - L)) added to both terms of the objective, none to the constraint)."""
    alpha = np.maximum(xmin, 0.9 * L + 0.1 * x)
    beta = np.minimum(xmax, 0.9 * U + 0.1 * x)
    reg = 0.001 * np.abs(df) + 0.5 * FEPS / (U - L)
    p0 = (U - x) ** 2 * (np.maximum(df, 0.0) + reg)
    q0 = (x - L) ** 2 * (np.maximum(-df, 0.0) + reg)
    p1 = (U - x) ** 2 * np.maximum(dg, 0.0)
    q1 = (x - L) ** 2 * np.maximum(-dg, 0.0)
    b = np.sum(p1 / (U - x) + q1 / (x - L)) - g
    return alpha, beta, p0, q0, p1, q1, b


def _x_of_lambda(lam, L, U, alpha, beta, p0, q0, p1, q1):
    sp, sq = np.sqrt(p0 + lam * p1), np.sqrt(q0 + lam * q1)
    return np.clip((sp * L + sq * U) / (sp + sq), alpha, beta)


def _dual_bisection(L, U, alpha, beta, p0, q0, p1, q1, b, c=1000.0):
    """lambda* of the m = 1 dual (a = 0, d = 0, so z = 0 and y = max(0, lambda - c)): the dual
    gradient h(lambda) = sum p1/(U-x) + q1/(x-L) - b - y is decreasing in lambda; lambda* = 0 if
    h(0) <= 0, else its root."""
    def h(lam):
        xl = _x_of_lambda(lam, L, U, alpha, beta, p0, q0, p1, q1)
        return np.sum(p1 / (U - xl) + q1 / (xl - L)) - b - max(0.0, lam - c)

    if h(0.0) <= 0.0:
        return 0.0
    lo, hi = 0.0, 1.0
    while h(hi) > 0.0:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if h(mid) > 0.0 else (lo, mid)
    return 0.5 * (lo + hi)


def _compliance_like(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.05, 1.5, n)
    return t


@pytest.mark.parametrize("asym", [(0.5, 0.7, 1.2), (0.1, 0.6, 1.05)])
def test_mma_update_solves_the_svanberg_subproblem(asym):
    """Each update's x equals the subproblem minimiser from the dual bisection (to 1e-7 of the
    box), for 8 iterations of a compliance-like problem with an active volume constraint."""
    n, V = 60, 0.3
    t = _compliance_like(n, 5)
    x = np.full(n, 0.5)
    mma = O.MMA(n, 1, x)
    mma.set_asymptotes(*asym)
    n_active = 0
    for it in range(8):
        df = -t / (0.1 + x) ** 2 / n  # a decreasing, convex objective term per variable
        g = np.array([x.mean() - V])
        dg = np.full(n, 1.0 / n)
        xmin, xmax = O.MMA.outer_movelimit(0.0, 1.0, 0.2, x)
        xk = x.copy()
        x = mma.update(x, df, g, dg[None, :], xmin, xmax)
        L, U, lam = mma.state()
        alpha, beta, p0, q0, p1, q1, b = _subproblem(xk, df, g[0], dg, xmin, xmax, L, U)
        lstar = _dual_bisection(L, U, alpha, beta, p0, q0, p1, q1, b)
        xs = _x_of_lambda(lstar, L, U, alpha, beta, p0, q0, p1, q1)
        assert np.abs(x - xs).max() < 1e-7, f"iteration {it}"
        if lstar > 0:
            n_active += 1
            assert lam[0] == pytest.approx(lstar, rel=1e-6)
        # asymptotes enclose the point and the box (MMA.cc:545-560, 583-584)
        assert np.all(L < xk) and np.all(U > xk)
        assert np.all(x >= alpha - 1e-15) and np.all(x <= beta + 1e-15)
    assert n_active >= 4  # the constraint was active in most iterations


def test_mma_subproblem_check_rejects_a_wrong_regularisation():
    """The bisection check is sharp: the same check with the 0.001 |df| term at 0.01 fails."""
    n, V = 60, 0.3
    t = _compliance_like(n, 5)
    x = np.full(n, 0.5)
    mma = O.MMA(n, 1, x)
    mma.set_asymptotes(0.5, 0.7, 1.2)
    df = -t / (0.1 + x) ** 2 / n
    g = np.array([x.mean() - V + 0.1])  # active from the first step
    dg = np.full(n, 1.0 / n)
    xmin, xmax = O.MMA.outer_movelimit(0.0, 1.0, 0.2, x)
    xn = mma.update(x, df, g, dg[None, :], xmin, xmax)
    L, U, _ = mma.state()
    alpha = np.maximum(xmin, 0.9 * L + 0.1 * x)
    beta = np.minimum(xmax, 0.9 * U + 0.1 * x)
    reg = 0.01 * np.abs(df) + 0.5 * FEPS / (U - L)
    p0 = (U - x) ** 2 * (np.maximum(df, 0.0) + reg)
    q0 = (x - L) ** 2 * (np.maximum(-df, 0.0) + reg)
    p1 = (U - x) ** 2 * dg
    q1 = np.zeros(n)
    b = np.sum(p1 / (U - x)) - g[0]
    lw = _dual_bisection(L, U, alpha, beta, p0, q0, p1, q1, b)
    assert np.abs(xn - _x_of_lambda(lw, L, U, alpha, beta, p0, q0, p1, q1)).max() > 1e-5


# ---------------------------------------------------------------- MMA: asymptotes and move limits by hand
# Two variables, main.cc's asymptote factors (init 0.1, decrease 0.6, increase 1.05, main.cc:58),
# an inactive constraint (g = -1, dg = 0.5), and objective gradients of magnitude 1, so that each new
# x sits on the move limit alpha = max(xmin, 0.9 L + 0.1 x) or beta = min(xmax, 0.9 U + 0.1 x)
# (MMA.cc:583-584; the unconstrained minimiser lies 0.0308 (U - L) inside the asymptote, beyond it).
#
# Iteration 1 (k = 1 < 3: L = x - 0.1 (xmax - xmin), U = x + 0.1 (xmax - xmin), MMA.cc:516-520):
#   x = (0.5, 0.5), xmin = (0.3, 0.3), xmax = (0.7, 0.7), df = (+1, -1)
#   L = (0.46, 0.46), U = (0.54, 0.54); alpha_1 = max(0.3, 0.414 + 0.05) = 0.464,
#   beta_2 = min(0.7, 0.486 + 0.05) = 0.536  ->  x = (0.464, 0.536)
# Iteration 2 (k = 2): xmin = x - 0.2, xmax = x + 0.2, df = (+1, +1)
#   L = x - 0.04 = (0.424, 0.496), U = x + 0.04 = (0.504, 0.576)
#   alpha = (max(0.264, 0.3816 + 0.0464), max(0.336, 0.4464 + 0.0536)) = (0.428, 0.5)  ->  x = (0.428, 0.5)
# Iteration 3 (k = 3, MMA.cc:523-560): df = (+1, +1); variable 1 gets the tight box [0.427, 0.429]
# (xmi = 0.002), variable 2 the box [0.3, 0.7] (xmi = 0.4)
#   var 1: (x - xo1)(xo1 - xo2) = (-0.036)(-0.036) > 0 -> gamma = 1.05:
#          L = 0.428 - 1.05 (0.464 - 0.424) = 0.386 -> clamped to x - 10 xmi = 0.408
#          U = 0.428 + 1.05 (0.504 - 0.464) = 0.470 -> clamped to x + 10 xmi = 0.448
#          alpha = max(0.427, 0.3672 + 0.0428) = 0.427 -> x = 0.427
#   var 2: (0.5 - 0.536)(0.536 - 0.5) < 0 -> gamma = 0.6:
#          L = 0.5 - 0.6 (0.536 - 0.496) = 0.476, U = 0.5 + 0.6 (0.576 - 0.536) = 0.524 (no clamp)
#          alpha = max(0.3, 0.4284 + 0.05) = 0.4784 -> x = 0.4784
HAND = [
    # (L, U, x after the update)
    ((0.46, 0.46), (0.54, 0.54), (0.464, 0.536)),
    ((0.424, 0.496), (0.504, 0.576), (0.428, 0.5)),
    ((0.408, 0.476), (0.448, 0.524), (0.427, 0.4784)),
]


def _hand_case(asym):
    x = np.array([0.5, 0.5])
    mma = O.MMA(2, 1, x)
    mma.set_asymptotes(*asym)
    dfs = [np.array([1.0, -1.0]), np.array([1.0, 1.0]), np.array([1.0, 1.0])]
    out = []
    for it in range(3):
        if it < 2:
            xmin, xmax = O.MMA.outer_movelimit(0.0, 1.0, 0.2, x)
        else:
            xmin, xmax = np.array([0.427, 0.3]), np.array([0.429, 0.7])
        x = mma.update(x, dfs[it], np.array([-1.0]), np.full((1, 2), 0.5), xmin, xmax)
        L, U, _ = mma.state()
        out.append((L.copy(), U.copy(), x.copy()))
    return out


def _matches_hand(out, tol=1e-12):
    for (L, U, x), (Lh, Uh, xh) in zip(out, HAND):
        if not (np.allclose(L, Lh, rtol=0, atol=tol) and np.allclose(U, Uh, rtol=0, atol=tol)
                and np.allclose(x, xh, rtol=0, atol=1e-9)):
            return False
    return True


def test_mma_asymptotes_and_move_limits_match_hand_computed_values():
    out = _hand_case((0.1, 0.6, 1.05))
    for it, ((L, U, x), (Lh, Uh, xh)) in enumerate(zip(out, HAND)):
        assert L == pytest.approx(Lh, abs=1e-12), f"L, iteration {it + 1}"
        assert U == pytest.approx(Uh, abs=1e-12), f"U, iteration {it + 1}"
        assert x == pytest.approx(xh, abs=1e-9), f"x, iteration {it + 1}"


@pytest.mark.parametrize("asym", [(0.1, 1.05, 0.6), (0.2, 0.5, 1.5), (0.9, 0.9, 1.01)])
def test_mma_hand_case_rejects_other_asymptote_factors(asym):
    """Swapped decrease/increase factors, or other factor sets, give other asymptotes or iterates."""
    assert not _matches_hand(_hand_case(asym))
