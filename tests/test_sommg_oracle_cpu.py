"""Pins the oracle's SOMMG reduced-order model (SOMMG.cc:40-265 and the reduced branches of
LinearElasticity.cc) against what the mathematics fixes (CPU only).

The reference ships no reduced-order tests or vectors.  Each check is chosen so that a plausible
mistake in the restatement fails it:

* the basis is recomputed with scipy's direct solver on the assembled matrices (the Arnoldi-type
  sequence Q_0 ~ P0^-1 N.F, Q_j ~ (I - Q Q^T) P0^-1 N.(2 om M Q_{j-1}), P0 = K - om^2 M);
* the reduced matrices are recomputed as Q A Q^T from the assembled matrices;
* the reduced states are the Galerkin solution of the full Newmark equations: the full-mesh residual
  of the expanded states is orthogonal to the basis, and the corrector holds;
* the reduced gradient is the derivative of the reduced objective with the basis held fixed (central
  differences), which is what the reference's reduced sensitivities compute;
* the reduced objective equals the full-order objective evaluated on the expanded states.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O

NB, OM = 5, 0.02


def _problem(objective=3):
    p = O.cantilever_params(6, 3, 3, nsteps=24, dt=0.05, pulse_steps=8, objective=objective)
    x = 0.3 + 0.7 * np.random.default_rng(5).random(O.nel(p))
    return p, x


def _assembled(p, x):
    K = O.assemble(p, x, 1.0, 0.0)
    M = O.assemble(p, x, 0.0, 1.0)
    N = O.bc_mask(p)
    return K, M, N


def test_basis_is_the_moment_matching_sequence():
    p, x = _problem()
    Q, Ut = O.sommg_basis(p, x, NB, OM)
    K, M, N = _assembled(p, x)
    Nd = sp.diags(N)
    P0 = (Nd @ (K - OM * OM * M) @ Nd + sp.diags(1.0 - N)).tocsc()
    f = O.load_location(p) * N
    ref = []
    for j in range(NB):
        if j > 0:
            f = N * (2.0 * OM * (M @ ref[j - 1]))
        v = spla.spsolve(P0, f)
        for q in ref:
            v = v - (q @ v) * q
        ref.append(v / np.linalg.norm(v))
    ref = np.array(ref)
    assert np.max(np.abs(Q - ref)) < 1e-8 * np.max(np.abs(ref))
    assert np.allclose(Q @ Q.T, np.eye(NB), atol=1e-12)
    assert np.all(Q[:, N == 0] == 0.0)
    assert np.all(np.diag(Ut) > 0) and np.allclose(np.tril(Ut, -1), 0.0)


def test_reduced_matrices_are_projections():
    p, x = _problem()
    Q, _ = O.sommg_basis(p, x, NB, OM)
    Kr, Mr, Cr, Pr = O.sommg_reduce(p, x, Q)
    K, M, N = _assembled(p, x)
    C = p.dalpha * M + p.dbeta * K
    for red, full in ((Kr, K), (Mr, M), (Cr, C)):
        ref = Q @ (full @ Q.T)
        assert np.max(np.abs(red - ref)) < 1e-12 * np.max(np.abs(ref))
    assert np.allclose(Pr, Q @ (O.load_location(p) * N), rtol=1e-13, atol=0)
    # symmetric positive definite reduced stiffness and mass
    assert np.all(np.linalg.eigvalsh(0.5 * (Kr + Kr.T)) > 0) and np.all(np.linalg.eigvalsh(0.5 * (Mr + Mr.T)) > 0)


def test_reduced_states_are_the_galerkin_solution():
    p, x = _problem()
    r = O.reduced_model(p, x, NB, OM)
    Q = r["Q"]
    w = O.reduced_expand(p, Q, r["rstates"])
    K, M, N = _assembled(p, x)
    C = p.dalpha * M + p.dbeta * K
    a = O.newmark_consts(p.ti, p.t1 / p.nsteps)
    F = O.load_location(p) * N
    # t = 0: M a0 = f(0) on the basis
    res0 = Q @ (M @ w[0, 2] - F * O.load_intensity(p, 0))
    assert np.linalg.norm(res0) <= 1e-12 * np.linalg.norm(Q @ F) * max(abs(O.load_intensity(p, 0)), 1e-30) + 1e-300
    S = K + a[2] * C + a[5] * M
    for t in range(1, p.nsteps + 1):
        u, v, acc = w[t]
        up, vp, ap = w[t - 1]
        rhs = (F * O.load_intensity(p, t) + (a[5] * M + a[2] * C) @ up + (a[3] * M - a[0] * C) @ vp
               + (a[4] * M - a[1] * C) @ ap)
        res = Q @ (S @ u - rhs)
        assert np.linalg.norm(res) <= 1e-10 * np.linalg.norm(Q @ rhs) + 1e-14, t
        du = u - up
        # the corrector's terms cancel: tolerance relative to the largest term
        terms_v = np.abs(a[0] * vp) + np.abs(a[1] * ap) + np.abs(a[2] * du)
        terms_a = np.abs(a[3] * vp) + np.abs(a[4] * ap) + np.abs(a[5] * du)
        assert np.max(np.abs(v - (a[0] * vp + a[1] * ap + a[2] * du))) <= 1e-12 * terms_v.max()
        assert np.max(np.abs(acc - (-a[3] * vp - a[4] * ap + a[5] * du))) <= 1e-12 * terms_a.max()
    assert np.linalg.norm(w[:, 0]) > 0


@pytest.mark.parametrize("objective", [3, 1, 6])
def test_reduced_gradient_is_the_fixed_basis_derivative(objective):
    p, x = _problem(objective)
    r = O.reduced_model(p, x, NB, OM)
    Q = r["Q"]

    def f(xx):
        Kr, Mr, Cr, Pr = O.sommg_reduce(p, xx, Q)
        rs = O.reduced_forward(p, Kr, Mr, Cr, Pr)
        return O.reduced_objective_gradient(p, xx, Q, Kr, Mr, Cr, rs)[0]

    assert abs(f(x) - r["fx"]) <= 1e-12 * abs(r["fx"])  # (threaded BLAS reductions: not bitwise, ~1e-13 apart)
    h = 1e-6
    for e in (0, 17, 40, O.nel(p) - 1):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (f(xp) - f(xm)) / (2 * h)
        assert abs(fd - r["dfdx"][e]) <= 1e-6 * np.max(np.abs(r["dfdx"])), (e, fd, r["dfdx"][e])


def test_reduced_objective_is_the_full_objective_of_the_expanded_states():
    p, x = _problem()
    r = O.reduced_model(p, x, NB, OM)
    w = O.reduced_expand(p, r["Q"], r["rstates"])
    assert abs(O.objective(p, x, w) - r["fx"]) <= 1e-13 * abs(r["fx"])


def test_more_basis_vectors_approach_the_full_model():
    """The reduced response converges towards the full-order one as the basis grows."""
    p, x = _problem()
    full = O.forward(p, x)
    errs = []
    for nb in (2, 4, 8):
        r = O.reduced_model(p, x, nb, OM)
        w = O.reduced_expand(p, r["Q"], r["rstates"])
        errs.append(np.linalg.norm(w[:, 0] - full[:, 0]) / np.linalg.norm(full[:, 0]))
    assert errs[0] > errs[1] > errs[2]
