// SOMMG reduced-order model (-reduction).
//
// Basis: SOMMG::ConstructBasis (SOMMG.cc:40-230).  Q_0 is the solution of P0 q = N.F with
// P0 = K - omega0^2 M on the free dofs; each later vector solves P0 q = N.(2 omega0 M Q_{j-1}) and is
// orthonormalised against the earlier ones by modified Gram-Schmidt (:206-218).  The reference adds a
// moment-matching correction to that right-hand side as fvec += (-1.0)*(-1,0)*massVec (:176); the
// comma expression (-1,0) is 0, so the correction never changes fvec and is not formed here.  The
// basis solves are the multigrid-preconditioned flexible CG of the full model with the operator
// scalars (1, -omega0^2), like the reference's FGMRES KSP with multigrid (:63-66; LinearElasticity.cc:2767);
// the multigrid here is that of the shifted operator K + omega0^2 M, which stays positive definite when
// omega0 lies above the design's first eigenfrequency and P0 is indefinite.
//
// Reduced model: Kr = Q^T K Q, Mr = Q^T M Q, Cr = Q^T C Q, Pr = Q^T N.F (:236-262), then the same
// Newmark and adjoint recurrences as the full model on these nb x nb matrices, solved by LU
// (LinearElasticity.cc:977-1019, SetUpSolver2 :3007-3014), with the state gradients of the objective
// projected onto the basis (:2318-2335).  The objective and the sensitivities use the states and
// multipliers expanded back to the full mesh (:1089-1107, 1683-1764, 1850-1863), so the element loops
// are the full model's kernels; the basis is held fixed (its dependence on the design is not
// differentiated, as in the reference).
//
// The device work per design is nb multigrid solves, nb^2 / 2 Gram-Schmidt dot products and 3 nb
// applies with nb^2 dot products for the reduced matrices; per time step of the backward sweep it is
// the expansion of two states and one multiplier (nb-term linear combinations), the objective's state
// gradient, its projection (nb three-column dot products) and the sensitivity kernel.  The reduced
// recurrences themselves are a few nb x nb products per step on the host.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "tt_context.h"

namespace tt {

// x = A^-1 b, Gaussian elimination with partial pivoting (n x n row-major)
static void dense_solve(int n, const std::vector<double>& A, const double* b, double* x) {
    std::vector<double> w((size_t)n * (n + 1));
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) w[(size_t)i * (n + 1) + j] = A[(size_t)i * n + j];
        w[(size_t)i * (n + 1) + n] = b[i];
    }
    auto W = [&](int i, int j) -> double& { return w[(size_t)i * (n + 1) + j]; };
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(W(i, k)) > std::fabs(W(piv, k))) piv = i;
        if (W(piv, k) == 0.0) throw Error{TT_ERR_NONFINITE, "singular reduced system (basis vectors dependent?)"};
        if (piv != k)
            for (int j = 0; j <= n; j++) std::swap(W(k, j), W(piv, j));
        for (int i = k + 1; i < n; i++) {
            const double f = W(i, k) / W(k, k);
            for (int j = k; j <= n; j++) W(i, j) -= f * W(k, j);
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = W(i, n);
        for (int j = i + 1; j < n; j++) s -= W(i, j) * x[j];
        x[i] = s / W(i, i);
    }
    for (int i = 0; i < n; i++)
        if (!std::isfinite(x[i])) throw Error{TT_ERR_NONFINITE, "non-finite reduced solution"};
}

static void dense_mv(int n, const std::vector<double>& X, const double* w, double* y) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += X[(size_t)i * n + j] * w[j];
        y[i] = s;
    }
}

void Context::expand(const double* coef, double* out) {
    Expand3 e{};
    e.Q = Qb.p;
    e.nb = nbasis;
    e.nout = 1;
    e.out[0] = out;
    for (int k = 0; k < nbasis; k++) e.c[0][k] = coef[k];
    launch_expand3(e, g.nd, stream);
}

void Context::expand_state(int t, double* out) {
    // u, v and a in one pass over the basis
    Expand3 e{};
    e.Q = Qb.p;
    e.nb = nbasis;
    e.nout = 3;
    for (int c = 0; c < 3; c++) {
        e.out[c] = out + (size_t)c * g.nd;
        for (int k = 0; k < nbasis; k++) e.c[c][k] = rst[((size_t)t * 3 + c) * nbasis + k];
    }
    launch_expand3(e, g.nd, stream);
}

void Context::project3(const double* h0, const double* h1, const double* h2, double* out) {
    // one pass over h0, h1, h2 per 8 basis vectors; every vector's block partials into its own slice,
    // one copy and one synchronisation, then the sums in block order
    int nblk = 0;
    for (int k0 = 0; k0 < nbasis; k0 += 8)
        nblk = launch_proj3(Qb.p, k0, std::min(8, nbasis - k0), h0, h1, h2, g.nd, projpart.p, stream);
    TT_HIP(hipMemcpyAsync(projhost, projpart.p, sizeof(double) * nbasis * 3 * 1024, hipMemcpyDeviceToHost, stream));
    sync();
    for (int k = 0; k < nbasis; k++) {
        const double* pk = projhost + (size_t)k * 3 * 1024;
        double s[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < nblk; b++)
            for (int i = 0; i < 3; i++) s[i] += pk[(size_t)b * 3 + i];
        for (int i = 0; i < 3; i++) out[3 * k + i] = s[i];
    }
}

// Flexible GMRES(kRestart) with right preconditioning z_j = B v_j, B the multigrid V-cycle as set up
// (modified Gram-Schmidt Arnoldi, Givens rotations; the true residual is formed at every restart and
// decides convergence: ||b - A x|| <= rtol ||b||).
static constexpr int kRestart = 30;
SolveResult Context::fgmres(double cK, double cM, const double* b, double* x) {
    const long nd = g.nd;
    const int m = kRestart;
    if (!gmV.p) {
        gmV.alloc((size_t)(m + 1) * nd);
        gmZ.alloc((size_t)m * nd);
    }
    auto V = [&](int i) { return gmV.p + (size_t)i * nd; };
    auto Z = [&](int i) { return gmZ.p + (size_t)i * nd; };
    auto dot = [&](const double* p, const double* q) {
        launch_dot(p, q, nd, red, PB_MISC, stream);
        return reduce1(PB_MISC);
    };
    const double rtol = prm.rtol;
    const double bn = std::sqrt(dot(b, b));
    SolveResult res;
    stats.solves++;
    if (bn == 0.0) {
        launch_fill(x, 0.0, nd, true, stream);
        return res;
    }
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    int it = 0;
    double rel = 1.0;
    while (true) {
        launch_residual(op_fine64, cK, cM, b, x, V(0), true, stream);
        const double beta = std::sqrt(dot(V(0), V(0)));
        rel = beta / bn;
        if (rel <= rtol) break;
        if (it >= prm.maxit || !std::isfinite(rel)) break;
        launch_scale(V(0), nd, 1.0 / beta, stream);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = beta;
        int k = 0;
        for (; k < m && it < prm.maxit; k++) {
            mg.apply(V(k), Z(k));
            launch_apply_bc(op_fine64, cK, cM, Z(k), V(k + 1), true, stream);
            for (int i = 0; i <= k; i++) {
                const double h = dot(V(k + 1), V(i));
                H[(size_t)i * m + k] = h;
                launch_lincomb(1.0, V(k + 1), -h, V(i), V(k + 1), nd, stream);
            }
            const double hn = std::sqrt(dot(V(k + 1), V(k + 1)));
            H[(size_t)(k + 1) * m + k] = hn;
            if (hn > 0.0) launch_scale(V(k + 1), nd, 1.0 / hn, stream);
            for (int i = 0; i < k; i++) {  // earlier rotations on the new column
                const double a0 = H[(size_t)i * m + k], a1 = H[(size_t)(i + 1) * m + k];
                H[(size_t)i * m + k] = cs[i] * a0 + sn[i] * a1;
                H[(size_t)(i + 1) * m + k] = -sn[i] * a0 + cs[i] * a1;
            }
            const double a0 = H[(size_t)k * m + k], a1 = H[(size_t)(k + 1) * m + k];
            const double r = std::hypot(a0, a1);
            cs[k] = r > 0.0 ? a0 / r : 1.0;
            sn[k] = r > 0.0 ? a1 / r : 0.0;
            H[(size_t)k * m + k] = r;
            H[(size_t)(k + 1) * m + k] = 0.0;
            gv[k + 1] = -sn[k] * gv[k];
            gv[k] = cs[k] * gv[k];
            it++;
            if (std::fabs(gv[k + 1]) <= rtol * bn || hn == 0.0) {
                k++;
                break;
            }
        }
        // y = R^-1 g (upper triangular k x k), x += Z y
        for (int i = k - 1; i >= 0; i--) {
            double s = gv[i];
            for (int j = i + 1; j < k; j++) s -= H[(size_t)i * m + j] * y[j];
            y[i] = s / H[(size_t)i * m + i];
        }
        for (int i = 0; i < k; i++) launch_lincomb(1.0, x, y[i], Z(i), x, nd, stream);
    }
    stats.krylov_iters += it;
    stats.last_iters = it;
    stats.last_relres = rel;
    stats.max_relres = std::max(stats.max_relres, rel);
    res.iters = it;
    res.relres = rel;
    if (!(rel <= rtol)) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "flexible GMRES did not reach rtol %.3g in %d iterations (relres %.3g)", rtol,
                      prm.maxit, rel);
        throw Error{TT_ERR_NOT_CONVERGED, msg};
    }
    return res;
}

void Context::reduced_initial_state() {
    const long nd = g.nd;
    const int nb = nbasis;
    const double om = prm.omega0;
    const double cK0 = 1.0, cM0 = -om * om;
    mg.setup_operator(1.0, om * om);  // preconditioner: the positive-definite shifted operator
    // The basis solves smooth with Chebyshev(cheb_deg) on every level, not with the Newmark solves'
    // one damped-Jacobi step on the intermediate levels: flexible GMRES on the indefinite K - om^2 M
    // needs about twice the iterations with the lighter schedule, and its Arnoldi work grows with them.
    struct Schedule {
        Multigrid& m;
        const int pre, post;
        ~Schedule() {
            m.coarse_pre = pre;
            m.coarse_post = post;
        }
    } restore{mg, mg.coarse_pre, mg.coarse_post};
    mg.coarse_pre = mg.coarse_post = mg.cheb_deg;
    // f0 = N.F: the BC form with zero coefficients returns (1 - N).F
    double* f0 = rhs.p;
    double* f = Kl.p;
    double* qb = Ml.p;
    launch_apply_bc(op_fine64, 0.0, 0.0, F.p, f, true, stream);
    launch_lincomb(1.0, F.p, -1.0, f, f0, nd, stream);
    Ut.assign((size_t)nb * nb, 0.0);
    for (int j = 0; j < nb; j++) {
        const double* rhs_j = f0;
        if (j > 0) {
            // fvec = N.(2 om M Q_{j-1}): the BC form of M on a vector that vanishes on clamped dofs
            launch_apply_bc(op_fine64, 0.0, 1.0, Qb.p + (size_t)(j - 1) * nd, f, true, stream);
            launch_scale(f, nd, -1.0 * (-2.0 * om), stream);
            rhs_j = f;
        }
        launch_fill(qb, 0.0, nd, true, stream);
        fgmres(cK0, cM0, rhs_j, qb);
        for (int i = 0; i < j; i++) {
            const double* qi = Qb.p + (size_t)i * nd;
            launch_dot(qi, qb, nd, red, PB_MISC, stream);
            const double d = reduce1(PB_MISC);
            Ut[(size_t)i * nb + j] = d;
            launch_lincomb(1.0, qb, -d, qi, qb, nd, stream);
        }
        launch_dot(qb, qb, nd, red, PB_MISC, stream);
        const double nrm = std::sqrt(reduce1(PB_MISC));
        if (!(nrm > 0.0) || !std::isfinite(nrm))
            throw Error{TT_ERR_NONFINITE, "SOMMG basis vector " + std::to_string(j + 1) + " vanished"};
        Ut[(size_t)j * nb + j] = nrm;
        launch_lincomb(1.0 / nrm, qb, 0.0, qb, Qb.p + (size_t)j * nd, nd, stream);
    }
    // reduced matrices and load: columns K Q_j, M Q_j, C Q_j against every Q_i
    Kr.assign((size_t)nb * nb, 0.0);
    Mr.assign((size_t)nb * nb, 0.0);
    Cr.assign((size_t)nb * nb, 0.0);
    Pr.assign(nb, 0.0);
    double* tK = hdu.p;
    double* tM = hddu.p;
    double* tC = hdddu.p;
    std::vector<double> col(3 * (size_t)nb);
    for (int j = 0; j < nb; j++) {
        const double* qj = Qb.p + (size_t)j * nd;
        launch_apply_bc(op_fine64, 1.0, 0.0, qj, tK, true, stream);
        launch_apply_bc(op_fine64, 0.0, 1.0, qj, tM, true, stream);
        launch_apply_bc(op_fine64, prm.dbeta, prm.dalpha, qj, tC, true, stream);
        project3(tK, tM, tC, col.data());
        for (int i = 0; i < nb; i++) {
            Kr[(size_t)i * nb + j] = col[3 * i];
            Mr[(size_t)i * nb + j] = col[3 * i + 1];
            Cr[(size_t)i * nb + j] = col[3 * i + 2];
        }
    }
    project3(f0, f0, f0, col.data());
    for (int i = 0; i < nb; i++) Pr[i] = col[3 * i];
    // state 0: u0 = v0 = 0, Mr a0 = Pr f(0) - Cr v0 - Kr u0 (:1050-1078 with N = 1)
    const int T = prm.nsteps;
    rst.assign((size_t)(T + 1) * 3 * nb, 0.0);
    std::vector<double> R(nb);
    const double s0 = load_intensity(0);
    for (int i = 0; i < nb; i++) R[i] = Pr[i] * s0;
    dense_solve(nb, Mr, R.data(), rst.data() + 2 * nb);
}

void Context::reduced_forward() {
    const int nb = nbasis, T = prm.nsteps;
    const size_t n2 = (size_t)nb * nb;
    std::vector<double> S(n2), B1(n2), B2(n2), B3(n2), R(nb), w(nb);
    for (size_t i = 0; i < n2; i++) {
        S[i] = a[5] * Mr[i] + a[2] * Cr[i] + Kr[i];  // -A (:1021-1024)
        B1[i] = a[5] * Mr[i] + a[2] * Cr[i];
        B2[i] = a[3] * Mr[i] - a[0] * Cr[i];
        B3[i] = a[4] * Mr[i] - a[1] * Cr[i];
    }
    for (int t = 1; t <= T; t++) {
        const double* up = rst.data() + (size_t)(t - 1) * 3 * nb;
        const double* vp = up + nb;
        const double* ap = up + 2 * nb;
        double* un = rst.data() + (size_t)t * 3 * nb;
        double* vn = un + nb;
        double* an = un + 2 * nb;
        const double s = load_intensity(t);
        for (int i = 0; i < nb; i++) R[i] = Pr[i] * s;
        dense_mv(nb, B1, up, w.data());
        for (int i = 0; i < nb; i++) R[i] += w[i];
        dense_mv(nb, B2, vp, w.data());
        for (int i = 0; i < nb; i++) R[i] += w[i];
        dense_mv(nb, B3, ap, w.data());
        for (int i = 0; i < nb; i++) R[i] += w[i];
        dense_solve(nb, S, R.data(), un);
        for (int i = 0; i < nb; i++) {
            const double du = un[i] - up[i];
            vn[i] = a[0] * vp[i] + a[1] * ap[i] + a[2] * du;
            an[i] = -a[3] * vp[i] - a[4] * ap[i] + a[5] * du;
        }
        stats.forward_steps++;
    }
}

// The backward sweep of the reduced model (SolveLagrangen :808-954 and SolveLagrange0 :678-804 on the
// reduced matrices).  grad_state fills hdu / hddu / hdddu from a full-mesh state; sens adds the
// sensitivities and objective partials of one step.
void Context::reduced_adjoint(const std::function<void(const double*)>& grad_state,
                              const std::function<void(int, const double*, const double*, const double*, int)>& sens) {
    const int nb = nbasis, T = prm.nsteps;
    const size_t n2 = (size_t)nb * nb;
    std::vector<double> S(n2), B1(n2), B2(n2), B3(n2);
    for (size_t i = 0; i < n2; i++) {
        S[i] = a[5] * Mr[i] + a[2] * Cr[i] + Kr[i];
        B1[i] = a[5] * Mr[i] + a[2] * Cr[i];
        B2[i] = a[3] * Mr[i] - a[0] * Cr[i];
        B3[i] = a[4] * Mr[i] - a[1] * Cr[i];
    }
    std::vector<double> h(3 * (size_t)nb), lr(nb), lamn(nb, 0.0), mu(nb), nu(nb), mun(nb, 0.0), nun(nb, 0.0),
        b1(nb), b2(nb), b3(nb), R1(nb);
    double* hi = st_cur.p;  // state t
    double* lo = st_next.p; // state t - 1
    expand_state(T, hi);
    for (int t = T; t >= 1; t--) {
        expand_state(t - 1, lo);
        grad_state(hi);
        project3(hdu.p, hddu.p, hdddu.p, h.data());
        if (t == T) {
            for (int i = 0; i < nb; i++) {
                mu[i] = -h[3 * i + 1];
                nu[i] = -h[3 * i + 2];
                R1[i] = -(h[3 * i] - a[2] * mu[i] - a[5] * nu[i]);
            }
        } else {
            dense_mv(nb, B1, lamn.data(), b1.data());
            dense_mv(nb, B2, lamn.data(), b2.data());
            dense_mv(nb, B3, lamn.data(), b3.data());
            for (int i = 0; i < nb; i++) {
                const double m = h[3 * i + 1] + b2[i] - a[0] * mun[i] + a[3] * nun[i];
                const double n_ = h[3 * i + 2] + b3[i] + a[4] * nun[i];
                const double r = -(h[3 * i] + b1[i] + a[2] * mun[i] + a[5] * nun[i]);
                mu[i] = -m;
                nu[i] = -n_;
                R1[i] = r + a[2] * mu[i] + a[5] * nu[i];
            }
        }
        // A lambda = R1 with A = -S (:1118-1125)
        for (int i = 0; i < nb; i++) R1[i] = -R1[i];
        dense_solve(nb, S, R1.data(), lr.data());
        expand(lr.data(), lam.p);
        sens(0, lam.p, hi, lo, t);
        lamn = lr;
        mun = mu;
        nun = nu;
        std::swap(hi, lo);
        stats.adjoint_steps++;
    }
    // t = 0: Mr nu0 = -(hdddu + B3 lambda_1 - a1 mu_1 + a4 nu_1)  (:763-769 with N = 1)
    grad_state(hi);
    project3(hdu.p, hddu.p, hdddu.p, h.data());
    dense_mv(nb, B3, lamn.data(), b3.data());
    for (int i = 0; i < nb; i++) R1[i] = -(h[3 * i + 2] + b3[i] - a[1] * mun[i] + a[4] * nun[i]);
    std::vector<double> n0(nb);
    dense_solve(nb, Mr, R1.data(), n0.data());
    expand(n0.data(), nu0.p);
    sens(1, nu0.p, hi, nullptr, 0);
}

}  // namespace tt
