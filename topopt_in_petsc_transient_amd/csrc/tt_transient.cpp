// Newmark forward sweep, binomial checkpointing and the backward adjoint sweep.
//
// Forward: CompleteInitialState (LinearElasticity.cc:956-1206) and
// SolveStateBlock (:1298-1397).  Per step the right-hand side
//   f(t) + B1 u + B2 v + B3 a,  B1 = a5 M + a2 C, B2 = a3 M - a0 C, B3 = a4 M - a1 C
// is one fused matrix-free pass (K and M applied to two combinations of u, v, a),
// followed by one MG-preconditioned flexible-CG solve with (K + a2 C + a5 M)
// and the Newmark corrector.
//
// Adjoint: SolveLagrangen (:808-954) and SolveLagrange0 (:678-804), with the
// sensitivities of ComputeSensitivitiesElementLoopn/0 (:1911-2204) and the
// objective of ComputeObjectiveElementLoop (:2359-2452) evaluated in the same
// backward pass.  The reference keeps ncp*tsc states on disk/in RAM; here the
// states live in HBM, either all of them or a binomial set of checkpoints with
// recomputation between them (Griewank's revolve schedule).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

#include "tt_context.h"

namespace tt {

double Context::load_intensity(int step) const {
    double s;
    if (prm.loadsignal == 0) {
        // LinearElasticity.cc:598-609 at t = step*dt
        const double t = step * dt, t0 = 8.0, c = 1.1;
        s = 1.0 * std::exp(-std::pow(t - t0, 2) / std::pow(2 * c, 2)) * std::cos(4.0 * (t - t0)) *
            std::cos(5.0 * (t - t0)) * std::cos(6.0 * (t - t0)) * std::cos(7.0 * (t - t0));
    } else {
        s = (step <= prm.pulse_steps) ? std::sin(M_PI * (double)step / (double)prm.pulse_steps) : 0.0;
    }
    return s / (double)prm.nely;  // LoadIntensity = signal/(nn[1]-1) (:611)
}

// ------------------------------------------------------------------ forward
void Context::complete_initial_state() {
    TT_CHECK(density_set, TT_ERR_STATE, "tt_set_density must be called first");
    given.clear();
    if (reduced) {
        reduced_initial_state();
        initial_done = true;
        forward_done = false;
        return;
    }
    const long nd = g.nd;
    mg.setup_operator(cK_A, cM_A);
    launch_diag_inv(op_fine64, 0.0, 1.0, dinvM.p, true, stream);
    // state 0: u0 = v0 = 0 (LinearElasticity.cc:521-525), or the u0 of tt_set_initial_displacement;
    // M~ a0 = N.(f(0) - C v0 - K u0) (:1050-1078)
    double* s0 = slots[0].p;
    // u0 = v0 = 0 and no acceleration history; the solve starts from a0 = 0, or once from the a0 that
    // tt_set_restart_vectors supplied (the reference's DDU0 read from RestartSol0x.dat, :2716-2720)
    launch_fill(s0, 0.0, 2 * nd, true, stream);
    if (u0_set) launch_copy(u0in.p, s0, nd, true, true, stream);  // -solutionAsU0 (:523, 584, 2621-2676)
    launch_fill(s0 + 3 * nd, 0.0, (kStateVecs - 3) * nd, true, stream);
    if (!a0_guess) launch_fill(s0 + 2 * nd, 0.0, nd, true, stream);
    a0_guess = false;
    const double wK[3] = {-1.0, -prm.dbeta, 0.0};
    const double wM[3] = {0.0, -prm.dalpha, 0.0};
    launch_newmark_rhs(op_fine64, s0, s0 + nd, s0 + 2 * nd, F.p, load_intensity(0), wK, wM, rhs.p, stream);
    solve_M(rhs.p, s0 + 2 * nd);
    mg.finish_inverse();  // the coarsest inverse, computed beside the smoother set-up and this solve
    for (auto& v : slot_step) v = -1;
    slot_step[0] = 0;
    initial_done = true;
    forward_done = false;
}

// Initial guess of step t from state t-1 (s_in): the displacement for which the corrector returns the
// acceleration predicted by the polynomial through the last m = min(t, kFwdHist) accelerations a_{t-1} ..
// a_{t-m} (held in the state), u_{t-1} + (a3 v_{t-1} + a4 a_{t-1} + a^_t) / a5; with m = 1 it is the
// constant-acceleration step u + dt v + dt^2/2 a (Newmark).  The reference starts from u_{t-1} (:1358); the
// converged answer is the same to the solver tolerance.  Measured on the 128x64x64 sweep: 9.1 Krylov
// iterations per step from u_{t-1}, 5.2 with m = 1, 2.1 with m = 4.
Extrap Context::forward_guess(int t, const double* s_in) const {
    const long nd = g.nd;
    const int m = std::min(t, kFwdHist);
    Extrap e{};
    e.m = 2 + m;
    e.v[0] = s_in;
    e.c[0] = 1.0;
    e.v[1] = s_in + nd;
    e.c[1] = a[3] / a[5];
    long bin = 1;
    for (int k = 1; k <= m; k++) {
        bin = bin * (m - k + 1) / k;
        e.v[1 + k] = s_in + (size_t)(1 + k) * nd;  // a_{t-k}
        e.c[1 + k] = (k % 2 ? 1.0 : -1.0) * (double)bin / a[5];
    }
    e.c[2] += a[4] / a[5];
    return e;
}

void Context::newmark_step(const double* s_in, double* s_out, int t, double* next_out) {
    const long nd = g.nd;
    const double* u = s_in;
    const double* v = s_in + nd;
    const double* ac = s_in + 2 * nd;
    double* un = s_out;
    double* vn = s_out + nd;
    double* an = s_out + 2 * nd;
    const double al = prm.dalpha, be = prm.dbeta;
    const double wK[3] = {a[2] * be, -a[0] * be, -a[1] * be};
    const double wM[3] = {a[5] + a[2] * al, a[3] - a[0] * al, a[4] - a[1] * al};
    launch_newmark_rhs(op_fine64, u, v, ac, F.p, load_intensity(t), wK, wM, rhs.p, stream);
    // the initial guess, unless the previous step's corrector has already written it here
    if (!(guess_t == t && guess_out == un)) launch_extrapolate(forward_guess(t, s_in), un, nd, stream);
    guess_t = -1;
    guess_out = nullptr;
    solve_A(rhs.p, un, kSolveForward);
    // corrector, the acceleration history of the new state (a_{t-1} .. a_{t-kFwdHist+1}) and, with
    // next_out, the next step's initial guess from the new state, written into next_out (the same
    // arithmetic as k_extrapolate, so the result does not depend on which kernel formed it)
    Extrap ng{};
    if (next_out) {
        ng = forward_guess(t + 1, s_out);
        guess_t = t + 1;
        guess_out = next_out;
    }
    launch_newmark_update(u, v, ac, un, vn, an, a, nd, s_in + 3 * nd, s_out + 3 * nd, kFwdHist - 1, stream,
                          next_out ? &ng : nullptr, next_out);
    stats.forward_steps++;
}

static long binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    long r = 1;
    for (int i = 1; i <= k; i++) {
        r = r * (n - k + i) / i;
        if (r > (1L << 40)) return 1L << 40;
    }
    return r;
}

// right-part length of a binomial split of n steps with s free slots
static int revolve_right(int n, int s) {
    int t = 1;
    while (binom(s + t, s) < n) t++;
    long right = binom(s - 1 + t, s - 1);  // beta(s-1, t)
    if (right > n - 1) right = n - 1;
    if (right < 1) right = 1;
    return (int)right;
}

struct Revolve {
    Context* c;
    // adjoint step t; state t-1 is at prev (a checkpoint slot or st_cur), read only
    std::function<void(int, const double*)> adjoint_at;
    // called before the slot of a checkpointed step is freed (the caller may still point into it)
    std::function<void(int)> on_release;
    int find_slot(int step) const {
        for (size_t i = 0; i < c->slot_step.size(); i++)
            if (c->slot_step[i] == step) return (int)i;
        return -1;
    }
    int free_slot() const {
        for (size_t i = 1; i < c->slot_step.size(); i++)
            if (c->slot_step[i] < 0) return (int)i;
        return -1;
    }
    const double* slot_of(int step) const {
        const int k = find_slot(step);
        TT_CHECK(k >= 0, TT_ERR_INTERNAL, "revolve: checkpoint missing");
        return c->slots[k].p;
    }
    // states from+1 .. to, starting from the checkpoint of `from` read in place (no restore copy);
    // the last one goes straight into `last` (a free checkpoint slot) when given, else it ends in
    // st_cur.  Returns where state `to` is.
    const double* advance(int from, int to, double* last) {
        const double* src = slot_of(from);
        for (int t = from + 1; t <= to; t++) {
            if (t == to && last) {
                c->newmark_step(src, last, t);
                return last;
            }
            c->newmark_step(src, c->st_next.p, t);
            std::swap(c->st_cur.p, c->st_next.p);
            src = c->st_cur.p;
        }
        return src;
    }
    void release(int step) {
        const int k = find_slot(step);
        if (k > 0) {
            if (on_release) on_release(step);
            c->slot_step[k] = -1;
        }
    }
    // reverse steps r, r-1, ..., l+1; state l is checkpointed; s free slots
    void reverse(int l, int r, int s) {
        if (r - l == 1) {
            adjoint_at(r, slot_of(l));
            return;
        }
        if (s == 0) {
            for (int t = r; t > l; t--) adjoint_at(t, t - 1 == l ? slot_of(l) : advance(l, t - 1, nullptr));
            return;
        }
        const int m = r - revolve_right(r - l, s);
        if (find_slot(m) < 0) {
            // from the nearest stored state at or below m, the last step written into a free slot
            int best = l;
            for (size_t i = 0; i < c->slot_step.size(); i++)
                if (c->slot_step[i] > best && c->slot_step[i] <= m) best = c->slot_step[i];
            const int k = free_slot();
            TT_CHECK(k >= 0, TT_ERR_INTERNAL, "revolve: no free checkpoint slot");
            advance(best, m, c->slots[k].p);
            c->slot_step[k] = m;
        }
        reverse(m, r, s - 1);
        release(m);
        reverse(l, m, s);
    }
    // states stored by the forward sweep: the split points of the first descent
    static void first_descent(int n, int s, std::vector<int>& pts) {
        int l = 0, r = n;
        while (r - l > 1 && s > 0) {
            const int m = r - revolve_right(r - l, s);
            pts.push_back(m);
            l = m;
            s--;
        }
    }
};

// State t of the last forward sweep: a stored one (checkpoint slot, or state T), else recomputed in
// st_cur from the nearest stored state below t, or from the state recomputed last when that is
// nearer, so that asking for t = 1, 2, ..., T in turn costs one Newmark step each.
const double* Context::state_at(int t) {
    TT_CHECK(forward_done || (t == 0 && initial_done), TT_ERR_STATE, "tt_forward must be called first");
    TT_CHECK(t >= 0 && t <= prm.nsteps, TT_ERR_ARG, "time step out of range");
    if (reduced) {  // the reduced state expanded on the full mesh
        expand_state(t, st_cur.p);
        scratch_step = -1;
        return st_cur.p;
    }
    int base = -1, kb = -1;
    for (size_t i = 0; i < slot_step.size(); i++) {
        if (slot_step[i] == t) return slots[i].p;
        if (slot_step[i] >= 0 && slot_step[i] < t && slot_step[i] > base) {
            base = slot_step[i];
            kb = (int)i;
        }
    }
    if (!store_all && t == prm.nsteps && given.empty()) return st_last.p;
    TT_CHECK(kb >= 0, TT_ERR_INTERNAL, "no stored state below the requested one");
    const double* src = slots[kb].p;
    if (scratch_step > base && scratch_step <= t) {
        if (scratch_step == t) return st_cur.p;
        base = scratch_step;
        src = st_cur.p;
    }
    for (int s = base + 1; s <= t; s++) {
        newmark_step(src, st_next.p, s);
        std::swap(st_cur.p, st_next.p);
        src = st_cur.p;
    }
    scratch_step = t;
    return st_cur.p;
}

void Context::set_checkpoint(int t, const double* u, const double* v, const double* ac) {
    TT_CHECK(!reduced, TT_ERR_STATE, "no checkpoint states with the reduced-order model");
    TT_CHECK(initial_done, TT_ERR_STATE, "tt_complete_initial_state must be called first");
    TT_CHECK(t >= 0 && t <= prm.nsteps, TT_ERR_ARG, "time step out of range");
    const long nd = g.nd;
    if (given.empty()) {  // the first given state replaces the forward sweep's record; state 0 stays
        for (size_t i = 1; i < slot_step.size(); i++) slot_step[i] = -1;
        slot_step[0] = 0;
        given.insert(0);
        scratch_step = -1;
    }
    int k = -1;
    if (t == 0) {
        k = 0;
    } else if (store_all) {
        k = t;
    } else {
        for (size_t i = 1; i < slot_step.size() && k < 0; i++)
            if (slot_step[i] == t) k = (int)i;
        for (size_t i = 1; i < slot_step.size() && k < 0; i++)
            if (slot_step[i] < 0) k = (int)i;
        TT_CHECK(k > 0, TT_ERR_ARG, "more checkpoint states than checkpoint slots");
    }
    double* s = slots[k].p;
    launch_copy(u, s, nd, true, true, stream);
    launch_copy(v, s + nd, nd, true, true, stream);
    launch_copy(ac, s + 2 * nd, nd, true, true, stream);
    // no acceleration history comes with a file: the next solve's guess extrapolates a constant acceleration
    for (int h = 3; h < kStateVecs; h++) launch_copy(ac, s + (size_t)h * nd, nd, true, true, stream);
    slot_step[k] = t;
    given.insert(t);
    forward_done = true;
    sync();
}

void Context::forward() {
    TT_CHECK(initial_done, TT_ERR_STATE, "tt_complete_initial_state must be called first");
    given.clear();
    scratch_step = -1;
    guess_t = -1;
    guess_out = nullptr;
    if (reduced) {
        reduced_forward();
        forward_done = true;
        return;
    }
    const int T = prm.nsteps;
    const long nd3 = kStateVecs * g.nd;  // whole states
    for (size_t i = 1; i < slot_step.size(); i++) slot_step[i] = -1;
    launch_copy(slots[0].p, st_cur.p, nd3, true, true, stream);
    if (store_all) {
        for (int t = 1; t <= T; t++) {
            newmark_step(slots[t - 1].p, slots[t].p, t, t < T ? slots[t + 1].p : nullptr);
            slot_step[t] = t;
        }
        launch_copy(slots[T].p, st_cur.p, nd3, true, true, stream);
    } else {
        std::vector<int> pts;
        Revolve::first_descent(T, (int)slots.size() - 1, pts);
        int cur = 0;
        size_t k = 0;
        for (int t = 1; t <= T; t++) {
            // (the next step writes into this step's input: its guess goes there)
            newmark_step(st_cur.p, st_next.p, t, t < T ? st_cur.p : nullptr);
            std::swap(st_cur.p, st_next.p);
            cur = t;
            if (k < pts.size() && pts[k] == cur) {
                launch_copy(st_cur.p, slots[k + 1].p, nd3, true, true, stream);
                slot_step[k + 1] = cur;
                k++;
            }
        }
        launch_copy(st_cur.p, st_last.p, nd3, true, true, stream);  // state T, read by the first adjoint step
    }
    sync();
    forward_done = true;
}

// ------------------------------------------------------------------ adjoint
// Multipliers the adjoint solves extrapolate from: the initial guess is the degree-6 polynomial through
// the latest 7 of them.  Measured adjoint iterations per solve on the 128x64x64 gradient (200 steps,
// rtol 1e-8): 5.4 with 4 multipliers, 4.4 with 5, 3.0 with 6, 2.5 with 7, 2.8 with 8 (a higher degree
// amplifies the solver-tolerance noise of the stored multipliers more than it reduces the truncation).
constexpr int kAdjHist = 7;
static_assert(kAdjHist <= (int)(sizeof(Context::lamh) / sizeof(Context::lamh[0])) && kAdjHist <= 8, "history size");
static_assert(Context::kFwdHist >= 1 && 2 + Context::kFwdHist <= 8, "forward guess: at most 8 extrapolated vectors");

void Context::objective_gradient(double* fx_host, double* dfdx_dev) {
    TT_CHECK(forward_done, TT_ERR_STATE, "tt_forward must be called first");
    scratch_step = -1;  // the backward sweep uses st_cur as its work space
    const int T = prm.nsteps;
    const long nd = g.nd, nn = g.nn;
    const double al = prm.dalpha, be = prm.dbeta;
    const int obj = prm.objective;
    const double coef[8] = {a[0], a[1], a[2], a[3], a[4], a[5], al, be};
    const long nbs = (g.nel + 255) / 256;  // objective partials per time step (one per sensitivity block)
    launch_fill(df.p, 0.0, g.nel, true, stream);
    launch_fill(hdu.p, 0.0, nd, true, stream);
    launch_fill(hddu.p, 0.0, nd, true, stream);
    launch_fill(hdddu.p, 0.0, nd, true, stream);
    launch_fill(mun.p, 0.0, nd, true, stream);
    launch_fill(nun.p, 0.0, nd, true, stream);

    // d(objective)/d(state) (ComputeGradientObjectiveElementLoop + GradObjectiveWrtState, :2206-2356, :1516-1576)
    auto grad_state = [&](const double* s) {
        const double* u = s;
        const double* v = s + nd;
        const double* ac = s + 2 * nd;
        switch (obj) {
        case 0: launch_hdiag(u, rcnt.p, dt, hdu.p, nn, stream); break;
        case 1: launch_hdiag(v, rcnt.p, dt, hddu.p, nn, stream); break;
        case 2: launch_hdiag(ac, rcnt.p, dt, hdddu.p, nn, stream); break;
        case 3: launch_apply_raw(op_fine64, Pkf.p, Pmf.p, dt, 0.0, u, hdu.p, stream); break;
        case 4: launch_apply_raw(op_fine64, Pkf.p, Pmf.p, 0.0, dt, v, hddu.p, stream); break;
        case 5:
            launch_apply_raw(op_fine64, Pkf.p, Pmf.p, dt, 0.0, u, hdu.p, stream);
            launch_apply_raw(op_fine64, Pkf.p, Pmf.p, 0.0, dt, v, hddu.p, stream);
            break;
        default: launch_apply_raw(op_fine64, Pkf.p, Pmf.p, dt * be, dt * al, v, hddu.p, stream); break;
        }
    };

    auto sens = [&](int mode, const double* lamv, const double* st, const double* stb, int t) {
        SensArgs sa;
        sa.g = g;
        sa.KE = KE64d.p;
        sa.M8 = M864d.p;
        sa.kf = kf.p; sa.mf = mf.p; sa.dkf = dkf.p; sa.dmf = dmf.p; sa.P = P.p;
        sa.lam = lamv;
        sa.u = st; sa.v = st + nd; sa.a = st + 2 * nd;
        sa.ub = stb ? stb : st;
        sa.vb = stb ? stb + nd : st + nd;
        sa.ab = stb ? stb + 2 * nd : st + 2 * nd;
        if (mode == 0) {
            const double wk[4] = {-(1.0 + a[2] * be), a[2] * be, -a[0] * be, -a[1] * be};
            const double wm[4] = {-(a[5] + a[2] * al), a[5] + a[2] * al, a[3] - a[0] * al, a[4] - a[1] * al};
            std::memcpy(sa.wK, wk, sizeof(wk));
            std::memcpy(sa.wM, wm, sizeof(wm));
        } else {
            const double wk[4] = {1.0, be, 0.0, 0.0};
            const double wm[4] = {0.0, al, 1.0, 0.0};
            std::memcpy(sa.wK, wk, sizeof(wk));
            std::memcpy(sa.wM, wm, sizeof(wm));
        }
        sa.dt = dt; sa.dalpha = al; sa.dbeta = be;
        sa.objective = obj;
        sa.mode = mode;
        sa.df = df.p;
        launch_sens(sa, red, fxs.p + (size_t)t * nbs, stream);
    };

    // one backward step at t: needs state t and state t-1
    int nhist = 0;  // multipliers of later steps held in lamh[0..] (lamh[0] = lambda_{t+1})
    auto adjoint_step = [&](int t, const double* st, const double* stb) {
        grad_state(st);
        if (t == T) {
            launch_adjoint_combine(nullptr, nullptr, hdu.p, hddu.p, hdddu.p, mun.p, nun.p, mask.p, coef, true, mu.p,
                                   nu.p, rhs.p, nn, stream);
            // the terminal multiplier of the previous call is the initial guess (:877-883)
        } else {
            launch_apply_split(op_fine64, 1.0, 1.0, lamh[0].p, Kl.p, Ml.p, stream);
            launch_adjoint_combine(Kl.p, Ml.p, hdu.p, hddu.p, hdddu.p, mun.p, nun.p, mask.p, coef, false, mu.p, nu.p,
                                   rhs.p, nn, stream);
            // Initial guess: the polynomial through the last m = min(nhist, kAdjHist) multipliers,
            // evaluated one step on: sum_k (-1)^(k+1) C(m, k) lambda_{t+k}.  The reference starts from
            // lambda_{t+1} (:917); the converged multiplier is the same to the solver tolerance.
            Extrap e{};
            e.m = std::min(nhist, kAdjHist);
            long bin = 1;
            for (int k = 1; k <= e.m; k++) {
                bin = bin * (e.m - k + 1) / k;
                e.v[k - 1] = lamh[k - 1].p;
                e.c[k - 1] = (k % 2 ? 1.0 : -1.0) * (double)bin;
            }
            launch_extrapolate(e, lam.p, nd, stream);
        }
        solve_A(rhs.p, lam.p, kSolveAdjoint);
        sens(0, lam.p, st, stb, t);
        // lambda_t becomes lambda_{t+1} of the next (earlier) step
        double* oldest = lamh[kAdjHist - 1].p;
        for (int k = kAdjHist - 1; k > 0; k--) lamh[k].p = lamh[k - 1].p;
        lamh[0].p = lam.p;
        lam.p = oldest;
        nhist = std::min(nhist + 1, kAdjHist);
        std::swap(mu.p, mun.p);
        std::swap(nu.p, nun.p);
        stats.adjoint_steps++;
        // keep the terminal multiplier as next call's starting guess
        if (t == T) launch_copy(lamh[0].p, lamT.p, nd, true, true, stream);
    };

    if (reduced) {
        reduced_adjoint(grad_state, sens);
    } else {
        launch_copy(lamT.p, lam.p, nd, true, true, stream);
        if (!given.empty()) {
            // States given by the caller (the reference's checkpoint files, read back as
            // ComputeObjectiveSensitivities does, LinearElasticity.cc:1651-1680): the blocks between
            // consecutive given steps are recomputed from their first state (SolveStateBlock) and the
            // adjoint runs over each block, from the last block back.
            std::vector<int> cps(given.begin(), given.end());
            Revolve rv;
            rv.c = this;
            if (store_all) {
                for (size_t b = 0; b < cps.size(); b++) {
                    const int l = cps[b], r = b + 1 < cps.size() ? cps[b + 1] : T;
                    for (int t = l + 1; t <= r; t++) {
                        if (slot_step[t] == t) continue;  // a given state is kept as given
                        newmark_step(slots[t - 1].p, slots[t].p, t);
                        slot_step[t] = t;
                    }
                }
                for (int t = T; t >= 1; t--) adjoint_step(t, slots[t].p, slots[t - 1].p);
            } else {
                const double* later = nullptr;
                int later_step = -1;
                rv.adjoint_at = [&](int t, const double* prev) {
                    adjoint_step(t, later, prev);
                    if (prev == st_cur.p) {
                        std::swap(st_hi.p, st_cur.p);
                        later = st_hi.p;
                        later_step = -1;
                    } else {
                        later = prev;
                        later_step = t - 1;
                    }
                };
                rv.on_release = [&](int step) {
                    if (step != later_step) return;
                    launch_copy(later, st_hi.p, 3 * nd, true, true, stream);
                    later = st_hi.p;
                    later_step = -1;
                };
                int nfree = 0;
                for (size_t i = 1; i < slot_step.size(); i++) nfree += slot_step[i] < 0 ? 1 : 0;
                for (int b = (int)cps.size() - 1; b >= 0; b--) {
                    const int l = cps[b], r = b + 1 < (int)cps.size() ? cps[b + 1] : T;
                    if (r == l) continue;
                    if (b + 1 == (int)cps.size()) {  // state T: the last block recomputed once
                        const double* sT = rv.advance(l, T, nullptr);
                        launch_copy(sT, st_last.p, 3 * nd, true, true, stream);
                        later = st_last.p;
                    } else {
                        later = rv.slot_of(r);  // a given state, never released
                    }
                    later_step = -1;
                    rv.reverse(l, r, nfree);
                }
            }
        } else if (store_all) {
            for (int t = T; t >= 1; t--) adjoint_step(t, slots[t].p, slots[t - 1].p);
        } else {
            // The adjoint step at t reads states t and t-1 only.  The sweep runs t = T, T-1, ..., 1, so
            // state t is the state t-1 of the previous adjoint step: it is kept (in its checkpoint slot,
            // or moved to st_hi) instead of being recomputed from state t-1, which would cost one Newmark
            // step per adjoint step.  The schedule then needs the forward sweep plus the binomial
            // recomputation only: 379 Newmark steps for 200 steps and 20 slots, against 579 when every
            // adjoint step redoes its forward step.  Checkpoints are read in place, never restored.
            const double* later = st_last.p;  // state T
            int later_step = -1;              // the checkpointed step `later` points into, or -1
            Revolve rv;
            rv.c = this;
            rv.adjoint_at = [&](int t, const double* prev) {
                adjoint_step(t, later, prev);
                if (prev == st_cur.p) {  // out of the way of the next recomputation
                    std::swap(st_hi.p, st_cur.p);
                    later = st_hi.p;
                    later_step = -1;
                } else {
                    later = prev;  // a checkpoint slot, valid until it is released
                    later_step = t - 1;
                }
            };
            rv.on_release = [&](int step) {
                if (step != later_step) return;
                launch_copy(later, st_hi.p, 3 * nd, true, true, stream);  // u, v, a: what an adjoint step reads
                later = st_hi.p;
                later_step = -1;
            };
            rv.reverse(0, T, (int)slots.size() - 1);
        }
        // t = 0 (SolveLagrange0, :678-804; ComputeSensitivitiesElementLoop0, :2080-2204)
        const double* s0 = slots[0].p;
        grad_state(s0);
        launch_apply_split(op_fine64, 1.0, 1.0, lamh[0].p, Kl.p, Ml.p, stream);
        launch_adjoint0_rhs(Kl.p, Ml.p, hdddu.p, mun.p, nun.p, mask.p, coef, rhs.p, nn, stream);
        solve_M(rhs.p, nu0.p);
        sens(1, nu0.p, s0, nullptr, 0);
    }
    // objective: sum over t = 1..T, then t = 0 (:1778-1802)
    std::vector<double> fp((size_t)(T + 1) * nbs);
    TT_HIP(hipMemcpyAsync(fp.data(), fxs.p, sizeof(double) * fp.size(), hipMemcpyDeviceToHost, stream));
    if (dfdx_dev) launch_copy(df.p, dfdx_dev, g.nel, true, true, stream);
    sync();
    std::vector<double> f(T + 1, 0.0);
    for (int t = 0; t <= T; t++)
        for (long b = 0; b < nbs; b++) f[t] += fp[(size_t)t * nbs + b];
    double fx = 0.0;
    for (int t = 1; t <= T; t++) fx += f[t];
    fx += f[0];
    if (!std::isfinite(fx)) throw Error{TT_ERR_NONFINITE, "non-finite objective"};
    *fx_host = fx;
}

}  // namespace tt
