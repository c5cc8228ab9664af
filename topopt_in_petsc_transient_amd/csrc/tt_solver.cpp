// Multigrid preconditioner and flexible CG.
//
// Replaces the reference's KSP set-up (LinearElasticity.cc:2679-2939: FGMRES +
// PCMG V-cycle with Galerkin coarse operators and GMRES/SOR smoothers) by
//   * a flexible CG outer iteration in FP64 (Polak-Ribiere form, so that a
//     variable preconditioner is allowed), stopping on the true
//     unpreconditioned residual ||b - A x|| <= rtol ||b||;
//   * a geometric V-cycle on the same DMDA-style hierarchy (nodes coarsened
//     by 2 per direction, trilinear interpolation, restriction = P^T) with
//     rediscretised coarse operators (element coefficients averaged over the
//     8 children), Chebyshev smoothing on D^-1 A, and an exact coarsest solve
//     with a dense FP64 inverse;
//   * in the mixed mode the V-cycle runs in FP32 with BF16-stored coarse
//     coefficients; the outer iteration, its residual and dot products stay
//     FP64, so the converged answer is an FP64 answer.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>

#include <rocblas/rocblas.h>

#include "tt_context.h"

namespace tt {

static void build_mask(const Grid& g, int boundaries, std::vector<uint8_t>& m) {
    m.assign(g.nn, 0);
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
                uint8_t b = 0;
                if (boundaries == 1) {
                    if (i == 0) b = 7;  // cantilever: x = xmin face fully clamped (LinearElasticity.cc:565-572)
                } else {
                    if (i == 0) b |= 3;                   // MBB: x,y at x = xmin (:553-557)
                    if (i == g.nx - 1 && k == 0) b |= 6;  // y,z on the edge x = xmax, z = zmin (:558-562)
                }
                m[i + (long)g.nx * (j + (long)g.ny * k)] = b;
            }
}

// Split of an order-n block for the Schur-complement inverse: the leading block a multiple of 64 near n / 2.
static int spd_split(int n) { return std::min(((n / 2 + 63) / 64) * 64, n - 1); }
// workspace (doubles) of spd_inverse at order n: T of the top split, and the trailing block's own
static size_t spd_workspace(int n) {
    if (n <= 64) return 0;
    const int n1 = spd_split(n), n2 = n - n1;
    return std::max(spd_workspace(n1), (size_t)n1 * n2 + spd_workspace(n2));
}

void Multigrid::build(Context* c, int nlvls, bool mixed_prec, int deg) {
    ctx = c;
    mixed = mixed_prec;
    cheb_deg = deg;
    coarse_pre = c->prm.mg_coarse_pre > 0 ? c->prm.mg_coarse_pre : 1;
    coarse_post = c->prm.mg_coarse_post;
    const Grid& g0 = c->g;
    const int f = 1 << (nlvls - 1);
    TT_CHECK(nlvls >= 1 && nlvls <= 8, TT_ERR_ARG, "nlvls must be in 1..8");
    TT_CHECK(g0.ex % f == 0 && g0.ey % f == 0 && g0.ez % f == 0, TT_ERR_ARG,
             "mesh dimension not compatible with the number of multigrid levels");
    L.clear();
    for (int l = 0; l < nlvls; l++) {
        auto lev = std::make_unique<MGLevel>();
        const int s = 1 << l;
        lev->g = make_grid(g0.ex / s, g0.ey / s, g0.ez / s, g0.hx * g0.ex, g0.hy * g0.ey, g0.hz * g0.ez);
        lev->fp64 = !mixed;
        const Grid& g = lev->g;
        std::vector<uint8_t> m;
        build_mask(g, c->prm.boundaries, m);
        lev->mask.alloc(g.nn);
        TT_HIP(hipMemcpy(lev->mask.p, m.data(), g.nn, hipMemcpyHostToDevice));
        {
            std::vector<uint8_t> f;
            tile_bc_flags(g, m, 4, f);
            lev->tbc4.alloc(f.size());
            TT_HIP(hipMemcpy(lev->tbc4.p, f.data(), f.size(), hipMemcpyHostToDevice));
            tile_bc_flags(g, m, 2, f);
            lev->tbc2.alloc(f.size());
            TT_HIP(hipMemcpy(lev->tbc2.p, f.data(), f.size(), hipMemcpyHostToDevice));
        }
        double ke[576], m8[64];
        hex8_stiffness(g.hx, g.hy, g.hz, c->prm.nu, ke);
        hex8_mass(g.hx, g.hy, g.hz, m8);
        float ke32[576], m832[64];
        for (int i = 0; i < 576; i++) ke32[i] = (float)ke[i];
        for (int i = 0; i < 64; i++) m832[i] = (float)m8[i];
        lev->KE64.alloc(576);
        lev->M864.alloc(64);
        lev->KE32.alloc(576);
        lev->M832.alloc(64);
        TT_HIP(hipMemcpy(lev->KE64.p, ke, sizeof(ke), hipMemcpyHostToDevice));
        TT_HIP(hipMemcpy(lev->M864.p, m8, sizeof(m8), hipMemcpyHostToDevice));
        TT_HIP(hipMemcpy(lev->KE32.p, ke32, sizeof(ke32), hipMemcpyHostToDevice));
        TT_HIP(hipMemcpy(lev->M832.p, m832, sizeof(m832), hipMemcpyHostToDevice));
        {
            double af64[18 * 64];
            float af32[AF32_WORDS];
            mfma_fragments(ke, m8, af64, af32, reinterpret_cast<uint32_t*>(af32 + 18 * 64), lev->op.hs);
            lev->AF64.alloc(18 * 64);
            lev->AF32.alloc(AF32_WORDS);
            TT_HIP(hipMemcpy(lev->AF64.p, af64, sizeof(af64), hipMemcpyHostToDevice));
            TT_HIP(hipMemcpy(lev->AF32.p, af32, sizeof(af32), hipMemcpyHostToDevice));
        }
        if (l > 0) {
            lev->kf64.alloc(g.nel);
            lev->mf64.alloc(g.nel);
            if (mixed) {
                lev->kfbf.alloc(g.nel);
                lev->mfbf.alloc(g.nel);
            }
        } else if (mixed) {
            lev->kf32.alloc(g.nel);
            lev->mf32.alloc(g.nel);
        }
        const size_t es = lev->esz();
        for (DBuf<char>* b : {&lev->b, &lev->x, &lev->x2, &lev->r, &lev->d, &lev->d2, &lev->dinv})
            b->alloc(g.nd * es);
        lev->op.g = g;
        lev->op.mask = lev->mask.p;
        lev->op.tbc4 = lev->tbc4.p;
        lev->op.tbc2 = lev->tbc2.p;
        build_batch_tables(g, lev->btab);
        lev->op.tab4 = lev->btab.t4;
        lev->op.tab2 = lev->btab.t2;
        lev->op.KE64 = lev->KE64.p;
        lev->op.M864 = lev->M864.p;
        lev->op.KE32 = lev->KE32.p;
        lev->op.M832 = lev->M832.p;
        lev->op.AF64 = lev->AF64.p;
        lev->op.AF32 = lev->AF32.p;
        lev->op.AFH = reinterpret_cast<const uint32_t*>(lev->AF32.p + 18 * 64);
        L.push_back(std::move(lev));
    }
    const Grid& gc = L.back()->g;
    nc = (int)gc.nd;
    // explicit FP64 inverse up to 320 MB (6,476 dofs), or up to the caller's limit
    const long dmax = c->prm.coarse_direct_max > 0 ? c->prm.coarse_direct_max : 6476;
    dense = nc <= dmax && (long)nc * nc * 8 <= (1l << 34);
    cb.alloc(nc);
    cx.alloc(nc);
    if (dense) {
        Ainv.alloc((size_t)nc * nc);
        AinvP.alloc(sym_packed_size(nc));
        {
            const long nt = (nc + 63) / 64, ntiles = nt * (nt + 1) / 2;
            sprow.alloc(ntiles * 64);
            spcol.alloc(ntiles * 64);
        }
        dinfo.alloc(2);
        Winv.alloc(spd_workspace(nc));
        // the inversion runs on a stream of its own, beside the smoother set-up and the initial state
        TT_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        TT_HIP(hipEventCreateWithFlags(&ev_asm, hipEventDisableTiming));
        TT_HIP(hipEventCreateWithFlags(&ev_inv, hipEventDisableTiming));
        TT_HIP(hipHostMalloc((void**)&info_host, 2 * sizeof(int), hipHostMallocDefault));
        rocblas_handle h = nullptr;
        TT_CHECK(rocblas_create_handle(&h) == rocblas_status_success, TT_ERR_HIP, "rocblas_create_handle failed");
        blas = h;
        TT_CHECK(rocblas_set_stream(h, s2) == rocblas_status_success, TT_ERR_HIP, "rocblas_set_stream failed");
    } else {
        cdinv.alloc(nc);
        cx2.alloc(nc);
        cd.alloc(nc);
        cd2.alloc(nc);
    }
}

void Multigrid::set_coefficients(const double* kf_fine, const double* mf_fine) {
    hipStream_t s = ctx->stream;
    MGLevel& l0 = *L[0];
    if (mixed) {
        launch_convert_coef(kf_fine, l0.kf32.p, l0.g.nel, COEF_F32, s);
        launch_convert_coef(mf_fine, l0.mf32.p, l0.g.nel, COEF_F32, s);
        l0.op.coef_kind = COEF_F32;
        l0.op.kf = l0.kf32.p;
        l0.op.mf = l0.mf32.p;
    } else {
        l0.op.coef_kind = COEF_F64;
        l0.op.kf = kf_fine;
        l0.op.mf = mf_fine;
    }
    const double* pk = kf_fine;
    const double* pm = mf_fine;
    for (size_t l = 1; l < L.size(); l++) {
        MGLevel& lv = *L[l];
        launch_coarsen_coef(L[l - 1]->g, lv.g, pk, pm, lv.kf64.p, lv.mf64.p, s);
        pk = lv.kf64.p;
        pm = lv.mf64.p;
        if (mixed) {
            launch_convert_coef(lv.kf64.p, lv.kfbf.p, lv.g.nel, COEF_BF16, s);
            launch_convert_coef(lv.mf64.p, lv.mfbf.p, lv.g.nel, COEF_BF16, s);
            lv.op.coef_kind = COEF_BF16;
            lv.op.kf = lv.kfbf.p;
            lv.op.mf = lv.mfbf.p;
        } else {
            lv.op.coef_kind = COEF_F64;
            lv.op.kf = lv.kf64.p;
            lv.op.mf = lv.mf64.p;
        }
    }
}

// largest eigenvalue of D^-1 A by power iteration (deterministic start vector)
double Multigrid::power_lmax(int l, int iters) {
    MGLevel& lv = *L[l];
    hipStream_t s = ctx->stream;
    const long nd = lv.g.nd;
    launch_fill_hash(lv.x.p, nd, 0x9E3779B97F4A7C15ull, lv.fp64, s);  // deterministic start vector
    for (int it = 0; it < iters; it++) {
        launch_apply_bc(lv.op, cK, cM, lv.x.p, lv.r.p, lv.fp64, s);
        launch_norm_power(lv.r.p, lv.dinv.p, lv.x.p, nd, ctx->red, lv.fp64, s);
        launch_scale_by_norm(lv.x.p, nd, ctx->red, lv.fp64, s);
    }
    return std::sqrt(ctx->reduce1(PB_MISC));
}

Multigrid::~Multigrid() {
    if (s2) (void)hipStreamSynchronize(s2);
    if (blas) (void)rocblas_destroy_handle((rocblas_handle)blas);
    if (info_host) (void)hipHostFree(info_host);
    if (ev_asm) (void)hipEventDestroy(ev_asm);
    if (ev_inv) (void)hipEventDestroy(ev_inv);
    if (s2) (void)hipStreamDestroy(s2);
}

void Multigrid::assemble_coarse_inverse() {
    MGLevel& lc = *L.back();
    const Grid& g = lc.g;
    const double* kd = L.size() > 1 ? lc.kf64.p : (const double*)lc.op.kf;
    const double* md = L.size() > 1 ? lc.mf64.p : (const double*)lc.op.mf;
    if (L.size() == 1 && mixed) {  // single-level mixed run: use the context's FP64 coefficients
        kd = ctx->kf.p;
        md = ctx->mf.p;
    }
    const int n = nc;
    hipStream_t s = ctx->stream;
    finish_inverse();  // (an inversion still in flight reads the matrix assembled below)
    // the dense coarsest operator, assembled on the device from the level's FP64 coefficients
    launch_dense_coarse(g, kd, md, cK, cM, lc.KE64.p, lc.M864.p, lc.mask.p, Ainv.p, s);
    // A^-1 on the second stream, while the main stream sets up the smoothers and computes the initial
    // state (spd_inverse), then packed; the status word goes to pinned host memory and is checked where
    // the inverse is first needed (finish_inverse)
    TT_HIP(hipEventRecord(ev_asm, s));
    TT_HIP(hipStreamWaitEvent(s2, ev_asm, 0));
    TT_HIP(hipMemsetAsync(dinfo.p, 0, 2 * sizeof(int), s2));
    spd_inverse(Ainv.p, n, n, Winv.p);
    launch_sym_pack(Ainv.p, n, AinvP.p, s2);
    TT_HIP(hipMemcpyAsync(info_host, dinfo.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s2));
    TT_HIP(hipEventRecord(ev_inv, s2));
    inv_pending = true;
}

// In-place inverse of the SPD matrix A (order n, column-major, leading dimension lda) on the inversion
// stream, by Schur complements:
//     [A11 A12; A21 A22]^-1 = [X11 + T X22 T', -T X22; -X22 T', X22],  X11 = A11^-1, T = X11 A12,
//     X22 = S^-1 with S = A22 - A21 T (SPD when A is),
// recursively down to blocks of at most 64, which one wave inverts in LDS (launch_spd_inv_small); the rest
// is 4 rocBLAS GEMMs per split (4/3 n^3 flops in all) and a transpose for the lower block.
void Multigrid::spd_inverse(double* A, int n, int lda, double* W) {
    if (n <= 64) {
        launch_spd_inv_small(A, n, lda, dinfo.p, s2);
        return;
    }
    const int n1 = spd_split(n), n2 = n - n1;
    double* A11 = A;
    double* A21 = A + n1;
    double* A12 = A + (size_t)n1 * lda;
    double* A22 = A12 + n1;
    spd_inverse(A11, n1, lda, W);  // (its workspace is free again before T is written: same stream)
    rocblas_handle h = (rocblas_handle)blas;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    double* T = W;  // n1 x n2, leading dimension n1
    auto gemm = [&](rocblas_operation tb, int m, int nn, int k, const double* alpha, const double* a, int la,
                    const double* b, int lb, const double* beta, double* c, int lc) {
        TT_CHECK(rocblas_dgemm(h, rocblas_operation_none, tb, m, nn, k, alpha, a, la, b, lb, beta, c, lc) ==
                     rocblas_status_success,
                 TT_ERR_HIP, "rocblas_dgemm failed");
    };
    gemm(rocblas_operation_none, n1, n2, n1, &one, A11, lda, A12, lda, &zero, T, n1);     // T = X11 A12
    gemm(rocblas_operation_none, n2, n2, n1, &mone, A21, lda, T, n1, &one, A22, lda);     // S = A22 - A21 T
    spd_inverse(A22, n2, lda, W + (size_t)n1 * n2);                                          // X22 = S^-1
    gemm(rocblas_operation_none, n1, n2, n2, &mone, T, n1, A22, lda, &zero, A12, lda);    // X12 = -T X22
    gemm(rocblas_operation_transpose, n1, n1, n2, &mone, A12, lda, T, n1, &one, A11, lda);  // X11 -= X12 T'
    launch_transpose(A12, n1, n2, lda, A21, lda, s2);                                       // X21 = X12'
}

void Multigrid::finish_inverse() {
    if (!inv_pending) return;
    inv_pending = false;
    TT_HIP(hipStreamWaitEvent(ctx->stream, ev_inv, 0));
    TT_HIP(hipEventSynchronize(ev_inv));
    TT_CHECK(info_host[0] == 0, TT_ERR_INTERNAL, "coarsest operator is not positive definite");
}

std::vector<char**> Multigrid::rotating() {
    std::vector<char**> v;
    for (size_t l = 0; l + 1 < L.size(); l++)
        for (DBuf<char>* b : {&L[l]->x, &L[l]->x2, &L[l]->d, &L[l]->d2}) v.push_back(&b->p);
    return v;
}

void Multigrid::setup_operator(double cK_, double cM_) {
    ctx->clear_graphs();
    cK = cK_;
    cM = cM_;
    hipStream_t s = ctx->stream;
    const int nl = (int)L.size();
    // the coarsest inverse first: it runs on its own stream beside the power iterations below
    if (dense) assemble_coarse_inverse();
    for (int l = 0; l < nl - 1; l++) {
        MGLevel& lv = *L[l];
        launch_diag_inv(lv.op, cK, cM, lv.dinv.p, lv.fp64, s);
        lv.lmax = power_lmax(l, 12);
        if (l < 8) ctx->stats.lambda_max[l] = lv.lmax;
    }
    ctx->stats.nlvls = nl;
    if (!dense) {
        // Chebyshev coarse solve in FP64 on [clmax/30, 1.05 clmax] of D^-1 A
        OpLevel op = coarse_op64();
        launch_diag_inv(op, cK, cM, cdinv.p, true, s);
        launch_fill_hash(cx.p, nc, 0x2545F4914F6CDD1Dull, true, s);
        for (int it = 0; it < 20; it++) {
            launch_apply_bc(op, cK, cM, cx.p, cd.p, true, s);
            launch_norm_power(cd.p, cdinv.p, cx.p, nc, ctx->red, true, s);
            launch_scale_by_norm(cx.p, nc, ctx->red, true, s);
        }
        clmax = 1.05 * std::sqrt(ctx->reduce1(PB_MISC));
        clmin = clmax / 30.0;
    }
}

OpLevel Multigrid::coarse_op64() const {
    const MGLevel& lc = *L.back();
    OpLevel op = lc.op;
    op.coef_kind = COEF_F64;
    if (L.size() > 1) {
        op.kf = lc.kf64.p;
        op.mf = lc.mf64.p;
    } else {
        op.kf = ctx->kf.p;
        op.mf = ctx->mf.p;
    }
    return op;
}

// direct64: the FP32 level above restricted straight into cb and prolongs straight from cx.
// The coarsest solve is FP64 in both precision modes.
void Multigrid::coarse_solve(bool direct64) {
    MGLevel& lc = *L.back();
    hipStream_t s = ctx->stream;
    const double* b64 = (const double*)lc.b.p;
    double* x64 = (double*)lc.x.p;
    if (!lc.fp64) {
        b64 = cb.p;
        x64 = cx.p;
        if (direct64) {
            launch_sym_solve(AinvP.p, b64, x64, nc, sprow.p, spcol.p, s);
            return;
        }
        launch_copy(lc.b.p, cb.p, nc, false, true, s);
    }
    if (dense) {
        launch_sym_solve(AinvP.p, b64, x64, nc, sprow.p, spcol.p, s);
    } else {
        OpLevel op = coarse_op64();
        const double theta = 0.5 * (clmax + clmin), delta = 0.5 * (clmax - clmin), sigma = theta / delta;
        double* x = cx.p;
        double* x2 = cx2.p;
        double* d = cd.p;
        double* d2 = cd2.p;
        launch_cheb_first(b64, cdinv.p, 1.0 / theta, x, nc, true, s);
        double rho0 = 1.0 / sigma;
        for (int k = 1; k < coarse_deg; k++) {
            const double rho1 = 1.0 / (2.0 * sigma - rho0);
            launch_cheb_step(op, cK, cM, b64, x, k == 1 ? x : d, cdinv.p, rho1 * rho0, 2.0 * rho1 / delta, x2, d2,
                             true, s);
            std::swap(x, x2);
            std::swap(d, d2);
            rho0 = rho1;
        }
        if (lc.fp64) launch_copy(x, x64, nc, true, true, s);
        else x64 = x;
    }
    if (!lc.fp64) launch_copy(x64, lc.x.p, nc, true, false, s);
}

void Multigrid::vcycle(int l, const double* r64, double* z64, const double* rp64, bool first) {
    const int nl = (int)L.size();
    if (l == nl - 1) {
        coarse_solve(l > 0 && dense && !L[l]->fp64);  // as decided by the level above (direct64)
        return;
    }
    MGLevel& lv = *L[l];
    hipStream_t s = ctx->stream;
    const bool f64 = lv.fp64;
    const long nd = lv.g.nd;
    const double lo = cheb_lo * lv.lmax, hi = cheb_hi * lv.lmax;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    // Smoothing steps: cheb_deg on each side of the finest level; coarse_pre / coarse_post (default 1 / 0)
    // on the levels below it.  Measured on the 128x64x64 Newmark sweep (Chebyshev(2) on the finest level,
    // 4 levels): 2 / 2 on every level 1226 mixed / 879 FP64 steps/s, 2.055 Krylov iterations per solve;
    // 2 / 1 below the finest 1252 / 903 (2.08); 1 / 1 1299 / 943 (2.055); 1 / 0 1364 / 992 (2.06); fewer
    // steps on the finest level cost iterations (1 / 2 there: 2.69 per solve).  The intermediate levels
    // only need to pass the mid frequencies on: the one damped-Jacobi step and the residual are enough.
    const int npre = l == 0 ? cheb_deg : coarse_pre;
    const int npost = l == 0 ? cheb_deg : coarse_post;
    // pre-smoothing from x = 0 (on the finest level of a Krylov preconditioner call, b = (T) r is
    // formed by the same kernel)
    double rho0 = 1.0 / sigma;
    bool d_is_x = true;
    int k0 = 1;
    if (npre >= 2) {
        // x0 = D^-1 b / theta is formed inside the staging of the first step's operator apply
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        launch_cheb_first_step(lv.op, cK, cM, r64, lv.b.p, lv.dinv.p, 1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta,
                               lv.x2.p, lv.d2.p, f64, s);
        std::swap(lv.x.p, lv.x2.p);
        std::swap(lv.d.p, lv.d2.p);
        d_is_x = false;
        rho0 = rho1;
        k0 = 2;
    } else if (r64) {
        launch_cheb_first_from64(r64, lv.b.p, lv.dinv.p, 1.0 / theta, lv.x.p, nd, f64, s);
    }  // (else the single step x = D^-1 b / theta is formed by the residual's kernel below)
    for (int k = k0; k < npre; k++) {
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        launch_cheb_step(lv.op, cK, cM, lv.b.p, lv.x.p, d_is_x ? lv.x.p : lv.d.p, lv.dinv.p, rho1 * rho0,
                         2.0 * rho1 / delta, lv.x2.p, lv.d2.p, f64, s);
        std::swap(lv.x.p, lv.x2.p);
        std::swap(lv.d.p, lv.d2.p);
        d_is_x = false;
        rho0 = rho1;
    }
    // coarse-grid correction
    if (npre == 1 && !r64) launch_resid_first(lv.op, cK, cM, lv.b.p, lv.dinv.p, 1.0 / theta, lv.x.p, lv.r.p, f64, s);
    else launch_residual(lv.op, cK, cM, lv.b.p, lv.x.p, lv.r.p, f64, s);
    MGLevel& lc = *L[l + 1];
    // an FP32 level above the dense FP64 coarsest solve restricts into and prolongs from its FP64
    // vectors directly
    const bool direct64 = l + 2 == nl && dense && !lc.fp64;
    launch_restrict(lv.g, lc.g, lv.r.p, lc.mask.p, direct64 ? (void*)cb.p : lc.b.p, f64, direct64, s);
    vcycle(l + 1);
    launch_prolong_add(lv.g, lc.g, direct64 ? (const void*)cx.p : lc.x.p, lv.mask.p, lv.x.p, f64, direct64, s);
    // post-smoothing
    rho0 = 1.0 / sigma;
    for (int k = 0; k < npost; k++) {
        double c0 = 0.0, c1 = 1.0 / theta;
        if (k > 0) {
            const double rho1 = 1.0 / (2.0 * sigma - rho0);
            c0 = rho1 * rho0;
            c1 = 2.0 * rho1 / delta;
            rho0 = rho1;
        }
        if (z64 && k == npost - 1) {  // last step: z = x_new in FP64 plus the flexible-CG dots
            launch_cheb_last(lv.op, cK, cM, lv.b.p, lv.x.p, k > 0 ? lv.d.p : nullptr, lv.dinv.p, c0, c1, z64, r64,
                             rp64, first, f64, ctx->red, s);
            break;
        }
        launch_cheb_step(lv.op, cK, cM, lv.b.p, lv.x.p, k > 0 ? lv.d.p : nullptr, lv.dinv.p, c0, c1, lv.x2.p,
                         lv.d2.p, f64, s);
        std::swap(lv.x.p, lv.x2.p);
        std::swap(lv.d.p, lv.d2.p);
    }
}

void Multigrid::apply_fcg(const double* r, const double* rp, double* z, bool first) {
    finish_inverse();
    if (L.size() == 1) {
        apply(r, z);
        launch_fcg_dots(z, r, rp, L[0]->g.nd, ctx->red, first, ctx->stream);
        return;
    }
    vcycle(0, r, z, rp, first);
}

void Multigrid::apply(const double* r, double* z) {
    finish_inverse();
    MGLevel& l0 = *L[0];
    hipStream_t s = ctx->stream;
    const long nd = l0.g.nd;
    launch_copy(r, l0.b.p, nd, true, l0.fp64, s);
    vcycle(0);
    launch_copy(l0.x.p, z, nd, l0.fp64, true, s);
}

// ------------------------------------------------------------------ context-level Krylov
void Context::sync() { TT_HIP(hipStreamSynchronize(stream)); }

void Context::wait_look() {
    TT_HIP(hipEventRecord(look_ev, stream));
    for (;;) {
        const hipError_t e = hipEventQuery(look_ev);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) TT_HIP(e);
    }
}

double Context::read_scal(int slot) {
    TT_HIP(hipMemcpyAsync(red.host + slot, red.scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, stream));
    TT_HIP(hipStreamSynchronize(stream));
    return red.host[slot];
}

double Context::reduce1(int region) {
    double v[4];
    red.host_sum(region, stream, v);
    return v[0];
}

OpLevel Context::fine_op(bool fp64) const {
    OpLevel o = op_fine64;
    if (!fp64) {
        o.coef_kind = COEF_F32;
        o.kf = kf32.p;
        o.mf = mf32.p;
    }
    return o;
}

SolveResult Context::fcg(double cK, double cM, bool use_mg, SolveKind kind, const double* b, double* x) {
    const long nd = g.nd;
    const OpLevel& op = op_fine64;
    const double rtol = prm.rtol;
    // r = b - A x and the partials of b.b and r.r; the device derives the tolerance, the b = 0 case
    // and an already converged x from them, so the loop starts without a host round trip.
    launch_residual_dots(op, cK, cM, b, x, kr.p, red, PB_MISC, PB_MISC2, stream);
    launch_fcg_init(red, PB_MISC, PB_MISC2, rtol * rtol, true, stream);
    // The residual is read back (a host round trip, ~25 us of idle GPU when the solve goes on) only
    // from one iteration before the fewest any of the last kHintLen solves of this kind took.  A look
    // that comes too late costs idle iterations instead (~0.4 ms each at 128x64x64), never a
    // different answer: the device freezes x at the first iteration whose recursive residual meets
    // the tolerance (k_fcg_alpha).  Measured on the 128x64x64 forward sweep (2-3 iterations per
    // solve, a few solves converging after one): looking first at the fewest instead of one before
    // it ran 3 % slower.
    int* hist = fcg_hist[kind];
    int first_look = 1 << 30;
    for (int k = 0; k < kHintLen; k++) first_look = std::min(first_look, hist[k]);
    first_look = std::max(first_look - 1, 1);
    // The recursive residual of this FP64 recursion drifts from b - A x by rounding only (~1e-14 of
    // ||b|| here); the true residual is confirmed when the tolerance comes within reach of that.
    const bool confirm = rtol < 1e-9;
    SolveResult res;
    stats.solves++;
    double rel = 0.0, bb = 0.0;
    int moved = 0;  // iterations that moved x, as the device counted them at the last look
    bool first = true, conv = false;
    double* pold = kp0.p;
    double* pnew = kp1.p;
    double* qold = kq2.p;
    double* qnew = kq.p;
    // A multigrid-preconditioned iteration is replayed as a HIP graph; its solution vector is then
    // the fixed kx (copied in and out) so that one graph serves every solve.
    const bool graphed = use_mg && prm.graphs;
    double* const xs = graphed ? kx.p : x;
    if (graphed) launch_copy(x, xs, nd, true, true, stream);
    auto iteration = [&]() {
        if (use_mg) {
            mg.apply_fcg(kr.p, krp.p, kz.p, first);  // z = B r with the dots z.r, z.r_prev
        } else {
            launch_jacobi(dinvM.p, kr.p, kz.p, nd, stream);
            launch_fcg_dots(kz.p, kr.p, krp.p, nd, red, first, stream);
        }
        launch_fcg_apply(op, cK, cM, kz.p, pold, pnew, qold, qnew, red, first, stream);
        launch_fcg_update(xs, pnew, kr.p, krp.p, qnew, nd, red, first, stream);
    };
    int it = 0;
    while (!conv && it < prm.maxit) {
        if (graphed) {
            const std::vector<char**> rot = mg.rotating();
            std::vector<uintptr_t> key = {(uintptr_t)first, (uintptr_t)pold, (uintptr_t)pnew, (uintptr_t)qold,
                                          (uintptr_t)qnew};
            uint64_t cb[2];
            std::memcpy(cb, &cK, 8);
            std::memcpy(cb + 1, &cM, 8);
            key.push_back((uintptr_t)cb[0]);
            key.push_back((uintptr_t)cb[1]);
            for (char** q : rot) key.push_back((uintptr_t)*q);
            auto gi = iter_graphs.find(key);
            if (gi == iter_graphs.end()) {
                TT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                try {
                    iteration();
                } catch (...) {
                    hipGraph_t dropped = nullptr;
                    (void)hipStreamEndCapture(stream, &dropped);
                    if (dropped) (void)hipGraphDestroy(dropped);
                    throw;
                }
                hipGraph_t gr = nullptr;
                TT_HIP(hipStreamEndCapture(stream, &gr));
                IterGraph ig;
                const hipError_t e = hipGraphInstantiate(&ig.exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                TT_HIP(e);
                for (char** q : rot) ig.after.push_back(*q);  // rotation done by the captured launches
                std::memcpy(ig.nblk, red.nblk, sizeof(ig.nblk));
                std::memcpy(ig.nv, red.nv, sizeof(ig.nv));
                gi = iter_graphs.emplace(std::move(key), ig).first;
            } else {
                for (size_t k = 0; k < rot.size(); k++) *rot[k] = gi->second.after[k];
                std::memcpy(red.nblk, gi->second.nblk, sizeof(red.nblk));
                std::memcpy(red.nv, gi->second.nv, sizeof(red.nv));
            }
            TT_HIP(hipGraphLaunch(gi->second.exec, stream));
        } else {
            iteration();
        }
        std::swap(pold, pnew);
        std::swap(qold, qnew);
        first = false;
        it++;
        if (it < first_look) continue;
        // r.r partials, the device's count of iterations that moved x and b.b, as the kernels stored
        // them into mapped host memory: no copy, just the wait
        double rr = 0.0;
        wait_look();
        {
            const volatile double* lk = red.look;
            const int nb = red.nblk[PB_RR];
            for (int i = 0; i < nb; i++) rr += lk[i];
            moved = (int)lk[nb];
            bb = lk[Reducer::kLookBB];
        }
        if (!std::isfinite(rr) || !std::isfinite(bb))
            throw Error{TT_ERR_NONFINITE, "non-finite right-hand side or residual in flexible CG"};
        if (bb == 0.0) {  // b = 0: the device has set x = 0
            rel = 0.0;
            conv = true;
            break;
        }
        const double tol2 = rtol * rtol * bb;
        rel = std::sqrt(rr / bb);
        // (the slack covers the device's different summation order of the same partials, so a
        // solve the device has frozen is always accepted here)
        if (rr <= tol2 * (1.0 + 1e-10)) {
            if (!confirm) {
                conv = true;
                break;
            }
            // confirm with the true residual; restart the recursion if it drifted
            launch_residual(op, cK, cM, b, xs, kr.p, true, stream);
            launch_dot(kr.p, kr.p, nd, red, PB_MISC2, stream);
            rel = std::sqrt(reduce1(PB_MISC2) / bb);
            if (rel <= rtol) {
                conv = true;
            } else {
                first = true;
                launch_fcg_init(red, PB_MISC, PB_MISC2, rtol * rtol, false, stream);
            }
        }
    }
    if (graphed) launch_copy(xs, x, nd, true, true, stream);
    // the count of the device, not `it`: a solve frozen before the first look would otherwise keep
    // the first look where it is even when later solves need fewer iterations
    for (int k = kHintLen - 1; k > 0; k--) hist[k] = hist[k - 1];
    hist[0] = conv ? std::max(moved, 1) : it;
    stats.krylov_iters += it;
    stats.last_iters = it;
    stats.last_relres = rel;
    stats.max_relres = std::max(stats.max_relres, rel);
    res.iters = it;
    res.relres = rel;
    if (!conv) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "flexible CG did not reach rtol %.3g in %d iterations (relres %.3g)", rtol,
                      prm.maxit, rel);
        throw Error{TT_ERR_NOT_CONVERGED, msg};
    }
    return res;
}

SolveResult Context::solve_A(const double* b, double* x, SolveKind kind) { return fcg(cK_A, cM_A, true, kind, b, x); }
SolveResult Context::solve_M(const double* b, double* x) { return fcg(0.0, 1.0, false, kSolveMass, b, x); }

}  // namespace tt
