// HIP kernels for gfx950: element sensitivities and objective, PDE (Helmholtz)
// filter, Heaviside projection, and the element-wise parts of MMA.
#include "../../include/topopt_gpu.h"
#include "tt_internal.h"

namespace tt {

constexpr int BS2 = 256;
static inline int nblk(long n, int per = BS2) { return (int)((n + per - 1) / per); }
static inline int rblk(long n) {
    int b = nblk(n);
    return b > 256 ? 256 : (b < 1 ? 1 : b);
}

// ---- deterministic reduction (same protocol as tt_kernels.hip)
template <int NV>
__device__ __forceinline__ void bsum(double (&v)[NV], double* sh) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[wid * NV + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w * NV + k];
            v[k] = s;
        }
    }
}
// block partial sums -> partials[block*NV + k]; summed later in block order by the host
// (Reducer::host_sum) or by the next kernel
template <int NV>
__device__ __forceinline__ void rfinish(double (&v)[NV], double* partials, unsigned*, double*) {
    __shared__ double sh[16 * NV];
    bsum<NV>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(size_t)blockIdx.x * NV + k] = v[k];
    }
}
static inline void set_region(Reducer& red, int region, int nb, int nv) {
    red.nblk[region] = nb;
    red.nv[region] = nv;
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

struct GA {
    int nx, ny, nz, ex, ey, ez;
};
static inline GA ga(const Grid& g) { return GA{g.nx, g.ny, g.nz, g.ex, g.ey, g.ez}; }

__device__ __forceinline__ void elem_nodes(const GA& g, long e, long (&nd)[8]) {
    int i, j, k;
    grid_ijk((unsigned)e, g.ex, g.ey, i, j, k);
    const long n0 = i + (long)g.nx * (j + (long)g.ny * k), sy = g.nx, sz = (long)g.nx * g.ny;
    nd[0] = n0;
    nd[1] = n0 + 1;
    nd[2] = n0 + 1 + sy;
    nd[3] = n0 + sy;
    nd[4] = n0 + sz;
    nd[5] = n0 + 1 + sz;
    nd[6] = n0 + 1 + sy + sz;
    nd[7] = n0 + sy + sz;
}

// ------------------------------------------------------------------ sensitivities
// Per element and time step (ComputeSensitivitiesElementLoopn, LinearElasticity.cc:1911-2077):
//   df_e += P_e*dfx0 + lambda_e^T (dA u_t + dB1 u_{t-1} + dB2 v_{t-1} + dB3 a_{t-1})
// which factorises into dkfac * lambda^T KE w_K + dmfac * lambda^T ME w_M with
// w_K, w_M linear combinations of (u_t, u_{t-1}, v_{t-1}, a_{t-1}) given by wK/wM.
// mode 1 is the t = 0 loop (ComputeSensitivitiesElementLoop0, :2080-2204) with
// the slots (u0, v0, a0): w_K = u0 + dbeta v0, w_M = dalpha v0 + a0.
struct SensDev {
    GA g;
    const double* KE;
    const double* M8;
    const double *kf, *mf, *dkf, *dmf, *P;
    const double *lam, *u, *v, *a, *s1, *s2, *s3;
    double wK[4], wM[4];
    double dt, al, be;
    int objective;
    double* df;
};

// On the matrix cores (v_mfma_f64_16x16x4f64).  A wave takes batches of 16
// consecutive elements, one per MFMA column, and forms per batch
//     KE w_K and ME w_M, plus the objective's element products: KE u (objectives 3, 5), ME v (4-6),
//     KE v (6)   (each: 2 row tiles x 6 k-steps of 24 x 24 by 24 x 16).
// The element's 24 dofs are taken in the order sens_dof(k): lane group q owns the two x-adjacent
// corners of one element edge (corners 0-1, 3-2, 4-5, 7-6), whose 6 dofs are consecutive in memory
// (3 wide loads per vector instead of 6 scattered ones), as k = 4 t + q, t = 0..5.  Operand layout:
// lane j + 16 q holds the B value of k row q of k-step t, i.e. sens_dof(4 t + q) of element j, and
// receives D rows 4 v + q of each row tile (measured on gfx950), i.e. k = 16 R + 4 v + q: the same
// six dofs.  The A fragments hold KE and ME with rows and columns in the same order, so every lane
// loads just its six values of each input vector, and lambda^T KE w_K, lambda^T ME w_M and the
// objective's quadratic forms (u^T u, v^T v, a^T a for objectives 0-2) are sums over the four lane
// groups (two shuffles, in a fixed order).  The element matrices stay in registers as A fragments.
// Measured at 128x64x64 (objective 3): 72 us per launch (148 VGPRs, 3 waves per SIMD), against 82 us
// (204 VGPRs) with the dofs = q (mod 4) in lane group q, which takes 6 scattered 8-byte loads per vector,
// and 64-bit element index arithmetic (75 us with the contiguous loads alone).  sens_dof is in
// tt_internal.h.
typedef double sens_d4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double lane_group_sum(double x) {
    x += __shfl_xor(x, 16, 64);
    return x + __shfl_xor(x, 32, 64);
}
template <int OBJ>
__global__ __launch_bounds__(256) void k_sens_mfma(SensDev s, double* partials) {
    constexpr bool KU = OBJ == 3 || OBJ == 5, MV = OBJ >= 4, KV = OBJ == 6;
    constexpr bool LV = OBJ == 1 || OBJ >= 4, LA = OBJ == 2;  // v or a of state t needed
    const int lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4, wave = threadIdx.x >> 6;
    const long nel = (long)s.g.ex * s.g.ey * s.g.ez;
    // A fragments: lane j + 16 q holds row j, k column q of each 16 x 4 block (rows 24..31 zero)
    double aK[2][6], aM[2][6];
#pragma unroll
    for (int R = 0; R < 2; R++)
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const int r = 16 * R + j, rd = sens_dof(r < 24 ? r : 23), cd = sens_dof(4 * t + q);
            const double k = s.KE[rd * 24 + cd], m = s.M8[(rd / 3) * 8 + cd / 3];
            aK[R][t] = r < 24 ? k : 0.0;
            aM[R][t] = (r < 24 && rd % 3 == cd % 3) ? m : 0.0;
        }
    const long sy = s.g.nx, sz = (long)s.g.nx * s.g.ny;
    const double hdt = 0.5 * s.dt;
    double h[1] = {0.0};
#pragma unroll 1
    for (int b = 0; b < 4; b++) {
        const long e = (long)blockIdx.x * 256 + (wave * 4 + b) * 16 + j;
        const bool valid = e < nel;
        const long ec = valid ? e : nel - 1;
        // element -> (i, jj, kk) in 32-bit unsigned arithmetic (two divisions; element and node
        // indices fit in 32 bits, as everywhere in the apply kernels)
        const unsigned eu = (unsigned)ec, exu = (unsigned)s.g.ex, eyu = (unsigned)s.g.ey;
        const unsigned row = eu / exu, kk = row / eyu, i = eu - row * exu, jj = row - kk * eyu;
        const long n0 = (long)i + sy * (long)jj + sz * (long)kk;
        // the low-x node of the lane group's edge: dofs sens_dof(4 t + q), t = 0..5, at x0 + t
        const long x0 = 3 * (n0 + sy * (q & 1) + sz * (q >> 1));
        double lam[6], u[6], wk[6], wm[6], w2[6];
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const long x = x0 + t;
            const double uu = s.u[x], x1 = s.s1[x], x2 = s.s2[x], x3 = s.s3[x];
            lam[t] = s.lam[x];
            u[t] = uu;
            wk[t] = fma(s.wK[3], x3, fma(s.wK[2], x2, fma(s.wK[1], x1, s.wK[0] * uu)));
            wm[t] = fma(s.wM[3], x3, fma(s.wM[2], x2, fma(s.wM[1], x1, s.wM[0] * uu)));
            w2[t] = LV ? s.v[x] : LA ? s.a[x] : 0.0;  // the state-t vector of the objective
        }
        sens_d4 accK[2], accM[2], accU[2], accVM[2], accVK[2];
#pragma unroll
        for (int R = 0; R < 2; R++) accK[R] = accM[R] = accU[R] = accVM[R] = accVK[R] = sens_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 6; t++)
#pragma unroll
            for (int R = 0; R < 2; R++) {
                accK[R] = __builtin_amdgcn_mfma_f64_16x16x4f64(aK[R][t], wk[t], accK[R], 0, 0, 0);
                accM[R] = __builtin_amdgcn_mfma_f64_16x16x4f64(aM[R][t], wm[t], accM[R], 0, 0, 0);
                if (KU) accU[R] = __builtin_amdgcn_mfma_f64_16x16x4f64(aK[R][t], u[t], accU[R], 0, 0, 0);
                if (MV) accVM[R] = __builtin_amdgcn_mfma_f64_16x16x4f64(aM[R][t], w2[t], accVM[R], 0, 0, 0);
                if (KV) accVK[R] = __builtin_amdgcn_mfma_f64_16x16x4f64(aK[R][t], w2[t], accVK[R], 0, 0, 0);
            }
        // D row 4 v + q of row tile R is k = 4 (4 R + v) + q: the lane's value t = 4 R + v (t < 6)
        double tK = 0.0, tM = 0.0, fU = 0.0, fVM = 0.0, fVK = 0.0;
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const int R = t >> 2, v = t & 3;
            tK = fma(accK[R][v], lam[t], tK);
            tM = fma(accM[R][v], lam[t], tM);
            if (KU) fU = fma(accU[R][v], u[t], fU);
            if (OBJ == 0) fU = fma(u[t], u[t], fU);
            if (OBJ == 1 || OBJ == 2) fU = fma(w2[t], w2[t], fU);
            if (MV) fVM = fma(accVM[R][v], w2[t], fVM);
            if (KV) fVK = fma(accVK[R][v], w2[t], fVK);
        }
        tK = lane_group_sum(tK);
        tM = lane_group_sum(tM);
        fU = lane_group_sum(fU);
        if (MV) fVM = lane_group_sum(fVM);
        if (KV) fVK = lane_group_sum(fVK);
        if (valid && q == 0) {
            const double dkf = s.dkf[e], dmf = s.dmf[e], kf = s.kf[e], mf = s.mf[e], P = s.P[e];
            // objective terms of state t (EvaluateObjective / EvaluateObjectiveGradient, :1416-1514)
            double h0 = 0.0, dfx0 = 0.0;
            if (OBJ <= 2) h0 = hdt * fU;
            if (OBJ == 3) { h0 = hdt * (kf * fU); dfx0 = hdt * (dkf * fU); }
            if (OBJ == 4) { h0 = hdt * (mf * fVM); dfx0 = hdt * (dmf * fVM); }
            if (OBJ == 5) { h0 = hdt * (kf * fU + mf * fVM); dfx0 = hdt * (dkf * fU + dmf * fVM); }
            if (OBJ == 6) {
                h0 = hdt * (s.al * mf * fVM + s.be * kf * fVK);
                dfx0 = hdt * (s.al * dmf * fVM + s.be * dkf * fVK);
            }
            s.df[e] += P * dfx0 + (dkf * tK + dmf * tM);
            h[0] += h0 * P;
        }
    }
    rfinish<1>(h, partials, nullptr, nullptr);
}

void launch_sens(const SensArgs& a, Reducer& red, double* fx_out, hipStream_t st) {
    SensDev s;
    s.g = ga(a.g);
    s.KE = a.KE;
    s.M8 = a.M8;
    s.kf = a.kf; s.mf = a.mf; s.dkf = a.dkf; s.dmf = a.dmf; s.P = a.P;
    s.lam = a.lam;
    s.v = a.v;
    s.a = a.a;
    s.dt = a.dt; s.al = a.dalpha; s.be = a.dbeta;
    s.objective = a.objective;
    s.df = a.df;
    if (a.mode == 0) {
        s.u = a.u; s.s1 = a.ub; s.s2 = a.vb; s.s3 = a.ab;
    } else {
        s.u = a.u; s.s1 = a.v; s.s2 = a.a; s.s3 = a.a;
    }
    for (int k = 0; k < 4; k++) { s.wK[k] = a.wK[k]; s.wM[k] = a.wM[k]; }
    // fx_out receives one partial objective value per block (nblk(nel) of them)
    const int nb = nblk(a.g.nel);
    (void)red;
    static_assert(BS2 == 256, "k_sens_mfma: one block per 256 elements");
    switch (a.objective) {
    case 0: hipLaunchKernelGGL(k_sens_mfma<0>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    case 1: hipLaunchKernelGGL(k_sens_mfma<1>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    case 2: hipLaunchKernelGGL(k_sens_mfma<2>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    case 3: hipLaunchKernelGGL(k_sens_mfma<3>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    case 4: hipLaunchKernelGGL(k_sens_mfma<4>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    case 5: hipLaunchKernelGGL(k_sens_mfma<5>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    default: hipLaunchKernelGGL(k_sens_mfma<6>, dim3(nb), dim3(256), 0, st, s, fx_out); break;
    }
    TT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ PDE filter (PDEFilter.cc)
// y_n = sum_{e at n} sum_b KF[a][b] U_{node(e,b)}  (constant-coefficient scalar hex8 operator)
__global__ void k_filter_apply(GA g, const double* __restrict__ KF, const double* __restrict__ u,
                               double* __restrict__ y, double* partials, unsigned* counter, double* out) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long nn = (long)g.nx * g.ny * g.nz;
    double acc[1] = {0.0};
    if (n < nn) {
        int i, j, k;
        grid_ijk((unsigned)n, g.nx, g.ny, i, j, k);
        double s = 0.0;
#pragma unroll
        for (int ez = 0; ez < 2; ez++)
#pragma unroll
            for (int ey = 0; ey < 2; ey++)
#pragma unroll
                for (int ex = 0; ex < 2; ex++) {
                    const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
                    if (ei < 0 || ej < 0 || ek < 0 || ei >= g.ex || ej >= g.ey || ek >= g.ez) continue;
                    const int a = (1 - ez) * 4 + ((1 - ey) == 0 ? ((1 - ex) == 0 ? 0 : 1) : ((1 - ex) == 1 ? 2 : 3));
#pragma unroll
                    for (int b = 0; b < 8; b++) {
                        const int bx = ((b & 3) == 1 || (b & 3) == 2), by = (b & 3) >= 2, bz = b >> 2;
                        const long m = (ei + bx) + (long)g.nx * ((ej + by) + (long)g.ny * (ek + bz));
                        s = fma(KF[a * 8 + b], u[m], s);
                    }
                }
        y[n] = s;
        acc[0] = s * u[n];
    }
    if (partials) rfinish<1>(acc, partials, counter, out);
}
void launch_filter_apply(const Grid& g, const double* kf, const double* u, double* y, Reducer* red, int dot_slot,
                         hipStream_t s) {
    const int nb = nblk(g.nn);
    if (red) {
        TT_CHECK(nb <= red->maxblocks, -3, "reducer too small for filter apply");
        hipLaunchKernelGGL(k_filter_apply, dim3(nb), dim3(BS2), 0, s, ga(g), kf, u, y, red->part(dot_slot),
                           (unsigned*)nullptr, (double*)nullptr);
        set_region(*red, dot_slot, nb, 1);
    } else {
        hipLaunchKernelGGL(k_filter_apply, dim3(nb), dim3(BS2), 0, s, ga(g), kf, u, y, (double*)nullptr,
                           (unsigned*)nullptr, (double*)nullptr);
    }
    TT_HIP(hipGetLastError());
}

// rhs_n = scale * sum_{e at n} x_e / 8   (T x, PDEFilter.cc:580-588, scaled by the element volume at :208)
__global__ void k_filter_T(GA g, const double* __restrict__ x, double scale, double* __restrict__ rhs) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    grid_ijk((unsigned)n, g.nx, g.ny, i, j, k);
    double s = 0.0;
    for (int ez = 0; ez < 2; ez++)
        for (int ey = 0; ey < 2; ey++)
            for (int ex = 0; ex < 2; ex++) {
                const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
                if (ei < 0 || ej < 0 || ek < 0 || ei >= g.ex || ej >= g.ey || ek >= g.ez) continue;
                s += 0.125 * x[ei + (long)g.ex * (ej + (long)g.ey * ek)];
            }
    rhs[n] = s * scale;
}
void launch_filter_T(const Grid& g, const double* x, double scale, double* rhs, hipStream_t s) {
    hipLaunchKernelGGL(k_filter_T, dim3(nblk(g.nn)), dim3(BS2), 0, s, ga(g), x, scale, rhs);
    TT_HIP(hipGetLastError());
}
// out_e = sum_{n in e} U_n / 8  (T' U, PDEFilter.cc:216)
__global__ void k_filter_Tt(GA g, const double* __restrict__ U, double* __restrict__ out) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)g.ex * g.ey * g.ez) return;
    long nd[8];
    elem_nodes(g, e, nd);
    double s = 0.0;
    for (int a = 0; a < 8; a++) s += 0.125 * U[nd[a]];
    out[e] = s;
}
void launch_filter_Tt(const Grid& g, const double* U, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_filter_Tt, dim3(nblk(g.nel)), dim3(BS2), 0, s, ga(g), U, out);
    TT_HIP(hipGetLastError());
}
__global__ void k_filter_diag(GA g, const double* __restrict__ KF, double* __restrict__ dinv) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)g.nx * g.ny * g.nz) return;
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / ((long)g.nx * g.ny);
    double s = 0.0;
    for (int ez = 0; ez < 2; ez++)
        for (int ey = 0; ey < 2; ey++)
            for (int ex = 0; ex < 2; ex++) {
                const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
                if (ei < 0 || ej < 0 || ek < 0 || ei >= g.ex || ej >= g.ey || ek >= g.ez) continue;
                const int a = (1 - ez) * 4 + ((1 - ey) == 0 ? ((1 - ex) == 0 ? 0 : 1) : ((1 - ex) == 1 ? 2 : 3));
                s += KF[a * 9];
            }
    dinv[n] = 1.0 / s;
}
void launch_filter_diag(const Grid& g, const double* kf, double* dinv, hipStream_t s) {
    hipLaunchKernelGGL(k_filter_diag, dim3(nblk(g.nn)), dim3(BS2), 0, s, ga(g), kf, dinv);
    TT_HIP(hipGetLastError());
}

// clamp to [0,1] (Filter.cc:83-102) and Heaviside projection (Filter.h:80-83)
__global__ void k_clamp_project(const double* __restrict__ in, double* __restrict__ xt, double* __restrict__ xp,
                                long n, int projection, double beta, double eta) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = in[i];
    if (x < 0.0) x = 0.0;
    if (x > 1.0) x = 1.0;
    xt[i] = x;
    xp[i] = projection ? (tanh(beta * eta) + tanh(beta * (x - eta))) / (tanh(beta * eta) + tanh(beta * (1.0 - eta)))
                       : x;
}
void launch_clamp_project(const double* xt_in, double* xt, double* xp, long n, int projection, double beta,
                          double eta, hipStream_t s) {
    hipLaunchKernelGGL(k_clamp_project, dim3(nblk(n)), dim3(BS2), 0, s, xt_in, xt, xp, n, projection, beta, eta);
    TT_HIP(hipGetLastError());
}
// chain rule of the projection (Filter.h:85-88, Filter.cc:127-160)
__global__ void k_project_chain(const double* __restrict__ xt, double* __restrict__ df, double* __restrict__ dg,
                                long n, double beta, double eta) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d =
        beta * (1.0 - pow(tanh(beta * (xt[i] - eta)), 2.0)) / (tanh(beta * eta) + tanh(beta * (1.0 - eta)));
    df[i] = df[i] * d;
    if (dg) dg[i] = dg[i] * d;
}
void launch_project_chain(const double* xt, double* df, double* dg, long n, double beta, double eta, hipStream_t s) {
    hipLaunchKernelGGL(k_project_chain, dim3(nblk(n)), dim3(BS2), 0, s, xt, df, dg, n, beta, eta);
    TT_HIP(hipGetLastError());
}

__global__ void k_sum(const double* __restrict__ x, long n, int mnd, double* partials, unsigned* counter,
                      double* out) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc[0] += mnd ? 4 * x[i] * (1.0 - x[i]) : x[i];
    rfinish<1>(acc, partials, counter, out);
}
void launch_sum(const double* x, long n, Reducer& red, int region, hipStream_t s) {
    hipLaunchKernelGGL(k_sum, dim3(rblk(n)), dim3(BS2), 0, s, x, n, 0, red.part(region), (unsigned*)nullptr,
                       (double*)nullptr);
    set_region(red, region, rblk(n), 1);
    TT_HIP(hipGetLastError());
}
int launch_sum_partials(const double* x, long n, double* partials, int capacity, hipStream_t s) {
    const int nb = rblk(n);
    TT_CHECK(nb <= capacity, TT_ERR_INTERNAL, "sum: more block partials than the partial buffer holds");
    hipLaunchKernelGGL(k_sum, dim3(nb), dim3(BS2), 0, s, x, n, 0, partials, (unsigned*)nullptr, (double*)nullptr);
    TT_HIP(hipGetLastError());
    return nb;
}
void launch_mnd(const double* x, long n, Reducer& red, int region, hipStream_t s) {
    hipLaunchKernelGGL(k_sum, dim3(rblk(n)), dim3(BS2), 0, s, x, n, 1, red.part(region), (unsigned*)nullptr,
                       (double*)nullptr);
    set_region(red, region, rblk(n), 1);
    TT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ MMA (MMA.cc), m = 1
struct MD {
    long n;
    double *L, *U, *alpha, *beta, *p0, *q0, *pij, *qij, *xo1, *xo2;
};
static inline MD md(const MMADev& m) { return MD{m.n, m.L, m.U, m.alpha, m.beta, m.p0, m.q0, m.pij, m.qij, m.xo1, m.xo2}; }

// GenSub (MMA.cc:487-600) element part; b partial sum -> out[0]; also xo2 = xo1, xo1 = x (Update, :506-509)
__global__ void k_mma_gensub(MD m, const double* __restrict__ x, const double* __restrict__ df,
                             const double* __restrict__ dg, const double* __restrict__ xmin,
                             const double* __restrict__ xmax, int k, double ainit, double adec, double ainc,
                             double* partials, unsigned* counter, double* out) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m.n; i += (long)gridDim.x * blockDim.x) {
        const double xv = x[i];
        double L, U;
        if (k < 3) {
            L = xv - ainit * xmax[i];
            L += ainit * xmin[i];
            U = xv + ainit * xmax[i];
            U += -ainit * xmin[i];
        } else {
            L = m.L[i];
            U = m.U[i];
            const double x1 = m.xo1[i], x2 = m.xo2[i];
            const double hv = (xv - x1) * (x1 - x2);
            const double gam = hv < 0.0 ? adec : (hv > 0.0 ? ainc : 1.0);
            L = xv - gam * (x1 - L);
            U = xv + gam * (U - x1);
            const double xmi = fmax(1.0e-5, xmax[i] - xmin[i]);
            L = fmax(L, xv - 10.0 * xmi);
            L = fmin(L, xv - 0.01 * xmi);
            U = fmax(U, xv + 0.01 * xmi);
            U = fmin(U, xv + 10.0 * xmi);
        }
        m.L[i] = L;
        m.U[i] = U;
        m.alpha[i] = fmax(xmin[i], 0.9 * L + 0.1 * xv);
        m.beta[i] = fmin(xmax[i], 0.9 * U + 0.1 * xv);
        const double feps = 1.0e-6, d = df[i];
        const double dfp = fmax(0.0, d), dfm = fmax(0.0, -1.0 * d);
        m.p0[i] = pow(U - xv, 2.0) * (dfp + 0.001 * fabs(d) + 0.5 * feps / (U - L));
        m.q0[i] = pow(xv - L, 2.0) * (dfm + 0.001 * fabs(d) + 0.5 * feps / (U - L));
        const double g = dg[i];
        const double pij = pow(U - xv, 2.0) * fmax(0.0, g);
        const double qij = pow(xv - L, 2.0) * fmax(0.0, -1.0 * g);
        m.pij[i] = pij;
        m.qij[i] = qij;
        acc[0] += pij / (U - xv) + qij / (xv - L);
        m.xo2[i] = m.xo1[i];
        m.xo1[i] = xv;
    }
    rfinish<1>(acc, partials, counter, out);
}
void launch_mma_gensub(const MMADev& m, const double* x, const double* df, const double* dg, const double* xmin,
                       const double* xmax, int k, double asyminit, double asymdec, double asyminc, Reducer& red,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_mma_gensub, dim3(rblk(m.n)), dim3(BS2), 0, s, md(m), x, df, dg, xmin, xmax, k, asyminit,
                       asymdec, asyminc, red.part(PB_MISC), (unsigned*)nullptr, (double*)nullptr);
    set_region(red, PB_MISC, rblk(m.n), 1);
    TT_HIP(hipGetLastError());
}

// XYZofLAMBDA (MMA.cc:684-727) x part, fused with DualGrad (:729-765) and DualHess (:767-840) sums
__global__ void k_mma_xgh(MD m, double lam, double* __restrict__ x, double* partials, unsigned* counter, double* out) {
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m.n; i += (long)gridDim.x * blockDim.x) {
        const double pj = m.p0[i] + m.pij[i] * lam, qj = m.q0[i] + m.qij[i] * lam;
        const double L = m.L[i], U = m.U[i];
        const double xp = (sqrt(pj) * L + sqrt(qj) * U) / (sqrt(pj) + sqrt(qj));
        double xv = xp;
        if (xv < m.alpha[i]) xv = m.alpha[i];
        if (xv > m.beta[i]) xv = m.beta[i];
        x[i] = xv;
        acc[0] += m.pij[i] / (U - xv) + m.qij[i] / (xv - L);
        const double PQ = m.pij[i] / pow(U - xv, 2.0) - m.qij[i] / pow(xv - L, 2.0);
        double df2 = -1.0 / (2.0 * pj / pow(U - xv, 3.0) + 2.0 * qj / pow(xv - L, 3.0));
        if (xp < m.alpha[i]) df2 = 0.0;
        if (xp > m.beta[i]) df2 = 0.0;
        acc[1] += (PQ * df2) * PQ;
    }
    rfinish<2>(acc, partials, counter, out);
}
void launch_mma_xyz_grad_hess(const MMADev& m, double lam, double* x, Reducer& red, hipStream_t s) {
    hipLaunchKernelGGL(k_mma_xgh, dim3(rblk(m.n)), dim3(BS2), 0, s, md(m), lam, x, red.part(PB_MISC),
                       (unsigned*)nullptr, (double*)nullptr);
    set_region(red, PB_MISC, rblk(m.n), 2);
    TT_HIP(hipGetLastError());
}
// XYZofLAMBDA + DualResidual sum (MMA.cc:893-934)
__global__ void k_mma_xres(MD m, double lam, double* __restrict__ x, double* partials, unsigned* counter, double* out) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m.n; i += (long)gridDim.x * blockDim.x) {
        const double pj = m.p0[i] + m.pij[i] * lam, qj = m.q0[i] + m.qij[i] * lam;
        const double L = m.L[i], U = m.U[i];
        double xv = (sqrt(pj) * L + sqrt(qj) * U) / (sqrt(pj) + sqrt(qj));
        if (xv < m.alpha[i]) xv = m.alpha[i];
        if (xv > m.beta[i]) xv = m.beta[i];
        x[i] = xv;
        acc[0] += m.pij[i] / (U - xv) + m.qij[i] / (xv - L);
    }
    rfinish<1>(acc, partials, counter, out);
}
void launch_mma_xyz_res(const MMADev& m, double lam, double* x, Reducer& red, hipStream_t s) {
    hipLaunchKernelGGL(k_mma_xres, dim3(rblk(m.n)), dim3(BS2), 0, s, md(m), lam, x, red.part(PB_MISC),
                       (unsigned*)nullptr, (double*)nullptr);
    set_region(red, PB_MISC, rblk(m.n), 1);
    TT_HIP(hipGetLastError());
}

// SetOuterMovelimit (MMA.cc:374-393)
__global__ void k_mma_movelimit(const double* __restrict__ x, double* __restrict__ xmin, double* __restrict__ xmax,
                                long n, double Xmin, double Xmax, double movlim) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i] + movlim, b = x[i] - movlim;
    xmax[i] = Xmax < a ? Xmax : a;
    xmin[i] = Xmin > b ? Xmin : b;
}
void launch_mma_movelimit(const double* x, double* xmin, double* xmax, long n, double Xmin, double Xmax,
                          double movlim, hipStream_t s) {
    hipLaunchKernelGGL(k_mma_movelimit, dim3(nblk(n)), dim3(BS2), 0, s, x, xmin, xmax, n, Xmin, Xmax, movlim);
    TT_HIP(hipGetLastError());
}

// DesignChange (MMA.cc:395-414): max |x - xold| -> out (pre-zeroed), xold = x
__global__ void k_mma_change(const double* __restrict__ x, double* __restrict__ xold, long n, double* out) {
    double m = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        m = fmax(m, fabs(x[i] - xold[i]));
        xold[i] = x[i];
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, m);
}
void launch_mma_change(const double* x, double* xold, long n, Reducer& red, int slot, hipStream_t s) {
    TT_HIP(hipMemsetAsync(red.scal.p + slot, 0, sizeof(double), s));
    hipLaunchKernelGGL(k_mma_change, dim3(rblk(n)), dim3(BS2), 0, s, x, xold, n, red.scal.p + slot);
    TT_HIP(hipGetLastError());
}

// KKTresidual element part (MMA.cc:416-470): sum of squares -> scal[TMP1], max abs -> scal[TMP2]
__global__ void k_mma_kkt(const double* __restrict__ x, const double* __restrict__ df, const double* __restrict__ dg,
                          const double* __restrict__ xmin, const double* __restrict__ xmax, long n, double lam,
                          double* partials, unsigned* counter, double* out, double* outmax) {
    double acc[1] = {0.0};
    double mx = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double ri = df[i] + lam * dg[i];
        double mmin = 0.0, mmax = 0.0;
        if (x[i] < xmin[i] + 1.0e-5 && ri > 0.0) mmin = ri;
        if (x[i] > xmax[i] - 1.0e-5 && ri < 0.0) mmax = -ri;
        ri += -mmin + mmax;
        acc[0] += pow(ri, 2.0);
        mx = fmax(fabs(ri), mx);
        double r2 = mmin * (x[i] - xmin[i]);
        acc[0] += pow(r2, 2.0);
        mx = fmax(fabs(r2), mx);
        r2 = mmax * (xmax[i] - x[i]);
        acc[0] += pow(r2, 2.0);
        mx = fmax(fabs(r2), mx);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(outmax, mx);
    rfinish<1>(acc, partials, counter, out);
}
void launch_mma_kkt(const double* x, const double* df, const double* dg, const double* xmin, const double* xmax,
                    long n, double lam, Reducer& red, hipStream_t s) {
    TT_HIP(hipMemsetAsync(red.scal.p + SC_TMP2, 0, sizeof(double), s));
    hipLaunchKernelGGL(k_mma_kkt, dim3(rblk(n)), dim3(BS2), 0, s, x, df, dg, xmin, xmax, n, lam, red.part(PB_MISC),
                       (unsigned*)nullptr, (double*)nullptr, red.scal.p + SC_TMP2);
    set_region(red, PB_MISC, rblk(n), 1);
    TT_HIP(hipGetLastError());
}

__global__ void k_lincomb(double a, const double* x, double b, const double* y, double* out, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + b * y[i];
}
void launch_lincomb(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb, dim3(nblk(n)), dim3(BS2), 0, s, a, x, b, y, out, n);
    TT_HIP(hipGetLastError());
}
__global__ void k_lincomb3(double a, const double* x, double b, const double* y, double c, const double* z,
                           double* out, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = fma(c, z[i], fma(b, y[i], a * x[i]));
}
void launch_lincomb3(double a, const double* x, double b, const double* y, double c, const double* z, double* out,
                     long n, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb3, dim3(nblk(n)), dim3(BS2), 0, s, a, x, b, y, c, z, out, n);
    TT_HIP(hipGetLastError());
}
__global__ void k_extrapolate(Extrap e, double* out, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = e.c[0] * e.v[0][i];
    for (int k = 1; k < e.m; k++) s = fma(e.c[k], e.v[k][i], s);
    out[i] = s;
}
void launch_extrapolate(const Extrap& e, double* out, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate, dim3(nblk(n)), dim3(BS2), 0, s, e, out, n);
    TT_HIP(hipGetLastError());
}

}  // namespace tt
