// HIP kernels for gfx950: matrix-free SIMP hex8 operator, multigrid transfers,
// Krylov vector kernels and deterministic reductions.
//
// The operator y = sum_e (cK*kfac_e*KE + cM*mfac_e*ME) u_e of the reference's
// assembled AIJ matrices (LinearElasticity.cc:2502-2619) is applied node by
// node: each thread owns one node and gathers the contributions of its (up to)
// eight elements from an LDS tile holding the node box (8x8x4 owned + halo 1)
// and the element coefficients.  No scatter, no atomics: results are
// bit-reproducible.  Epilogue functors fuse the Krylov / Chebyshev vector work
// into the same pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "tt_internal.h"

namespace tt {

// ------------------------------------------------------------------ helpers
// Element i of p through a 32-bit byte offset from a uniform base (scalar-base addressing; every
// array here is below 4 GiB, which the context checks at creation).
template <typename T>
__device__ __forceinline__ T ld(const T* p, unsigned i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(T));
}
__device__ __forceinline__ double coef_ld(const double* p, unsigned i) { return ld(p, i); }
__device__ __forceinline__ double coef_ld(const float* p, unsigned i) { return (double)ld(p, i); }
__device__ __forceinline__ double coef_ld(const bf16_t* p, unsigned i) { return (double)__uint_as_float(((uint32_t)ld(p, i)) << 16); }
__device__ __forceinline__ double coef_ld(const double* p, long i) { return p[i]; }
__device__ __forceinline__ double coef_ld(const float* p, long i) { return (double)p[i]; }
__device__ __forceinline__ double coef_ld(const bf16_t* p, long i) {
    return (double)__uint_as_float(((uint32_t)p[i]) << 16);
}

constexpr int BS = 256;  // threads per block for all 1-D kernels

// ------------------------------------------------------------------ deterministic block reduction
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
    // sh: BS/64 * NV doubles
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

// Writes the block's NV partial sums to part[block*NV + k] (plain stores).
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* part) {
    __shared__ double sh[16 * NV];
    block_sum<NV>(v, sh);
    if (threadIdx.x == 0) {
        const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
        for (int k = 0; k < NV; k++) part[(size_t)bid * NV + k] = v[k];
    }
}

// Sum of partials written by an earlier kernel; every thread of the block gets the totals.
// Thread t sums blocks t, t + blockDim, ... (a fixed order for a given block size), then the
// waves are combined by a fixed butterfly and a fixed wave order, so every block of a launch
// obtains bit-identical totals.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nblk, double (&out)[NV]) {
    __shared__ double sh[16 * NV];
    // four independent accumulators per thread so that four loads are in flight; fixed order
    double acc[NV], a4[3][NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = a4[0][k] = a4[1][k] = a4[2][k] = 0.0;
    const int st = blockDim.x;
    int b = threadIdx.x;
    for (; b + 3 * st < nblk; b += 4 * st)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            acc[k] += part[(size_t)b * NV + k];
            a4[0][k] += part[(size_t)(b + st) * NV + k];
            a4[1][k] += part[(size_t)(b + 2 * st) * NV + k];
            a4[2][k] += part[(size_t)(b + 3 * st) * NV + k];
        }
    for (; b < nblk; b += st)
#pragma unroll
        for (int k = 0; k < NV; k++) acc[k] += part[(size_t)b * NV + k];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] += (a4[0][k] + a4[1][k]) + a4[2][k];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sh[wid * NV + k] = x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = 0.0;
        for (int w = 0; w < nw; w++) x += sh[w * NV + k];
        out[k] = x;
    }
    __syncthreads();  // sh may be reused by a later call in the same kernel
}

void Reducer::init(int maxb) {
    maxblocks = maxb;
    partials.alloc((size_t)PB_COUNT * maxb * 4);
    scal.alloc(SC_COUNT);
    TT_HIP(hipMemset(partials.p, 0, sizeof(double) * PB_COUNT * maxb * 4));
    TT_HIP(hipMemset(scal.p, 0, sizeof(double) * SC_COUNT));
    TT_HIP(hipHostMalloc((void**)&host, sizeof(double) * SC_COUNT, hipHostMallocDefault));
    TT_HIP(hipHostMalloc((void**)&hpart, sizeof(double) * maxb * 4, hipHostMallocDefault));
    TT_HIP(hipHostMalloc((void**)&look, sizeof(double) * (kLookBB + 1), hipHostMallocMapped | hipHostMallocCoherent));
    TT_HIP(hipHostGetDevicePointer((void**)&look_dev, look, 0));
    for (int i = 0; i <= kLookBB; i++) look[i] = 0.0;
}
void Reducer::host_sum(int region, hipStream_t s, double* out) {
    const int nb = nblk[region], v = nv[region];
    TT_HIP(hipMemcpyAsync(hpart, part(region), sizeof(double) * nb * v, hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < v; k++) out[k] = 0.0;
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < v; k++) out[k] += hpart[(size_t)b * v + k];
}
Reducer::~Reducer() {
    if (host) (void)hipHostFree(host);
    if (hpart) (void)hipHostFree(hpart);
    if (look) (void)hipHostFree(look);
}

static inline int nblocks(long n, int per = BS) { return (int)((n + per - 1) / per); }
static inline int red_blocks(long n) {
    int b = nblocks(n);
    return b > 2048 ? 2048 : (b < 1 ? 1 : b);
}

// ------------------------------------------------------------------ operator apply
constexpr int TX = 8, TY = 8, TZ = 4, NT = TX * TY * TZ;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HN = HX * HY * HZ;
constexpr int EXs = TX + 1, EYs = TY + 1, EZs = TZ + 1, EN = EXs * EYs * EZs;
static_assert(NT == BS, "tile must match block size");

struct GridArgs {
    int nx, ny, nz, ex, ey, ez;
};
static inline GridArgs gargs(const Grid& g) { return GridArgs{g.nx, g.ny, g.nz, g.ex, g.ey, g.ez}; }

// Per-node contraction of one input: yK = sum_e kfe * KE[a rows] u_e, yM = sum_e mfe * (M8 (x) I3)[a rows] u_e.
// FENCE keeps each element's matrix rows from being hoisted across the node's 8 elements (the
// compiler otherwise holds the 576 + 64 uniform values in scalar registers and spills them).  It is
// applied where K and M results leave the kernel separately (the split K / M outputs); with one
// summed output the same fence makes the code larger.
template <typename T, bool FENCE>
__device__ __forceinline__ void node_contract(const T* __restrict__ sU, const T* __restrict__ sk,
                                              const T* __restrict__ sm,
                                              const T* __restrict__ KE, const T* __restrict__ M8, int tx, int ty,
                                              int tz, T (&yK)[3], T (&yM)[3]) {
#pragma unroll
    for (int ez = 0; ez < 2; ez++)
#pragma unroll
        for (int ey = 0; ey < 2; ey++)
#pragma unroll
            for (int ex = 0; ex < 2; ex++) {
                if (FENCE) __asm__ __volatile__("" ::: "memory");
                const int a = corner_id(1 - ex, 1 - ey, 1 - ez);
                const int eidx = ((tz + ez) * EYs + (ty + ey)) * EXs + (tx + ex);
                const T kfe = sk[eidx], mfe = sm[eidx];
                T k0 = 0, k1 = 0, k2 = 0, m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const int nidx = ((tz + ez + cz_of(b)) * HY + (ty + ey + cy_of(b))) * HX + (tx + ex + cx_of(b));
                    const T u0 = sU[3 * nidx], u1 = sU[3 * nidx + 1], u2 = sU[3 * nidx + 2];
                    const T* kr = KE + (3 * a) * 24 + 3 * b;
                    k0 = fma(kr[0], u0, fma(kr[1], u1, fma(kr[2], u2, k0)));
                    k1 = fma(kr[24], u0, fma(kr[25], u1, fma(kr[26], u2, k1)));
                    k2 = fma(kr[48], u0, fma(kr[49], u1, fma(kr[50], u2, k2)));
                    const T mab = M8[a * 8 + b];
                    m0 = fma(mab, u0, m0);
                    m1 = fma(mab, u1, m1);
                    m2 = fma(mab, u2, m2);
                }
                yK[0] = fma(kfe, k0, yK[0]);
                yK[1] = fma(kfe, k1, yK[1]);
                yK[2] = fma(kfe, k2, yK[2]);
                yM[0] = fma(mfe, m0, yM[0]);
                yM[1] = fma(mfe, m1, yM[1]);
                yM[2] = fma(mfe, m2, yM[2]);
            }
}

// Per-node contraction of two inputs, yK = sum_e kfe * KE[a rows] u_e and yM = sum_e mfe * (M8 (x) I3)[a rows]
// w_e, in neighbour order: the node's 27 neighbours in the outer loop, the 1-8 elements that hold both
// the node and the neighbour in the inner one.  Each neighbour's values are read from LDS once (162
// reads per node instead of 384 in element order); the 8 element coefficients (element e at offset
// (e & 1, (e >> 1) & 1, e >> 2) - 1 from the node) come in registers.  The fence per neighbour keeps
// the uniform matrix entries from being hoisted into scalar registers and spilled.  On the 128 x 64 x 64
// right-hand side this order takes 68 us against 74 us per launch.
template <typename T>
__device__ __forceinline__ void node_contract_nb(const T* __restrict__ sU, const T* __restrict__ sW,
                                                 const T (&kfe)[8], const T (&mfe)[8],
                                                 const T* __restrict__ KE, const T* __restrict__ M8, int tx, int ty,
                                                 int tz, T (&yK)[3], T (&yM)[3]) {
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                __asm__ __volatile__("" ::: "memory");
                const int nidx = ((tz + 1 + dz) * HY + (ty + 1 + dy)) * HX + (tx + 1 + dx);
                const T u0 = sU[3 * nidx], u1 = sU[3 * nidx + 1], u2 = sU[3 * nidx + 2];
                const T w0 = sW[3 * nidx], w1 = sW[3 * nidx + 1], w2 = sW[3 * nidx + 2];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
                    const int bx = 1 + dx - ex, by = 1 + dy - ey, bz = 1 + dz - ez;  // b as a corner of e
                    if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                    const int a = corner_id(1 - ex, 1 - ey, 1 - ez), b = corner_id(bx, by, bz);
                    const T* kr = KE + (3 * a) * 24 + 3 * b;
                    const T t0 = fma(kr[0], u0, fma(kr[1], u1, kr[2] * u2));
                    const T t1 = fma(kr[24], u0, fma(kr[25], u1, kr[26] * u2));
                    const T t2 = fma(kr[48], u0, fma(kr[49], u1, kr[50] * u2));
                    yK[0] = fma(kfe[e], t0, yK[0]);
                    yK[1] = fma(kfe[e], t1, yK[1]);
                    yK[2] = fma(kfe[e], t2, yK[2]);
                    const T cm = mfe[e] * M8[a * 8 + b];
                    yM[0] = fma(cm, w0, yM[0]);
                    yM[1] = fma(cm, w1, yM[1]);
                    yM[2] = fma(cm, w2, yM[2]);
                }
            }
}

// Input staging policies: produce the (masked) value of dof (node n, comp c) for the K and M inputs
template <typename T>
struct StageOne {  // w = u
    const T* u;
    __device__ __forceinline__ void prologue() {}
    __device__ __forceinline__ T k(unsigned i) const { return ld(u, i); }
    __device__ __forceinline__ T m(unsigned i) const { return ld(u, i); }
};
// First Chebyshev iterate x0 = c1 D^-1 b formed on the fly, with b = (T) r (FP64) on the finest level
template <typename T, bool FROM64>
struct StageChebFirst {
    const double* r64;
    const T* b;
    const T* dinv;
    T c1;
    __device__ __forceinline__ void prologue() {}
    __device__ __forceinline__ T bval(unsigned i) const { return FROM64 ? (T)ld(r64, i) : ld(b, i); }
    __device__ __forceinline__ T k(unsigned i) const { return c1 * ld(dinv, i) * bval(i); }
    __device__ __forceinline__ T m(unsigned i) const { return k(i); }
};
struct StageThree {  // wK = sum_i cK_i in_i, wM = sum_i cM_i in_i (FP64)
    const double *u, *v, *a;
    double k0, k1, k2, m0, m1, m2;
    __device__ __forceinline__ void prologue() {}
    __device__ __forceinline__ double k(unsigned i) const { return k0 * ld(u, i) + k1 * ld(v, i) + k2 * ld(a, i); }
    __device__ __forceinline__ double m(unsigned i) const { return m0 * ld(u, i) + m1 * ld(v, i) + m2 * ld(a, i); }
};

struct EpiSplit;
// The two-input form (the right-hand side) reads each node's 8 element coefficients into registers
// instead of staging them in LDS: 28.8 instead of 35.3 KB of LDS per block, and with at most 96 VGPRs
// (5 waves per SIMD requested) 5 instead of 4 blocks per CU.
template <typename T, typename CT, bool TWO, bool MASKIN, class Stage, class Epi>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TWO ? 5 : 1))) void k_apply(
    GridArgs g, const T* __restrict__ KE, const T* __restrict__ M8, const CT* __restrict__ kf,
    const CT* __restrict__ mf, T cK, T cM, const uint8_t* __restrict__ mask, Stage st_in, Epi epi) {
    __shared__ T sU[HN * 3];
    __shared__ T sW[TWO ? HN * 3 : 1];
    __shared__ T sk[TWO ? 1 : EN], sm[TWO ? 1 : EN];
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, k0 = blockIdx.z * TZ;
    const int tid = threadIdx.x;
    // the thread's own node
    const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
    const int oi = i0 + tx, oj = j0 + ty, ok = k0 + tz;
    const bool inside = oi < g.nx && oj < g.ny && ok < g.nz;
    Stage st = st_in;
    st.prologue();
    // Stage the node box and element coefficients.  Every global load of the
    // block is issued before the first LDS store, so the block waits for one
    // memory latency instead of one per loop trip.
    constexpr int NSV = (HN * 3 + NT - 1) / NT;
    constexpr int NSE = TWO ? 1 : (EN + NT - 1) / NT;
    T val[NSV], valw[TWO ? NSV : 1];
    T ca[NSE], cb[NSE];
    T kfe[TWO ? 8 : 1], mfe[TWO ? 8 : 1];
    if constexpr (TWO) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int ei = oi - 1 + (e & 1), ej = oj - 1 + ((e >> 1) & 1), ek = ok - 1 + (e >> 2);
            const bool in = inside && ei >= 0 && ei < g.ex && ej >= 0 && ej < g.ey && ek >= 0 && ek < g.ez;
            const long ge = in ? ei + (long)g.ex * (ej + (long)g.ey * ek) : 0;
            kfe[e] = in ? cK * (T)coef_ld(kf, ge) : (T)0;
            mfe[e] = in ? cM * (T)coef_ld(mf, ge) : (T)0;
        }
    }
#pragma unroll
    for (int q = 0; q < NSV; q++) {
        const int v = tid + q * NT;
        const int node = v / 3, c = v - 3 * node;
        const int hx = node % HX, hy = (node / HX) % HY, hz = node / (HX * HY);
        const int gi = i0 - 1 + hx, gj = j0 - 1 + hy, gk = k0 - 1 + hz;
        const bool in = v < HN * 3 && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz;
        const long n = in ? gi + (long)g.nx * (gj + (long)g.ny * gk) : 0;
        T x = in ? st.k(3 * n + c) : (T)0;
        T w = 0;
        if (TWO) w = in ? st.m(3 * n + c) : (T)0;
        if (MASKIN) {
            const uint8_t cl = in ? mask[n] : (uint8_t)0;
            if ((cl >> c) & 1) {
                x = 0;
                w = 0;
            }
        }
        val[q] = x;
        if (TWO) valw[q] = w;
    }
#pragma unroll
    for (int q = 0; q < (TWO ? 0 : NSE); q++) {
        const int e = tid + q * NT;
        const int hx = e % EXs, hy = (e / EXs) % EYs, hz = e / (EXs * EYs);
        const int gi = i0 - 1 + hx, gj = j0 - 1 + hy, gk = k0 - 1 + hz;
        const bool in = e < EN && gi >= 0 && gi < g.ex && gj >= 0 && gj < g.ey && gk >= 0 && gk < g.ez;
        const long ge = in ? gi + (long)g.ex * (gj + (long)g.ey * gk) : 0;
        ca[q] = in ? cK * (T)coef_ld(kf, ge) : (T)0;
        cb[q] = in ? cM * (T)coef_ld(mf, ge) : (T)0;
    }
#pragma unroll
    for (int q = 0; q < NSV; q++) {
        const int v = tid + q * NT;
        if (v < HN * 3) {
            sU[v] = val[q];
            if (TWO) sW[v] = valw[q];
        }
    }
#pragma unroll
    for (int q = 0; q < (TWO ? 0 : NSE); q++) {
        const int e = tid + q * NT;
        if (e < EN) {
            sk[e] = ca[q];
            sm[e] = cb[q];
        }
    }
    __syncthreads();
    T yK[3] = {0, 0, 0}, yM[3] = {0, 0, 0};
    long n = 0;
    uint8_t cl = 0;
    if (inside) {
        if constexpr (TWO)
            node_contract_nb<T>(sU, sW, kfe, mfe, KE, M8, tx, ty, tz, yK, yM);
        else
            node_contract<T, std::is_same<Epi, EpiSplit>::value>(sU, sk, sm, KE, M8, tx, ty, tz, yK, yM);
        n = oi + (long)g.nx * (oj + (long)g.ny * ok);
        cl = mask[n];
    }
    epi(inside, n, cl, yK, yM, st);
}

#include "tt_apply_mfma.inc"

// ---- epilogues.  fetch() loads the per-node operands an epilogue reads (its Pre) for one owned
// node; the tile kernels issue it together with the staging loads, so that those loads are in
// flight while the element contraction runs.  put() then handles that node (inside the grid) and
// may add to the thread's reduction share; done() is called once by every thread of the block at
// the end (reducing epilogues write their block partial there).  operator() is the
// one-node-per-thread form.
#define TT_EPI_CALL                                                                                      \
    template <class S, typename V>                                                                     \
    __device__ void operator()(bool in, int n, uint8_t cl, const V (&yK)[3], const V (&yM)[3],          \
                               const S& st) const {                                                     \
        double acc[2] = {0.0, 0.0};                                                                    \
        if (in) {                                                                                      \
            Pre pre;                                                                                   \
            fetch(n, st, pre);                                                                         \
            const V uo[3] = {st.k(3 * n), st.k(3 * n + 1), st.k(3 * n + 2)};                           \
            put(n, cl, yK, yM, st, uo, acc, pre);                                                      \
        }                                                                                              \
        done(acc);                                                                                     \
    }
#define TT_NO_PRE                                                     \
    struct Pre {};                                                    \
    template <class S>                                                \
    __device__ __forceinline__ void fetch(int, const S&, Pre&) const {}

// Input value of dof i for an epilogue: the staged value uo (free dofs), re-read through the stage
// only on clamped dofs, where the staged copy is masked.
template <class S, typename T>
__device__ __forceinline__ T input_at(const S& st, uint8_t cl, int c, int i, const T (&uo)[3]) {
    T x = uo[c];
    if ((cl >> c) & 1) x = st.k(i);
    return x;
}

template <typename T>
struct EpiBC {  // y = A u on free dofs, u on clamped
    T* y;
    TT_NO_PRE
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&)[3],
                        double (&)[2], const Pre&) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T v = yK[c] + yM[c];
            if ((cl >> c) & 1) v = st.k(3 * n + c);
            y[3 * n + c] = v;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

template <typename T>
struct EpiBCDot {  // y = A u (BC form) and partials of u.y
    T* y;
    double* part;
    TT_NO_PRE
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&uo)[3],
                        double (&acc)[2], const Pre&) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T u = input_at(st, cl, c, 3 * n + c, uo);
            const T v = ((cl >> c) & 1) ? u : yK[c] + yM[c];
            y[3 * n + c] = v;
            acc[0] += (double)u * (double)v;
        }
    }
    __device__ void done(double (&acc)[2]) const {
        double a[1] = {acc[0]};
        write_partials<1>(a, part);
    }
    TT_EPI_CALL
};

template <typename T>
struct EpiResid {  // r = b - A x
    const T* b;
    T* r;
    struct Pre {
        T b[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) p.b[c] = b[3 * n + c];
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&)[3],
                        double (&)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T ax = yK[c] + yM[c];
            if ((cl >> c) & 1) ax = st.k(3 * n + c);
            r[3 * n + c] = p.b[c] - ax;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

// One damped-Jacobi step from x = 0 and its residual in one pass (an intermediate multigrid level's
// pre-smoothing): x = c1 D^-1 b is formed in staging (StageChebFirst) and written, with r = b - A x.
template <typename T>
struct EpiResidFirst {
    T* x;
    T* r;
    struct Pre {
        T b[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S& st, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) p.b[c] = st.bval(3 * n + c);
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&uo)[3],
                        double (&)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            const T xi = input_at(st, cl, c, i, uo);
            const T ax = ((cl >> c) & 1) ? xi : yK[c] + yM[c];
            x[i] = xi;
            r[i] = p.b[c] - ax;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

// r = b - A x (FP64) with the partials of b.b and r.r, one per block into each of two regions
struct EpiResidDots {
    const double* b;
    double* r;
    double* pbb;
    double* prr;
    struct Pre {
        double b[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) p.b[c] = b[3 * n + c];
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const double (&yK)[3], const double (&yM)[3], const S& st,
                        const double (&)[3], double (&acc)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double ax = yK[c] + yM[c];
            if ((cl >> c) & 1) ax = st.k(3 * n + c);
            const double bi = p.b[c], ri = bi - ax;
            r[3 * n + c] = ri;
            acc[0] = fma(bi, bi, acc[0]);
            acc[1] = fma(ri, ri, acc[1]);
        }
    }
    __device__ void done(double (&acc)[2]) const {
        __shared__ double sh[16 * 2];
        block_sum<2>(acc, sh);
        if (threadIdx.x == 0) {
            const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
            pbb[bid] = acc[0];
            prr[bid] = acc[1];
        }
    }
    TT_EPI_CALL
};

template <typename T>
struct EpiCheb {  // r = b - A x; dn = c0 d + c1 dinv r; xn = x + dn
    const T* b;
    const T* d;
    const T* dinv;
    T c0, c1;
    T* xn;
    T* dn;
    struct Pre {
        T b[3], d[3], di[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            p.b[c] = b[i];
            p.di[c] = dinv[i];
            p.d[c] = d ? d[i] : (T)0;
        }
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&uo)[3],
                        double (&)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            const T x = input_at(st, cl, c, i, uo);
            const T ax = ((cl >> c) & 1) ? x : yK[c] + yM[c];
            const T r = p.b[c] - ax;
            T dd = c1 * p.di[c] * r;
            if (d) dd = fma(c0, p.d[c], dd);
            xn[i] = x + dd;
            if (dn) dn[i] = dd;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

// Second Chebyshev step from x0 = c1 D^-1 b (staged by StageChebFirst, so d = x0):
// dn = c0 x0 + c1 D^-1 (b - A x0), xn = x0 + dn; on the finest level it also writes b = (T) r.
template <typename T, bool FROM64>
struct EpiChebFirst {
    T* bout;
    const T* dinv;
    T c0, c1;
    T* xn;
    T* dn;
    struct Pre {
        T b[3], di[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S& st, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            p.b[c] = st.bval(i);
            p.di[c] = dinv[i];
        }
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&uo)[3],
                        double (&)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            const T bi = p.b[c];
            if (FROM64) bout[i] = bi;
            const T x = input_at(st, cl, c, i, uo);
            const T ax = ((cl >> c) & 1) ? x : yK[c] + yM[c];
            const T dd = fma(c0, x, c1 * p.di[c] * (bi - ax));
            xn[i] = x + dd;
            dn[i] = dd;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

// Last post-smoothing step of the finest level: z = x + c0 d + c1 D^-1 (b - A x) written in FP64
// as the preconditioned residual of flexible CG, with partials of z.r and z.r_prev.
template <typename T>
struct EpiChebOut {
    const T* b;
    const T* d;
    const T* dinv;
    T c0, c1;
    double* z;
    const double* r;
    const double* rp;
    int first;
    double* part;
    struct Pre {
        T b[3], d[3], di[3];
        double r[3], rp[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            p.b[c] = b[i];
            p.di[c] = dinv[i];
            p.d[c] = d ? d[i] : (T)0;
            p.r[c] = r[i];
            p.rp[c] = first ? 0.0 : rp[i];
        }
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const T (&yK)[3], const T (&yM)[3], const S& st, const T (&uo)[3],
                        double (&acc)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            const T x = input_at(st, cl, c, i, uo);
            const T ax = ((cl >> c) & 1) ? x : yK[c] + yM[c];
            T dd = c1 * p.di[c] * (p.b[c] - ax);
            if (d) dd = fma(c0, p.d[c], dd);
            const double zi = (double)(x + dd);
            z[i] = zi;
            acc[0] = fma(zi, p.r[c], acc[0]);
            if (!first) acc[1] = fma(zi, p.rp[c], acc[1]);
        }
    }
    __device__ void done(double (&acc)[2]) const { write_partials<2>(acc, part); }
    TT_EPI_CALL
};

struct EpiSplit {  // yk = N.K(N u), ym = N.M(N u)
    double* yk;
    double* ym;
    TT_NO_PRE
    template <class S>
    __device__ void put(int n, uint8_t cl, const double (&yK)[3], const double (&yM)[3], const S&,
                        const double (&)[3], double (&)[2], const Pre&) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const bool clamped = (cl >> c) & 1;
            yk[3 * n + c] = clamped ? 0.0 : yK[c];
            ym[3 * n + c] = clamped ? 0.0 : yM[c];
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

struct EpiRaw {  // y = yK + yM, no BC
    double* y;
    TT_NO_PRE
    template <class S>
    __device__ void put(int n, uint8_t, const double (&yK)[3], const double (&yM)[3], const S&, const double (&)[3],
                        double (&)[2], const Pre&) const {
#pragma unroll
        for (int c = 0; c < 3; c++) y[3 * n + c] = yK[c] + yM[c];
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

struct EpiRHS {  // R = N.(fs F + yK + yM)
    const double* F;
    double fs;
    double* R;
    struct Pre {
        double F[3];
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) p.F[c] = F[3 * n + c];
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const double (&yK)[3], const double (&yM)[3], const S&, const double (&)[3],
                        double (&)[2], const Pre& p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            double r = fs * p.F[c];
            r += yK[c];
            r += yM[c];
            R[i] = ((cl >> c) & 1) ? 0.0 : r;
        }
    }
    __device__ void done(double (&)[2]) const {}
    TT_EPI_CALL
};

// Flexible-CG direction and its product, with z staged (A z from the contraction): p = z + beta p_old,
// q = A p = A z + beta q_old (p_old, q_old read at the owned node), partials of p.q.  beta from k_fcg_beta
// (0 on a (re)start).
struct EpiFcg {
    double* pnew;
    double* q;
    const double* pold;
    const double* qold;
    const double* scal;
    int first;
    double* part;
    struct Pre {
        double p[3], q[3], beta;
    };
    template <class S>
    __device__ __forceinline__ void fetch(int n, const S&, Pre& pr) const {
        pr.beta = first ? 0.0 : scal[SC_BETA];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pr.p[c] = first ? 0.0 : pold[3 * n + c];
            pr.q[c] = first ? 0.0 : qold[3 * n + c];
        }
    }
    template <class S>
    __device__ void put(int n, uint8_t cl, const double (&yK)[3], const double (&yM)[3], const S& st,
                        const double (&uo)[3], double (&acc)[2], const Pre& pr) const {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = 3 * n + c;
            const double z = input_at(st, cl, c, i, uo);
            const double az = ((cl >> c) & 1) ? z : yK[c] + yM[c];
            const double p = first ? z : fma(pr.beta, pr.p[c], z);
            const double v = first ? az : fma(pr.beta, pr.q[c], az);
            pnew[i] = p;
            q[i] = v;
            acc[0] = fma(p, v, acc[0]);
        }
    }
    __device__ void done(double (&acc)[2]) const {
        double a[1] = {acc[0]};
        write_partials<1>(a, part);
    }
    TT_EPI_CALL
};
#undef TT_EPI_CALL
#undef TT_NO_PRE

static inline dim3 apply_grid(const Grid& g) {
    return dim3((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, (g.nz + TZ - 1) / TZ);
}

// operator kernel family: 1 = element GEMM on the matrix cores (default), 0 = per-node vector FMA
static int apply_impl() {
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("TT_APPLY_IMPL");
        v = e ? atoi(e) : 1;
    }
    return v;
}

static int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        TT_HIP(hipGetDevice(&dev));
        TT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus <= 0) cus = 256;
    }
    return cus;
}

void tile_bc_flags(const Grid& g, const std::vector<uint8_t>& mask, int tz, std::vector<uint8_t>& f) {
    const int tx = (g.nx + 7) / 8, ty = (g.ny + 7) / 8, tzn = (g.nz + tz - 1) / tz;
    f.assign((size_t)tx * ty * tzn, 0);
    // tile a (extent e) stages nodes e a - 1 .. e a + e
    auto range = [](int i, int e, int nt, int& lo, int& hi) {
        lo = std::max(0, (i - e) / e - 1);
        hi = std::min(nt - 1, (i + 1) / e);
    };
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
                if (!mask[i + (size_t)g.nx * (j + (size_t)g.ny * k)]) continue;
                int a0, a1, b0, b1, c0, c1;
                range(i, 8, tx, a0, a1);
                range(j, 8, ty, b0, b1);
                range(k, tz, tzn, c0, c1);
                for (int c = c0; c <= c1; c++)
                    for (int b = b0; b <= b1; b++)
                        for (int a = a0; a <= a1; a++)
                            if (i >= 8 * a - 1 && i <= 8 * a + 8 && j >= 8 * b - 1 && j <= 8 * b + 8 &&
                                k >= tz * c - 1 && k <= tz * c + tz)
                                f[a + (size_t)tx * (b + (size_t)ty * c)] = 1;
            }
}

// Launch of k_apply_tiles: one block per tile, rounded up to a multiple of 8 for the XCD-aware
// tile order; returns the block count (= number of reduction partials).
template <typename T, typename CT, int TZ_, bool MASKIN, class Stage, class Epi>
static int launch_tiles_ct(const OpLevel& L, const T* AF, int tx, int ty, int tz, T cK, T cM, const Stage& st,
                           const Epi& epi, hipStream_t s) {
    const int ntiles = tx * ty * tz;
    const int nb = 8 * ((ntiles + 7) / 8);
    hipLaunchKernelGGL((k_apply_tiles<T, CT, TZ_, MASKIN, Stage, Epi>), dim3(nb), dim3(PTile<TZ_>::NTH), 0, s,
                       gargs(L.g), tx, ty, ntiles, AF, (const CT*)L.kf, (const CT*)L.mf, cK, cM, L.mask,
                       TZ_ == 4 ? L.tbc4 : L.tbc2, TZ_ == 4 ? L.tab4 : L.tab2, st, epi);
    TT_HIP(hipGetLastError());
    return nb;
}
template <typename T, int TZ_, bool MASKIN, class Stage, class Epi>
static int launch_tiles(const OpLevel& L, const T* AF, int tx, int ty, int tz, T cK, T cM, const Stage& st,
                        const Epi& epi, hipStream_t s) {
    switch (L.coef_kind) {
    case COEF_F64: return launch_tiles_ct<T, double, TZ_, MASKIN>(L, AF, tx, ty, tz, cK, cM, st, epi, s);
    case COEF_F32: return launch_tiles_ct<T, float, TZ_, MASKIN>(L, AF, tx, ty, tz, cK, cM, st, epi, s);
    default: return launch_tiles_ct<T, bf16_t, TZ_, MASKIN>(L, AF, tx, ty, tz, cK, cM, st, epi, s);
    }
}

// dispatch on coefficient kind and precision
template <typename T, bool TWO, bool MASKIN, class Stage, class Epi>
static int run_apply(const OpLevel& L, double cK, double cM, const Stage& st, const Epi& epi, hipStream_t s,
                     bool allow_mfma = true) {
    const T* KE = std::is_same<T, double>::value ? (const T*)L.KE64 : (const T*)L.KE32;
    const T* M8 = std::is_same<T, double>::value ? (const T*)L.M864 : (const T*)L.M832;
    GridArgs ga = gargs(L.g);
    if (!TWO && allow_mfma && apply_impl() == 1) {
        // FP32 runs on the FP16-split fragments, whose K / M rows carry the scales hs
        const bool f32 = std::is_same<T, float>::value;
        const T* AF = f32 ? (const T*)L.AFH : (const T*)L.AF64;
        if (f32) {
            cK *= L.hs[0];
            cM *= L.hs[1];
        }
        // 8x8x4 tiles on grids with enough of them to fill the chip, 8x8x2 on the coarse levels
        const int tx = (L.g.nx + 7) / 8, ty = (L.g.ny + 7) / 8;
        const bool deep = (long)tx * ty * ((L.g.nz + 3) / 4) >= 4L * device_cus();
        if (deep) return launch_tiles<T, 4, MASKIN>(L, AF, tx, ty, (L.g.nz + 3) / 4, (T)cK, (T)cM, st, epi, s);
        return launch_tiles<T, 2, MASKIN>(L, AF, tx, ty, (L.g.nz + 1) / 2, (T)cK, (T)cM, st, epi, s);
    }
    dim3 grid = apply_grid(L.g);
    switch (L.coef_kind) {
    case COEF_F64:
        hipLaunchKernelGGL((k_apply<T, double, TWO, MASKIN, Stage, Epi>), grid, dim3(NT), 0, s, ga, KE, M8,
                           (const double*)L.kf, (const double*)L.mf, (T)cK, (T)cM, L.mask, st, epi);
        break;
    case COEF_F32:
        hipLaunchKernelGGL((k_apply<T, float, TWO, MASKIN, Stage, Epi>), grid, dim3(NT), 0, s, ga, KE, M8,
                           (const float*)L.kf, (const float*)L.mf, (T)cK, (T)cM, L.mask, st, epi);
        break;
    default:
        hipLaunchKernelGGL((k_apply<T, bf16_t, TWO, MASKIN, Stage, Epi>), grid, dim3(NT), 0, s, ga, KE, M8,
                           (const bf16_t*)L.kf, (const bf16_t*)L.mf, (T)cK, (T)cM, L.mask, st, epi);
        break;
    }
    TT_HIP(hipGetLastError());
    return (int)(grid.x * grid.y * grid.z);
}

// upper bound of the blocks of any apply launch on grid g (8x8x2 tiles, rounded up to a multiple of 8)
static inline int apply_nblocks(const Grid& g) {
    return 8 * ((((g.nx + 7) / 8) * ((g.ny + 7) / 8) * ((g.nz + 1) / 2) + 7) / 8);
}

void launch_apply_bc(const OpLevel& L, double cK, double cM, const void* u, void* y, bool fp64, hipStream_t s,
                     Reducer* red, int dot_slot) {
    if (fp64) {
        StageOne<double> st{(const double*)u};
        if (red) {
            TT_CHECK(apply_nblocks(L.g) <= red->maxblocks, -3, "reducer too small");
            EpiBCDot<double> e{(double*)y, red->part(dot_slot)};
            red->nblk[dot_slot] = run_apply<double, false, true>(L, cK, cM, st, e, s);
            red->nv[dot_slot] = 1;
        } else {
            run_apply<double, false, true>(L, cK, cM, st, EpiBC<double>{(double*)y}, s);
        }
    } else {
        StageOne<float> st{(const float*)u};
        if (red) {
            TT_CHECK(apply_nblocks(L.g) <= red->maxblocks, -3, "reducer too small");
            EpiBCDot<float> e{(float*)y, red->part(dot_slot)};
            red->nblk[dot_slot] = run_apply<float, false, true>(L, cK, cM, st, e, s);
            red->nv[dot_slot] = 1;
        } else {
            run_apply<float, false, true>(L, cK, cM, st, EpiBC<float>{(float*)y}, s);
        }
    }
}

void launch_residual(const OpLevel& L, double cK, double cM, const void* b, const void* x, void* r, bool fp64,
                     hipStream_t s) {
    if (fp64)
        run_apply<double, false, true>(L, cK, cM, StageOne<double>{(const double*)x},
                                       EpiResid<double>{(const double*)b, (double*)r}, s);
    else
        run_apply<float, false, true>(L, cK, cM, StageOne<float>{(const float*)x},
                                      EpiResid<float>{(const float*)b, (float*)r}, s);
}

void launch_resid_first(const OpLevel& L, double cK, double cM, const void* b, const void* dinv, double c1, void* x,
                        void* r, bool fp64, hipStream_t s) {
    if (fp64)
        run_apply<double, false, true>(
            L, cK, cM, StageChebFirst<double, false>{nullptr, (const double*)b, (const double*)dinv, c1},
            EpiResidFirst<double>{(double*)x, (double*)r}, s);
    else
        run_apply<float, false, true>(
            L, cK, cM, StageChebFirst<float, false>{nullptr, (const float*)b, (const float*)dinv, (float)c1},
            EpiResidFirst<float>{(float*)x, (float*)r}, s);
}

void launch_residual_dots(const OpLevel& L, double cK, double cM, const double* b, const double* x, double* r,
                          Reducer& red, int bb_region, int rr_region, hipStream_t s) {
    TT_CHECK(apply_nblocks(L.g) <= red.maxblocks, -3, "reducer too small");
    EpiResidDots e{b, r, red.part(bb_region), red.part(rr_region)};
    const int nb = run_apply<double, false, true>(L, cK, cM, StageOne<double>{x}, e, s);
    red.nblk[bb_region] = red.nblk[rr_region] = nb;
    red.nv[bb_region] = red.nv[rr_region] = 1;
}

void launch_cheb_step(const OpLevel& L, double cK, double cM, const void* b, const void* x, const void* d,
                      const void* dinv, double c0, double c1, void* xn, void* dn, bool fp64, hipStream_t s) {
    if (fp64) {
        EpiCheb<double> e{(const double*)b, (const double*)d, (const double*)dinv, c0, c1, (double*)xn, (double*)dn};
        run_apply<double, false, true>(L, cK, cM, StageOne<double>{(const double*)x}, e, s);
    } else {
        EpiCheb<float> e{(const float*)b,  (const float*)d, (const float*)dinv, (float)c0,
                         (float)c1,        (float*)xn,      (float*)dn};
        run_apply<float, false, true>(L, cK, cM, StageOne<float>{(const float*)x}, e, s);
    }
}

void launch_cheb_first_step(const OpLevel& L, double cK, double cM, const double* r64, void* b, const void* dinv,
                            double c1_first, double c0, double c1, void* xn, void* dn, bool fp64, hipStream_t s) {
    if (fp64) {
        const double* dv = (const double*)dinv;
        if (r64)
            run_apply<double, false, true>(L, cK, cM, StageChebFirst<double, true>{r64, nullptr, dv, c1_first},
                                           EpiChebFirst<double, true>{(double*)b, dv, c0, c1, (double*)xn,
                                                                      (double*)dn}, s);
        else
            run_apply<double, false, true>(L, cK, cM,
                                           StageChebFirst<double, false>{nullptr, (const double*)b, dv, c1_first},
                                           EpiChebFirst<double, false>{nullptr, dv, c0, c1, (double*)xn,
                                                                       (double*)dn}, s);
    } else {
        const float* dv = (const float*)dinv;
        if (r64)
            run_apply<float, false, true>(L, cK, cM, StageChebFirst<float, true>{r64, nullptr, dv, (float)c1_first},
                                          EpiChebFirst<float, true>{(float*)b, dv, (float)c0, (float)c1,
                                                                    (float*)xn, (float*)dn}, s);
        else
            run_apply<float, false, true>(L, cK, cM,
                                          StageChebFirst<float, false>{nullptr, (const float*)b, dv, (float)c1_first},
                                          EpiChebFirst<float, false>{nullptr, dv, (float)c0, (float)c1, (float*)xn,
                                                                     (float*)dn}, s);
    }
}

void launch_cheb_last(const OpLevel& L, double cK, double cM, const void* b, const void* x, const void* d,
                      const void* dinv, double c0, double c1, double* z, const double* r, const double* rp, bool first,
                      bool fp64, Reducer& red, hipStream_t s) {
    TT_CHECK(apply_nblocks(L.g) <= red.maxblocks, -3, "reducer too small");
    if (fp64) {
        EpiChebOut<double> e{(const double*)b, (const double*)d, (const double*)dinv, c0, c1, z, r, rp,
                             first ? 1 : 0, red.part(PB_DOTS)};
        red.nblk[PB_DOTS] = run_apply<double, false, true>(L, cK, cM, StageOne<double>{(const double*)x}, e, s);
    } else {
        EpiChebOut<float> e{(const float*)b, (const float*)d, (const float*)dinv, (float)c0, (float)c1, z, r, rp,
                            first ? 1 : 0, red.part(PB_DOTS)};
        red.nblk[PB_DOTS] = run_apply<float, false, true>(L, cK, cM, StageOne<float>{(const float*)x}, e, s);
    }
    red.nv[PB_DOTS] = 2;
}

void launch_apply_split(const OpLevel& L, double cK, double cM, const double* u, double* yk, double* ym,
                        hipStream_t s) {
    run_apply<double, false, true>(L, cK, cM, StageOne<double>{u}, EpiSplit{yk, ym}, s, false);
}

void launch_apply_raw(const OpLevel& L, const double* kf, const double* mf, double cK, double cM, const double* u,
                      double* y, hipStream_t s) {
    OpLevel L2 = L;
    L2.kf = kf;
    L2.mf = mf;
    L2.coef_kind = COEF_F64;
    run_apply<double, false, false>(L2, cK, cM, StageOne<double>{u}, EpiRaw{y}, s);
}

void launch_newmark_rhs(const OpLevel& L, const double* u, const double* v, const double* a, const double* F,
                        double fs, const double wK[3], const double wM[3], double* R, hipStream_t s) {
    StageThree st{u, v, a, wK[0], wK[1], wK[2], wM[0], wM[1], wM[2]};
    run_apply<double, true, true>(L, 1.0, 1.0, st, EpiRHS{F, fs, R}, s);
}

// single-block reductions of partials: 1024 threads, so each sums ~5 of the ~5000 partials of a
// fine-grid launch (latency bound: 256 threads took ~4.6 us per call)
constexpr int BSR = 1024;
__global__ __launch_bounds__(BSR) void k_fcg_beta(const double* __restrict__ dots_part, int dots_nblk, double* scal,
                                                  int first);
void launch_fcg_apply(const OpLevel& L, double cK, double cM, const double* z, const double* p_old, double* p_new,
                      const double* q_old, double* q_new, Reducer& red, bool first, hipStream_t s, bool beta) {
    TT_CHECK(apply_nblocks(L.g) <= red.maxblocks, -3, "reducer too small");
    if (beta) {
        hipLaunchKernelGGL(k_fcg_beta, dim3(1), dim3(BSR), 0, s, red.part(PB_DOTS), red.nblk[PB_DOTS], red.scal.p,
                           first ? 1 : 0);
        TT_HIP(hipGetLastError());
    }
    EpiFcg e{p_new, q_new, p_old, q_old, red.scal.p, first ? 1 : 0, red.part(PB_PQ)};
    red.nblk[PB_PQ] = run_apply<double, false, true>(L, cK, cM, StageOne<double>{z}, e, s);
    red.nv[PB_PQ] = 1;
}

// ------------------------------------------------------------------ diagonal
template <typename T, typename CT>
__global__ void k_diag_inv(GridArgs g, const double* __restrict__ KE, const double* __restrict__ M8,
                           const CT* __restrict__ kf, const CT* __restrict__ mf, double cK, double cM,
                           const uint8_t* __restrict__ mask, T* __restrict__ dinv) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)g.nx * g.ny * g.nz) return;
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / ((long)g.nx * g.ny);
    double d[3] = {0, 0, 0};
    for (int ez = 0; ez < 2; ez++)
        for (int ey = 0; ey < 2; ey++)
            for (int ex = 0; ex < 2; ex++) {
                const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
                if (ei < 0 || ej < 0 || ek < 0 || ei >= g.ex || ej >= g.ey || ek >= g.ez) continue;
                const long e = ei + (long)g.ex * (ej + (long)g.ey * ek);
                const int a = corner_id(1 - ex, 1 - ey, 1 - ez);
                const double kk = cK * coef_ld(kf, e), mm = cM * coef_ld(mf, e);
                for (int c = 0; c < 3; c++) d[c] += kk * KE[(3 * a + c) * 25] + mm * M8[a * 9];
            }
    const uint8_t cl = mask[n];
    for (int c = 0; c < 3; c++) dinv[3 * n + c] = (T)(((cl >> c) & 1) ? 1.0 : 1.0 / d[c]);
}

void launch_diag_inv(const OpLevel& L, double cK, double cM, void* dinv, bool fp64, hipStream_t s) {
    GridArgs ga = gargs(L.g);
    int nb = nblocks(L.g.nn);
#define TT_DIAG(T, CT) \
    hipLaunchKernelGGL((k_diag_inv<T, CT>), dim3(nb), dim3(BS), 0, s, ga, L.KE64, L.M864, (const CT*)L.kf, (const CT*)L.mf, cK, cM, L.mask, (T*)dinv)
    if (fp64) {
        if (L.coef_kind == COEF_F64) TT_DIAG(double, double);
        else if (L.coef_kind == COEF_F32) TT_DIAG(double, float);
        else TT_DIAG(double, bf16_t);
    } else {
        if (L.coef_kind == COEF_F64) TT_DIAG(float, double);
        else if (L.coef_kind == COEF_F32) TT_DIAG(float, float);
        else TT_DIAG(float, bf16_t);
    }
#undef TT_DIAG
    TT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ vector kernels
template <typename T>
__global__ void k_fill(T* x, T v, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
void launch_fill(void* x, double v, long n, bool fp64, hipStream_t s) {
    if (n <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_fill<double>, dim3(nblocks(n)), dim3(BS), 0, s, (double*)x, v, n);
    else hipLaunchKernelGGL(k_fill<float>, dim3(nblocks(n)), dim3(BS), 0, s, (float*)x, (float)v, n);
    TT_HIP(hipGetLastError());
}
void launch_fill_scal(void* x, long n, double v, hipStream_t s) { launch_fill(x, v, n, true, s); }

// x_i uniform in [-1, 1) from a counter-based hash of (seed, i) (splitmix64 finaliser): a deterministic
// start vector made on the device, with no host loop or copy
template <typename T>
__global__ void k_fill_hash(T* __restrict__ x, long n, unsigned long long seed) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}
void launch_fill_hash(void* x, long n, unsigned long long seed, bool fp64, hipStream_t s) {
    if (n <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_fill_hash<double>, dim3(nblocks(n)), dim3(BS), 0, s, (double*)x, n, seed);
    else hipLaunchKernelGGL(k_fill_hash<float>, dim3(nblocks(n)), dim3(BS), 0, s, (float*)x, n, seed);
    TT_HIP(hipGetLastError());
}

// x = N.x: dof c of node n is zeroed where bit c of mask[n] is set
__global__ void k_mask_dofs(double* __restrict__ x, const uint8_t* __restrict__ mask, long nd) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < nd && ((mask[i / 3] >> (i % 3)) & 1)) x[i] = 0.0;
}
void launch_mask_dofs(double* x, const uint8_t* mask, long nd, hipStream_t s) {
    if (nd <= 0) return;
    hipLaunchKernelGGL(k_mask_dofs, dim3(nblocks(nd)), dim3(BS), 0, s, x, mask, nd);
    TT_HIP(hipGetLastError());
}

template <typename S, typename D>
__global__ void k_copy(const S* __restrict__ a, D* __restrict__ b, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (D)a[i];
}
void launch_copy(const void* src, void* dst, long n, bool s64, bool d64, hipStream_t s) {
    if (n <= 0) return;
    dim3 gr(nblocks(n)), bl(BS);
    if (s64 && d64) hipLaunchKernelGGL((k_copy<double, double>), gr, bl, 0, s, (const double*)src, (double*)dst, n);
    else if (s64) hipLaunchKernelGGL((k_copy<double, float>), gr, bl, 0, s, (const double*)src, (float*)dst, n);
    else if (d64) hipLaunchKernelGGL((k_copy<float, double>), gr, bl, 0, s, (const float*)src, (double*)dst, n);
    else hipLaunchKernelGGL((k_copy<float, float>), gr, bl, 0, s, (const float*)src, (float*)dst, n);
    TT_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_cheb_first(const T* __restrict__ b, const T* __restrict__ dinv, T c1, T* __restrict__ x, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = c1 * dinv[i] * b[i];
}
// b = (T) r (FP64 in) and x = c1 D^-1 b: the start of a V-cycle on the finest level
template <typename T>
__global__ void k_cheb_first_from64(const double* __restrict__ r, T* __restrict__ b, const T* __restrict__ dinv, T c1,
                                    T* __restrict__ x, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) {
        const T bi = (T)r[i];
        b[i] = bi;
        x[i] = c1 * dinv[i] * bi;
    }
}
void launch_cheb_first_from64(const double* r, void* b, const void* dinv, double c1, void* x, long n, bool fp64,
                              hipStream_t s) {
    if (fp64)
        hipLaunchKernelGGL(k_cheb_first_from64<double>, dim3(nblocks(n)), dim3(BS), 0, s, r, (double*)b,
                           (const double*)dinv, c1, (double*)x, n);
    else
        hipLaunchKernelGGL(k_cheb_first_from64<float>, dim3(nblocks(n)), dim3(BS), 0, s, r, (float*)b,
                           (const float*)dinv, (float)c1, (float*)x, n);
    TT_HIP(hipGetLastError());
}

void launch_cheb_first(const void* b, const void* dinv, double c1, void* x, long n, bool fp64, hipStream_t s) {
    if (fp64)
        hipLaunchKernelGGL(k_cheb_first<double>, dim3(nblocks(n)), dim3(BS), 0, s, (const double*)b,
                           (const double*)dinv, c1, (double*)x, n);
    else
        hipLaunchKernelGGL(k_cheb_first<float>, dim3(nblocks(n)), dim3(BS), 0, s, (const float*)b,
                           (const float*)dinv, (float)c1, (float*)x, n);
    TT_HIP(hipGetLastError());
}

// restriction R = P^T of trilinear interpolation (PETSc DMDA Q1 interpolation, LinearElasticity.cc:2840-2844).
// Branch-free and fully unrolled so that the 27 neighbour loads are in flight together: neighbours
// outside the fine grid get weight 0 at a clamped (valid) index, which leaves the sum unchanged.
// The coarse output may be wider than the fine input (FP32 levels feeding an FP64 coarsest solve).
template <typename Tf, typename Tc>
__global__ void k_restrict(GridArgs f, GridArgs c, const Tf* __restrict__ rf, const uint8_t* __restrict__ cmask,
                           Tc* __restrict__ bc) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)c.nx * c.ny * c.nz) return;
    int I, J, K;
    grid_ijk((unsigned)n, c.nx, c.ny, I, J, K);
    Tf s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int dk = -1; dk <= 1; dk++) {
        const int k = 2 * K + dk;
        const bool kin = k >= 0 && k < f.nz;
        const Tf wk = kin ? (dk == 0 ? (Tf)1 : (Tf)0.5) : (Tf)0;
        const int kc = kin ? k : 2 * K;
#pragma unroll
        for (int dj = -1; dj <= 1; dj++) {
            const int j = 2 * J + dj;
            const bool jin = j >= 0 && j < f.ny;
            const Tf wj = jin ? (dj == 0 ? wk : wk * (Tf)0.5) : (Tf)0;
            const int jc = jin ? j : 2 * J;
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                const int i = 2 * I + di;
                const bool iin = i >= 0 && i < f.nx;
                const Tf w = iin ? (di == 0 ? wj : wj * (Tf)0.5) : (Tf)0;
                const int ic = iin ? i : 2 * I;
                const long m = ic + (long)f.nx * (jc + (long)f.ny * kc);
                s0 = fma(w, rf[3 * m], s0);
                s1 = fma(w, rf[3 * m + 1], s1);
                s2 = fma(w, rf[3 * m + 2], s2);
            }
        }
    }
    const uint8_t cl = cmask[n];
    bc[3 * n] = (cl & 1) ? (Tc)0 : (Tc)s0;
    bc[3 * n + 1] = (cl & 2) ? (Tc)0 : (Tc)s1;
    bc[3 * n + 2] = (cl & 4) ? (Tc)0 : (Tc)s2;
}
void launch_restrict(const Grid& fine, const Grid& coarse, const void* rf, const uint8_t* cmask, void* bc, bool fp64,
                     bool out64, hipStream_t s) {
    const dim3 gr(nblocks(coarse.nn)), bl(BS);
    if (fp64)
        hipLaunchKernelGGL((k_restrict<double, double>), gr, bl, 0, s, gargs(fine), gargs(coarse), (const double*)rf,
                           cmask, (double*)bc);
    else if (out64)
        hipLaunchKernelGGL((k_restrict<float, double>), gr, bl, 0, s, gargs(fine), gargs(coarse), (const float*)rf,
                           cmask, (double*)bc);
    else
        hipLaunchKernelGGL((k_restrict<float, float>), gr, bl, 0, s, gargs(fine), gargs(coarse), (const float*)rf,
                           cmask, (float*)bc);
    TT_HIP(hipGetLastError());
}

// xf += P xc on free dofs.  Each fine node interpolates from 1, 2, 4 or 8 coarse nodes; the
// unrolled 2 x 2 x 2 loop gives a second node of a direction weight 0 (at a valid index) where the
// fine index is even, so the sum is the same as over the contributing nodes only.
template <typename Tf, typename Tc>
__global__ void k_prolong_add(GridArgs f, GridArgs c, const Tc* __restrict__ xc, const uint8_t* __restrict__ fmask,
                              Tf* __restrict__ xf) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)f.nx * f.ny * f.nz) return;
    int i, j, k;
    grid_ijk((unsigned)n, f.nx, f.ny, i, j, k);
    const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
    const int oi = i & 1, oj = j & 1, ok = k & 1;
    Tf s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const Tf wk = ok ? (Tf)0.5 : (kk ? (Tf)0 : (Tf)1);
        const int K = K0 + (kk & ok);
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const Tf wj = oj ? wk * (Tf)0.5 : (jj ? (Tf)0 : wk);
            const int J = J0 + (jj & oj);
#pragma unroll
            for (int ii = 0; ii < 2; ii++) {
                const Tf w = oi ? wj * (Tf)0.5 : (ii ? (Tf)0 : wj);
                const int I = I0 + (ii & oi);
                const long m = I + (long)c.nx * (J + (long)c.ny * K);
                s0 = fma(w, (Tf)xc[3 * m], s0);
                s1 = fma(w, (Tf)xc[3 * m + 1], s1);
                s2 = fma(w, (Tf)xc[3 * m + 2], s2);
            }
        }
    }
    const uint8_t cl = fmask[n];
    if (!(cl & 1)) xf[3 * n] += s0;
    if (!(cl & 2)) xf[3 * n + 1] += s1;
    if (!(cl & 4)) xf[3 * n + 2] += s2;
}
void launch_prolong_add(const Grid& fine, const Grid& coarse, const void* xc, const uint8_t* fmask, void* xf,
                        bool fp64, bool in64, hipStream_t s) {
    const dim3 gr(nblocks(fine.nn)), bl(BS);
    if (fp64)
        hipLaunchKernelGGL((k_prolong_add<double, double>), gr, bl, 0, s, gargs(fine), gargs(coarse),
                           (const double*)xc, fmask, (double*)xf);
    else if (in64)
        hipLaunchKernelGGL((k_prolong_add<float, double>), gr, bl, 0, s, gargs(fine), gargs(coarse),
                           (const double*)xc, fmask, (float*)xf);
    else
        hipLaunchKernelGGL((k_prolong_add<float, float>), gr, bl, 0, s, gargs(fine), gargs(coarse),
                           (const float*)xc, fmask, (float*)xf);
    TT_HIP(hipGetLastError());
}

// rediscretised coarse coefficients: mean of the 8 child elements
__global__ void k_coarsen_coef(GridArgs f, GridArgs c, const double* __restrict__ kff, const double* __restrict__ mff,
                               double* __restrict__ kfc, double* __restrict__ mfc) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)c.ex * c.ey * c.ez) return;
    const int I = e % c.ex, J = (e / c.ex) % c.ey, K = e / ((long)c.ex * c.ey);
    double sk = 0.0, sm = 0.0;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                const long fe = (2 * I + di) + (long)f.ex * ((2 * J + dj) + (long)f.ey * (2 * K + dk));
                sk += kff[fe];
                sm += mff[fe];
            }
    kfc[e] = 0.125 * sk;
    mfc[e] = 0.125 * sm;
}
void launch_coarsen_coef(const Grid& fine, const Grid& coarse, const double* kf_f, const double* mf_f, double* kf_c,
                         double* mf_c, hipStream_t s) {
    hipLaunchKernelGGL(k_coarsen_coef, dim3(nblocks(coarse.nel)), dim3(BS), 0, s, gargs(fine), gargs(coarse), kf_f,
                       mf_f, kf_c, mf_c);
    TT_HIP(hipGetLastError());
}

__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}
__global__ void k_convert_coef(const double* __restrict__ src, void* dst, long n, int kind) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (kind == COEF_F32) ((float*)dst)[i] = (float)src[i];
    else if (kind == COEF_BF16) ((bf16_t*)dst)[i] = f2bf((float)src[i]);
    else ((double*)dst)[i] = src[i];
}
void launch_convert_coef(const double* src, void* dst, long n, int kind, hipStream_t s) {
    hipLaunchKernelGGL(k_convert_coef, dim3(nblocks(n)), dim3(BS), 0, s, src, dst, n, kind);
    TT_HIP(hipGetLastError());
}

// SIMP / RAMP / linear interpolation (LinearElasticity.cc:1580-1623)
__device__ __forceinline__ void interp1(int scheme, double vmin, double vmax, double p, double x, double& f,
                                        double& df) {
    if (scheme == 0) {
        f = vmin + x * (vmax - vmin);
        df = vmax - vmin;
    } else if (scheme == 1) {
        f = vmin + pow(x, p) * (vmax - vmin);
        df = p * pow(x, p - 1) * (vmax - vmin);
    } else {
        f = vmin + x / (1 + p * (1 - x)) * (vmax - vmin);
        df = (1 + p) / pow(1 + p * (1 - x), 2) * (vmax - vmin);
    }
}
__global__ void k_interp(const double* __restrict__ x, long n, int sInt, double Emin, double Emax, double pK, int mInt,
                         double Rmin, double Rmax, double pM, const double* __restrict__ P, double* kf, double* mf,
                         double* dkf, double* dmf, double* Pkf, double* Pmf) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    double k, dk, m, dm;
    interp1(sInt, Emin, Emax, pK, x[e], k, dk);
    interp1(mInt, Rmin, Rmax, pM, x[e], m, dm);
    kf[e] = k;
    mf[e] = m;
    dkf[e] = dk;
    dmf[e] = dm;
    Pkf[e] = P[e] * k;
    Pmf[e] = P[e] * m;
}
void launch_interp(const double* x, long n, int sInt, double Emin, double Emax, double pK, int mInt, double Rmin,
                   double Rmax, double pM, const double* P, double* kf, double* mf, double* dkf, double* dmf,
                   double* Pkf, double* Pmf, hipStream_t s) {
    hipLaunchKernelGGL(k_interp, dim3(nblocks(n)), dim3(BS), 0, s, x, n, sInt, Emin, Emax, pK, mInt, Rmin, Rmax, pM,
                       P, kf, mf, dkf, dmf, Pkf, Pmf);
    TT_HIP(hipGetLastError());
}

// In-place inverse of an m x m (m <= 64) symmetric positive definite block (column-major, leading
// dimension lda) by the symmetric sweep: step k, with p = 1 / a_kk, sets a_ij -= a_ik a_jk p (i, j != k),
// a_ik *= p (i != k) and a_kk = -p; the matrix stays symmetric, and after all m steps it holds -A^-1.
// 256 threads: thread t holds rows 16 (t / 64) .. + 15 of column t % 64 in registers; the owner lanes
// put column k in LDS for each step (double-buffered, one barrier per step).  A pivot that is not
// positive sets *info.
__global__ __launch_bounds__(256) void k_spd_inv_small(double* A, int m, int lda, int* info) {
    __shared__ double colk[2][64];
    const int t = threadIdx.x, j = t & 63;
    const int rb = __builtin_amdgcn_readfirstlane(t >> 6);  // (wave-uniform: the pivot row's wave branches)
    double c[16];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
        const int i = 16 * rb + ii;
        c[ii] = (i < m && j < m) ? A[i + (size_t)j * lda] : 0.0;
    }
    bool bad = false;
    for (int kb = 0; kb < 4; kb++) {
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const int k = 16 * kb + kk;
            if (k < m) {
                double* ck = colk[k & 1];
                if (j == k)
#pragma unroll
                    for (int ii = 0; ii < 16; ii++) ck[16 * rb + ii] = c[ii];
                __syncthreads();
                const double akk = ck[k];
                bad = bad || !(akk > 0.0);
                const double p = 1.0 / akk;
                const double ajk = ck[j] * p;  // a_jk / a_kk (= a_kj / a_kk)
                // c_i -= a_ik a_jk / a_kk; column k itself becomes a_ik p; then the pivot row's element
#pragma unroll
                for (int ii = 0; ii < 16; ii++) c[ii] = fma(-ck[16 * rb + ii], ajk, c[ii]);
                if (j == k)
#pragma unroll
                    for (int ii = 0; ii < 16; ii++) c[ii] = ck[16 * rb + ii] * p;
                if (rb == kb) c[kk] = j == k ? -p : ajk;
            }
        }
    }
    if (bad && t == 0) *info = 1;
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
        const int i = 16 * rb + ii;
        if (i < m && j < m) A[i + (size_t)j * lda] = -c[ii];
    }
}
void launch_spd_inv_small(double* A, int m, int lda, int* info, hipStream_t s) {
    TT_CHECK(m >= 1 && m <= 64, -3, "launch_spd_inv_small: block of 1..64");  // TT_ERR_INTERNAL
    hipLaunchKernelGGL(k_spd_inv_small, dim3(1), dim3(256), 0, s, A, m, lda, info);
    TT_HIP(hipGetLastError());
}

// dst (cols x rows) = src (rows x cols)^T, both column-major with their leading dimensions
__global__ void k_transpose(const double* __restrict__ src, int rows, int cols, int lds, double* __restrict__ dst,
                            int ldd) {
    __shared__ double t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;   // src tile: rows i0.., cols j0..
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        if (i < rows && j < cols) t[r][tx] = src[i + (size_t)j * lds];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;  // dst (row j, col i) = src (row i, col j)
        if (i < rows && j < cols) dst[j + (size_t)i * ldd] = t[tx][r];
    }
}
void launch_transpose(const double* src, int rows, int cols, int lds, double* dst, int ldd, hipStream_t s) {
    hipLaunchKernelGGL(k_transpose, dim3((rows + 31) / 32, (cols + 31) / 32), dim3(256), 0, s, src, rows, cols, lds,
                       dst, ldd);
    TT_HIP(hipGetLastError());
}
// ---- dense coarsest operator, assembled on the device: entry (i, j) of N (cK K + cM M) N + I - N, i = 3 p + c,
// j = 3 q + d, for the 27 neighbours q of node p (all other entries stay 0 from the clear before).  One
// thread per (i, neighbour, d) sums the element contributions in increasing element order, so the result
// does not depend on scheduling (deterministic set-up).
__global__ void k_dense_coarse(GridArgs g, const double* __restrict__ kf, const double* __restrict__ mf, double cK,
                               double cM, const double* __restrict__ KE, const double* __restrict__ M8,
                               const uint8_t* __restrict__ mask, int n, double* __restrict__ A) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)n * 81) return;
    const int i = (int)(t / 81), r = (int)(t % 81);
    const int nb = r / 3, d = r % 3;
    const int p = i / 3, c = i % 3;
    const int pi = p % g.nx, pj = (p / g.nx) % g.ny, pk = p / (g.nx * g.ny);
    const int qi = pi + nb % 3 - 1, qj = pj + (nb / 3) % 3 - 1, qk = pk + nb / 9 - 1;
    if (qi < 0 || qi >= g.nx || qj < 0 || qj >= g.ny || qk < 0 || qk >= g.nz) return;
    const int q = qi + g.nx * (qj + g.ny * qk), j = 3 * q + d;
    double val = 0.0;
    if (((mask[p] >> c) & 1) || ((mask[q] >> d) & 1)) {
        val = i == j ? 1.0 : 0.0;
    } else {
        // elements holding both nodes: origin (ei, ej, ek) with max(p, q) - 1 <= e <= min(p, q) per axis
        for (int ek = max(max(pk, qk) - 1, 0); ek <= min(min(pk, qk), g.ez - 1); ek++)
            for (int ej = max(max(pj, qj) - 1, 0); ej <= min(min(pj, qj), g.ey - 1); ej++)
                for (int ei = max(max(pi, qi) - 1, 0); ei <= min(min(pi, qi), g.ex - 1); ei++) {
                    const long e = ei + (long)g.ex * (ej + (long)g.ey * ek);
                    const int la = corner_id(pi - ei, pj - ej, pk - ek), lb = corner_id(qi - ei, qj - ej, qk - ek);
                    double v = cK * kf[e] * KE[(3 * la + c) * 24 + 3 * lb + d];
                    if (c == d) v += cM * mf[e] * M8[la * 8 + lb];
                    val += v;
                }
    }
    A[(size_t)i * n + j] = val;
}
void launch_dense_coarse(const Grid& g, const double* kf, const double* mf, double cK, double cM, const double* KE,
                         const double* M8, const uint8_t* mask, double* A, hipStream_t s) {
    const int n = (int)g.nd;
    TT_HIP(hipMemsetAsync(A, 0, (size_t)n * n * sizeof(double), s));
    hipLaunchKernelGGL(k_dense_coarse, dim3(nblocks((long)n * 81)), dim3(BS), 0, s, gargs(g), kf, mf, cK, cM, KE, M8,
                       mask, n, A);
    TT_HIP(hipGetLastError());
}

// ---- symmetric coarsest solve on the packed lower triangle: 64 x 64 tiles (I, J), I >= J, stored
// contiguously in row-major order, tile index I (I + 1) / 2 + J.  Half the bytes of the full matrix.
constexpr int ST = 64;
static inline int sym_tiles(int n) { return (n + ST - 1) / ST; }
long sym_packed_size(int n) {
    const long nt = sym_tiles(n);
    return nt * (nt + 1) / 2 * ST * ST;
}
// pack the lower triangle of the full symmetric matrix A (n x n) into tiles (zero padded)
__global__ void k_sym_pack(const double* __restrict__ A, int n, double* __restrict__ P) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J > I) return;
    double* T = P + ((size_t)I * (I + 1) / 2 + J) * ST * ST;
    for (int e = threadIdx.x; e < ST * ST; e += blockDim.x) {
        const int r = e / ST, c = e % ST, i = I * ST + r, j = J * ST + c;
        T[e] = (i < n && j < n) ? A[(size_t)i * n + j] : 0.0;
    }
}
void launch_sym_pack(const double* A, int n, double* P, hipStream_t s) {
    const int nt = sym_tiles(n);
    hipLaunchKernelGGL(k_sym_pack, dim3(nt, nt), dim3(256), 0, s, A, n, P);
    TT_HIP(hipGetLastError());
}
// per tile: row sums A_IJ x_J and (off the diagonal) column sums A_IJ^T x_I, written as FP64
// partials.  Each thread first loads its 16 tile values (16-byte loads, all in flight), then the
// tile goes through LDS once.
template <typename PT>
__global__ __launch_bounds__(256) void k_sym_tiles(const PT* __restrict__ P, const double* __restrict__ x, int n,
                                                   double* __restrict__ prow, double* __restrict__ pcol) {
    __shared__ double tile[ST][ST + 1];
    __shared__ double xs[2][ST];
    __shared__ double red[4][ST];
    const int id = blockIdx.x;
    int I = (int)((sqrt(8.0 * id + 1.0) - 1.0) * 0.5);  // I (I + 1) / 2 <= id < (I + 1)(I + 2) / 2
    while ((I + 1) * (I + 2) / 2 <= id) I++;
    while (I * (I + 1) / 2 > id) I--;
    const int J = id - I * (I + 1) / 2;
    constexpr int VW = 16 / sizeof(PT), NL = ST * ST / (256 * VW);  // values per load, loads per thread
    using V = typename std::conditional<VW == 2, double2, float4>::type;
    const V* T2 = reinterpret_cast<const V*>(P + (size_t)id * ST * ST);
    const int tid = threadIdx.x;
    V v[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) v[k] = T2[tid + 256 * k];
    if (tid < ST) {
        const int j = J * ST + tid, i = I * ST + tid;
        xs[0][tid] = j < n ? x[j] : 0.0;
        xs[1][tid] = i < n ? x[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = VW * (tid + 256 * k);
        const PT* pv = reinterpret_cast<const PT*>(&v[k]);
#pragma unroll
        for (int q = 0; q < VW; q++) tile[e / ST][e % ST + q] = (double)pv[q];
    }
    __syncthreads();
    const int part = tid >> 6, k = tid & 63;  // 4 x 64 threads: rows (parts 0, 1) and columns (parts 2, 3)
    double sacc = 0.0;
    if (part < 2) {
#pragma unroll 8
        for (int c = part * 32; c < part * 32 + 32; c++) sacc = fma(tile[k][c], xs[0][c], sacc);
    } else {
#pragma unroll 8
        for (int r = (part - 2) * 32; r < (part - 2) * 32 + 32; r++) sacc = fma(tile[r][k], xs[1][r], sacc);
    }
    red[part][k] = sacc;
    __syncthreads();
    if (tid < ST) prow[(size_t)id * ST + tid] = red[0][tid] + red[1][tid];
    else if (tid < 2 * ST) pcol[(size_t)id * ST + (tid - ST)] = I == J ? 0.0 : red[2][tid - ST] + red[3][tid - ST];
}
// y_i = sum_{J <= I} prow(I, J) + sum_{K > I} pcol(K, I): one block of 256 threads per tile row I,
// 4 threads per output summing every 4th partial, combined in a fixed order
__global__ __launch_bounds__(256) void k_sym_sum(const double* __restrict__ prow, const double* __restrict__ pcol,
                                                 int n, int nt, double* __restrict__ y) {
    __shared__ double red[4][ST];
    const int I = blockIdx.x, q = threadIdx.x >> 6, r = threadIdx.x & 63;
    double s = 0.0;
    for (int m = q; m < nt; m += 4) {  // m <= I: tile (I, m) rows; m > I: tile (m, I) columns
        s += m <= I ? prow[((size_t)I * (I + 1) / 2 + m) * ST + r] : pcol[((size_t)m * (m + 1) / 2 + I) * ST + r];
    }
    red[q][r] = s;
    __syncthreads();
    const int i = I * ST + threadIdx.x;
    if (threadIdx.x < ST && i < n) y[i] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}
template <typename PT>
static void sym_solve(const PT* P, const double* b, double* x, int n, double* prow, double* pcol, hipStream_t s) {
    const int nt = sym_tiles(n);
    hipLaunchKernelGGL(k_sym_tiles<PT>, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, P, b, n, prow, pcol);
    TT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sym_sum, dim3(nt), dim3(256), 0, s, prow, pcol, n, nt, x);
    TT_HIP(hipGetLastError());
}
void launch_sym_solve(const double* P, const double* b, double* x, int n, double* prow, double* pcol,
                      hipStream_t s) {
    sym_solve(P, b, x, n, prow, pcol, s);
}

// ------------------------------------------------------------------ reductions / Krylov
// Reductions over long vectors use at most RB blocks (grid-stride), so that the
// consumers of their partials (next kernel or host) sum few values.
constexpr int RB = 1024;
static inline int rb_blocks(long n) {
    int b = nblocks(n);
    return b > RB ? RB : (b < 1 ? 1 : b);
}

__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ b, long n, double* part) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc[0] = fma(a[i], b[i], acc[0]);
    write_partials<1>(acc, part);
}
void launch_dot(const double* a, const double* b, long n, Reducer& red, int region, hipStream_t s) {
    const int nb = rb_blocks(n);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(BS), 0, s, a, b, n, red.part(region));
    red.nblk[region] = nb;
    red.nv[region] = 1;
    TT_HIP(hipGetLastError());
}

__global__ void k_expand3(Expand3 e, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < e.nb; k++) {
        const double q = e.Q[(size_t)k * n + i];
        s0 = fma(e.c[0][k], q, s0);
        s1 = fma(e.c[1][k], q, s1);
        s2 = fma(e.c[2][k], q, s2);
    }
    e.out[0][i] = s0;
    if (e.nout > 1) e.out[1][i] = s1;
    if (e.nout > 2) e.out[2][i] = s2;
}
void launch_expand3(const Expand3& e, long n, hipStream_t s) {
    TT_CHECK(e.nb >= 1 && e.nb <= 64 && e.nout >= 1 && e.nout <= 3, -3, "launch_expand3: 1..64 vectors, 1..3 outputs");
    hipLaunchKernelGGL(k_expand3, dim3(nblocks(n)), dim3(BS), 0, s, e, n);
    TT_HIP(hipGetLastError());
}

template <int NQ>
__global__ void k_proj3(const double* __restrict__ Q, int k0, const double* __restrict__ a,
                        const double* __restrict__ b, const double* __restrict__ c, long n, double* part) {
    double acc[3 * NQ];
#pragma unroll
    for (int j = 0; j < 3 * NQ; j++) acc[j] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double ai = a[i], bi = b[i], ci = c[i];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double qi = Q[(size_t)(k0 + q) * n + i];
            acc[3 * q] = fma(qi, ai, acc[3 * q]);
            acc[3 * q + 1] = fma(qi, bi, acc[3 * q + 1]);
            acc[3 * q + 2] = fma(qi, ci, acc[3 * q + 2]);
        }
    }
    __shared__ double sh[(BS / 64) * 3 * NQ];
    block_sum<3 * NQ>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++)
            for (int j = 0; j < 3; j++) part[((size_t)(k0 + q) * RB + blockIdx.x) * 3 + j] = acc[3 * q + j];
}
int launch_proj3(const double* Q, int k0, int nq, const double* a, const double* b, const double* c, long n,
                 double* part, hipStream_t s) {
    const int nb = rb_blocks(n);
    switch (nq) {
#define TT_PROJ3(NQ) \
    case NQ: hipLaunchKernelGGL(k_proj3<NQ>, dim3(nb), dim3(BS), 0, s, Q, k0, a, b, c, n, part); break;
        TT_PROJ3(1) TT_PROJ3(2) TT_PROJ3(3) TT_PROJ3(4) TT_PROJ3(5) TT_PROJ3(6) TT_PROJ3(7) TT_PROJ3(8)
#undef TT_PROJ3
    default: throw Error{-3, "launch_proj3: 1..8 vectors per launch"};  // TT_ERR_INTERNAL
    }
    TT_HIP(hipGetLastError());
    return nb;
}

// gamma = z.r, delta = z.r_prev (flexible CG, Polak-Ribiere form)
__global__ void k_fcg_dots(const double* __restrict__ z, const double* __restrict__ r, const double* __restrict__ rp,
                           long n, double* part, int first) {
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        acc[0] = fma(z[i], r[i], acc[0]);
        if (!first) acc[1] = fma(z[i], rp[i], acc[1]);
    }
    write_partials<2>(acc, part);
}
// gamma = z.r and beta = (z.r - z.r_prev) / gamma_old (0 on a (re)start, and once the solve is
// frozen, when gamma_old may be 0) from the partials of k_fcg_dots
__global__ __launch_bounds__(BSR) void k_fcg_beta(const double* __restrict__ dots_part, int dots_nblk, double* scal, int first) {
    double d[2];
    sum_partials<2>(dots_part, dots_nblk, d);
    if (threadIdx.x == 0) {
        scal[SC_GAMMA] = d[0];
        scal[SC_BETA] = (first || scal[SC_DONE] != 0.0) ? 0.0 : (d[0] - d[1]) / scal[SC_GAMMA_OLD];
    }
}
void launch_fcg_dots(const double* z, const double* r, const double* rprev, long n, Reducer& red, bool first,
                     hipStream_t s) {
    const int nb = rb_blocks(n);
    hipLaunchKernelGGL(k_fcg_dots, dim3(nb), dim3(BS), 0, s, z, r, rprev, n, red.part(PB_DOTS), (int)first);
    red.nblk[PB_DOTS] = nb;
    red.nv[PB_DOTS] = 2;
    TT_HIP(hipGetLastError());
}

// Start (or restart) of a flexible-CG solve, from the partials of b.b and of r.r for the current
// x: the tolerance (rtol ||b||)^2, whether b = 0, and whether x already meets the tolerance.  The
// b.b total replaces the first partial of its region (which then counts one block): the host reads
// it back at every convergence look, and a copy of a few thousand partials takes a slower path
// (measured: ~20 us more per look for 21 KB than for 8 KB).
__global__ __launch_bounds__(BSR) void k_fcg_init(double* __restrict__ bb_part, int bb_nblk, const double* __restrict__ rr_part,
                           int rr_nblk, double rtol2, int reset_iters, double* scal, double* look) {
    double bb[1], rr[1];
    sum_partials<1>(bb_part, bb_nblk, bb);
    sum_partials<1>(rr_part, rr_nblk, rr);
    if (threadIdx.x == 0) {
        bb_part[0] = bb[0];
        look[Reducer::kLookBB] = bb[0];
        const double tol2 = rtol2 * bb[0];
        scal[SC_TOL2] = tol2;
        scal[SC_ZERO] = bb[0] == 0.0 ? 1.0 : 0.0;
        scal[SC_DONE] = (bb[0] == 0.0 || rr[0] <= tol2) ? 1.0 : 0.0;
        if (reset_iters) scal[SC_ITERS] = 0.0;
    }
}
void launch_fcg_init(Reducer& red, int bb_region, int rr_region, double rtol2, bool reset_iters, hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_init, dim3(1), dim3(BSR), 0, s, red.part(bb_region), red.nblk[bb_region] * red.nv[bb_region],
                       red.part(rr_region), red.nblk[rr_region] * red.nv[rr_region], rtol2, reset_iters ? 1 : 0,
                       red.scal.p, red.look_dev);
    TT_HIP(hipGetLastError());
    red.nblk[bb_region] = 1;
    red.nv[bb_region] = 1;
}

// alpha = gamma / (p.q) from the partials of the FCG apply (one block); gamma becomes gamma_old.
// After the first iteration of a (re)start, the previous update's r.r (its partials) is checked
// too: once it is <= (rtol ||b||)^2, SC_DONE stays set and alpha is 0 for the rest of the solve,
// so x is frozen at the first iteration that met the tolerance.
__global__ __launch_bounds__(BSR) void k_fcg_alpha(const double* __restrict__ pq_part, int pq_nblk, const double* __restrict__ rr_part,
                            int rr_nblk, double* scal, int first) {
    double pq[1], rr[1] = {0.0};
    sum_partials<1>(pq_part, pq_nblk, pq);
    if (!first) sum_partials<1>(rr_part, rr_nblk, rr);
    if (threadIdx.x == 0) {
        const bool done = scal[SC_DONE] != 0.0 || (!first && rr[0] <= scal[SC_TOL2]);
        const double gam = scal[SC_GAMMA];
        scal[SC_DONE] = done ? 1.0 : 0.0;
        scal[SC_ALPHA] = done ? 0.0 : gam / pq[0];
        scal[SC_GAMMA_OLD] = gam;
        if (!done) scal[SC_ITERS] += 1.0;
    }
}

// x += alpha p (x = 0 if b = 0); r_prev = r; r -= alpha q; partials of r.r.  A frozen solve keeps
// x and r as they are.
__global__ void k_fcg_update(double* __restrict__ x, const double* __restrict__ p, double* __restrict__ r,
                             double* __restrict__ rp, const double* __restrict__ q, long n,
                             const double* __restrict__ scal, double* part, double* look) {
    const double al = scal[SC_ALPHA];
    const bool zero = scal[SC_ZERO] != 0.0, done = scal[SC_DONE] != 0.0;
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (zero) x[i] = 0.0;
        else if (!done) x[i] = fma(al, p[i], x[i]);
        const double ri = r[i];
        rp[i] = ri;
        const double rn = done ? ri : fma(-al, q[i], ri);
        r[i] = rn;
        acc[0] = fma(rn, rn, acc[0]);
    }
    __shared__ double sh[16];
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) {  // the partial for k_fcg_alpha, and its copy for the host's look
        part[blockIdx.x] = acc[0];
        look[blockIdx.x] = acc[0];
        if (blockIdx.x == 0) look[gridDim.x] = scal[SC_ITERS];
    }
}
void launch_fcg_update(double* x, const double* p, double* r, double* rprev, const double* q, long n, Reducer& red,
                       bool first, hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_alpha, dim3(1), dim3(BSR), 0, s, red.part(PB_PQ), red.nblk[PB_PQ], red.part(PB_RR),
                       red.nblk[PB_RR], red.scal.p, first ? 1 : 0);
    TT_HIP(hipGetLastError());
    const int nb = rb_blocks(n);
    TT_CHECK(nb < Reducer::kLookBB, -3, "look buffer too small");
    hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(BS), 0, s, x, p, r, rprev, q, n, red.scal.p, red.part(PB_RR),
                       red.look_dev);
    red.nblk[PB_RR] = nb;
    red.nv[PB_RR] = 1;
    TT_HIP(hipGetLastError());
}

__global__ void k_fcg_direction(const double* __restrict__ z, const double* __restrict__ pold,
                                double* __restrict__ pnew, long n, const double* __restrict__ scal, int first) {
    const double beta = first ? 0.0 : scal[SC_BETA];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        pnew[i] = first ? z[i] : fma(beta, pold[i], z[i]);
}
void launch_fcg_direction(const double* z, const double* p_old, double* p_new, long n, Reducer& red, bool first,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_fcg_beta, dim3(1), dim3(BSR), 0, s, red.part(PB_DOTS), red.nblk[PB_DOTS], red.scal.p,
                       first ? 1 : 0);
    TT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fcg_direction, dim3(nblocks(n)), dim3(BS), 0, s, z, p_old, p_new, n, red.scal.p,
                       first ? 1 : 0);
    TT_HIP(hipGetLastError());
}

__global__ void k_jacobi(const double* __restrict__ dinv, const double* __restrict__ r, double* __restrict__ z,
                         long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) z[i] = dinv[i] * r[i];
}
void launch_jacobi(const double* dinv, const double* r, double* z, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi, dim3(nblocks(n)), dim3(BS), 0, s, dinv, r, z, n);
    TT_HIP(hipGetLastError());
}

// power iteration helper: x = dinv.Ax, partials of ||x||^2 in PB_MISC
template <typename T>
__global__ void k_norm_power(const T* __restrict__ Ax, const T* __restrict__ dinv, T* __restrict__ x, long n,
                             double* part) {
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const T v = dinv[i] * Ax[i];
        x[i] = v;
        acc[0] += (double)v * (double)v;
    }
    write_partials<1>(acc, part);
}
void launch_norm_power(const void* Ax, const void* dinv, void* x, long n, Reducer& red, bool fp64, hipStream_t s) {
    const int nb = rb_blocks(n);
    if (fp64)
        hipLaunchKernelGGL(k_norm_power<double>, dim3(nb), dim3(BS), 0, s, (const double*)Ax, (const double*)dinv,
                           (double*)x, n, red.part(PB_MISC));
    else
        hipLaunchKernelGGL(k_norm_power<float>, dim3(nb), dim3(BS), 0, s, (const float*)Ax, (const float*)dinv,
                           (float*)x, n, red.part(PB_MISC));
    red.nblk[PB_MISC] = nb;
    red.nv[PB_MISC] = 1;
    TT_HIP(hipGetLastError());
}

// x /= ||x|| with ||x||^2 summed from the partials of k_norm_power
template <typename T>
__global__ void k_scale_by_norm(T* x, long n, const double* part, int nblk) {
    double s[1];
    sum_partials<1>(part, nblk, s);
    const double inv = 1.0 / sqrt(s[0]);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = (T)((double)x[i] * inv);
}
void launch_scale_by_norm(void* x, long n, Reducer& red, bool fp64, hipStream_t s) {
    const int nb = rb_blocks(n);
    if (fp64)
        hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(nb), dim3(BS), 0, s, (double*)x, n, red.part(PB_MISC),
                           red.nblk[PB_MISC]);
    else
        hipLaunchKernelGGL(k_scale_by_norm<float>, dim3(nb), dim3(BS), 0, s, (float*)x, n, red.part(PB_MISC),
                           red.nblk[PB_MISC]);
    TT_HIP(hipGetLastError());
}

__global__ void k_scale(double* x, long n, double a) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}
void launch_scale(double* x, long n, double a, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(BS), 0, s, x, n, a);
    TT_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ Newmark / adjoint vector kernels
// corrector (LinearElasticity.cc:1363-1375): c = a[0..5].  The acceleration history of the new state
// is shifted in the same pass: hout[0] = a, hout[j] = hin[j - 1] (j < nh), vectors n apart.  With gout,
// the next step's initial guess sum_k gc.c[k] w_k over w = (un, vn, an, a, hin[0], hin[1]) is written as
// well, term by term in the order k_extrapolate sums the same vectors (bit-identical results).
struct GuessCoef {
    double c[6];
    int m;
};
__global__ void k_newmark_update(const double* u, const double* __restrict__ v,
                                 const double* __restrict__ a, const double* __restrict__ un, double* __restrict__ vn,
                                 double* __restrict__ an, double a0, double a1, double a2, double a3, double a4,
                                 double a5, long n, const double* __restrict__ hin, double* __restrict__ hout, int nh,
                                 GuessCoef gc, double* gout) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ai = a[i];
    if (nh > 0) hout[i] = ai;
    for (int j = 1; j < nh; j++) hout[j * n + i] = hin[(j - 1) * n + i];
    const double uni = un[i];
    const double du = uni - u[i];
    double vv = 0.0;
    vv += a0 * v[i];
    vv += a1 * ai;
    vv += a2 * du;
    double aa = 0.0;
    aa += -a3 * v[i];
    aa += -a4 * ai;
    aa += a5 * du;
    vn[i] = vv;
    an[i] = aa;
    if (gout) {  // the next step's guess (gout may be this step's input u: u[i] is already read)
        double s = gc.c[0] * uni;
        if (gc.m > 1) s = fma(gc.c[1], vv, s);
        if (gc.m > 2) s = fma(gc.c[2], aa, s);
        if (gc.m > 3) s = fma(gc.c[3], ai, s);
        if (gc.m > 4) s = fma(gc.c[4], hin[i], s);
        if (gc.m > 5) s = fma(gc.c[5], hin[n + i], s);
        gout[i] = s;
    }
}
void launch_newmark_update(const double* u, const double* v, const double* a, const double* un, double* vn,
                           double* an, const double* c, long n, const double* hin, double* hout, int nh,
                           hipStream_t s, const Extrap* next, double* gout) {
    GuessCoef gc{};
    if (next) {
        TT_CHECK(next->m >= 1 && next->m <= 6 && nh >= 3, -3, "launch_newmark_update: guess of 1..6 terms");
        for (int k = 0; k < next->m; k++) gc.c[k] = next->c[k];
        gc.m = next->m;
    }
    hipLaunchKernelGGL(k_newmark_update, dim3(nblocks(n)), dim3(BS), 0, s, u, v, a, un, vn, an, c[0], c[1], c[2],
                       c[3], c[4], c[5], n, hin, hout, nh, gc, next ? gout : (double*)nullptr);
    TT_HIP(hipGetLastError());
}

// Adjoint step combination (LinearElasticity.cc:864-925).  coef = {a0..a5, dalpha, dbeta}.
// Kl = N K N lambda_{t+1}, Ml = N M N lambda_{t+1}; B1 = a5 M + a2 C, B2 = a3 M - a0 C, B3 = a4 M - a1 C.
// rhs = N.(-R1): the SPD form of the reference's solve with A = -(K + a2 C + a5 M).
struct AdjCoef {
    double a[6], al, be;
};
__global__ void k_adjoint_combine(const double* __restrict__ Kl, const double* __restrict__ Ml,
                                  const double* __restrict__ hdu, const double* __restrict__ hddu,
                                  const double* __restrict__ hdddu, const double* __restrict__ mun,
                                  const double* __restrict__ nun, const uint8_t* __restrict__ mask, AdjCoef c,
                                  int terminal, double* __restrict__ mu, double* __restrict__ nu,
                                  double* __restrict__ rhs, long nd) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const double* a = c.a;
    double m, v, r;
    if (terminal) {
        m = -hddu[i];
        v = -hdddu[i];
        r = hdu[i];
        r += -a[2] * m;
        r += -a[5] * v;
        r = -r;
    } else {
        const double kl = Kl[i], ml = Ml[i];
        const double b1 = (a[5] + a[2] * c.al) * ml + a[2] * c.be * kl;
        const double b2 = (a[3] - a[0] * c.al) * ml - a[0] * c.be * kl;
        const double b3 = (a[4] - a[1] * c.al) * ml - a[1] * c.be * kl;
        double mm = hddu[i];
        mm += b2;
        mm += -a[0] * mun[i];
        mm += a[3] * nun[i];
        double vv = hdddu[i];
        vv += b3;
        vv += a[4] * nun[i];
        r = hdu[i];
        r += b1;
        r += a[2] * mun[i];
        r += a[5] * nun[i];
        r = -r;
        m = -mm;
        v = -vv;
        r += a[2] * m;
        r += a[5] * v;
    }
    mu[i] = m;
    nu[i] = v;
    const bool clamped = (mask[i / 3] >> (i % 3)) & 1;
    rhs[i] = clamped ? 0.0 : -r;
}
void launch_adjoint_combine(const double* Kl, const double* Ml, const double* hdu, const double* hddu,
                            const double* hdddu, const double* mun, const double* nun, const uint8_t* mask,
                            const double* coef, bool terminal, double* mu, double* nu, double* rhs, long nn,
                            hipStream_t s) {
    AdjCoef c;
    for (int k = 0; k < 6; k++) c.a[k] = coef[k];
    c.al = coef[6];
    c.be = coef[7];
    long nd = 3 * nn;
    hipLaunchKernelGGL(k_adjoint_combine, dim3(nblocks(nd)), dim3(BS), 0, s, Kl, Ml, hdu, hddu, hdddu, mun, nun, mask,
                       c, (int)terminal, mu, nu, rhs, nd);
    TT_HIP(hipGetLastError());
}

// t = 0 adjoint right-hand side (LinearElasticity.cc:753-776): R3 = N.-(HDDDU + B3 l1 - a1 mu1 + a4 nu1)
__global__ void k_adjoint0_rhs(const double* __restrict__ Kl, const double* __restrict__ Ml,
                               const double* __restrict__ hdddu, const double* __restrict__ mun,
                               const double* __restrict__ nun, const uint8_t* __restrict__ mask, AdjCoef c,
                               double* __restrict__ rhs, long nd) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const double* a = c.a;
    const double b3 = (a[4] - a[1] * c.al) * Ml[i] - a[1] * c.be * Kl[i];
    double r = hdddu[i];
    r += b3;
    r += -a[1] * mun[i];
    r += a[4] * nun[i];
    const bool clamped = (mask[i / 3] >> (i % 3)) & 1;
    rhs[i] = clamped ? 0.0 : -r;
}
void launch_adjoint0_rhs(const double* Kl, const double* Ml, const double* hdddu, const double* mun,
                         const double* nun, const uint8_t* mask, const double* coef, double* rhs, long nn,
                         hipStream_t s) {
    AdjCoef c;
    for (int k = 0; k < 6; k++) c.a[k] = coef[k];
    c.al = coef[6];
    c.be = coef[7];
    long nd = 3 * nn;
    hipLaunchKernelGGL(k_adjoint0_rhs, dim3(nblocks(nd)), dim3(BS), 0, s, Kl, Ml, hdddu, mun, nun, mask, c, rhs, nd);
    TT_HIP(hipGetLastError());
}

// objectives 0..2: d/du of sum_e P_e dt/2 sum_k w_k^2 = dt * w_d * (sum of P over elements at the node)
__global__ void k_hdiag(const double* __restrict__ w, const double* __restrict__ cnt, double dt, double* __restrict__ h,
                        long nd) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < nd) h[i] = dt * w[i] * cnt[i / 3];
}
void launch_hdiag(const double* w, const double* cnt, double dt, double* h, long nn, hipStream_t s) {
    hipLaunchKernelGGL(k_hdiag, dim3(nblocks(3 * nn)), dim3(BS), 0, s, w, cnt, dt, h, 3 * nn);
    TT_HIP(hipGetLastError());
}
__global__ void k_region_count(GridArgs g, const double* __restrict__ P, double* __restrict__ cnt) {
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)g.nx * g.ny * g.nz) return;
    const int i = n % g.nx, j = (n / g.nx) % g.ny, k = n / ((long)g.nx * g.ny);
    double s = 0.0;
    for (int ez = 0; ez < 2; ez++)
        for (int ey = 0; ey < 2; ey++)
            for (int ex = 0; ex < 2; ex++) {
                const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
                if (ei < 0 || ej < 0 || ek < 0 || ei >= g.ex || ej >= g.ey || ek >= g.ez) continue;
                s += P[ei + (long)g.ex * (ej + (long)g.ey * ek)];
            }
    cnt[n] = s;
}
void launch_region_count(const Grid& g, const double* P, double* cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_region_count, dim3(nblocks(g.nn)), dim3(BS), 0, s, gargs(g), P, cnt);
    TT_HIP(hipGetLastError());
}

}  // namespace tt
