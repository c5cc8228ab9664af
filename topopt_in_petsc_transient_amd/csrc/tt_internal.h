// Internal declarations of the MI355X transient-elasticity library.
//
// Layout in HBM (natural ordering of the reference's single-rank DMDA):
//   node  n = i + nx*(j + ny*k)      vectors: 3 dofs per node, interleaved (3n+c)
//   elem  e = i + ex*(j + ey*k)      per-element coefficients kfac/mfac
// Each multigrid level l holds its own element coefficients, Dirichlet mask
// (one byte per node, bit c set when dof c is clamped), inverse diagonal and
// vectors.  Precision of level vectors is FP64 in the FP64 mode and FP32 in
// the mixed mode; coarse-level coefficients are BF16 in the mixed mode.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tt {

typedef uint16_t bf16_t;

// ------------------------------------------------------------------ errors
struct Error {
    int code;
    std::string msg;
};

#define TT_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::tt::Error{-2, std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                                      __FILE__ + ":" + std::to_string(__LINE__)};               \
    } while (0)

#define TT_CHECK(cond, code, msg)                      \
    do {                                               \
        if (!(cond)) throw ::tt::Error{(code), (msg)}; \
    } while (0)

// ------------------------------------------------------------------ device buffers
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t n_) { alloc(n_); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t n_) {
        release();
        n = n_;
        if (n) TT_HIP(hipMalloc((void**)&p, sizeof(T) * n));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
};

// message of a failed call made without a context (tt_last_error(NULL) returns it)
void set_contextless_error(const std::string& msg);

// ------------------------------------------------------------------ grid
struct Grid {
    int ex, ey, ez;  // elements
    int nx, ny, nz;  // nodes
    long nel, nn, nd;
    double hx, hy, hz;
};

inline Grid make_grid(int ex, int ey, int ez, double lx, double ly, double lz) {
    Grid g;
    g.ex = ex; g.ey = ey; g.ez = ez;
    g.nx = ex + 1; g.ny = ey + 1; g.nz = ez + 1;
    g.nel = (long)ex * ey * ez;
    g.nn = (long)g.nx * g.ny * g.nz;
    g.nd = 3 * g.nn;
    g.hx = lx / ex; g.hy = ly / ey; g.hz = lz / ez;
    return g;
}

// ------------------------------------------------------------------ host element matrices
void hex8_stiffness(double hx, double hy, double hz, double nu, double* ke);  // 24x24, E = 1
void hex8_mass(double hx, double hy, double hz, double* m8);                 // 8x8 scalar, density 1
void hex8_helmholtz(double hx, double hy, double hz, double R, double* kf);  // 8x8, R^2 Lap + mass

// ------------------------------------------------------------------ reductions
// Deterministic device reductions without fences or atomics: a reducing kernel
// writes one partial per block into a partial buffer; the next kernel on the
// stream (or the host) sums the partials in block order.  Kernel boundaries
// give the visibility, so no block pays a device-scope fence (which on a
// multi-XCD GPU means an L2 write-back).
enum { PB_DOTS = 0, PB_PQ, PB_RR, PB_MISC, PB_MISC2, PB_COUNT };
struct Reducer {
    DBuf<double> partials;   // PB_COUNT regions of [maxblocks * 4]
    DBuf<double> scal;       // device scalars (see SC_* indices)
    double* host = nullptr;  // pinned mirror of scal
    double* hpart = nullptr; // pinned host copy of one partial region
    // Host memory the kernels store into directly (pinned, mapped, coherent), so that the flexible-CG
    // convergence look needs no copy: [0, nb) the r.r partials of the last k_fcg_update, [nb] the
    // iterations that moved x; [kLookBB] the b.b of the solve (k_fcg_init).
    static constexpr int kLookBB = 1100;
    double* look = nullptr;
    double* look_dev = nullptr;  // the same memory as the kernels address it
    int maxblocks = 0;
    int nblk[PB_COUNT] = {};  // blocks that wrote each region
    int nv[PB_COUNT] = {};
    void init(int maxb);
    double* part(int region) const { return partials.p + (size_t)region * maxblocks * 4; }
    // sum a region on the host in block order (synchronises the stream)
    void host_sum(int region, hipStream_t s, double* out);
    ~Reducer();
};

enum {
    SC_GAMMA = 0,   // z.r of current iteration
    SC_GAMMA_OLD,   // z.r of previous iteration
    SC_DELTA,       // z.r_prev
    SC_BETA,
    SC_PQ,
    SC_ALPHA,
    SC_RR,          // r.r
    SC_TMP0,
    SC_TMP1,
    SC_TMP2,
    SC_TMP3,
    SC_DONE,        // flexible CG: 1 once the recursive residual met the tolerance (x is then frozen)
    SC_TOL2,        // flexible CG: (rtol ||b||)^2
    SC_ZERO,        // flexible CG: 1 if b = 0 (the solution is x = 0)
    SC_ITERS,       // flexible CG: iterations that moved x in this solve
    SC_FX0,         // objective accumulators start here
    SC_COUNT = 32
};

// ------------------------------------------------------------------ kernel launchers (tt_kernels.hip)
// Elasticity operator on one level.  Coefficient storage CT is double, float or BF16.
enum CoefKind { COEF_F64 = 0, COEF_F32 = 1, COEF_BF16 = 2 };

// Batch tables of the node tiles of one grid (see tt_apply_mfma.inc): per tile position (bits 0/1: the
// first / last tile in x, bits 2/3 in y, bits 4/5 in z) the offset and batch count of its table in
// `codes`.  A tile's table holds the elements that exist and touch one of its nodes: the first tile in
// a direction has no elements on its low side, and the last tile owns fewer nodes than a full tile
// when the grid's node count is not a multiple of the tile extent.
// Element-local corner numbering of the reference (LinearElasticity.cc:3067-3074)
__host__ __device__ constexpr int corner_id(int x, int y, int z) { return z * 4 + (y == 0 ? (x == 0 ? 0 : 1) : (x == 1 ? 2 : 3)); }
// (i, j, k) of node or element n on an nx x ny x * grid, x fastest: two 32-bit unsigned divisions
// (indices fit in 32 bits; 64-bit division is a long branching sequence on the GPU)
__host__ __device__ inline void grid_ijk(unsigned n, unsigned nx, unsigned ny, int& i, int& j, int& k) {
    const unsigned row = n / nx, kk = row / ny;
    i = (int)(n - row * nx);
    j = (int)(row - kk * ny);
    k = (int)kk;
}
// The sensitivity kernel's element dof order: k = 4 t + q (t = 0..5) is dof 3 a + c of corner
// a = sens_corner(q, t >= 3), component c = t mod 3.  Lane group q owns the two x-adjacent corners of one
// element edge (corners 0-1, 3-2, 4-5, 7-6), whose six dofs are consecutive in memory from the edge's
// low-x node.
__host__ __device__ constexpr int sens_corner(int q, int hi) {  // low-x / high-x corner of edge q
    return q == 0 ? hi : q == 1 ? 3 - hi : q == 2 ? 4 + hi : 7 - hi;
}
__host__ __device__ constexpr int sens_dof(int k) {  // k = 4 t + q -> element dof 3 a + c
    return 3 * sens_corner(k & 3, (k >> 2) >= 3) + (k >> 2) % 3;
}
__host__ __device__ constexpr int cx_of(int b) { return (b & 3) == 1 || (b & 3) == 2 ? 1 : 0; }
__host__ __device__ constexpr int cy_of(int b) { return (b & 3) >= 2 ? 1 : 0; }
__host__ __device__ constexpr int cz_of(int b) { return b >> 2; }

struct BatchTab {
    const int4* codes;
    const int2* at;  // per position: (offset, batch count), in device memory (keeps the launch arguments small)
};
struct BatchTabStore {
    DBuf<int4> buf;
    DBuf<int2> at;
    BatchTab t4{}, t2{};  // 8 x 8 x 4 and 8 x 8 x 2 node tiles
};
// tables of both tile depths for grid g, uploaded into st.buf
void build_batch_tables(const Grid& g, BatchTabStore& st);

struct OpLevel {
    Grid g;
    int coef_kind;         // CoefKind
    const void* kf;        // element stiffness factors (kfac)
    const void* mf;        // element mass factors (mfac)
    const double* KE64;    // 576
    const double* M864;    // 64
    const float* KE32;
    const float* M832;
    const uint8_t* mask;   // per node clamped bits
    const double* AF64;    // MFMA A fragments of [KE; M8 (x) I3] in lane order (18 x 64)
    const float* AF32;
    const uint32_t* AFH;   // FP16-split fragments (3 x 2 x 64 x 4 words), K / M rows scaled by 1 / hs
    float hs[2];
    const uint8_t* tbc4;   // per 8x8x4 / 8x8x2 tile: 1 if its staged node box holds a clamped node
    const uint8_t* tbc2;
    BatchTab tab4, tab2;   // element batch tables of the 8x8x4 / 8x8x2 tiles
};
// Per-tile flags of 8 x 8 x tz node tiles: 1 if the tile's node box (one halo layer) holds a node
// with a nonzero mask.
void tile_bc_flags(const Grid& g, const std::vector<uint8_t>& mask, int tz, std::vector<uint8_t>& f);

// MFMA A fragments of the interleaved [KE; M8 (x) I3] matrix, lane order (see tt_kernels.hip)
// Symmetric coarsest solve on the packed lower triangle (64 x 64 tiles): size of the packed array,
// packing of a full symmetric matrix, and x = A b with per-tile partials (nt (nt + 1) / 2 x 64 each).
long sym_packed_size(int n);
void launch_sym_pack(const double* A, int n, double* P, hipStream_t s);
void launch_sym_solve(const double* P, const double* b, double* x, int n, double* prow, double* pcol, hipStream_t s);
// Dense N (cK K + cM M) N + I - N of a (coarsest) level from its FP64 element coefficients, on the device
// (row-major n x n, n = g.nd; cleared first).
void launch_dense_coarse(const Grid& g, const double* kf, const double* mf, double cK, double cM, const double* KE,
                         const double* M8, const uint8_t* mask, double* A, hipStream_t s);
// In-place inverse of an m x m (m <= 64) symmetric positive definite block, column-major with leading
// dimension lda (one block, Gauss-Jordan in registers); *info = 1 if a pivot is not positive.
void launch_spd_inv_small(double* A, int m, int lda, int* info, hipStream_t s);
// dst (cols x rows) = src (rows x cols) transposed, column-major with leading dimensions lds, ldd.
void launch_transpose(const double* src, int rows, int cols, int lds, double* dst, int ldd, hipStream_t s);
constexpr int AF32_WORDS = 18 * 64 + 3 * 2 * 64 * 4;  // FP32 fragments followed by the FP16-split ones
void mfma_fragments(const double* KE, const double* M8, double* af64, float* af32, uint32_t* afh, float* hs);

// y = N.(cK*K + cM*M)(N.u) on free dofs, identity on clamped dofs.  With red: partials of u.y in
// partial region dot_slot (PB_*).  Reductions below likewise take a partial region.
void launch_apply_bc(const OpLevel& L, double cK, double cM, const void* u, void* y, bool fp64, hipStream_t s,
                     Reducer* red = nullptr, int dot_slot = -1);
// r = b - A x
void launch_residual(const OpLevel& L, double cK, double cM, const void* b, const void* x, void* r, bool fp64,
                     hipStream_t s);
// x = c1 D^-1 b and r = b - A x in one pass
void launch_resid_first(const OpLevel& L, double cK, double cM, const void* b, const void* dinv, double c1, void* x,
                        void* r, bool fp64, hipStream_t s);
// r = b - A x (FP64) with the partials of b.b and r.r in two partial regions
void launch_residual_dots(const OpLevel& L, double cK, double cM, const double* b, const double* x, double* r,
                          Reducer& red, int bb_region, int rr_region, hipStream_t s);
// Chebyshev step: r = b - A x; dn = c0*d + c1*dinv.r; xn = x + dn  (d may be null => c0 ignored)
void launch_cheb_step(const OpLevel& L, double cK, double cM, const void* b, const void* x, const void* d,
                      const void* dinv, double c0, double c1, void* xn, void* dn, bool fp64, hipStream_t s);
// split outputs: yk = N.(cK K (N.u)), ym = N.(cM M (N.u))  (FP64)
void launch_apply_split(const OpLevel& L, double cK, double cM, const double* u, double* yk, double* ym,
                        hipStream_t s);
// unmasked K/M apply with arbitrary coefficient arrays (FP64): y = cK*K(kf)u + cM*M(mf)u
void launch_apply_raw(const OpLevel& L, const double* kf, const double* mf, double cK, double cM, const double* u,
                      double* y, hipStream_t s);
// Newmark right-hand side: R = N.( fs*F + sum_i (K-weights, M-weights) applied to u,v,a )
void launch_newmark_rhs(const OpLevel& L, const double* u, const double* v, const double* a, const double* F,
                        double fs, const double wK[3], const double wM[3], double* R, hipStream_t s);
// diagonal of (cK K + cM M) with identity on clamped dofs, inverted
void launch_diag_inv(const OpLevel& L, double cK, double cM, void* dinv, bool fp64, hipStream_t s);

// FCG fused apply: p = z + beta p_old (beta from scal) and q = A p = A z + beta q_old, partial p.q ->
// alpha = gamma / pq.  Only z is staged with its halo; p_old and q_old are read at the owned nodes.
// beta = false: the tile kernel alone, with the beta already in scal (measurement)
void launch_fcg_apply(const OpLevel& L, double cK, double cM, const double* z, const double* p_old, double* p_new,
                      const double* q_old, double* q_new, Reducer& red, bool first, hipStream_t s, bool beta = true);

// vector kernels
void launch_fill(void* x, double v, long n, bool fp64, hipStream_t s);
void launch_copy(const void* src, void* dst, long n, bool src64, bool dst64, hipStream_t s);
void launch_cheb_first(const void* b, const void* dinv, double c1, void* x, long n, bool fp64, hipStream_t s);
// b = (T) r and x = c1 D^-1 b (finest level, r in FP64)
void launch_cheb_first_from64(const double* r, void* b, const void* dinv, double c1, void* x, long n, bool fp64,
                              hipStream_t s);
// first two Chebyshev iterates in one operator apply: x0 = c1_first D^-1 b formed on the fly
// (b = (T) r64 written out when r64 is given), then dn = c0 x0 + c1 D^-1 (b - A x0), xn = x0 + dn
void launch_cheb_first_step(const OpLevel& L, double cK, double cM, const double* r64, void* b, const void* dinv,
                            double c1_first, double c0, double c1, void* xn, void* dn, bool fp64, hipStream_t s);
// last post-smoothing step of the finest level: z (FP64) = x + dn, partials of z.r and z.r_prev into PB_DOTS
void launch_cheb_last(const OpLevel& L, double cK, double cM, const void* b, const void* x, const void* d,
                      const void* dinv, double c0, double c1, double* z, const double* r, const double* rp, bool first,
                      bool fp64, Reducer& red, hipStream_t s);
// bc = R rf (fine precision fp64; an FP32 fine level may write an FP64 coarse vector, out64)
void launch_restrict(const Grid& fine, const Grid& coarse, const void* rf, const uint8_t* cmask, void* bc,
                     bool fp64, bool out64, hipStream_t s);
// xf += P xc (fine precision fp64; an FP32 fine level may read an FP64 coarse vector, in64)
void launch_prolong_add(const Grid& fine, const Grid& coarse, const void* xc, const uint8_t* fmask, void* xf,
                        bool fp64, bool in64, hipStream_t s);
void launch_coarsen_coef(const Grid& fine, const Grid& coarse, const double* kf_f, const double* mf_f, double* kf_c,
                         double* mf_c, hipStream_t s);
void launch_convert_coef(const double* src, void* dst, long n, int kind, hipStream_t s);
void launch_interp(const double* x, long n, int sInt, double Emin, double Emax, double pK, int mInt, double Rmin,
                   double Rmax, double pM, const double* P, double* kf, double* mf, double* dkf, double* dmf,
                   double* Pkf, double* Pmf, hipStream_t s);

// dots (FP64): results land in red.scal[slot..]
void launch_dot(const double* a, const double* b, long n, Reducer& red, int slot, hipStream_t s);
// Q_k.a, Q_k.b, Q_k.c for the nq <= 8 vectors Q_k = Q + k*n, k0 <= k < k0 + nq, in one pass over a, b, c:
// block partials into part[(k*1024 + block)*3 + i]; returns the number of blocks (at most 1024)
int launch_proj3(const double* Q, int k0, int nq, const double* a, const double* b, const double* c, long n,
                 double* part, hipStream_t s);
// out[o] = sum_k c[o][k] Q_k for nout <= 3 outputs and nb <= 64 vectors Q_k = Q + k*n, one pass over Q
struct Expand3 {
    const double* Q;
    double* out[3];
    double c[3][64];
    int nb, nout;
};
void launch_expand3(const Expand3& e, long n, hipStream_t s);
void launch_fcg_dots(const double* z, const double* r, const double* rprev, long n, Reducer& red, bool first,
                     hipStream_t s);
// Flexible-CG (re)start from the b.b and r.r partial regions: device tolerance (rtol ||b||)^2, the
// b = 0 flag, and whether the current x already meets the tolerance.
// reset_iters: a new solve (SC_ITERS = 0) rather than a restart of the same one.  After every
// launch_fcg_update the PB_RR region holds nblk partials of r.r followed by SC_ITERS.
void launch_fcg_init(Reducer& red, int bb_region, int rr_region, double rtol2, bool reset_iters, hipStream_t s);
// x += alpha p, r -= alpha q with alpha = gamma / p.q, unless x already met the tolerance (set at the
// (re)start, or by the previous update's r.r after its first iteration): then alpha = 0 from there on,
// so the result does not depend on when the host looks at the residual.  x = 0 if b = 0.
void launch_fcg_update(double* x, const double* p, double* r, double* rprev, const double* q, long n, Reducer& red,
                       bool first, hipStream_t s);
void launch_jacobi(const double* dinv, const double* r, double* z, long n, hipStream_t s);
// Flexible-CG direction for a matrix without a fused apply kernel: beta from the z.r / z.r_prev partials
// of launch_fcg_dots (0 on a (re)start), then p_new = z + beta p_old.
void launch_fcg_direction(const double* z, const double* p_old, double* p_new, long n, Reducer& red, bool first,
                          hipStream_t s);
// power iteration: x = dinv.Ax with partials of ||x||^2 in PB_MISC; then x /= ||x||
void launch_norm_power(const void* Ax, const void* dinv, void* x, long n, Reducer& red, bool fp64, hipStream_t s);
void launch_scale_by_norm(void* x, long n, Reducer& red, bool fp64, hipStream_t s);

// Newmark corrector: vn = a0 v + a1 a + a2 (un - u); an = -a3 v - a4 a + a5 (un - u).  With nh > 0 it
// also writes the nh vectors hout[0] = a, hout[j] = hin[j - 1] (each n long, stored n apart).
struct Extrap;
// With next (its coefficients; its vectors are ignored) and gout, the kernel also writes the next step's
// initial guess sum_k c_k w_k over w = (un, vn, an, a, hin[0], hin[1]) (the new state's u, v, a and history),
// in k_extrapolate's order of operations.
void launch_newmark_update(const double* u, const double* v, const double* a, const double* un, double* vn,
                           double* an, const double* c, long n, const double* hin, double* hout, int nh,
                           hipStream_t s, const Extrap* next = nullptr, double* gout = nullptr);
// adjoint combination (see transient.cpp)
void launch_adjoint_combine(const double* Kl, const double* Ml, const double* hdu, const double* hddu,
                            const double* hdddu, const double* mun, const double* nun, const uint8_t* mask,
                            const double* coef, bool terminal, double* mu, double* nu, double* rhs, long nn,
                            hipStream_t s);
void launch_adjoint0_rhs(const double* Kl, const double* Ml, const double* hdddu, const double* mun,
                         const double* nun, const uint8_t* mask, const double* coef, double* rhs, long nn,
                         hipStream_t s);
// objective gradient wrt state for objectives 0..2 (diagonal with region counts)
void launch_hdiag(const double* w, const double* cnt, double dt, double* h, long nn, hipStream_t s);
void launch_region_count(const Grid& g, const double* P, double* cnt, hipStream_t s);

// sensitivities and objective (per time step); accumulates into df and red.scal[slot] (objective)
struct SensArgs {
    Grid g;
    const double* KE;  // 576
    const double* M8;  // 64
    const double *kf, *mf, *dkf, *dmf, *P;
    const double *lam, *u, *v, *a, *ub, *vb, *ab;
    double wK[4];  // coefficients of (u, ub, vb, ab) in w_K
    double wM[4];  // coefficients of (u, ub, vb, ab) in w_M
    double dt, dalpha, dbeta;
    int objective;
    int mode;      // 0: general step (lambda), 1: t = 0 step (nu0 with dK u0 + dC v0 + dM a0)
    double* df;
};
void launch_sens(const SensArgs& a, Reducer& red, double* fx_out, hipStream_t s);

// PDE filter (scalar Helmholtz on nodes)
void launch_filter_apply(const Grid& g, const double* kf, const double* u, double* y, Reducer* red, int dot_slot,
                         hipStream_t s);
void launch_filter_T(const Grid& g, const double* x, double scale, double* rhs, hipStream_t s);
void launch_filter_Tt(const Grid& g, const double* U, double* out, hipStream_t s);
void launch_filter_diag(const Grid& g, const double* kf, double* dinv, hipStream_t s);
void launch_clamp_project(const double* xt_in, double* xt, double* xp, long n, int projection, double beta,
                          double eta, hipStream_t s);
void launch_project_chain(const double* xt, double* df, double* dg, long n, double beta, double eta,
                          hipStream_t s);
// reductions over element arrays
void launch_sum(const double* x, long n, Reducer& red, int slot, hipStream_t s);
// the same block partial sums of x written to `partials` (room for `capacity` values; at most 256 are
// needed, whatever n is); returns the number of blocks
int launch_sum_partials(const double* x, long n, double* partials, int capacity, hipStream_t s);
void launch_mnd(const double* x, long n, Reducer& red, int slot, hipStream_t s);

// MMA kernels (m = 1)
struct MMADev {
    long n;
    double *L, *U, *alpha, *beta, *p0, *q0, *pij, *qij, *xo1, *xo2;
};
void launch_mma_gensub(const MMADev& m, const double* x, const double* df, const double* dg, const double* xmin,
                       const double* xmax, int k, double asyminit, double asymdec, double asyminc, Reducer& red,
                       hipStream_t s);
void launch_mma_xyz_grad_hess(const MMADev& m, double lam, double* x, Reducer& red, hipStream_t s);
void launch_mma_xyz_res(const MMADev& m, double lam, double* x, Reducer& red, hipStream_t s);
void launch_mma_movelimit(const double* x, double* xmin, double* xmax, long n, double Xmin, double Xmax,
                          double movlim, hipStream_t s);
void launch_mma_change(const double* x, double* xold, long n, Reducer& red, int slot, hipStream_t s);
void launch_mma_kkt(const double* x, const double* df, const double* dg, const double* xmin, const double* xmax,
                    long n, double lam, Reducer& red, hipStream_t s);
void launch_fill_scal(void* x, long n, double v, hipStream_t s);
// x_i in [-1, 1) from a counter-based hash of (seed, i): deterministic, generated on the device
void launch_fill_hash(void* x, long n, unsigned long long seed, bool fp64, hipStream_t s);
// x = N.x (clamped dofs of the node mask set to 0)
void launch_mask_dofs(double* x, const uint8_t* mask, long nd, hipStream_t s);
void launch_scale(double* x, long n, double a, hipStream_t s);
// out = a*x + b*y (out may alias x or y)
void launch_lincomb(double a, const double* x, double b, const double* y, double* out, long n, hipStream_t s);
// out = a x + b y + c z
void launch_lincomb3(double a, const double* x, double b, const double* y, double c, const double* z, double* out,
                     long n, hipStream_t s);
// out = sum_k c[k] v[k] over the m <= 8 vectors of e (polynomial extrapolation of a sequence)
struct Extrap {
    const double* v[8];
    double c[8];
    int m;
};
void launch_extrapolate(const Extrap& e, double* out, long n, hipStream_t s);

}  // namespace tt
