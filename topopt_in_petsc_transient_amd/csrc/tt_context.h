// Library context: device data of one transient-elasticity problem on one GPU.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/topopt_gpu.h"
#include "tt_internal.h"

namespace tt {

struct Context;

// One multigrid level of the elasticity operator.
struct MGLevel {
    Grid g;
    bool fp64 = true;                  // precision of this level's vectors and element matrices
    DBuf<uint8_t> mask, tbc4, tbc2;    // node mask; per-tile flags of clamped nodes (8x8x4, 8x8x2 tiles)
    BatchTabStore btab;                // element batch tables of the tiles (full and edge tiles)
    DBuf<double> kf64, mf64;           // FP64 coefficients (coarse levels: averages of children)
    DBuf<float> kf32, mf32;            // mixed mode, finest level
    DBuf<bf16_t> kfbf, mfbf;           // mixed mode, coarse levels
    DBuf<double> KE64, M864, AF64;
    DBuf<float> KE32, M832, AF32;
    OpLevel op{};
    DBuf<char> b, x, x2, r, d, d2, dinv;  // level vectors (element size = 8 or 4 bytes)
    double lmax = 1.0;
    size_t esz() const { return fp64 ? 8 : 4; }
};

// Geometric multigrid preconditioner (V-cycle, Chebyshev smoothing on D^-1 A,
// rediscretised coarse operators, dense FP64 coarsest solve through an explicit inverse).
struct Multigrid {
    Context* ctx = nullptr;
    std::vector<std::unique_ptr<MGLevel>> L;
    bool mixed = false;
    int cheb_deg = 2;                       // smoothing steps on each side of the finest level
    int coarse_pre = 1, coarse_post = 0;    // before / after the correction on the intermediate levels
    double cK = 1.0, cM = 0.0;
    double cheb_lo = 0.1, cheb_hi = 1.1;
    // coarsest level
    int nc = 0;
    bool dense = true;
    DBuf<double> Ainv, AinvP, Winv, sprow, spcol, cb, cx, cdinv, cx2, cd, cd2;  // FP64 coarsest level
    DBuf<int> dinfo;
    void* blas = nullptr;  // rocblas_handle of the coarsest operator's inversion (GEMMs)
    // the inversion's own stream, its events (assembly done, inverse done) and rocSOLVER's status words
    hipStream_t s2 = nullptr;
    hipEvent_t ev_asm = nullptr, ev_inv = nullptr;
    int* info_host = nullptr;
    bool inv_pending = false;
    // waits (device and host) for the inverse queued by the last set-up and checks its status
    void finish_inverse();
    void spd_inverse(double* A, int n, int lda, double* W);
    int coarse_deg = 12;
    double clmin = 0.0, clmax = 1.0;

    Multigrid() = default;
    Multigrid(const Multigrid&) = delete;
    Multigrid& operator=(const Multigrid&) = delete;
    ~Multigrid();

    void build(Context* c, int nlvls, bool mixed_prec, int deg);
    // the level work buffers a V-cycle rotates (x, x2, d, d2 of every smoothing level)
    std::vector<char**> rotating();
    void set_coefficients(const double* kf_fine, const double* mf_fine);
    void setup_operator(double cK_, double cM_);
    // z = B r (FP64 in and out)
    void apply(const double* r, double* z);
    // z = B r as the flexible-CG preconditioner, also leaving the partials of z.r and z.r_prev
    // (r_prev ignored when first) in the reducer's PB_DOTS region
    void apply_fcg(const double* r, const double* rp, double* z, bool first);

   private:
    void vcycle(int l, const double* r64 = nullptr, double* z64 = nullptr, const double* rp64 = nullptr,
                bool first = false);
    void coarse_solve(bool direct64);
    double power_lmax(int l, int iters);
    void assemble_coarse_inverse();
    OpLevel coarse_op64() const;
};

struct SolveResult {
    int iters = 0;
    double relres = 0.0;
};

struct Context {
    tt_params prm{};
    Grid g{};
    hipStream_t stream = nullptr;
    int device = 0;
    Reducer red;
    std::string last_error;

    double KE[576], M8[64];
    double a[6], dt = 0.0;
    double cK_A = 1.0, cM_A = 1.0;  // Newmark operator scalars

    // element data
    DBuf<double> xphys, kf, mf, dkf, dmf, P, Pkf, Pmf;
    DBuf<double> rcnt;  // per node: sum of region indicators of the adjacent elements
    DBuf<double> F;     // load location vector (3 nn)
    DBuf<uint8_t> mask, tbc4, tbc2;
    BatchTabStore btab;
    std::vector<double> Fhost;
    OpLevel op_fine64{};  // FP64 fine operator view (for FCG residuals)
    DBuf<double> KE64d, M864d, AF64d;
    DBuf<float> KE32d, M832d, AF32d;
    float afh_scale[2];
    DBuf<float> kf32, mf32;  // FP32 fine coefficients for the FP32 microbenchmark apply
    bool density_set = false;
    bool initial_done = false;
    bool a0_guess = false;  // slot 0 holds a starting guess for a0 (tt_set_restart_vectors), used once
    DBuf<double> u0in;      // initial displacement N.u0 (tt_set_initial_displacement), used when u0_set
    bool u0_set = false;
    int scratch_step = -1;  // st_cur holds this state, recomputed by tt_get_state (-1: nothing)
    const double* state_at(int t);  // state t: stored, or recomputed from the nearest stored state
    bool forward_done = false;

    Multigrid mg;

    // Krylov work vectors (FP64, nd); kx holds the solution of a graph-replayed solve
    DBuf<double> kr, krp, kz, kp0, kp1, kq, kq2, dinvM, kx;

    // One multigrid-preconditioned flexible-CG iteration captured as a HIP graph, keyed by everything
    // its launches bake in (pointers of the rotating buffers, first-iteration flag, operator scalars),
    // with the rotation state and reducer bookkeeping the iteration leaves behind.
    struct IterGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<char*> after;
        int nblk[PB_COUNT], nv[PB_COUNT];
    };
    std::map<std::vector<uintptr_t>, IterGraph> iter_graphs;
    void clear_graphs();  // (the multigrid set-up changes baked-in Chebyshev scalars)

    // time states: slot storage, kStateVecs x nd contiguous each: u, v, a and the kFwdHist - 1 previous
    // accelerations (the Newmark solve's initial guess extrapolates the acceleration from them, so a
    // state restored from a checkpoint continues exactly as in the first sweep)
    static constexpr int kFwdHist = 4;
    static constexpr int kStateVecs = 2 + kFwdHist;
    std::vector<DBuf<double>> slots;
    std::vector<int> slot_step;  // time step held by each slot (-1 = free)
    bool store_all = false;
    DBuf<double> st_cur, st_next;
    // checkpointed runs: st_last = state T as the forward sweep left it; st_hi = the later of the two
    // states an adjoint step reads, kept from the previous (later) adjoint step
    DBuf<double> st_last, st_hi;
    // steps whose states the caller gave (tt_set_checkpoint) in place of a forward sweep
    std::set<int> given;
    void set_checkpoint(int t, const double* u, const double* v, const double* a);

    // adjoint work
    DBuf<double> lam, lamT,
        lamh[8],  // multipliers of the later steps t+1 .. (the adjoint solves' initial guess; 7 are used)
        mu, nu, mun, nun, hdu, hddu, hdddu, Kl, Ml, rhs, df, fxs, nu0;

    // PDE filter
    double KF[64];
    DBuf<double> KFd, fdinv, fU, frhs, fr, frp, fz, fp, fp2, fq, ftmp_e, ftmp_e2;
    int filter_iters = 0;  // iterations of the last filter solve (when the host first reads the residual)
    Grid fg{};

    // SOMMG reduced-order model (tt_params.reduction): basis Q (nbasis vectors of nd), its
    // Gram-Schmidt factor, the reduced stiffness, mass, damping and load, and the reduced states
    // rst[(t*3 + {0,1,2})*nbasis + k] (u, v, a) of the last forward sweep
    bool reduced = false;
    int nbasis = 0;
    DBuf<double> Qb;
    DBuf<double> projpart;          // projection partials: per basis vector 1024 blocks x 3
    double* projhost = nullptr;     // pinned host copy of projpart
    std::vector<double> Ut, Kr, Mr, Cr, Pr, rst;
    void reduced_initial_state();             // basis, reduced matrices, a0 (CompleteInitialState)
    void reduced_forward();                   // reduced Newmark sweep (SolveStateBlock)
    void expand(const double* coef, double* out);  // out = sum_k coef[k] Q_k
    void expand_state(int t, double* out);    // u, v, a of reduced state t on the full mesh
    void project3(const double* h0, const double* h1, const double* h2, double* out);  // out[3k+i] = Q_k.h_i
    DBuf<double> gmV, gmZ;  // flexible-GMRES Krylov and preconditioned vectors of the basis solves
    SolveResult fgmres(double cK, double cM, const double* b, double* x);
    void reduced_adjoint(const std::function<void(const double*)>& grad_state,
                         const std::function<void(int, const double*, const double*, const double*, int)>& sens);

    // stats
    tt_stats stats{};

    explicit Context(const tt_params& p);
    ~Context();

    // helpers
    OpLevel fine_op(bool fp64) const;
    // Solves of one kind (forward Newmark, adjoint, mass) have similar iteration counts; the last few
    // counts of each kind decide when the host starts reading the residual back.
    enum SolveKind { kSolveForward = 0, kSolveAdjoint = 1, kSolveMass = 2, kSolveKinds = 3 };
    static constexpr int kHintLen = 4;
    SolveResult solve_A(const double* b, double* x, SolveKind kind);  // Newmark operator, MG-preconditioned FCG
    SolveResult solve_M(const double* b, double* x);                  // mass operator N M N + I - N, Jacobi FCG
    SolveResult fcg(double cK, double cM, bool use_mg, SolveKind kind, const double* b, double* x);
    int fcg_hist[kSolveKinds][kHintLen] = {};  // iteration counts of the last solves of each kind (0: none)
    void sync();
    // waits for the stream by polling an event: the Krylov loop's residual looks, where the GPU idles
    // until the host has seen the result and queued the next launches
    hipEvent_t look_ev = nullptr;
    void wait_look();
    double read_scal(int slot);
    double reduce1(int region);  // host sum of a partial region (synchronises)

    // transient
    void complete_initial_state();
    void forward();
    void objective_gradient(double* fx_host, double* dfdx_dev);
    // one Newmark step from state s_in to s_out; with next_out, the corrector also writes the next step's
    // initial guess there (the next step's s_out), and that step skips its own extrapolation
    void newmark_step(const double* s_in, double* s_out, int t, double* next_out = nullptr);
    Extrap forward_guess(int t, const double* s_in) const;
    int guess_t = -1;               // the step whose initial guess a corrector has written ahead (-1: none)
    double* guess_out = nullptr;    // and where
    double load_intensity(int step) const;

    // filter
    void set_filter_radius(double rmin);
    void pde_filter(const double* in_dev, double* out_dev);
};

}  // namespace tt
