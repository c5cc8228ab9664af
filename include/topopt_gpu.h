/*
 * C-ABI of the MI355X (gfx950) transient-elasticity topology-optimisation path.
 *
 * Plain C types only: pointers, sizes, doubles.  Every function returns 0 on
 * success and a negative code on failure; tt_last_error() gives the message.
 * Krylov non-convergence and non-finite values are reported as errors
 * (TT_ERR_NOT_CONVERGED / TT_ERR_NONFINITE), not silently accepted.
 *
 * Pointer arguments marked "array" are host pointers when on_device == 0
 * (copied in/out synchronously) and device pointers on the context's device
 * when on_device == 1 (used in place; the caller keeps ownership).
 *
 * Each entry point names the reference interface it replaces
 * (file:line in topopt/TopOpt_in_PETSc_Transient).
 */
#ifndef TOPOPT_GPU_H
#define TOPOPT_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define TT_OK 0
#define TT_ERR_ARG (-1)
#define TT_ERR_HIP (-2)
#define TT_ERR_INTERNAL (-3)
#define TT_ERR_NOT_CONVERGED (-4)
#define TT_ERR_NONFINITE (-5)
#define TT_ERR_STATE (-6)

#define TT_PREC_FP64 0  /* FP64 flexible CG, FP64 multigrid */
#define TT_PREC_MIXED 1 /* FP64 flexible CG, FP32 multigrid, BF16 coarse coefficients, FP64 coarsest solve */

#define TT_OP_K 0 /* stiffness K(rho) */
#define TT_OP_M 1 /* mass M(rho) */
#define TT_OP_A 2 /* Newmark system K + a2 C + a5 M = (1 + a2 dbeta) K + (a5 + a2 dalpha) M */

typedef struct tt_ctx tt_ctx;
typedef struct tt_mma tt_mma;

/* Problem, material, time-integration and solver parameters.  Names follow
   the reference options (TopOpt.cc:71-165, 329-518, LinearElasticity.h:29). */
typedef struct tt_params {
    int nelx, nely, nelz;          /* elements; nodes are nel+1 (TopOpt.cc:179-181 -nx/-ny/-nz minus one) */
    double xc[6];                  /* domain box xmin,xmax,ymin,ymax,zmin,zmax (TopOpt.cc:185-190) */
    double nu;                     /* Poisson ratio (-nu) */
    int sInt, mInt;                /* stiffness / mass interpolation 0 linear, 1 SIMP, 2 RAMP */
    double Emin, Emax, Rmin, Rmax; /* Emin = Ef*Emax, Rmin = Rf*Rmax (TopOpt.cc:410-413) */
    double penalK, penalM;
    double dalpha, dbeta;          /* Rayleigh damping C = dalpha M + dbeta K (-dalph, -dbeta) */
    int ti;                        /* 0 backward Euler, 1 Newmark beta=1/4 gamma=1/2 (LinearElasticity.cc:185-207) */
    int nsteps;                    /* time steps after t = 0: Ts - 1 = ncp*tsc */
    double t1;                     /* final time; dt = t1/nsteps */
    int objective;                 /* 0..6 (LinearElasticity.cc:1416-1465) */
    int boundaries;                /* 0 MBB, 1 cantilever (LinearElasticity.cc:550-575) */
    int loadloc;                   /* 0,1 reference presets; 2 y-line load at x = xmax, z = zmin */
    int loadsignal;                /* 0 reference Gaussian-cosine pulse; 1 half-sine over pulse_steps */
    int pulse_steps;
    int region;                    /* optimizationRegion 0..3 (LinearElasticity.cc:222-335) */
    int nlvls;                     /* multigrid levels (-nlvls) */
    int precision;                 /* TT_PREC_FP64 or TT_PREC_MIXED */
    double rtol;                   /* relative tolerance of ||b - A x|| / ||b|| */
    int maxit;                     /* flexible-CG iteration cap per solve */
    int cheb_degree;               /* Chebyshev smoother degree on D^-1 A */
    int checkpoint_slots;          /* binomial checkpoint slots; >= nsteps stores every state */
    double rmin;                   /* PDE filter radius (-rmin); R = rmin/(2 sqrt 3) */
    double filter_rtol;            /* PDE filter solve tolerance */
    int verbose;
    int graphs;                    /* 1: each multigrid-preconditioned Krylov iteration is replayed as a
                                      captured HIP graph (bit-identical results; default 0: measured
                                      slower than eager launches on ROCm 7 for this ~30-kernel
                                      iteration) */
    int reduction;                 /* 1: SOMMG reduced-order model (-reduction): the Newmark and adjoint
                                      sweeps run on the projection onto `basis` moment-matching vectors
                                      (SOMMG.cc:40-265; LinearElasticity.cc:404-461, 994-1003) */
    int basis;                     /* basis vectors (-basis, 1..64) */
    double omega0;                 /* expansion frequency of the basis solves with K - omega0^2 M (-omega0,
                                      passed unconverted as in LinearElasticity.cc:996) */
    int coarse_direct_max;         /* largest coarsest multigrid level (dofs) solved through the dense FP64
                                      inverse; a larger one is smoothed by Chebyshev(12) on D^-1 A instead.
                                      0 = 6,476 dofs (a 320 MB dense matrix); 1 forces the Chebyshev path */
    int mg_coarse_pre;             /* smoothing steps before the coarse-grid correction on the levels between
                                      the finest and the coarsest (0 = 1: one damped-Jacobi step, which needs
                                      no operator apply); the finest level smooths cheb_degree steps on
                                      each side */
    int mg_coarse_post;            /* smoothing steps after the coarse-grid correction on those levels
                                      (default 0).  Both apply to the Newmark, adjoint and mass solves;
                                      the reduced-order basis solves (flexible GMRES) smooth cheb_degree
                                      steps on each side of every level */
} tt_params;

typedef struct tt_stats {
    long solves;          /* linear solves since the last reset */
    long krylov_iters;    /* flexible-CG iterations since the last reset */
    int last_iters;
    double last_relres;   /* relative residual of the last solve (recursive; the recomputed b - A x when rtol < 1e-9) */
    double max_relres;
    long forward_steps;   /* Newmark steps integrated (incl. recomputation) */
    long adjoint_steps;
    double lambda_max[8]; /* Chebyshev upper eigenvalue estimate per level */
    int nlvls;
} tt_stats;

/* Defaults of the cantilever pulse problem (3-level MG, FP64, rtol 1e-8). */
int tt_default_params(tt_params* prm);

/* Replaces the construction of LinearElasticity (LinearElasticity.cc:14-120, called at
   main.cc:35): meshes, BC mask (SetUpBC :337-591), load location (:616-675), region
   (:222-335), Newmark constants (:185-219), element matrices (:516-519). */
int tt_create(const tt_params* prm, tt_ctx** ctx);
void tt_destroy(tt_ctx* ctx);
const char* tt_last_error(const tt_ctx* ctx); /* ctx may be NULL: last creation error */

/* xPhys input of CompleteInitialState / ComputeObjectiveSensitivities (main.cc:82,90);
   replaces AssembleStiff/Mass/DampingMatrix (LinearElasticity.cc:2502-2619): element
   interpolation only, nothing is assembled.  nel doubles. */
int tt_set_density(tt_ctx* ctx, const double* xphys, int on_device);

/* Material and damping arguments of CompleteInitialState / ComputeObjectiveSensitivities
   (LinearElasticity.h: Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta; used at
   LinearElasticity.cc:1580-1623 and :185-219).  Replaces the values given to tt_create; the
   density must be set again afterwards. */
int tt_set_material(tt_ctx* ctx, double Emin, double Emax, double Rmin, double Rmax, double penalK, double penalM,
                    double dalpha, double dbeta);

/* PDEFilt construction (PDEFilter.cc:28-191) for filter radius rmin (R = rmin / (2 sqrt 3))
   on the context's node grid; replaces tt_params.rmin. */
int tt_set_filter_radius(tt_ctx* ctx, double rmin);

/* One matrix-free operator application y = N.(X (N.u)) + (1-N).u for X in {K, M, A}
   (the MatMult of LinearElasticity.cc:1342-1345); prec 0 = FP64 arrays, 1 = FP32 arrays.
   3*nn entries. */
int tt_op_apply(tt_ctx* ctx, int which, int prec, const void* u, void* y, int on_device);

/* CompleteInitialState (LinearElasticity.cc:956-1206): multigrid set-up for the Newmark
   operator and the initial-acceleration solve (N M N + I - N) a0 = N.(f(0) - C v0 - K u0). */
int tt_complete_initial_state(tt_ctx* ctx);

/* GenerateCheckPoints + SolveStateBlock (LinearElasticity.cc:1208-1280, 1298-1397): the
   whole Newmark forward sweep, keeping binomial checkpoints (or every state). */
int tt_forward(tt_ctx* ctx);

/* Initial displacement u0 (nd values; NULL: back to u0 = 0) of the next tt_complete_initial_state:
   -solutionAsU0 with -onlyLoadDesign, where LoadSolutionAsU0 reads U0 from -restartFileVecSol and the
   constructor masks it to N.U0 (LinearElasticity.cc:522-523, 584, 2621-2676); v0 stays 0.  The stored
   value is N.u0.  TT_ERR_STATE with the reduced-order model. */
int tt_set_initial_displacement(tt_ctx* ctx, const double* u0, int on_device);

/* State at step t (0..nsteps) of the last tt_forward (GetStateFieldU/DU/DDU, LinearElasticity.h:68-78);
   3*nn doubles each (may be NULL).  A state that the binomial checkpoints did not keep is recomputed
   from the nearest stored one; asking for t = 1, 2, ... in turn costs one Newmark step each.  With the
   reduced-order model the state is the reduced one expanded on the basis.  State 0 is available as soon
   as tt_complete_initial_state has run. */
int tt_get_state(tt_ctx* ctx, int t, double* u, double* v, double* a, int on_device);

/* A state of the forward sweep given by the caller, as ComputeObjectiveSensitivities reads the checkpoint
   files cp%04d.dat back (LinearElasticity.cc:1651-1672): u, v, a at step t (3*nn values each).  After
   tt_complete_initial_state, the given states (state 0 is the initial state unless given too) replace the
   forward sweep: tt_objective_gradient recomputes the states between consecutive given steps, and after
   the last one up to nsteps, from them (SolveStateBlock, :1298-1397) and runs the adjoint over each such
   block, from the last block back.  Without binomial checkpointing (checkpoint_slots >= nsteps) every
   state is kept; otherwise each given state takes a checkpoint slot and the rest recompute within a
   block.  A state from a file carries no acceleration history, so the first solve after it starts from a
   constant-acceleration guess.  tt_forward or tt_complete_initial_state drops the given states.
   TT_ERR_STATE with the reduced-order model. */
int tt_set_checkpoint(tt_ctx* ctx, int t, const double* u, const double* v, const double* a, int on_device);

/* ComputeObjectiveSensitivities (LinearElasticity.cc:1625-1908): objective fx and dfdx
   (nel) by the backward Newmark adjoint (SolveLagrangen :808-954, SolveLagrange0
   :678-804) with recomputation between checkpoints. */
int tt_objective_gradient(tt_ctx* ctx, double* fx, double* dfdx, int on_device);

/* VolumeConstraint (LinearElasticity.cc:1283-1296): gx = sum(xPhys)/nel - volfrac, dgdx = 1/nel. */
int tt_volume_constraint(tt_ctx* ctx, const double* xphys, double volfrac, double* gx, double* dgdx,
                         int on_device);

/* PDEFilt::FilterProject / Gradients (PDEFilter.cc:193-224): out = T' K_F^-1 (V T in). */
int tt_pde_filter(tt_ctx* ctx, const double* in, double* out, int on_device);

/* Filter::FilterProject with filter = 2 (Filter.cc:60-119): PDE filter, clamp to [0,1],
   optional Heaviside projection (Filter.h:80-83). */
int tt_filter_project(tt_ctx* ctx, const double* x, double* xtilde, double* xphys, int projection, double beta,
                      double eta, int on_device);

/* Filter::Gradients with filter = 2, m = 1 (Filter.cc:122-206), in place on dfdx and dgdx. */
int tt_filter_gradients(tt_ctx* ctx, const double* xtilde, double* dfdx, double* dgdx, int projection,
                        double beta, double eta, int on_device);

/* Filter::GetMND (Filter.cc:208-227). */
int tt_mnd(tt_ctx* ctx, const double* x, double* mnd, int on_device);

/* MMA with m = 1 constraint (MMA.h:12-53, MMA.cc).  a, c, d as in TopOpt.cc:527-531. */
int tt_mma_create(tt_ctx* ctx, double a, double c, double d, tt_mma** mma);
void tt_mma_destroy(tt_mma* mma);
int tt_mma_set_asymptotes(tt_mma* mma, double init, double decrease, double increase); /* MMA.cc:352-360 */
int tt_mma_set_outer_movelimit(tt_mma* mma, double Xmin, double Xmax, double movlim, const double* x,
                               double* xmin, double* xmax, int on_device);             /* MMA.cc:374-393 */
int tt_mma_update(tt_mma* mma, double* x, const double* dfdx, const double* gx, const double* dgdx,
                  const double* xmin, const double* xmax, int on_device);               /* MMA.cc:487-510 */
int tt_mma_design_change(tt_mma* mma, const double* x, double* xold, double* ch, int on_device); /* :395-414 */
int tt_mma_kkt_residual(tt_mma* mma, const double* x, const double* dfdx, const double* gx, const double* dgdx,
                        const double* xmin, const double* xmax, double* norm2, double* normInf,
                        int on_device);                                                  /* MMA.cc:416-484 */
int tt_mma_get_asymptotes(tt_mma* mma, double* L, double* U, double* lam); /* host arrays (nel, nel, 1) */
/* MMA restart state: iteration counter k, the two previous iterates and the asymptotes.  get replaces
   MMA::Restart (MMA.cc:307-350), set the restart constructor MMA(n, m, k, xo1, xo2, U, L, a, c, d)
   (MMA.cc:6-90).  Host arrays of nel values. */
int tt_mma_get_state(tt_mma* mma, int* k, double* xo1, double* xo2, double* U, double* L);
int tt_mma_set_state(tt_mma* mma, int k, const double* xo1, const double* xo2, const double* U, const double* L);

/* Restart vectors of the physics (LinearElasticity::WriteRestartFiles, LinearElasticity.cc:2457-2496,
   and their reading in SetUpSolver1, :2711-2723): the initial acceleration a0 (DDU0), the terminal
   multiplier lambda_T (LT) and the t = 0 adjoint multiplier nu0 (DDL0), each nd values in host
   memory.  get returns the latest ones (zeros before the first solve); set makes them the starting
   guesses of the next tt_complete_initial_state (a0) and tt_objective_gradient (lambda_T, nu0), as the
   reference's persistent vectors are.  A null pointer leaves that vector as it is.
   With the reduced-order model (reduction = 1) the reduced systems are solved directly, so there are
   no starting guesses: both return TT_ERR_STATE. */
int tt_get_restart_vectors(tt_ctx* ctx, double* a0, double* lamT, double* nu0);
int tt_set_restart_vectors(tt_ctx* ctx, const double* a0, const double* lamT, const double* nu0);

/* The SOMMG reduced model (reduction = 1) of the last tt_complete_initial_state: the members QQ, Kr, Mr,
   Cr, Pr of the reference's SOMMG object (SOMMG.h:16-22) and the Gram-Schmidt factor Utemp of
   ConstructBasis (SOMMG.cc:48-218), plus the reduced states U, DU, DDU of the last tt_forward.
   Q: basis x nd (on_device selects host or device memory); Ut, Kr, Mr, Cr: basis x basis row-major;
   Pr: basis; rstates: (nsteps + 1) x 3 x basis (u, v, a per step), host memory.  Any pointer may be
   NULL.  TT_ERR_STATE without the reduced model, or for rstates before tt_forward. */
int tt_get_reduced_model(tt_ctx* ctx, double* Q, int on_device, double* Ut, double* Kr, double* Mr, double* Cr,
                         double* Pr, double* rstates);

/* PETSc binary vector files, as PetscViewerBinary VecView / VecLoad write and read them for the
   reference's checkpoint and restart files (LinearElasticity.cc:1128-1137, 1261-1276, 2474-2486;
   TopOpt.cc:614-735) with 32-bit PetscInt: per vector the big-endian int32 class id 1211214, the
   big-endian int32 length, then the values as big-endian IEEE doubles; vectors follow each other.
   Natural ordering, which is this library's ordering.  Host memory only; no GPU work.  write
   creates or truncates the file; read reads the first nvec vectors and fails (TT_ERR_ARG, message
   from tt_last_error(NULL)) on a missing file, a wrong class id, a length other than lens[k] or a
   short file. */
int tt_petsc_write_vecs(const char* path, int nvec, const double* const* vecs, const long* lens);
int tt_petsc_read_vecs(const char* path, int nvec, double* const* vecs, const long* lens);

/* Measurement helper with no counterpart in the reference: nrep back-to-back launches of the
   tt_op_apply kernel on the context's stream, timed with HIP events on that stream; returns the
   average milliseconds per launch.  u and y are device arrays (FP64 or FP32 per prec). */
int tt_time_op_apply(tt_ctx* ctx, int which, int prec, const void* u, void* y, int nrep, double* ms_per_apply);
/* Measurement helper with no counterpart in the reference: nrep back-to-back launches of the FP64
   flexible-CG apply of a Krylov iteration (p = z + beta p_old formed while staging, q = A p with the
   Newmark operator, p.q partials) on the context's own Krylov vectors, timed with HIP events on the
   context's stream; average milliseconds per launch. */
int tt_time_fcg_apply(tt_ctx* ctx, int nrep, double* ms_per_apply);

/* Measurement, no counterpart in the reference: nrep multigrid preconditioner applications z = B r of
   the Krylov iteration (the V-cycle of the last operator set up, tt_complete_initial_state or a solve
   before), back to back on the context's stream.  ms_gpu: average milliseconds per application between
   HIP events; ms_host: average host milliseconds to enqueue one (when it is close to ms_gpu, the
   V-cycle is bound by kernel launches, not by the GPU). */
int tt_time_vcycle(tt_ctx* ctx, int nrep, double* ms_gpu, double* ms_host);

/* No counterpart in the reference: the allocation-free form of tt_volume_constraint
   (LinearElasticity.cc:1283-1296), for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops).  No context: xphys (nel doubles), gx (one double) and dgdx (nel doubles,
   may be NULL) are device arrays on the current device; the work is queued on `stream` (a hipStream_t;
   NULL = the null stream) and the call returns without synchronising.  The block partial sums live in
   a library-owned device array of 256 doubles (2 KB), which each GPU the library runs on holds until the
   process ends; calls that overlap on one GPU race on it.  No hipMalloc, hipFree or allocating library
   call.  1 <= nel <= 2^31-1, else TT_ERR_ARG; tt_last_error(NULL) gives the message.  gx is bit-identical
   to tt_volume_constraint's. */
int tt_volume_constraint_noalloc(const double* xphys, long nel, double volfrac, double* gx, double* dgdx,
                                 void* stream);

int tt_get_stats(tt_ctx* ctx, tt_stats* st);
int tt_reset_stats(tt_ctx* ctx);
int tt_synchronize(tt_ctx* ctx);
/* Sizes: nodes, elements. */
long tt_num_nodes(const tt_ctx* ctx);
long tt_num_elements(const tt_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
