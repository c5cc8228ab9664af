/*
 * FP64 CPU oracle for the transient-elasticity topology-optimisation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product library never links,
 * loads or calls it.
 *
 * It restates, in plain C on natural (single-rank PETSc) ordering, the
 * arithmetic of the reference files:
 *   LinearElasticity.cc  element matrices, BC, load, Newmark, adjoint,
 *                        objective, sensitivities
 *   PDEFilter.cc         Helmholtz density filter
 *   Filter.cc/.h         clamp, Heaviside projection and its chain rule
 *   MMA.cc               method of moving asymptotes (dual interior point)
 * Each function cites the reference lines it follows.  Linear systems that
 * the reference hands to PETSc KSP are solved here by Jacobi-preconditioned
 * CG on the same (SPD form of the) operator to a stated relative tolerance of
 * the unpreconditioned residual.
 */
#ifndef TOPOPT_ORACLE_H
#define TOPOPT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nelx, nely, nelz;  /* elements per direction */
    double xc[6];          /* domain box: xmin,xmax,ymin,ymax,zmin,zmax */
    double nu;             /* Poisson ratio */
    int sInt, mInt;        /* stiffness / mass interpolation: 0 lin, 1 SIMP, 2 RAMP */
    double Emin, Emax, Rmin, Rmax, penalK, penalM;
    double dalpha, dbeta;  /* Rayleigh damping C = dalpha*M + dbeta*K */
    int ti;                /* 0 backward Euler, 1 Newmark beta=1/4 gamma=1/2 */
    int nsteps;            /* Ts-1 time steps after t=0 */
    double t1;             /* final time; dt = t1/nsteps */
    int objective;         /* 0..6, LinearElasticity.cc:1416-1465 */
    int boundaries;        /* 0 MBB, 1 cantilever (x=xmin face clamped) */
    int loadloc;           /* 0,1 reference presets; 2 y-dof line at x=xmax,z=zmin */
    int loadsignal;        /* 0 reference Gaussian-cosine in time; 1 half-sine over pulse_steps */
    int pulse_steps;       /* half-sine length in steps (loadsignal 1) */
    int region;            /* optimizationRegion 0..3 */
    double rtol;           /* CG relative tolerance (unpreconditioned residual) */
    int maxit;             /* CG iteration cap */
} or_params;

/* geometry / indexing */
long or_nn(const or_params* p);   /* number of nodes */
long or_nel(const or_params* p);  /* number of elements */
void or_element_nodes(const or_params* p, long e, long nodes[8]);

/* element matrices (LinearElasticity.cc:3089-3346) */
void or_hex8_ke(double dx, double dy, double dz, double nu, double* ke);
void or_hex8_me(double dx, double dy, double dz, double* me);
void or_element_matrices(const or_params* p, double* ke, double* me);
/* PDE filter element matrix and T entries (PDEFilter.cc:485-588) */
void or_pdefilter_ke(double dx, double dy, double dz, double R, double* kf);

/* Newmark constants (LinearElasticity.cc:185-219) */
void or_newmark_consts(int ti, double dt, double a[6]);

/* interpolation (LinearElasticity.cc:1580-1623) */
void or_interp(int scheme, double vmin, double vmax, double penal, double x, double* f, double* df);

/* BC mask N (1 free, 0 clamped), load location vector and signal (LinearElasticity.cc:337-675) */
void or_bc_mask(const or_params* p, double* N);
void or_load_location(const or_params* p, double* F);
double or_load_intensity(const or_params* p, int step);
void or_region(const or_params* p, double* P);

/* y = sum_e (cK*kfac_e*KE + cM*mfac_e*ME) u_e over all elements, no BC */
void or_apply(const or_params* p, const double* xphys, double cK, double cM, const double* u, double* y);
/* BC form: y = N.(A(N.u)) + (1-N).u */
void or_apply_bc(const or_params* p, const double* xphys, double cK, double cM, const double* N,
                 const double* u, double* y);
/* Jacobi-PCG on the BC form; returns iterations (negative if not converged) */
int or_pcg(const or_params* p, const double* xphys, double cK, double cM, const double* N,
           const double* b, double* x, double rtol, int maxit, double* relres);

/* forward Newmark: states[(t*3 + {0,1,2})*nd + i] for t = 0..nsteps (u, v, a) */
int or_forward(const or_params* p, const double* xphys, double* states);
/* the same from a nonzero initial displacement N.u0 (-solutionAsU0, LinearElasticity.cc:523, 584, 2621-2676) */
int or_forward_u0(const or_params* p, const double* xphys, const double* u0, double* states);
/* objective and sensitivity (ComputeObjectiveSensitivities, LinearElasticity.cc:1625-1908) given states */
int or_objective_gradient(const or_params* p, const double* xphys, const double* states,
                          double* fx, double* dfdx);
/* objective only */
double or_objective(const or_params* p, const double* xphys, const double* states);

/* PDE filter (PDEFilter.cc:193-224): out = T' K_F^-1 (V T in) */
int or_pde_filter(const or_params* p, double rmin, const double* in, double* out, double rtol);
/* Heaviside projection and chain rule (Filter.h:80-88) */
double or_smooth_projection(double x, double beta, double eta);
double or_chainrule_projection(double x, double beta, double eta);
/* Filter::FilterProject with filter=2 (Filter.cc:60-119) */
int or_filter_project(const or_params* p, double rmin, const double* x, double* xtilde, double* xphys,
                      int projection, double beta, double eta, double rtol);
/* Filter::Gradients with filter=2 and m=1 (Filter.cc:122-206), in place */
int or_filter_gradients(const or_params* p, double rmin, const double* xtilde, double* dfdx, double* dgdx,
                        int projection, double beta, double eta, double rtol);

/* MMA (MMA.cc) */
typedef struct or_mma or_mma;
or_mma* or_mma_create(long n, int m, const double* x, const double* a, const double* c, const double* d);
void or_mma_destroy(or_mma* mma);
void or_mma_set_asymptotes(or_mma* mma, double init, double dec, double inc);
void or_mma_set_outer_movelimit(long n, double Xmin, double Xmax, double movlim, const double* x,
                                double* xmin, double* xmax);
void or_mma_update(or_mma* mma, double* x, const double* dfdx, const double* gx, const double* dgdx,
                   const double* xmin, const double* xmax);
double or_mma_design_change(long n, const double* x, double* xold);
void or_mma_kkt_residual(or_mma* mma, const double* x, const double* dfdx, const double* gx,
                         const double* dgdx, const double* xmin, const double* xmax, double* norm2,
                         double* normInf);
void or_mma_get_state(or_mma* mma, double* L, double* U, double* lam);

/* SOMMG reduced-order model (SOMMG.cc:40-265; LinearElasticity.cc reduced branches).
   Q: nb vectors of nd (row-major); Ut: nb x nb Gram-Schmidt factor; Kr, Mr, Cr: nb x nb; Pr: nb;
   rstates[(t*3 + {0,1,2})*nb + k] for t = 0..nsteps */
int or_sommg_basis(const or_params* p, const double* xphys, int nb, double om, double* Q, double* Ut);
void or_sommg_reduce(const or_params* p, const double* xphys, int nb, const double* Q, double* Kr, double* Mr,
                     double* Cr, double* Pr);
int or_reduced_forward(const or_params* p, int nb, const double* Kr, const double* Mr, const double* Cr,
                       const double* Pr, double* rstates);
void or_reduced_expand(long nd, int nb, int nsteps, const double* Q, const double* rstates, double* full);
int or_reduced_objective_gradient(const or_params* p, const double* xphys, int nb, const double* Q,
                                  const double* Kr, const double* Mr, const double* Cr, const double* rstates,
                                  double* fx, double* dfdx);

/* main.cc optimisation loop: runs maxit design iterations from x0 (V0).
   hist (maxit*5): fx_true, fx_scaled, gx, ch, mnd per iteration.  x/xphys out. */
int or_design_loop(const or_params* p, double rmin, double volfrac, double V0, int maxit, int projection,
                   double beta, double eta, double movlim, double* x, double* xphys, double* hist);

#ifdef __cplusplus
}
#endif
#endif
