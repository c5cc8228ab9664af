/* Drop-in C++ classes of the transient topology-optimisation path over the HIP C-ABI
 * (topopt_gpu.h).
 *
 * Same class names, constructor and method signatures, argument meaning and return
 * convention (PetscErrorCode 0 = ok) as the reference's LinearElasticity.h:29-85,
 * Filter.h:38-55, PDEFilter.h:32-36, MMA.h:12-53 and TopOpt.h, so that a main program
 * written against those headers (main.cc) drives the GPU path unchanged.  PETSc is not
 * linked: the few PETSc names these signatures use are supplied here with minimal
 * meaning -- Vec is a host array of doubles, DM a structured-grid descriptor that also
 * carries the GPU context of its mesh, and the options database holds the command line.
 * MPIIO writes the reference's single-rank field files.
 *
 * Differences a caller can observe:
 *   - Filter supports the PDE filter only (filterT = 2, Filter.cc:60-119 / 122-206);
 *   - MMA supports one constraint (m = 1);
 *   - restart files (Restart0x.dat, Restart0x_itr_f0.dat, RestartSol0x.dat) and the checkpoint files
 *     cp%04d.dat are written and read in PETSc's binary Vec format with the reference's options and
 *     names; with ncp = 1 the gradient uses the binomial checkpoints in device memory, with ncp > 1 it
 *     recomputes each block from the state read back from its file, as the reference does;
 *     -onlyLoadDesign -solutionAsU0 reads the initial displacement from -restartFileVecSol as the
 *     reference does;
 *   - the reduced-order path (reduction = TRUE, SOMMG) runs with the reduced systems solved directly,
 *     as in the reference; it reads and writes no solution restart file (RestartSol0x.dat) and no
 *     checkpoint file;
 *   - numerical non-convergence and non-finite values are reported as errors (< 0);
 *   - option names match without regard to case, as in PETSc's options database.
 */
#ifndef TOPOPT_DROPIN_H
#define TOPOPT_DROPIN_H

#include <memory>
#include <string>
#include <vector>

#include "topopt_gpu.h"

/* ---- the PETSc names used by the interfaces */
#define PETSC_MAX_PATH_LEN 4096
typedef double PetscScalar;
typedef double PetscReal;
typedef int PetscInt;
typedef int PetscErrorCode;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;

struct _p_Vec {
    std::vector<PetscScalar> a;
};
typedef _p_Vec* Vec;

struct TTMesh;  // shared GPU context of one node grid
struct _p_DM {
    PetscInt nxyz[3];   // nodes per direction (elements for an element DM)
    PetscScalar xc[6];  // domain box
    bool elemental;
    std::shared_ptr<TTMesh> mesh;
};
typedef _p_DM* DM;

/* MPIIO.h / MPIIO.cc: the single-rank form of the reference's field output (bin2vtu.py reads it).
   One domain of hex8 cells; the file is the info line, the header (unsigned long: domains, points,
   cells, point fields, cell fields, nodes per element), the field names, float32 point coordinates,
   the connectivity, offsets and VTK types (unsigned long), then per WriteVTK call the step number
   (unsigned long) and the float32 point and cell fields, one field after the other. */
class MPIIO {
   public:
    MPIIO(DM da_nodes, int nPfields, std::string pnames, int nCfields, std::string cnames, std::string filename);
    ~MPIIO();
    PetscErrorCode WriteVTK(DM da_nodes, Vec x, Vec xTilde, Vec xPhys, PetscInt itr);
    PetscErrorCode WriteVTK(DM da_nodes, Vec U, Vec x, Vec xTilde, Vec xPhys, PetscInt itr);
    PetscErrorCode WriteVTK(DM da_nodes, Vec U, Vec xPhys, PetscInt itr);
    PetscErrorCode WriteVTK(DM da_nodes, Vec U, Vec DU, Vec DDU, Vec x, Vec xtilde, Vec xPhys, PetscInt itr);

   private:
    std::string filename;
    unsigned long nPoints, nCells;
    int nPF, nCF;
    std::vector<float> workPoint, workCell;  // nPF x nPoints and nCF x nCells, one field after the other
    PetscErrorCode Append(PetscInt itr);
    void PutPoint(int field0, Vec v, int ncomp);
    void PutCell(int field, Vec v);
};

PetscErrorCode PetscInitialize(int* argc, char*** argv, const char* file, const char* help);
PetscErrorCode PetscFinalize();
PetscErrorCode PetscPrintf(int comm, const char* fmt, ...);
PetscErrorCode PetscOptionsGetInt(void*, const char* pre, const char* name, PetscInt* v, PetscBool* set);
PetscErrorCode PetscOptionsGetReal(void*, const char* pre, const char* name, PetscReal* v, PetscBool* set);
PetscErrorCode PetscOptionsGetBool(void*, const char* pre, const char* name, PetscBool* v, PetscBool* set);
PetscErrorCode PetscOptionsGetString(void*, const char* pre, const char* name, char* v, size_t len, PetscBool* set);
#define PETSC_COMM_WORLD 0
#define PETSC_NULL nullptr
double MPI_Wtime();

PetscErrorCode VecCreateSeq(int comm, PetscInt n, Vec* v);
PetscErrorCode VecDuplicate(Vec x, Vec* y);
PetscErrorCode VecDestroy(Vec* v);
PetscErrorCode VecSet(Vec v, PetscScalar a);
PetscErrorCode VecScale(Vec v, PetscScalar a);
PetscErrorCode VecCopy(Vec x, Vec y);
PetscErrorCode VecGetSize(Vec v, PetscInt* n);
PetscErrorCode VecGetArray(Vec v, PetscScalar** a);
PetscErrorCode VecRestoreArray(Vec v, PetscScalar** a);
PetscErrorCode DMDestroy(DM* dm);

/* The tt_* context shared by the classes built on one mesh (and the one MMA uses). */
tt_ctx* topopt_dropin_context(DM da_nodes);

/* The tt_params the LinearElasticity constructor (LinearElasticity.cc:14-120) creates its context with,
   from the constructor's arguments and the options database: the reference's time grid (dt = t1/(Ts-1),
   :96-99), load (location presets 0/1, Gaussian-cosine signal; -loadsignal is not read, as in the
   reference) and outer-solver tolerances (rtol 1e-10, 200 iterations, :2750-2754, or -ksp_rtol /
   -ksp_max_it), plus this library's own options (-precision, -slots, -cheb_degree, -mg_coarse_pre,
   -mg_coarse_post, -filter_rtol, and the
   y-line half-sine pulse preset -tt_loadloc 2 -tt_loadsignal 1 -tt_pulse_steps n).  No GPU work; returns
   nonzero for a load location without a preset. */
PetscErrorCode LinearElasticityParams(DM da_nodes, PetscScalar nu, PetscInt nlvls, PetscInt ncp, PetscInt tsc,
                                      PetscScalar t1, PetscInt ti, PetscInt objective, PetscInt sInt, PetscInt mInt,
                                      PetscScalar t0, PetscInt boundaries, PetscInt loadloc, PetscBool reduction,
                                      PetscInt basis, PetscScalar dalpha, PetscScalar dbeta, PetscScalar omega0,
                                      PetscInt optimizationRegion, tt_params* out);

class MMA;

/* TopOpt.h / TopOpt.cc:71-518: problem parameters (same option names and defaults),
   design vectors, meshes. */
class TopOpt {
   public:
    TopOpt(PetscInt nconstraint);
    TopOpt();
    ~TopOpt();
    PetscErrorCode AllocateMMAwithRestart(PetscInt* itr, MMA** mma);
    PetscErrorCode WriteRestartFiles(PetscInt* itr, MMA* mma);

    PetscScalar xc[6];
    PetscScalar dx, dy, dz;
    PetscInt nxyz[3];
    PetscInt nlvls;
    PetscScalar nu;
    PetscScalar V0;
    DM da_nodes;
    DM da_elem;
    PetscInt objective, minmax;
    PetscBool reduction;
    PetscBool outputPhysics0, outputPhysics1, outputDesign;
    PetscScalar omega0;
    PetscInt basis, boundaries, optimizationRegion, loadloc;
    PetscInt n, nloc, m;
    PetscScalar fx, fscale;
    PetscScalar* gx;
    PetscScalar Xmin, Xmax, movlim, volfrac;
    PetscScalar penalK, penalM, Emin, Emax, Rmin, Rmax, Ef, Rf, t0, t1;
    PetscInt ti, Ts, ncp, tsc;
    PetscScalar dalpha, dbeta;
    PetscInt maxItr;
    PetscScalar rmin;
    PetscInt filter, sInt, mInt;
    PetscBool projectionFilter;
    PetscReal beta, betaFinal, eta;
    Vec x, xTilde, xPhys, dfdx, xmin, xmax, xold;
    Vec* dgdx;
    PetscBool restart, flip;

   private:
    void Init(PetscInt nconstraint);
    PetscErrorCode SetUp();
    // restart (TopOpt.cc:533-735): MMA state vectors and the two alternating file sets
    Vec xo1, xo2, U, L;
    std::string filename00, filename00Itr, filename01, filename01Itr;
};

/* LinearElasticity.h:29-85 (LinearElasticity.cc): transient elasticity physics. */
class LinearElasticity {
   public:
    LinearElasticity(DM da_nodes, DM da_elem, PetscScalar nu_in, PetscInt nlvls_in, PetscInt ncp_in, PetscInt tsc_in,
                     PetscScalar t1_in, PetscInt ti_in, PetscInt objective_in, PetscInt stiffnessInterpolation_in,
                     PetscInt massInterpolation_in, PetscScalar t0_in, PetscInt boundaries_in, PetscInt loadloc_in,
                     PetscBool reduction_in, PetscInt basis_in, PetscScalar dalpha_in, PetscScalar dbeta_in,
                     PetscScalar omega0_in, PetscInt optimizationRegion, PetscScalar V0);
    ~LinearElasticity();
    PetscErrorCode ComputeObjectiveSensitivities(PetscScalar* fx, Vec dfdx, Vec xPhys, PetscScalar Emin,
                                                 PetscScalar Emax, PetscScalar Rmin, PetscScalar Rmax,
                                                 PetscScalar penalK, PetscScalar penalM, PetscScalar dalpha,
                                                 PetscScalar dbeta, PetscBool write, MPIIO* output);
    PetscErrorCode CompleteInitialState(Vec xPhys, PetscScalar Emin, PetscScalar Emax, PetscScalar Rmin,
                                        PetscScalar Rmax, PetscScalar penalK, PetscScalar penalM, PetscScalar dalpha,
                                        PetscScalar dbeta, PetscBool write, MPIIO* output2);
    PetscErrorCode GenerateCheckPoints(PetscBool write, Vec xPhys, MPIIO* output2);
    PetscErrorCode VolumeConstraint(PetscScalar* gx, Vec dgdx, Vec xPhys, PetscScalar volfrac);
    PetscErrorCode WriteRestartFiles();
    Vec GetStateFieldU(PetscInt j);
    Vec GetStateFieldDU(PetscInt j);
    Vec GetStateFieldDDU(PetscInt j);
    DM GetDM() { return da_nodal; }
    DM da_nodal;
    PetscInt Getts() { return Ts; }

   private:
    PetscInt Ts, ncp;
    PetscBool reduced;  // SOMMG reduced-order model (-reduction)
    tt_ctx* ctx;
    bool forward_done;
    PetscErrorCode WriteStates(MPIIO* output, Vec xPhys);
    // checkpoint files cp%04d.dat (LinearElasticity.cc:1128-1137, 1261-1276, 1651-1672)
    std::string cpdir;
    std::string CheckpointFile(PetscInt c) const;
    PetscErrorCode WriteCheckpoint(PetscInt c, PetscInt t);
    // restart (LinearElasticity.cc:2457-2496, 2679-2723): two alternating RestartSol files
    PetscBool restart, flip;
    std::string filename00, filename01;
    Vec state[3];
    PetscErrorCode SetMaterial(PetscScalar Emin, PetscScalar Emax, PetscScalar Rmin, PetscScalar Rmax,
                               PetscScalar penalK, PetscScalar penalM, PetscScalar dalpha, PetscScalar dbeta);
    Vec StateField(PetscInt j, int which);
    double mat[8];
};

/* PDEFilter.h:32-36 (PDEFilter.cc): Helmholtz-type PDE filter on the node grid. */
class PDEFilt {
   public:
    PDEFilt(DM da_nodes, PetscScalar rmin);
    ~PDEFilt();
    PetscErrorCode FilterProject(Vec XX, Vec F);
    PetscErrorCode Gradients(Vec OS, Vec FS);

   private:
    tt_ctx* ctx;
    PetscScalar R;
};

/* Filter.h:38-55 (Filter.cc): PDE filter, clamp and smooth Heaviside projection. */
class Filter {
   public:
    Filter(DM da_nodes, Vec x, PetscInt filterT, PetscScalar Rin);
    ~Filter();
    PetscErrorCode FilterProject(Vec x, Vec xTilde, Vec xPhys, PetscBool projectionFilter, PetscScalar beta,
                                 PetscScalar eta);
    PetscErrorCode Gradients(Vec x, Vec xTilde, Vec dfdx, PetscInt m, Vec* dgdx, PetscBool projectionFilter,
                             PetscScalar beta, PetscScalar eta);
    PetscBool IncreaseBeta(PetscReal* beta, PetscReal betaFinal, PetscScalar gx, PetscInt itr, PetscReal ch);
    PetscScalar GetMND(Vec x);

   private:
    tt_ctx* ctx;
    PetscInt filterType;
    PetscScalar R;
    PDEFilt* pdef;
};

/* MMA.h:12-53 (MMA.cc): method of moving asymptotes, one constraint. */
class MMA {
   public:
    MMA(PetscInt n, PetscInt m, Vec x);
    MMA(PetscInt n, PetscInt m, Vec x, PetscScalar* a, PetscScalar* c, PetscScalar* d);
    /* restart constructors (MMA.cc:6-177): iteration k, previous iterates and asymptotes */
    MMA(PetscInt n, PetscInt m, PetscInt k, Vec xo1, Vec xo2, Vec U, Vec L, PetscScalar* a, PetscScalar* c,
        PetscScalar* d);
    MMA(PetscInt n, PetscInt m, PetscInt k, Vec xo1, Vec xo2, Vec U, Vec L);
    ~MMA();
    PetscErrorCode Update(Vec xval, Vec dfdx, PetscScalar* gx, Vec* dgdx, Vec xmin, Vec xmax);
    PetscErrorCode Restart(Vec xo1, Vec xo2, Vec U, Vec L);
    PetscErrorCode SetAsymptotes(PetscScalar init, PetscScalar decrease, PetscScalar increase);
    PetscErrorCode ConstraintModification(PetscBool conMod);
    PetscErrorCode SetRobustAsymptotesType(PetscInt val);
    PetscErrorCode SetOuterMovelimit(PetscScalar Xmin, PetscScalar Xmax, PetscScalar movelim, Vec x, Vec xmin,
                                     Vec xmax);
    PetscErrorCode KKTresidual(Vec xval, Vec dfdx, PetscScalar* gx, Vec* dgdx, Vec xmin, Vec xmax, PetscScalar* norm2,
                               PetscScalar* normInf);
    PetscScalar DesignChange(Vec x, Vec xold);

   private:
    PetscInt n, m;
    tt_mma* h;
    void Create(PetscScalar a, PetscScalar c, PetscScalar d);
};

#endif
