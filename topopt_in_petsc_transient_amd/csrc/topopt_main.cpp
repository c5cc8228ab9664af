// Design-loop driver over the drop-in classes, in the order of the reference's main.cc:57-161:
// filter the initial design, then per iteration CompleteInitialState -> (GenerateCheckPoints)
// -> ComputeObjectiveSensitivities -> VolumeConstraint -> Filter::Gradients -> objective
// scaling on the first iteration -> outer move limits -> MMA update -> design change ->
// beta continuation -> constraint guard -> filter the new design; stop when ch <= 0.01 or
// after maxItr iterations.  Takes the reference's command-line options (TopOpt.cc) plus
// -precision, -slots, -ksp_rtol, -ksp_max_it, -cheb_degree, -mg_coarse_pre, -mg_coarse_post,
// -tt_loadloc, -tt_loadsignal, -tt_pulse_steps and
// -history <file> (one JSON line per iteration).
#include <cstdio>
#include <cstring>

#include "topopt_dropin.h"

int main(int argc, char* argv[]) {
    PetscInitialize(&argc, &argv, PETSC_NULL, "3D transient TopOpt on the MI355X GPU path");
    TopOpt* opt = new TopOpt();
    LinearElasticity* physics =
        new LinearElasticity(opt->da_nodes, opt->da_elem, opt->nu, opt->nlvls, opt->ncp, opt->tsc, opt->t1, opt->ti,
                             opt->objective, opt->sInt, opt->mInt, opt->t0, opt->boundaries, opt->loadloc,
                             opt->reduction, opt->basis, opt->dalpha, opt->dbeta, opt->omega0,
                             opt->optimizationRegion, opt->V0);
    // field output (main.cc:39-49; -outputDesign, -outputPhysics0, -outputPhysics1, all on by default)
    MPIIO* output1 = opt->outputDesign ? new MPIIO(opt->da_nodes, 0, "", 3, "x, xTilde, xPhys", "outputDesign_00000.dat")
                                       : nullptr;
    MPIIO* output2 = opt->outputPhysics0
                         ? new MPIIO(opt->da_nodes, 3, "ux, uy, uz", 1, "xPhys", "outputPhysics0_00000.dat")
                         : nullptr;
    MPIIO* output3 = opt->outputPhysics1
                         ? new MPIIO(opt->da_nodes, 3, "ux, uy, uz", 1, "xPhys", "outputPhysics1_00000.dat")
                         : nullptr;
    Filter* filter = new Filter(opt->da_nodes, opt->xPhys, opt->filter, opt->rmin);
    MMA* mma = nullptr;
    PetscInt itr = 0;
    opt->AllocateMMAwithRestart(&itr, &mma);
    mma->SetAsymptotes(0.1, 0.6, 1.05);
    char hist_path[4096] = {0};
    PetscBool has_hist = PETSC_FALSE;
    PetscOptionsGetString(NULL, NULL, "-history", hist_path, sizeof(hist_path), &has_hist);
    FILE* hist = has_hist ? std::fopen(hist_path, "w") : nullptr;

    int status = 0;
#define CHK(call)                 \
    do {                          \
        if ((call) != 0) {        \
            status = 1;           \
            goto done;            \
        }                         \
    } while (0)
    {
        CHK(filter->FilterProject(opt->x, opt->xTilde, opt->xPhys, opt->projectionFilter, opt->beta, opt->eta));
        PetscScalar ch = 1.0;
        const PetscScalar chlim = 0.01;
        PetscBool first = PETSC_TRUE;
        while (itr < opt->maxItr && ch > chlim) {
            itr++;
            const double t_start = MPI_Wtime();
            CHK(physics->CompleteInitialState(opt->xPhys, opt->Emin, opt->Emax, opt->Rmin, opt->Rmax, opt->penalK,
                                              opt->penalM, opt->dalpha, opt->dbeta, first, output2));
            if (opt->ncp > 1) CHK(physics->GenerateCheckPoints(first, opt->xPhys, output2));
            CHK(physics->ComputeObjectiveSensitivities(&opt->fx, opt->dfdx, opt->xPhys, opt->Emin, opt->Emax,
                                                       opt->Rmin, opt->Rmax, opt->penalK, opt->penalM, opt->dalpha,
                                                       opt->dbeta, first, output2));
            CHK(physics->VolumeConstraint(&opt->gx[0], opt->dgdx[0], opt->xPhys, opt->volfrac));
            CHK(filter->Gradients(opt->x, opt->xTilde, opt->dfdx, opt->m, opt->dgdx, opt->projectionFilter,
                                  opt->beta, opt->eta));
            first = PETSC_FALSE;
            if (output1) CHK(output1->WriteVTK(physics->da_nodal, opt->x, opt->xTilde, opt->xPhys, itr));
            const PetscScalar mnd = filter->GetMND(opt->xPhys);
            if (itr == 1) opt->fscale = 1.0 / opt->fx;
            opt->fx *= opt->fscale;
            VecScale(opt->dfdx, opt->fscale);
            CHK(mma->SetOuterMovelimit(opt->Xmin, opt->Xmax, opt->movlim, opt->x, opt->xmin, opt->xmax));
            if (opt->minmax == 1) VecScale(opt->dfdx, -1);
            CHK(mma->Update(opt->x, opt->dfdx, opt->gx, opt->dgdx, opt->xmin, opt->xmax));
            ch = mma->DesignChange(opt->x, opt->xold);
            const double t_end = MPI_Wtime();
            PetscPrintf(PETSC_COMM_WORLD,
                        "# It.: %03i, True fx: %4.10e, Scaled fx: %4.2f, gx[0]: %+02.2f, ch: %02.2e, mnd: %2.2f, "
                        "time: %02.2f\n",
                        itr, opt->fx / opt->fscale, opt->fx, opt->gx[0], ch, mnd, t_end - t_start);
            if (hist) {
                std::fprintf(hist,
                             "{\"itr\": %d, \"fx\": %.17g, \"fx_scaled\": %.17g, \"gx\": %.17g, \"ch\": %.17g, "
                             "\"mnd\": %.17g, \"seconds\": %.6f}\n",
                             itr, opt->fx / opt->fscale, opt->fx, opt->gx[0], ch, mnd, t_end - t_start);
                std::fflush(hist);
            }
            PetscBool change_beta = PETSC_FALSE;
            if (opt->projectionFilter && (itr >= 200 || ch < 0.01)) {
                change_beta = filter->IncreaseBeta(&opt->beta, opt->betaFinal, opt->gx[0], itr, ch);
                if (change_beta) ch = 1.0;
            }
            if (opt->gx[0] > 0) ch = 1;
            CHK(filter->FilterProject(opt->x, opt->xTilde, opt->xPhys, opt->projectionFilter, opt->beta, opt->eta));
            // restart files every 10 iterations and when beta changes (main.cc:156-160)
            if (itr % 10 == 0 || change_beta) {
                opt->WriteRestartFiles(&itr, mma);
                physics->WriteRestartFiles();
            }
        }
        // and at the end (main.cc:163-165)
        opt->WriteRestartFiles(&itr, mma);
        physics->WriteRestartFiles();
        // the final design's physics once more, into outputPhysics1 (main.cc:167-176)
        CHK(physics->CompleteInitialState(opt->xPhys, opt->Emin, opt->Emax, opt->Rmin, opt->Rmax, opt->penalK,
                                          opt->penalM, opt->dalpha, opt->dbeta, PETSC_TRUE, output3));
        if (opt->ncp > 1) CHK(physics->GenerateCheckPoints(PETSC_TRUE, opt->xPhys, output3));
        CHK(physics->ComputeObjectiveSensitivities(&opt->fx, opt->dfdx, opt->xPhys, opt->Emin, opt->Emax, opt->Rmin,
                                                   opt->Rmax, opt->penalK, opt->penalM, opt->dalpha, opt->dbeta,
                                                   PETSC_TRUE, output3));
        if (hist) {
            // final design (xPhys) for comparison
            std::fprintf(hist, "{\"xphys\": [");
            PetscInt n = 0;
            VecGetSize(opt->xPhys, &n);
            for (PetscInt i = 0; i < n; i++) std::fprintf(hist, i ? ",%.17g" : "%.17g", opt->xPhys->a[i]);
            std::fprintf(hist, "]}\n");
        }
    }
done:
    if (hist) std::fclose(hist);
    delete mma;
    delete output1;
    delete output2;
    delete output3;
    delete filter;
    delete physics;
    delete opt;
    PetscPrintf(PETSC_COMM_WORLD, status ? "# Optimization stopped on an error\n" : "# Optimization has terminated\n");
    PetscFinalize();
    return status;
}
