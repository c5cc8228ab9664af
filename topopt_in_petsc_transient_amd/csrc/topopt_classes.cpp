// Drop-in classes (include/topopt_dropin.h) over the tt_* C-ABI.
//
// Each class keeps the reference's signature and call semantics (LinearElasticity.h,
// Filter.h, PDEFilter.h, MMA.h, TopOpt.h); the arithmetic runs in libtopopt_gpu.so.
// Vectors are host arrays, so every call moves nel (or 3 nn) doubles over PCIe; the
// Newmark sweeps and adjoint stay on the device inside the C-ABI calls.
#include "topopt_dropin.h"

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

// ------------------------------------------------------------------ minimal PETSc names
// The options database matches names without regard to case, as PETSc's does (the reference's job
// scripts pass -Rmin for the -rmin that TopOpt.cc:363 reads): names are stored and looked up folded
// to lower case.  Values keep their case.
static std::string fold(const char* name) {
    std::string s(name);
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
static std::map<std::string, std::string>& options() {
    static std::map<std::string, std::string> db;
    return db;
}

PetscErrorCode PetscInitialize(int* argc, char*** argv, const char*, const char*) {
    auto& db = options();
    for (int i = 1; argc && argv && i < *argc; i++) {
        const char* a = (*argv)[i];
        if (a[0] != '-' || (a[1] >= '0' && a[1] <= '9') || a[1] == '.') continue;
        const bool has_val = i + 1 < *argc && ((*argv)[i + 1][0] != '-' || ((*argv)[i + 1][1] >= '0' &&
                                                                             (*argv)[i + 1][1] <= '9') ||
                                              (*argv)[i + 1][1] == '.');
        db[fold(a)] = has_val ? (*argv)[++i] : "1";
    }
    return 0;
}
PetscErrorCode PetscFinalize() { return 0; }
PetscErrorCode PetscPrintf(int, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    return 0;
}
static const std::string* opt(const char* name) {
    auto it = options().find(fold(name));
    return it == options().end() ? nullptr : &it->second;
}
PetscErrorCode PetscOptionsGetInt(void*, const char*, const char* name, PetscInt* v, PetscBool* set) {
    const std::string* s = opt(name);
    if (s) *v = (PetscInt)std::strtol(s->c_str(), nullptr, 10);
    if (set) *set = s ? PETSC_TRUE : PETSC_FALSE;
    return 0;
}
PetscErrorCode PetscOptionsGetReal(void*, const char*, const char* name, PetscReal* v, PetscBool* set) {
    const std::string* s = opt(name);
    if (s) *v = std::strtod(s->c_str(), nullptr);
    if (set) *set = s ? PETSC_TRUE : PETSC_FALSE;
    return 0;
}
PetscErrorCode PetscOptionsGetBool(void*, const char*, const char* name, PetscBool* v, PetscBool* set) {
    const std::string* s = opt(name);
    if (s) {
        const std::string b = fold(s->c_str());
        *v = (b == "0" || b == "false" || b == "no" || b == "off") ? PETSC_FALSE : PETSC_TRUE;
    }
    if (set) *set = s ? PETSC_TRUE : PETSC_FALSE;
    return 0;
}
PetscErrorCode PetscOptionsGetString(void*, const char*, const char* name, char* v, size_t len, PetscBool* set) {
    const std::string* s = opt(name);
    if (s && len > 0) {
        std::strncpy(v, s->c_str(), len - 1);
        v[len - 1] = 0;
    }
    if (set) *set = s ? PETSC_TRUE : PETSC_FALSE;
    return 0;
}
double MPI_Wtime() {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

PetscErrorCode VecCreateSeq(int, PetscInt n, Vec* v) {
    *v = new _p_Vec;
    (*v)->a.assign(n, 0.0);
    return 0;
}
PetscErrorCode VecDuplicate(Vec x, Vec* y) { return VecCreateSeq(0, (PetscInt)x->a.size(), y); }
PetscErrorCode VecDestroy(Vec* v) {
    delete *v;
    *v = nullptr;
    return 0;
}
PetscErrorCode VecSet(Vec v, PetscScalar a) {
    for (auto& e : v->a) e = a;
    return 0;
}
PetscErrorCode VecScale(Vec v, PetscScalar a) {
    for (auto& e : v->a) e *= a;
    return 0;
}
PetscErrorCode VecCopy(Vec x, Vec y) {
    y->a = x->a;
    return 0;
}
PetscErrorCode VecGetSize(Vec v, PetscInt* n) {
    *n = (PetscInt)v->a.size();
    return 0;
}
PetscErrorCode VecGetArray(Vec v, PetscScalar** a) {
    *a = v->a.data();
    return 0;
}
PetscErrorCode VecRestoreArray(Vec, PetscScalar** a) {
    *a = nullptr;
    return 0;
}
PetscErrorCode DMDestroy(DM* dm) {
    delete *dm;
    *dm = nullptr;
    return 0;
}

// ------------------------------------------------------------------ output writer
// MPIIO.cc:32-136, 409-586 (constructor, Allocate, writePoints, writeCells) for one rank and one domain:
// all of it is the file's first part, written here in one go.  With -workdir the file goes to
// "./" + workdir + "/" + filename, as MPIIO.cc:431-437 builds it; an absolute workdir (which that
// prefix would turn into a relative path) is used as it is.
MPIIO::MPIIO(DM da, int nPfields, std::string pnames, int nCfields, std::string cnames, std::string fname) {
    char wd[PETSC_MAX_PATH_LEN];
    PetscBool flg = PETSC_FALSE;
    PetscOptionsGetString(NULL, NULL, "-workdir", wd, sizeof(wd), &flg);
    filename = !flg ? fname : (wd[0] == '/' ? std::string(wd) : std::string("./") + wd) + "/" + fname;
    PetscPrintf(PETSC_COMM_WORLD, "# Outputfile is written to: %s \n", filename.c_str());
    const long nx = da->nxyz[0], ny = da->nxyz[1], nz = da->nxyz[2];
    nPoints = (unsigned long)(nx * ny * nz);
    nCells = (unsigned long)((nx - 1) * (ny - 1) * (nz - 1));
    nPF = nPfields;
    nCF = nCfields;
    workPoint.assign((size_t)std::max(nPF, 0) * nPoints, 0.0f);
    workCell.assign((size_t)std::max(nCF, 0) * nCells, 0.0f);
    std::string head = std::string("TopOpt result version 1.1") + "\n\x01";
    const unsigned long header[6] = {1, nPoints, nCells, (unsigned long)nPF, (unsigned long)nCF, 8};
    const std::string names = pnames + "\x01" + cnames + "\x01";
    // node coordinates (natural order, x fastest), as float32
    std::vector<float> pts(3 * nPoints);
    const double h[3] = {(da->xc[1] - da->xc[0]) / (nx - 1), (da->xc[3] - da->xc[2]) / (ny - 1),
                         (da->xc[5] - da->xc[4]) / (nz - 1)};
    for (long k = 0, n = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++, n++) {
                pts[3 * n] = (float)(da->xc[0] + i * h[0]);
                pts[3 * n + 1] = (float)(da->xc[2] + j * h[1]);
                pts[3 * n + 2] = (float)(da->xc[4] + k * h[2]);
            }
    // hex8 cells in the element corner order (LinearElasticity.cc:3067-3074), offsets, VTK type 12
    std::vector<unsigned long> conn(8 * nCells), offs(nCells), types(nCells, 12);
    for (long k = 0, e = 0; k < nz - 1; k++)
        for (long j = 0; j < ny - 1; j++)
            for (long i = 0; i < nx - 1; i++, e++) {
                const unsigned long n0 = i + nx * (j + ny * k), sy = nx, sz = nx * ny;
                const unsigned long c[8] = {n0, n0 + 1, n0 + 1 + sy, n0 + sy, n0 + sz, n0 + 1 + sz, n0 + 1 + sy + sz,
                                            n0 + sy + sz};
                for (int q = 0; q < 8; q++) conn[8 * e + q] = c[q];
                offs[e] = 8 * (e + 1);
            }
    FILE* f = std::fopen(filename.c_str(), "wb");
    bool ok = f != nullptr;
    ok = ok && std::fwrite(head.data(), 1, head.size(), f) == head.size();
    ok = ok && std::fwrite(header, sizeof(unsigned long), 6, f) == 6;
    ok = ok && std::fwrite(names.data(), 1, names.size(), f) == names.size();
    ok = ok && std::fwrite(pts.data(), sizeof(float), pts.size(), f) == pts.size();
    ok = ok && std::fwrite(conn.data(), sizeof(unsigned long), conn.size(), f) == conn.size();
    ok = ok && std::fwrite(offs.data(), sizeof(unsigned long), offs.size(), f) == offs.size();
    ok = ok && std::fwrite(types.data(), sizeof(unsigned long), types.size(), f) == types.size();
    if (f) ok = (std::fclose(f) == 0) && ok;
    if (!ok) {  // the reference aborts here (MPIIO.cc:536-538)
        std::fprintf(stderr, "[topopt] cannot write the output file %s\n", filename.c_str());
        std::exit(1);
    }
}
MPIIO::~MPIIO() {}
// one dataset (writePointFields + writeCellFields, MPIIO.cc:724-905): the step, nPF point fields,
// nCF cell fields
PetscErrorCode MPIIO::Append(PetscInt itr) {
    const unsigned long step = (unsigned long)itr;
    FILE* f = std::fopen(filename.c_str(), "ab");
    bool ok = f != nullptr;
    ok = ok && std::fwrite(&step, sizeof(step), 1, f) == 1;
    ok = ok && std::fwrite(workPoint.data(), sizeof(float), workPoint.size(), f) == workPoint.size();
    ok = ok && std::fwrite(workCell.data(), sizeof(float), workCell.size(), f) == workCell.size();
    if (f) ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        std::fprintf(stderr, "[topopt] cannot append to the output file %s\n", filename.c_str());
        return TT_ERR_ARG;
    }
    return 0;
}
// point fields field0 .. field0 + ncomp - 1 from the components of a nodal vector (fields past nPF
// are not written, as MPIIO.cc writes only nPFields of its work array)
void MPIIO::PutPoint(int field0, Vec v, int ncomp) {
    for (int c = 0; c < ncomp; c++) {
        const int fld = field0 + c;
        if (fld >= nPF) return;
        for (unsigned long i = 0; i < nPoints; i++) workPoint[fld * nPoints + i] = (float)v->a[ncomp * i + c];
    }
}
void MPIIO::PutCell(int field, Vec v) {
    if (field >= nCF) return;
    for (unsigned long i = 0; i < nCells; i++) workCell[field * nCells + i] = (float)v->a[i];
}
// MPIIO.cc:373-404: the design fields (x, xTilde, xPhys); the point fields keep their last values
PetscErrorCode MPIIO::WriteVTK(DM, Vec x, Vec xTilde, Vec xPhys, PetscInt itr) {
    PutCell(0, x);
    PutCell(1, xTilde);
    PutCell(2, xPhys);
    return Append(itr);
}
// MPIIO.cc:152-213: displacement (ux, uy, uz) and x, xTilde, xPhys
PetscErrorCode MPIIO::WriteVTK(DM, Vec U, Vec x, Vec xTilde, Vec xPhys, PetscInt itr) {
    PutPoint(0, U, 3);
    PutCell(0, x);
    PutCell(1, xTilde);
    PutCell(2, xPhys);
    return Append(itr);
}
// MPIIO.cc:215-268: displacement (ux, uy, uz) and xPhys
PetscErrorCode MPIIO::WriteVTK(DM, Vec U, Vec xPhys, PetscInt itr) {
    PutPoint(0, U, 3);
    PutCell(0, xPhys);
    return Append(itr);
}
// MPIIO.cc:270-371: displacement, velocity and acceleration, and x, xTilde, xPhys
PetscErrorCode MPIIO::WriteVTK(DM, Vec U, Vec DU, Vec DDU, Vec x, Vec xTilde, Vec xPhys, PetscInt itr) {
    PutPoint(0, U, 3);
    PutPoint(3, DU, 3);
    PutPoint(6, DDU, 3);
    PutCell(0, x);
    PutCell(1, xTilde);
    PutCell(2, xPhys);
    return Append(itr);
}

// ------------------------------------------------------------------ shared GPU context of a mesh
struct TTMesh {
    tt_ctx* ctx = nullptr;
    ~TTMesh() {
        if (ctx) tt_destroy(ctx);
    }
};
static tt_ctx* g_current = nullptr;  // the context MMA (which gets no DM) works with

static PetscErrorCode report(tt_ctx* ctx, int rc, const char* where) {
    if (rc != TT_OK) std::fprintf(stderr, "[topopt] %s failed (%d): %s\n", where, rc, tt_last_error(ctx));
    return rc;
}
#define TTCALL(ctx, call)                                   \
    do {                                                    \
        const int rc_ = (call);                             \
        if (rc_ != TT_OK) return report((ctx), rc_, #call); \
    } while (0)

static DM make_dm(const PetscInt nxyz[3], const PetscScalar xc[6], bool elemental, std::shared_ptr<TTMesh> mesh) {
    DM d = new _p_DM;
    for (int i = 0; i < 3; i++) d->nxyz[i] = nxyz[i];
    for (int i = 0; i < 6; i++) d->xc[i] = xc[i];
    d->elemental = elemental;
    d->mesh = std::move(mesh);
    return d;
}

// filter-only context for a mesh that has no physics yet (one step, default material)
static tt_ctx* mesh_context(DM da) {
    if (!da->mesh) da->mesh = std::make_shared<TTMesh>();
    if (!da->mesh->ctx) {
        tt_params p;
        tt_default_params(&p);
        p.nelx = da->nxyz[0] - 1;
        p.nely = da->nxyz[1] - 1;
        p.nelz = da->nxyz[2] - 1;
        for (int i = 0; i < 6; i++) p.xc[i] = da->xc[i];
        p.nsteps = 1;
        p.nlvls = 1;
        tt_ctx* c = nullptr;
        if (report(nullptr, tt_create(&p, &c), "tt_create") != TT_OK) std::exit(1);
        da->mesh->ctx = c;
    }
    g_current = da->mesh->ctx;
    return da->mesh->ctx;
}
tt_ctx* topopt_dropin_context(DM da_nodes) { return da_nodes && da_nodes->mesh ? da_nodes->mesh->ctx : nullptr; }

// ------------------------------------------------------------------ TopOpt (TopOpt.cc:71-518)
TopOpt::TopOpt(PetscInt nconstraint) { Init(nconstraint); }
TopOpt::TopOpt() { Init(1); }
void TopOpt::Init(PetscInt nconstraint) {
    m = nconstraint;
    gx = nullptr;
    x = xTilde = xPhys = dfdx = xmin = xmax = xold = nullptr;
    dgdx = nullptr;
    da_nodes = da_elem = nullptr;
    fx = 0.0;
    fscale = 1.0;
    restart = PETSC_TRUE;  // TopOpt.cc:534: restart files are written by default
    flip = PETSC_TRUE;
    xo1 = xo2 = U = L = nullptr;
    if (SetUp() != 0) std::exit(1);
}
TopOpt::~TopOpt() {
    for (Vec* v : {&x, &xTilde, &xPhys, &dfdx, &xmin, &xmax, &xold, &xo1, &xo2, &U, &L})
        if (*v) VecDestroy(v);
    if (dgdx) {
        for (PetscInt i = 0; i < m; i++) VecDestroy(&dgdx[i]);
        delete[] dgdx;
    }
    delete[] gx;
    if (da_nodes) DMDestroy(&da_nodes);
    if (da_elem) DMDestroy(&da_elem);
}

PetscErrorCode TopOpt::SetUp() {
    PetscBool flg;
    // defaults as TopOpt.cc:71-170
    boundaries = 0;
    PetscOptionsGetInt(NULL, NULL, "-boundaries", &boundaries, &flg);
    if (boundaries == 1) {
        nxyz[0] = 65; nxyz[1] = 33; nxyz[2] = 33;
        xc[0] = 0.0; xc[1] = 40.0; xc[2] = 0.0; xc[3] = 20.0; xc[4] = 0.0; xc[5] = 20.0;
        loadloc = 1;
    } else {
        nxyz[0] = 97; nxyz[1] = 33; nxyz[2] = 33;
        xc[0] = 0.0; xc[1] = 60.0; xc[2] = 0.0; xc[3] = 20.0; xc[4] = 0.0; xc[5] = 20.0;
        loadloc = 0;
    }
    nu = 0.3;
    nlvls = 4;
    volfrac = 0.3;
    V0 = volfrac;
    maxItr = 10;
    rmin = 1.5;
    penalK = 3.0;
    penalM = 1.0;
    Emax = 1e3;
    Rmax = 4.16666666667e-4;
    filter = 1;
    sInt = 1;
    mInt = 0;
    Xmin = 0.0;
    Xmax = 1.0;
    movlim = 0.2;
    Rf = 1e-3;
    Ef = 1e-9;
    ti = 1;
    t0 = 0.0;
    t1 = 50;
    ncp = 2;
    tsc = 25;
    dalpha = 1e-4;
    dbeta = 1e-4;
    objective = 0;
    minmax = -1;
    optimizationRegion = 0;
    projectionFilter = PETSC_TRUE;
    beta = 0.1;
    betaFinal = 48;
    eta = 0.15;
    reduction = PETSC_FALSE;
    basis = 5;
    omega0 = 1.0;
    outputPhysics0 = outputPhysics1 = outputDesign = PETSC_TRUE;  // TopOpt.cc:474-476
    // mesh (SetUpMESH, TopOpt.cc:175-330)
    PetscOptionsGetInt(NULL, NULL, "-nx", &nxyz[0], &flg);
    PetscOptionsGetInt(NULL, NULL, "-ny", &nxyz[1], &flg);
    PetscOptionsGetInt(NULL, NULL, "-nz", &nxyz[2], &flg);
    PetscOptionsGetReal(NULL, NULL, "-xcmin", &xc[0], &flg);
    PetscOptionsGetReal(NULL, NULL, "-xcmax", &xc[1], &flg);
    PetscOptionsGetReal(NULL, NULL, "-ycmin", &xc[2], &flg);
    PetscOptionsGetReal(NULL, NULL, "-ycmax", &xc[3], &flg);
    PetscOptionsGetReal(NULL, NULL, "-zcmin", &xc[4], &flg);
    PetscOptionsGetReal(NULL, NULL, "-zcmax", &xc[5], &flg);
    PetscOptionsGetInt(NULL, NULL, "-nlvls", &nlvls, &flg);
    const PetscInt div = 1 << (nlvls - 1);
    for (int d = 0; d < 3; d++)
        if ((nxyz[d] - 1) % div != 0) {
            PetscPrintf(PETSC_COMM_WORLD, "MESH DIMENSION NOT COMPATIBLE WITH NUMBER OF MULTIGRID LEVELS!\n");
            return TT_ERR_ARG;
        }
    dx = (xc[1] - xc[0]) / (nxyz[0] - 1);
    dy = (xc[3] - xc[2]) / (nxyz[1] - 1);
    dz = (xc[5] - xc[4]) / (nxyz[2] - 1);
    auto mesh = std::make_shared<TTMesh>();
    da_nodes = make_dm(nxyz, xc, false, mesh);
    const PetscInt nel[3] = {nxyz[0] - 1, nxyz[1] - 1, nxyz[2] - 1};
    da_elem = make_dm(nel, xc, true, mesh);
    // optimisation parameters (SetUpOPT, TopOpt.cc:340-470)
    PetscOptionsGetInt(NULL, NULL, "-boundaries", &boundaries, &flg);
    PetscOptionsGetInt(NULL, NULL, "-loadloc", &loadloc, &flg);
    PetscOptionsGetInt(NULL, NULL, "-filter", &filter, &flg);
    PetscOptionsGetReal(NULL, NULL, "-rmin", &rmin, &flg);
    PetscOptionsGetBool(NULL, NULL, "-projectionFilter", &projectionFilter, &flg);
    PetscOptionsGetReal(NULL, NULL, "-beta", &beta, &flg);
    PetscOptionsGetReal(NULL, NULL, "-betaFinal", &betaFinal, &flg);
    PetscOptionsGetReal(NULL, NULL, "-eta", &eta, &flg);
    // V0 keeps the default volume fraction it was given before -volfrac is read (TopOpt.cc:113-114, 373);
    // only -V0 changes it
    PetscOptionsGetReal(NULL, NULL, "-volfrac", &volfrac, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Xmax", &Xmax, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Xmin", &Xmin, &flg);
    PetscOptionsGetInt(NULL, NULL, "-maxItr", &maxItr, &flg);
    PetscOptionsGetReal(NULL, NULL, "-movlim", &movlim, &flg);
    PetscOptionsGetInt(NULL, NULL, "-optimizationRegion", &optimizationRegion, &flg);
    PetscOptionsGetReal(NULL, NULL, "-V0", &V0, &flg);
    PetscOptionsGetReal(NULL, NULL, "-nu", &nu, &flg);
    PetscOptionsGetReal(NULL, NULL, "-penalM", &penalM, &flg);
    PetscOptionsGetReal(NULL, NULL, "-penalK", &penalK, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Rf", &Rf, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Ef", &Ef, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Emax", &Emax, &flg);
    PetscOptionsGetReal(NULL, NULL, "-Rmax", &Rmax, &flg);
    Emin = Ef * Emax;
    Rmin = Rf * Rmax;
    PetscOptionsGetReal(NULL, NULL, "-dalph", &dalpha, &flg);
    PetscOptionsGetReal(NULL, NULL, "-dbeta", &dbeta, &flg);
    PetscOptionsGetInt(NULL, NULL, "-massInterpolation", &mInt, &flg);
    PetscOptionsGetInt(NULL, NULL, "-sInt", &sInt, &flg);
    PetscOptionsGetInt(NULL, NULL, "-ti", &ti, &flg);
    PetscOptionsGetReal(NULL, NULL, "-t0", &t0, &flg);
    PetscOptionsGetReal(NULL, NULL, "-t1", &t1, &flg);
    PetscOptionsGetInt(NULL, NULL, "-ncp", &ncp, &flg);
    PetscOptionsGetInt(NULL, NULL, "-tsc", &tsc, &flg);
    Ts = ncp * tsc + 1;
    PetscOptionsGetInt(NULL, NULL, "-objective", &objective, &flg);
    PetscOptionsGetInt(NULL, NULL, "-minmax", &minmax, &flg);
    PetscOptionsGetBool(NULL, NULL, "-reduction", &reduction, &flg);
    PetscOptionsGetInt(NULL, NULL, "-basis", &basis, &flg);
    PetscOptionsGetReal(NULL, NULL, "-omega0", &omega0, &flg);
    PetscOptionsGetBool(NULL, NULL, "-outputPhysics0", &outputPhysics0, &flg);
    PetscOptionsGetBool(NULL, NULL, "-outputPhysics1", &outputPhysics1, &flg);
    PetscOptionsGetBool(NULL, NULL, "-outputDesign", &outputDesign, &flg);
    n = nel[0] * nel[1] * nel[2];
    nloc = n;
    gx = new PetscScalar[m];
    for (PetscInt i = 0; i < m; i++) gx[i] = 0.0;
    // design vectors initialised to V0 (TopOpt.cc:495-515)
    VecCreateSeq(0, n, &xPhys);
    VecDuplicate(xPhys, &x);
    VecDuplicate(xPhys, &xTilde);
    VecSet(x, V0);
    VecSet(xTilde, V0);
    VecSet(xPhys, V0);
    VecDuplicate(x, &dfdx);
    dgdx = new Vec[m];
    for (PetscInt i = 0; i < m; i++) VecDuplicate(x, &dgdx[i]);
    VecDuplicate(x, &xmin);
    VecDuplicate(x, &xmax);
    VecDuplicate(x, &xold);
    VecSet(xold, V0);
    PetscPrintf(PETSC_COMM_WORLD, "# nodes (%d,%d,%d), elements %d, dofs %d, nlvls %d, Ts %d, volfrac %g, rmin %g\n",
                nxyz[0], nxyz[1], nxyz[2], n, 3 * nxyz[0] * nxyz[1] * nxyz[2], nlvls, Ts, volfrac, rmin);
    return 0;
}

static bool file_exists(const std::string& f) {
    if (f.empty()) return false;
    FILE* p = std::fopen(f.c_str(), "rb");
    if (p) std::fclose(p);
    return p != nullptr;
}
static PetscErrorCode write_vecs(const std::string& f, std::initializer_list<Vec> vs) {
    std::vector<const double*> p;
    std::vector<long> n;
    for (Vec v : vs) {
        p.push_back(v->a.data());
        n.push_back((long)v->a.size());
    }
    return report(nullptr, tt_petsc_write_vecs(f.c_str(), (int)p.size(), p.data(), n.data()), "tt_petsc_write_vecs");
}
static PetscErrorCode read_vecs(const std::string& f, std::initializer_list<Vec> vs) {
    std::vector<double*> p;
    std::vector<long> n;
    for (Vec v : vs) {
        p.push_back(v->a.data());
        n.push_back((long)v->a.size());
    }
    return report(nullptr, tt_petsc_read_vecs(f.c_str(), (int)p.size(), p.data(), n.data()), "tt_petsc_read_vecs");
}
static std::string option_string(const char* name, const std::string& def) {
    char buf[PETSC_MAX_PATH_LEN];
    PetscBool flg = PETSC_FALSE;
    PetscOptionsGetString(NULL, NULL, name, buf, sizeof(buf), &flg);
    return flg ? std::string(buf) : def;
}

// MMA with a = 0, c = 1000, d = 0 (TopOpt.cc:527-531) and the restart point (TopOpt.cc:533-695):
// -restart (default on: restart files are written), -onlyLoadDesign, -loadDesign (0: x, 2: xPhys),
// -restartFileVec / -restartFileItr (the files read), -workdir (where Restart0x.dat and
// Restart0x_itr_f0.dat are written)
PetscErrorCode TopOpt::AllocateMMAwithRestart(PetscInt* itr, MMA** mma) {
    std::vector<PetscScalar> a(m, 0.0), c(m, 1000.0), d(m, 0.0);
    restart = PETSC_TRUE;
    flip = PETSC_TRUE;
    PetscBool onlyLoadDesign = PETSC_FALSE, flg;
    PetscInt loadDesign = 2;
    PetscOptionsGetBool(NULL, NULL, "-restart", &restart, &flg);
    PetscOptionsGetBool(NULL, NULL, "-onlyLoadDesign", &onlyLoadDesign, &flg);
    PetscOptionsGetInt(NULL, NULL, "-loadDesign", &loadDesign, &flg);
    if (restart) {
        for (Vec* v : {&xo1, &xo2, &U, &L})
            if (!*v) VecDuplicate(x, v);
    }
    const std::string wd = option_string("-workdir", "./");
    filename00 = wd + "/Restart00.dat";
    filename00Itr = wd + "/Restart00_itr_f0.dat";
    filename01 = wd + "/Restart01.dat";
    filename01Itr = wd + "/Restart01_itr_f0.dat";
    const std::string fileVec = option_string("-restartFileVec", ""), fileItr = option_string("-restartFileItr", "");
    PetscPrintf(PETSC_COMM_WORLD, "# Continue from previous iteration (-restart): %i \n", (int)restart);
    PetscPrintf(PETSC_COMM_WORLD, "# Restart file (-restartFileVec): %s \n", fileVec.c_str());
    PetscPrintf(PETSC_COMM_WORLD, "# Restart file (-restartFileItr): %s \n", fileItr.c_str());
    const bool vecFile = file_exists(fileVec), itrFile = file_exists(fileItr);
    if (!vecFile) PetscPrintf(PETSC_COMM_WORLD, "File: %s NOT FOUND \n", fileVec.c_str());
    if (!itrFile) PetscPrintf(PETSC_COMM_WORLD, "File: %s NOT FOUND \n", fileItr.c_str());
    *itr = 0;
    if (restart && vecFile && itrFile) {
        // x, xPhys, xold1, xold2, U, L in the order TopOpt::WriteRestartFiles writes them
        if (read_vecs(fileVec, {x, xPhys, xo1, xo2, U, L}) != 0) return TT_ERR_ARG;
        if (onlyLoadDesign) {
            if (loadDesign == 2) VecCopy(xPhys, x);  // the physical design as the initial design
            PetscPrintf(PETSC_COMM_WORLD, "# Loading design from file: %s \n", fileVec.c_str());
            *mma = new MMA(n, m, x, a.data(), c.data(), d.data());
        } else {
            VecCopy(x, xold);
            FILE* f = std::fopen(fileItr.c_str(), "r");
            int it = 0;
            double fs = 0.0;
            const bool ok = f && std::fscanf(f, "%d %lf", &it, &fs) == 2;
            if (f) std::fclose(f);
            if (!ok) {
                std::fprintf(stderr, "[topopt] cannot read the iteration and fscale from %s\n", fileItr.c_str());
                return TT_ERR_ARG;
            }
            *itr = it;
            fscale = fs;
            PetscPrintf(PETSC_COMM_WORLD, "# Continue optimization from file: %s \n", fileVec.c_str());
            *mma = new MMA(n, m, *itr, xo1, xo2, U, L, a.data(), c.data(), d.data());
        }
        PetscPrintf(PETSC_COMM_WORLD, "# Successful restart from file: %s and %s \n", fileVec.c_str(),
                    fileItr.c_str());
    } else {
        *mma = new MMA(n, m, x, a.data(), c.data(), d.data());
    }
    return 0;
}

// TopOpt.cc:697-755: the iteration and fscale (ASCII) and x, xPhys, xold1, xold2, U, L (PETSc binary),
// alternating between the 00 and 01 files
PetscErrorCode TopOpt::WriteRestartFiles(PetscInt* itr, MMA* mma) {
    if (!restart) return -1;
    PetscErrorCode ierr = mma->Restart(xo1, xo2, U, L);
    if (ierr) return ierr;
    flip = flip ? PETSC_FALSE : PETSC_TRUE;
    const std::string& fv = flip ? filename01 : filename00;
    const std::string& fi = flip ? filename01Itr : filename00Itr;
    FILE* f = std::fopen(fi.c_str(), "w");
    if (!f || std::fprintf(f, "%d %1.16e\n", (int)itr[0], fscale) < 0) {
        if (f) std::fclose(f);
        std::fprintf(stderr, "[topopt] cannot write %s\n", fi.c_str());
        return TT_ERR_ARG;
    }
    std::fclose(f);
    return write_vecs(fv, {x, xPhys, xo1, xo2, U, L});
}

// ------------------------------------------------------------------ LinearElasticity
PetscErrorCode LinearElasticityParams(DM da_nodes, PetscScalar nu_in, PetscInt nlvls_in, PetscInt ncp_in,
                                      PetscInt tsc_in, PetscScalar t1_in, PetscInt ti_in, PetscInt objective_in,
                                      PetscInt sInt_in, PetscInt mInt_in, PetscScalar, PetscInt boundaries_in,
                                      PetscInt loadloc_in, PetscBool reduction_in, PetscInt basis_in,
                                      PetscScalar dalpha_in, PetscScalar dbeta_in, PetscScalar omega0_in,
                                      PetscInt optimizationRegion, tt_params* out) {
    tt_params p;
    tt_default_params(&p);
    // SOMMG reduced-order model (LinearElasticity.cc:404-461, 506-508, 994-1003)
    p.reduction = reduction_in ? 1 : 0;
    p.basis = basis_in;
    p.omega0 = omega0_in;
    p.nelx = da_nodes->nxyz[0] - 1;
    p.nely = da_nodes->nxyz[1] - 1;
    p.nelz = da_nodes->nxyz[2] - 1;
    for (int i = 0; i < 6; i++) p.xc[i] = da_nodes->xc[i];
    p.nu = nu_in;
    p.nlvls = nlvls_in;
    p.nsteps = ncp_in * tsc_in;  // Ts - 1 (LinearElasticity.cc:96)
    p.t1 = t1_in;                // dt = t1 / (Ts - 1) (:99); t0 takes no part in the time grid or the load
    p.ti = ti_in;
    p.objective = objective_in;
    p.sInt = sInt_in;
    p.mInt = mInt_in;
    p.boundaries = boundaries_in;
    p.region = optimizationRegion;
    p.dalpha = dalpha_in;
    p.dbeta = dbeta_in;
    // The reference's load: the location presets 0 and 1 of LoadLocation (:616-675) with the Gaussian-cosine
    // signal of Load (:593-615), which reads no option (the -loadsignal its job scripts pass is ignored).
    // The half-sine pulse on the y-line at x = xmax, z = zmin is this library's own preset, chosen with
    // options the reference does not know: -tt_loadloc 2 -tt_loadsignal 1 -tt_pulse_steps n.
    p.loadloc = loadloc_in;
    p.loadsignal = 0;
    PetscBool flg, own = PETSC_FALSE;
    PetscOptionsGetInt(NULL, NULL, "-tt_loadloc", &p.loadloc, &own);
    PetscOptionsGetInt(NULL, NULL, "-tt_loadsignal", &p.loadsignal, &flg);
    PetscOptionsGetInt(NULL, NULL, "-tt_pulse_steps", &p.pulse_steps, &flg);
    if (own ? (p.loadloc < 0 || p.loadloc > 2) : (p.loadloc != 0 && p.loadloc != 1)) {
        PetscPrintf(PETSC_COMM_WORLD, "# load location %d: the reference's presets are -loadloc 0 and 1 "
                    "(LinearElasticity.cc:627-656); -tt_loadloc 2 selects the y-line at x = xmax, z = zmin\n",
                    p.loadloc);
        return TT_ERR_ARG;
    }
    // the outer solver: the reference's fine-level KSP tolerances (rtol 1e-10, 200 iterations,
    // LinearElasticity.cc:2750-2754), changed by -ksp_rtol / -ksp_max_it as KSPSetFromOptions (:2794) does
    p.rtol = 1e-10;
    p.maxit = 200;
    PetscOptionsGetReal(NULL, NULL, "-ksp_rtol", &p.rtol, &flg);
    PetscOptionsGetInt(NULL, NULL, "-ksp_max_it", &p.maxit, &flg);
    // options with no reference counterpart: precision, state slots, smoothing steps, filter solve
    PetscOptionsGetInt(NULL, NULL, "-precision", &p.precision, &flg);
    PetscOptionsGetInt(NULL, NULL, "-slots", &p.checkpoint_slots, &flg);
    PetscOptionsGetInt(NULL, NULL, "-cheb_degree", &p.cheb_degree, &flg);
    PetscOptionsGetInt(NULL, NULL, "-mg_coarse_pre", &p.mg_coarse_pre, &flg);
    PetscOptionsGetInt(NULL, NULL, "-mg_coarse_post", &p.mg_coarse_post, &flg);
    PetscOptionsGetReal(NULL, NULL, "-filter_rtol", &p.filter_rtol, &flg);
    *out = p;
    return 0;
}

LinearElasticity::LinearElasticity(DM da_nodes, DM, PetscScalar nu_in, PetscInt nlvls_in, PetscInt ncp_in,
                                   PetscInt tsc_in, PetscScalar t1_in, PetscInt ti_in, PetscInt objective_in,
                                   PetscInt sInt_in, PetscInt mInt_in, PetscScalar t0_in, PetscInt boundaries_in,
                                   PetscInt loadloc_in, PetscBool reduction_in, PetscInt basis_in,
                                   PetscScalar dalpha_in, PetscScalar dbeta_in, PetscScalar omega0_in,
                                   PetscInt optimizationRegion, PetscScalar) {
    da_nodal = da_nodes;
    ctx = nullptr;
    forward_done = false;
    for (auto& s : state) s = nullptr;
    for (auto& v : mat) v = std::nan("");
    Ts = ncp_in * tsc_in + 1;  // LinearElasticity.cc:96
    ncp = ncp_in;
    reduced = reduction_in;
    tt_params p;
    if (LinearElasticityParams(da_nodes, nu_in, nlvls_in, ncp_in, tsc_in, t1_in, ti_in, objective_in, sInt_in,
                               mInt_in, t0_in, boundaries_in, loadloc_in, reduction_in, basis_in, dalpha_in,
                               dbeta_in, omega0_in, optimizationRegion, &p) != 0)
        std::exit(1);
    if (report(nullptr, tt_create(&p, &ctx), "tt_create") != TT_OK) std::exit(1);
    if (!da_nodes->mesh) da_nodes->mesh = std::make_shared<TTMesh>();
    if (da_nodes->mesh->ctx) tt_destroy(da_nodes->mesh->ctx);  // a filter-only context is superseded
    da_nodes->mesh->ctx = ctx;
    g_current = ctx;
    // restart point of the solves (SetUpSolver1, LinearElasticity.cc:2679-2723): with -restart (default
    // on) and without -onlyLoadDesign, DDU0, LT and DDL0 are read from -restartFileVecSol and become the
    // starting guesses of the initial-acceleration, terminal-adjoint and t = 0 adjoint solves
    restart = PETSC_TRUE;
    flip = PETSC_TRUE;
    PetscBool flg, onlyDesign = PETSC_FALSE, solutionAsU0 = PETSC_FALSE;
    PetscOptionsGetBool(NULL, NULL, "-restart", &restart, &flg);
    PetscOptionsGetBool(NULL, NULL, "-onlyLoadDesign", &onlyDesign, &flg);
    PetscOptionsGetBool(NULL, NULL, "-solutionAsU0", &solutionAsU0, &flg);
    if (onlyDesign && solutionAsU0) {
        // LoadSolutionAsU0 (LinearElasticity.cc:2621-2676): the first vector of -restartFileVecSol becomes
        // the initial displacement, masked to the free dofs (:584)
        const std::string fileU0 = option_string("-restartFileVecSol", "");
        PetscPrintf(PETSC_COMM_WORLD, "# Loaded solution from %s as initial displacements U0 (-restartFileVecSol):  \n",
                    fileU0.c_str());
        if (!file_exists(fileU0)) {
            PetscPrintf(PETSC_COMM_WORLD, "# File: %s NOT FOUND \n", fileU0.c_str());
        } else {
            const long nd = 3 * tt_num_nodes(ctx);
            std::vector<double> u0(nd);
            double* v[1] = {u0.data()};
            const long lens[1] = {nd};
            if (report(nullptr, tt_petsc_read_vecs(fileU0.c_str(), 1, v, lens), "tt_petsc_read_vecs") != TT_OK ||
                report(ctx, tt_set_initial_displacement(ctx, u0.data(), 0), "tt_set_initial_displacement") != TT_OK)
                std::exit(1);
        }
    }
    const std::string wd = option_string("-workdir", "./");
    filename00 = wd + "/RestartSol00.dat";
    filename01 = wd + "/RestartSol01.dat";
    cpdir = wd + "/";
    if (restart && !onlyDesign && reduced) {
        // the reduced systems are solved directly (SetUpSolver2, :3007-3014): no starting guesses to read
        PetscPrintf(PETSC_COMM_WORLD, "# Reduced-order model: -restartFileVecSol is not read (direct reduced solves)\n");
    } else if (restart && !onlyDesign) {
        const std::string fileVec = option_string("-restartFileVecSol", "");
        PetscPrintf(PETSC_COMM_WORLD, "# Restarting with solution (State Vector) from (-restartFileVecSol): %s \n",
                    fileVec.c_str());
        if (!file_exists(fileVec)) {
            PetscPrintf(PETSC_COMM_WORLD, "File: %s NOT FOUND \n", fileVec.c_str());
        } else {
            const long nd = 3 * tt_num_nodes(ctx);
            std::vector<double> a0(nd), lt(nd), dl0(nd);
            double* v[3] = {a0.data(), lt.data(), dl0.data()};
            const long lens[3] = {nd, nd, nd};
            if (report(nullptr, tt_petsc_read_vecs(fileVec.c_str(), 3, v, lens), "tt_petsc_read_vecs") != TT_OK ||
                report(ctx, tt_set_restart_vectors(ctx, a0.data(), lt.data(), dl0.data()), "tt_set_restart_vectors") !=
                    TT_OK)
                std::exit(1);
        }
    }
}
LinearElasticity::~LinearElasticity() {
    for (auto& s : state)
        if (s) VecDestroy(&s);
}

PetscErrorCode LinearElasticity::SetMaterial(PetscScalar Emin, PetscScalar Emax, PetscScalar Rmin, PetscScalar Rmax,
                                             PetscScalar penalK, PetscScalar penalM, PetscScalar dalpha,
                                             PetscScalar dbeta) {
    const double m[8] = {Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta};
    if (std::memcmp(m, mat, sizeof(m)) != 0) {
        TTCALL(ctx, tt_set_material(ctx, Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta));
        std::memcpy(mat, m, sizeof(m));
    }
    return 0;
}

// LinearElasticity.cc:956-1206
PetscErrorCode LinearElasticity::CompleteInitialState(Vec xPhys, PetscScalar Emin, PetscScalar Emax, PetscScalar Rmin,
                                                      PetscScalar Rmax, PetscScalar penalK, PetscScalar penalM,
                                                      PetscScalar dalpha, PetscScalar dbeta, PetscBool write,
                                                      MPIIO* output) {
    PetscErrorCode ierr = SetMaterial(Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta);
    if (ierr) return ierr;
    TTCALL(ctx, tt_set_density(ctx, xPhys->a.data(), 0));
    TTCALL(ctx, tt_complete_initial_state(ctx));
    forward_done = false;
    if (!reduced) {  // the checkpoint file of the initial state (:1128-1137)
        ierr = WriteCheckpoint(0, 0);
        if (ierr) return ierr;
    }
    if (write && output) {  // the initial state (:1086-1113): u0 (0 unless -solutionAsU0 gave one)
        Vec U0;
        VecCreateSeq(0, 3 * (PetscInt)tt_num_nodes(ctx), &U0);
        TTCALL(ctx, tt_get_state(ctx, 0, U0->a.data(), nullptr, nullptr, 0));
        ierr = output->WriteVTK(da_nodal, U0, xPhys, 0);
        VecDestroy(&U0);
        if (ierr) return ierr;
    }
    return 0;
}

// displacements of steps 1 .. Ts-1 to the field output (LinearElasticity.cc:1248-1250, 1737-1742);
// states the checkpoints did not keep are recomputed in order, one Newmark step each
PetscErrorCode LinearElasticity::WriteStates(MPIIO* output, Vec xPhys) {
    Vec U;
    VecCreateSeq(0, 3 * (PetscInt)tt_num_nodes(ctx), &U);
    PetscErrorCode ierr = 0;
    for (PetscInt t = 1; t < Ts && !ierr; t++) {
        ierr = report(ctx, tt_get_state(ctx, t, U->a.data(), nullptr, nullptr, 0), "tt_get_state");
        if (!ierr) ierr = output->WriteVTK(da_nodal, U, xPhys, t);
    }
    VecDestroy(&U);
    return ierr;
}

// cp%04d.dat (LinearElasticity.cc:1128-1137, 1261-1276): U, DU, DDU of step t as PETSc binary vectors in
// the work directory
std::string LinearElasticity::CheckpointFile(PetscInt c) const {
    char name[32];
    std::snprintf(name, sizeof(name), "cp%04d.dat", (unsigned short)c);
    return cpdir + name;
}
PetscErrorCode LinearElasticity::WriteCheckpoint(PetscInt c, PetscInt t) {
    const long nd = 3 * tt_num_nodes(ctx);
    std::vector<double> u(nd), v(nd), a(nd);
    TTCALL(ctx, tt_get_state(ctx, t, u.data(), v.data(), a.data(), 0));
    const double* vs[3] = {u.data(), v.data(), a.data()};
    const long lens[3] = {nd, nd, nd};
    const std::string f = CheckpointFile(c);
    return report(nullptr, tt_petsc_write_vecs(f.c_str(), 3, vs, lens), f.c_str());
}

// LinearElasticity.cc:1208-1280.  The sweep keeps its binomial checkpoints in device memory; as in the
// reference, the state at the end of each of the ncp blocks is written to cp%04d.dat (block c + 1), and
// with ncp > 1 the states are written to the field output as the sweep goes (:1222-1251).
PetscErrorCode LinearElasticity::GenerateCheckPoints(PetscBool write, Vec xPhys, MPIIO* output) {
    TTCALL(ctx, tt_forward(ctx));
    forward_done = true;
    if (!reduced) {
        const PetscInt tsc = (Ts - 1) / ncp;
        for (PetscInt c = 0; c < ncp; c++) {
            PetscErrorCode ierr = WriteCheckpoint(c + 1, tsc * (c + 1));
            if (ierr) return ierr;
        }
    }
    if (write && ncp > 1 && output) return WriteStates(output, xPhys);
    return 0;
}

// LinearElasticity.cc:1625-1908; dfdx is overwritten (the reference zeroes it first, :1649)
PetscErrorCode LinearElasticity::ComputeObjectiveSensitivities(PetscScalar* fx, Vec dfdx, Vec xPhys,
                                                               PetscScalar Emin, PetscScalar Emax, PetscScalar Rmin,
                                                               PetscScalar Rmax, PetscScalar penalK,
                                                               PetscScalar penalM, PetscScalar dalpha,
                                                               PetscScalar dbeta, PetscBool write, MPIIO* output) {
    const double m[8] = {Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta};
    if (std::memcmp(m, mat, sizeof(m)) != 0) {  // new material: redo the initial state as the reference would
        PetscErrorCode ierr = CompleteInitialState(xPhys, Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta,
                                                   PETSC_FALSE, nullptr);
        if (ierr) return ierr;
    }
    if (ncp > 1 && !reduced) {
        // as the reference (:1651-1680): the block start states are read back from cp%04d.dat, blocks
        // c = ncp-1 .. 0, and the gradient recomputes each block from its file state
        const long nd = 3 * tt_num_nodes(ctx);
        const PetscInt tsc = (Ts - 1) / ncp;
        std::vector<double> u(nd), v(nd), a(nd);
        double* vs[3] = {u.data(), v.data(), a.data()};
        const long lens[3] = {nd, nd, nd};
        for (PetscInt c = ncp - 1; c >= 0; c--) {
            const std::string f = CheckpointFile(c);
            if (report(nullptr, tt_petsc_read_vecs(f.c_str(), 3, vs, lens), f.c_str()) != TT_OK) return TT_ERR_ARG;
            TTCALL(ctx, tt_set_checkpoint(ctx, tsc * c, u.data(), v.data(), a.data(), 0));
        }
        forward_done = true;
    } else if (!forward_done) {
        TTCALL(ctx, tt_forward(ctx));
        forward_done = true;
    }
    if (write && ncp == 1 && output) {  // :1736-1742
        PetscErrorCode ierr = WriteStates(output, xPhys);
        if (ierr) return ierr;
    }
    TTCALL(ctx, tt_objective_gradient(ctx, fx, dfdx->a.data(), 0));
    return 0;
}

// LinearElasticity.cc:1283-1296
PetscErrorCode LinearElasticity::VolumeConstraint(PetscScalar* gx, Vec dgdx, Vec xPhys, PetscScalar volfrac) {
    TTCALL(ctx, tt_volume_constraint(ctx, xPhys->a.data(), volfrac, gx, dgdx->a.data(), 0));
    return 0;
}
// LinearElasticity.cc:2457-2496: DDU0, LT, DDL0 (PETSc binary), alternating between the 00 and 01 files
PetscErrorCode LinearElasticity::WriteRestartFiles() {
    if (!restart) return -1;
    if (reduced) {
        // the reference writes the reduced DDU0, LT, DDL0 here, starting guesses of direct solves that
        // need none, so no solution restart file is written with the reduced-order model
        PetscPrintf(PETSC_COMM_WORLD, "# Reduced-order model: no solution restart file\n");
        return 0;
    }
    flip = flip ? PETSC_FALSE : PETSC_TRUE;
    const long nd = 3 * tt_num_nodes(ctx);
    std::vector<double> a0(nd), lt(nd), dl0(nd);
    TTCALL(ctx, tt_get_restart_vectors(ctx, a0.data(), lt.data(), dl0.data()));
    const double* v[3] = {a0.data(), lt.data(), dl0.data()};
    const long lens[3] = {nd, nd, nd};
    const std::string& f = flip ? filename01 : filename00;
    if (report(nullptr, tt_petsc_write_vecs(f.c_str(), 3, v, lens), "tt_petsc_write_vecs") != TT_OK) return TT_ERR_ARG;
    PetscPrintf(PETSC_COMM_WORLD, flip ? "# Restart file written to 01-files.\n" : "# Restart file written to 00-files.\n");
    return 0;
}

Vec LinearElasticity::StateField(PetscInt j, int which) {
    const long nd = 3 * tt_num_nodes(ctx);
    if (!state[which]) VecCreateSeq(0, (PetscInt)nd, &state[which]);
    double* p[3] = {nullptr, nullptr, nullptr};
    p[which] = state[which]->a.data();
    if (report(ctx, tt_get_state(ctx, j, p[0], p[1], p[2], 0), "tt_get_state") != TT_OK) return nullptr;
    return state[which];
}
Vec LinearElasticity::GetStateFieldU(PetscInt j) { return StateField(j, 0); }
Vec LinearElasticity::GetStateFieldDU(PetscInt j) { return StateField(j, 1); }
Vec LinearElasticity::GetStateFieldDDU(PetscInt j) { return StateField(j, 2); }

// ------------------------------------------------------------------ PDEFilt (PDEFilter.cc)
PDEFilt::PDEFilt(DM da_nodes, PetscScalar rmin) : R(rmin) {
    ctx = mesh_context(da_nodes);
    if (report(ctx, tt_set_filter_radius(ctx, rmin), "tt_set_filter_radius") != TT_OK) std::exit(1);
}
PDEFilt::~PDEFilt() {}
PetscErrorCode PDEFilt::FilterProject(Vec XX, Vec F) {
    TTCALL(ctx, tt_pde_filter(ctx, XX->a.data(), F->a.data(), 0));
    return 0;
}
PetscErrorCode PDEFilt::Gradients(Vec OS, Vec FS) {  // self-adjoint (PDEFilter.cc:219-224)
    TTCALL(ctx, tt_pde_filter(ctx, OS->a.data(), FS->a.data(), 0));
    return 0;
}

// ------------------------------------------------------------------ Filter (Filter.cc)
Filter::Filter(DM da_nodes, Vec, PetscInt filterT, PetscScalar Rin) : filterType(filterT), R(Rin) {
    if (filterT != 2) {
        std::fprintf(stderr, "[topopt] only the PDE filter (-filter 2) is supported (got %d)\n", filterT);
        std::exit(1);
    }
    pdef = new PDEFilt(da_nodes, Rin);
    ctx = mesh_context(da_nodes);
}
Filter::~Filter() { delete pdef; }

// Filter.cc:60-119: PDE filter, clamp to [0, 1], optional smooth Heaviside projection
PetscErrorCode Filter::FilterProject(Vec x, Vec xTilde, Vec xPhys, PetscBool projectionFilter, PetscScalar beta,
                                     PetscScalar eta) {
    TTCALL(ctx, tt_filter_project(ctx, x->a.data(), xTilde->a.data(), xPhys->a.data(), projectionFilter ? 1 : 0,
                                  beta, eta, 0));
    return 0;
}

// Filter.cc:122-206: projection chain rule, then the PDE filter transpose, for dfdx and each dgdx
PetscErrorCode Filter::Gradients(Vec, Vec xTilde, Vec dfdx, PetscInt m, Vec* dgdx, PetscBool projectionFilter,
                                 PetscScalar beta, PetscScalar eta) {
    if (m < 1) return TT_ERR_ARG;
    TTCALL(ctx, tt_filter_gradients(ctx, xTilde->a.data(), dfdx->a.data(), dgdx[0]->a.data(),
                                    projectionFilter ? 1 : 0, beta, eta, 0));
    for (PetscInt i = 1; i < m; i++) {
        std::vector<double> scratch(dgdx[i]->a.size(), 0.0);
        TTCALL(ctx, tt_filter_gradients(ctx, xTilde->a.data(), scratch.data(), dgdx[i]->a.data(),
                                        projectionFilter ? 1 : 0, beta, eta, 0));
    }
    return 0;
}

// Filter.cc:274-300
PetscBool Filter::IncreaseBeta(PetscReal* beta, PetscReal betaFinal, PetscScalar gx, PetscInt itr, PetscReal ch) {
    PetscBool change = PETSC_FALSE;
    if ((ch < 0.01 || itr % 10 == 0) && beta[0] < betaFinal && gx <= 0.0) {
        change = PETSC_TRUE;
        beta[0] = beta[0] < 7 ? beta[0] + 1 : beta[0] * 1.2;
        if (beta[0] > betaFinal) {
            beta[0] = betaFinal;
            change = PETSC_FALSE;
        }
        PetscPrintf(PETSC_COMM_WORLD, "# Beta has been increased to: %f\n", beta[0]);
    }
    return change;
}

// Filter.cc:208-227
PetscScalar Filter::GetMND(Vec x) {
    double mnd = 0.0;
    if (report(ctx, tt_mnd(ctx, x->a.data(), &mnd, 0), "tt_mnd") != TT_OK) return std::nan("");
    return mnd;
}

// ------------------------------------------------------------------ MMA (MMA.cc)
void MMA::Create(PetscScalar a, PetscScalar c, PetscScalar d) {
    if (m != 1) {
        std::fprintf(stderr, "[topopt] the GPU MMA handles one constraint (m = %d requested)\n", m);
        std::exit(1);
    }
    if (!g_current || n != tt_num_elements(g_current)) {
        std::fprintf(stderr, "[topopt] MMA needs the mesh context of %d design variables\n", n);
        std::exit(1);
    }
    if (report(g_current, tt_mma_create(g_current, a, c, d, &h), "tt_mma_create") != TT_OK) std::exit(1);
}
MMA::MMA(PetscInt nn, PetscInt mm, Vec) : n(nn), m(mm), h(nullptr) { Create(0.0, 1000.0, 0.0); }
MMA::MMA(PetscInt nn, PetscInt mm, Vec, PetscScalar* a, PetscScalar* c, PetscScalar* d) : n(nn), m(mm), h(nullptr) {
    Create(a[0], c[0], d[0]);
}
// restart constructors (MMA.cc:6-177): the iteration counter, the two previous iterates, the asymptotes
MMA::MMA(PetscInt nn, PetscInt mm, PetscInt kk, Vec xo1, Vec xo2, Vec U, Vec L, PetscScalar* a, PetscScalar* c,
         PetscScalar* d)
    : n(nn), m(mm), h(nullptr) {
    Create(a[0], c[0], d[0]);
    if (kk < 3) PetscPrintf(PETSC_COMM_WORLD, "NOT A LEGAL RESTART POINT (k<3): EXPECT BREAKDOWN\n");
    if (report(g_current, tt_mma_set_state(h, (int)kk, xo1->a.data(), xo2->a.data(), U->a.data(), L->a.data()),
               "tt_mma_set_state") != TT_OK)
        std::exit(1);
}
MMA::MMA(PetscInt nn, PetscInt mm, PetscInt kk, Vec xo1, Vec xo2, Vec U, Vec L) : n(nn), m(mm), h(nullptr) {
    Create(0.0, 1000.0, 0.0);
    if (kk < 3) PetscPrintf(PETSC_COMM_WORLD, "NOT A LEGAL RESTART POINT (k<3): EXPECT BREAKDOWN\n");
    if (report(g_current, tt_mma_set_state(h, (int)kk, xo1->a.data(), xo2->a.data(), U->a.data(), L->a.data()),
               "tt_mma_set_state") != TT_OK)
        std::exit(1);
}
MMA::~MMA() {
    if (h) tt_mma_destroy(h);
}
PetscErrorCode MMA::Update(Vec xval, Vec dfdx, PetscScalar* gx, Vec* dgdx, Vec xmin, Vec xmax) {
    TTCALL(g_current, tt_mma_update(h, xval->a.data(), dfdx->a.data(), gx, dgdx[0]->a.data(), xmin->a.data(),
                                    xmax->a.data(), 0));
    return 0;
}
// MMA.cc:307-350: copies of the previous iterates and the asymptotes
PetscErrorCode MMA::Restart(Vec xo1, Vec xo2, Vec U, Vec L) {
    int k = 0;
    TTCALL(g_current, tt_mma_get_state(h, &k, xo1->a.data(), xo2->a.data(), U->a.data(), L->a.data()));
    return 0;
}
PetscErrorCode MMA::SetAsymptotes(PetscScalar init, PetscScalar decrease, PetscScalar increase) {
    TTCALL(g_current, tt_mma_set_asymptotes(h, init, decrease, increase));
    return 0;
}
PetscErrorCode MMA::ConstraintModification(PetscBool conMod) { return conMod ? TT_ERR_ARG : 0; }
PetscErrorCode MMA::SetRobustAsymptotesType(PetscInt val) { return val == 0 ? 0 : TT_ERR_ARG; }
PetscErrorCode MMA::SetOuterMovelimit(PetscScalar Xmin, PetscScalar Xmax, PetscScalar movelim, Vec x, Vec xmin,
                                      Vec xmax) {
    TTCALL(g_current, tt_mma_set_outer_movelimit(h, Xmin, Xmax, movelim, x->a.data(), xmin->a.data(),
                                                 xmax->a.data(), 0));
    return 0;
}
PetscErrorCode MMA::KKTresidual(Vec xval, Vec dfdx, PetscScalar* gx, Vec* dgdx, Vec xmin, Vec xmax,
                                PetscScalar* norm2, PetscScalar* normInf) {
    TTCALL(g_current, tt_mma_kkt_residual(h, xval->a.data(), dfdx->a.data(), gx, dgdx[0]->a.data(), xmin->a.data(),
                                          xmax->a.data(), norm2, normInf, 0));
    return 0;
}
PetscScalar MMA::DesignChange(Vec x, Vec xold) {
    double ch = 0.0;
    if (report(g_current, tt_mma_design_change(h, x->a.data(), xold->a.data(), &ch, 0), "tt_mma_design_change") !=
        TT_OK)
        return std::nan("");
    return ch;
}
