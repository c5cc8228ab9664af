/*
 * FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (see topopt_oracle.h).
 *
 * Restates the reference's arithmetic on natural ordering: element e =
 * i + nelx*(j + nely*k), node n = i + nx*(j + ny*k), dof 3n+c, element-local
 * corners (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k) then the same at k+1
 * (LinearElasticity.cc:3067-3074), edof[3a+c] = 3*node_a + c (:1958-1963).
 */
#include "topopt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* geometry                                                             */
/* ------------------------------------------------------------------ */
long or_nn(const or_params* p) { return (long)(p->nelx + 1) * (p->nely + 1) * (p->nelz + 1); }
long or_nel(const or_params* p) { return (long)p->nelx * p->nely * p->nelz; }

static double ddx(const or_params* p) { return (p->xc[1] - p->xc[0]) / p->nelx; }
static double ddy(const or_params* p) { return (p->xc[3] - p->xc[2]) / p->nely; }
static double ddz(const or_params* p) { return (p->xc[5] - p->xc[4]) / p->nelz; }

/* LinearElasticity.cc:3067-3074 (single rank: Xs = Ys = Zs = 0) */
void or_element_nodes(const or_params* p, long e, long nodes[8]) {
    long nx = p->nelx + 1, ny = p->nely + 1;
    long i = e % p->nelx, j = (e / p->nelx) % p->nely, k = e / ((long)p->nelx * p->nely);
    long n0 = i + nx * (j + ny * k);
    nodes[0] = n0;
    nodes[1] = n0 + 1;
    nodes[2] = n0 + 1 + nx;
    nodes[3] = n0 + nx;
    nodes[4] = n0 + nx * ny;
    nodes[5] = n0 + 1 + nx * ny;
    nodes[6] = n0 + 1 + nx + nx * ny;
    nodes[7] = n0 + nx + nx * ny;
}

/* ------------------------------------------------------------------ */
/* element matrices: LinearElasticity.cc:3089-3346, 3358-3503           */
/* ------------------------------------------------------------------ */
static void shape_derivs(double xi, double eta, double zeta, double* dNdxi, double* dNdeta, double* dNdzeta) {
    /* LinearElasticity.cc:3455-3486 */
    static const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int a = 0; a < 8; a++) {
        dNdxi[a] = 0.125 * sx[a] * (1.0 + sy[a] * eta) * (1.0 + sz[a] * zeta);
        dNdeta[a] = 0.125 * sy[a] * (1.0 + sx[a] * xi) * (1.0 + sz[a] * zeta);
        dNdzeta[a] = 0.125 * sz[a] * (1.0 + sx[a] * xi) * (1.0 + sy[a] * eta);
    }
}

static void shape_funcs(double xi, double eta, double zeta, double* S) {
    /* LinearElasticity.cc:3358-3372 */
    static const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int a = 0; a < 8; a++) S[a] = 0.125 * (1.0 + sx[a] * xi) * (1.0 + sy[a] * eta) * (1.0 + sz[a] * zeta);
}

static double dot8(const double* a, const double* b) {
    double r = 0.0;
    for (int i = 0; i < 8; i++) r = r + a[i] * b[i];
    return r;
}

static double inverse3(double J[3][3], double iJ[3][3]) {
    /* LinearElasticity.cc:3488-3503 */
    double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) - J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
    iJ[0][0] = (J[1][1] * J[2][2] - J[2][1] * J[1][2]) / det;
    iJ[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
    iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    iJ[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
    iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    iJ[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
    iJ[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    iJ[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
    iJ[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / det;
    return det;
}

/* Gauss point used by the reference (15 digits, LinearElasticity.cc:3150) */
#define OR_GP 0.577350269189626

static void jacobian(const double* X, const double* Y, const double* Z, const double* dNdxi, const double* dNdeta,
                     const double* dNdzeta, double J[3][3]) {
    J[0][0] = dot8(dNdxi, X);
    J[0][1] = dot8(dNdxi, Y);
    J[0][2] = dot8(dNdxi, Z);
    J[1][0] = dot8(dNdeta, X);
    J[1][1] = dot8(dNdeta, Y);
    J[1][2] = dot8(dNdeta, Z);
    J[2][0] = dot8(dNdzeta, X);
    J[2][1] = dot8(dNdzeta, Y);
    J[2][2] = dot8(dNdzeta, Z);
}

void or_hex8_ke(double dx, double dy, double dz, double nu, double* ke) {
    /* LinearElasticity.cc:3089-3250 (full integration, E = 1) */
    double X[8] = {0.0, dx, dx, 0.0, 0.0, dx, dx, 0.0};
    double Y[8] = {0.0, 0.0, dy, dy, 0.0, 0.0, dy, dy};
    double Z[8] = {0.0, 0.0, 0.0, 0.0, dz, dz, dz, dz};
    double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = 1.0 / (2.0 * (1.0 + nu));
    double C[6][6] = {{lam + 2.0 * mu, lam, lam, 0, 0, 0}, {lam, lam + 2.0 * mu, lam, 0, 0, 0},
                      {lam, lam, lam + 2.0 * mu, 0, 0, 0}, {0, 0, 0, mu, 0, 0},
                      {0, 0, 0, 0, mu, 0},               {0, 0, 0, 0, 0, mu}};
    /* strain rows (xx,yy,zz,xy,yz,zx) fed by d/dx, d/dy, d/dz of displacement components */
    double al[3][6][3];
    memset(al, 0, sizeof(al));
    al[0][0][0] = 1.0; al[0][3][1] = 1.0; al[0][5][2] = 1.0;
    al[1][1][1] = 1.0; al[1][3][0] = 1.0; al[1][4][2] = 1.0;
    al[2][2][2] = 1.0; al[2][4][1] = 1.0; al[2][5][0] = 1.0;
    double GP[2] = {-OR_GP, OR_GP};
    memset(ke, 0, sizeof(double) * 576);
    for (int ii = 0; ii < 2; ii++)
        for (int jj = 0; jj < 2; jj++)
            for (int kk = 0; kk < 2; kk++) {
                double dNdxi[8], dNdeta[8], dNdzeta[8], J[3][3], iJ[3][3];
                shape_derivs(GP[ii], GP[jj], GP[kk], dNdxi, dNdeta, dNdzeta);
                jacobian(X, Y, Z, dNdxi, dNdeta, dNdzeta, J);
                double w = inverse3(J, iJ);
                double B[6][24];
                memset(B, 0, sizeof(B));
                for (int ll = 0; ll < 3; ll++) {
                    const double* dN = (ll == 0) ? dNdxi : (ll == 1) ? dNdeta : dNdzeta;
                    double beta[6][3];
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 3; j++)
                            beta[i][j] = iJ[0][ll] * al[0][i][j] + iJ[1][ll] * al[1][i][j] + iJ[2][ll] * al[2][i][j];
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 24; j++) B[i][j] += beta[i][j % 3] * dN[j / 3];
                }
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++)
                        for (int k = 0; k < 6; k++)
                            for (int l = 0; l < 6; l++) ke[j + 24 * i] += w * (B[k][i] * C[k][l] * B[l][j]);
            }
}

void or_hex8_me(double dx, double dy, double dz, double* me) {
    /* LinearElasticity.cc:3252-3333 (density 1) */
    double X[8] = {0.0, dx, dx, 0.0, 0.0, dx, dx, 0.0};
    double Y[8] = {0.0, 0.0, dy, dy, 0.0, 0.0, dy, dy};
    double Z[8] = {0.0, 0.0, 0.0, 0.0, dz, dz, dz, dz};
    double GP[2] = {-OR_GP, OR_GP};
    memset(me, 0, sizeof(double) * 576);
    for (int ii = 0; ii < 2; ii++)
        for (int jj = 0; jj < 2; jj++)
            for (int kk = 0; kk < 2; kk++) {
                double S[8], vS[72], dNdxi[8], dNdeta[8], dNdzeta[8], J[3][3], iJ[3][3];
                shape_funcs(GP[ii], GP[jj], GP[kk], S);
                memset(vS, 0, sizeof(vS));
                for (int a = 0; a < 8; a++)
                    for (int c = 0; c < 3; c++) vS[c * 24 + 3 * a + c] = S[a];
                shape_derivs(GP[ii], GP[jj], GP[kk], dNdxi, dNdeta, dNdzeta);
                jacobian(X, Y, Z, dNdxi, dNdeta, dNdzeta, J);
                double w = inverse3(J, iJ);
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++)
                        for (int k = 0; k < 3; k++) me[j + 24 * i] += w * (vS[i + k * 24] * vS[j + 24 * k]);
            }
}

void or_element_matrices(const or_params* p, double* ke, double* me) {
    or_hex8_ke(ddx(p), ddy(p), ddz(p), p->nu, ke);
    or_hex8_me(ddx(p), ddy(p), ddz(p), me);
}

void or_pdefilter_ke(double dx, double dy, double dz, double R, double* kf) {
    /* R^2 * Laplacian + mass on a box element; PDEFilter.cc:485-578 give this
       matrix in closed form.  Two-point Gauss (exact abscissa) is exact here. */
    double X[8] = {0.0, dx, dx, 0.0, 0.0, dx, dx, 0.0};
    double Y[8] = {0.0, 0.0, dy, dy, 0.0, 0.0, dy, dy};
    double Z[8] = {0.0, 0.0, 0.0, 0.0, dz, dz, dz, dz};
    double g = 1.0 / sqrt(3.0);
    double GP[2] = {-g, g};
    memset(kf, 0, sizeof(double) * 64);
    for (int ii = 0; ii < 2; ii++)
        for (int jj = 0; jj < 2; jj++)
            for (int kk = 0; kk < 2; kk++) {
                double S[8], dNdxi[8], dNdeta[8], dNdzeta[8], J[3][3], iJ[3][3];
                shape_funcs(GP[ii], GP[jj], GP[kk], S);
                shape_derivs(GP[ii], GP[jj], GP[kk], dNdxi, dNdeta, dNdzeta);
                jacobian(X, Y, Z, dNdxi, dNdeta, dNdzeta, J);
                double w = inverse3(J, iJ);
                double gx[8], gy[8], gz[8];
                for (int a = 0; a < 8; a++) {
                    gx[a] = iJ[0][0] * dNdxi[a] + iJ[0][1] * dNdeta[a] + iJ[0][2] * dNdzeta[a];
                    gy[a] = iJ[1][0] * dNdxi[a] + iJ[1][1] * dNdeta[a] + iJ[1][2] * dNdzeta[a];
                    gz[a] = iJ[2][0] * dNdxi[a] + iJ[2][1] * dNdeta[a] + iJ[2][2] * dNdzeta[a];
                }
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++)
                        kf[8 * a + b] += w * (R * R * (gx[a] * gx[b] + gy[a] * gy[b] + gz[a] * gz[b]) + S[a] * S[b]);
            }
}

/* ------------------------------------------------------------------ */
/* Newmark constants and interpolation                                  */
/* ------------------------------------------------------------------ */
void or_newmark_consts(int ti, double dt, double a[6]) {
    /* LinearElasticity.cc:185-207 */
    if (ti == 0) {
        a[0] = 0.0; a[1] = 0.0; a[2] = 1.0 / dt; a[3] = 1.0 / dt; a[4] = 0.0; a[5] = 1.0 / (dt * dt);
    } else {
        double b = 0.25, g = 0.5;
        a[0] = 1.0 - g / b;
        a[1] = (1.0 - g / (2.0 * b)) * dt;
        a[2] = g / (b * dt);
        a[3] = 1 / (b * dt);
        a[4] = 1 / (2.0 * b) - 1.0;
        a[5] = 1.0 / (b * dt * dt);
    }
}

void or_interp(int scheme, double vmin, double vmax, double penal, double x, double* f, double* df) {
    /* LinearElasticity.cc:1580-1623 */
    switch (scheme) {
    case 0:
        *f = vmin + x * (vmax - vmin);
        *df = vmax - vmin;
        break;
    case 1:
        *f = vmin + pow(x, penal) * (vmax - vmin);
        *df = penal * pow(x, penal - 1) * (vmax - vmin);
        break;
    default:
        *f = vmin + x / (1 + penal * (1 - x)) * (vmax - vmin);
        *df = (1 + penal) / pow(1 + penal * (1 - x), 2) * (vmax - vmin);
        break;
    }
}

/* ------------------------------------------------------------------ */
/* BC, load, region: LinearElasticity.cc:222-335, 337-591, 593-675       */
/* ------------------------------------------------------------------ */
static double epsi_of(const or_params* p) { return fmin(ddx(p) * 0.05, fmin(ddy(p) * 0.05, ddz(p) * 0.05)); }
/* the physics mesh starts at the origin: xc = {0, ne*dx, 0, ne*dy, 0, ne*dz} (:380-385) */
static void phys_box(const or_params* p, double xc[6]) {
    xc[0] = 0.0; xc[1] = p->nelx * ddx(p);
    xc[2] = 0.0; xc[3] = p->nely * ddy(p);
    xc[4] = 0.0; xc[5] = p->nelz * ddz(p);
}
static void node_coord(const or_params* p, long n, double* x, double* y, double* z) {
    long nx = p->nelx + 1, ny = p->nely + 1;
    long i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
    double xc[6];
    phys_box(p, xc);
    *x = xc[0] + i * (xc[1] - xc[0]) / p->nelx;
    *y = xc[2] + j * (xc[3] - xc[2]) / p->nely;
    *z = xc[4] + k * (xc[5] - xc[4]) / p->nelz;
}

void or_bc_mask(const or_params* p, double* N) {
    long nn = or_nn(p);
    double xc[6], e = epsi_of(p);
    phys_box(p, xc);
    for (long i = 0; i < 3 * nn; i++) N[i] = 1.0;
    for (long n = 0; n < nn; n++) {
        double x, y, z;
        node_coord(p, n, &x, &y, &z);
        if (p->boundaries == 0) {
            if (fabs(x - xc[0]) < e) { N[3 * n] = 0.0; N[3 * n + 1] = 0.0; }
            if (fabs(x - xc[1]) < e && fabs(z - xc[4]) < e) { N[3 * n + 1] = 0.0; N[3 * n + 2] = 0.0; }
        } else {
            if (fabs(x - xc[0]) < e) { N[3 * n] = 0.0; N[3 * n + 1] = 0.0; N[3 * n + 2] = 0.0; }
        }
    }
}

void or_load_location(const or_params* p, double* F) {
    long nn = or_nn(p);
    double xc[6], e = epsi_of(p);
    phys_box(p, xc);
    memset(F, 0, sizeof(double) * 3 * nn);
    for (long n = 0; n < nn; n++) {
        double x, y, z;
        node_coord(p, n, &x, &y, &z);
        int end = fabs(y - xc[2]) < e || fabs(y - xc[3]) < e;
        if (p->loadloc == 0) {
            if (fabs(x - xc[0]) < e && fabs(z - xc[5]) < e) F[3 * n + 2] = end ? 0.5 : 1.0;
        } else if (p->loadloc == 1) {
            if (fabs(x - xc[1]) < e && fabs(z - xc[5] / 2.0) < e) F[3 * n + 2] = end ? 0.5 : 1.0;
        } else {
            /* line load in y on the edge x = xmax, z = zmin (configured cantilever pulse) */
            if (fabs(x - xc[1]) < e && fabs(z - xc[4]) < e) F[3 * n + 1] = end ? 0.5 : 1.0;
        }
    }
}

double or_load_intensity(const or_params* p, int step) {
    double dt = p->t1 / p->nsteps;
    double s;
    if (p->loadsignal == 0) {
        /* LinearElasticity.cc:598-609, evaluated at t = step*dt (:1328) */
        double t = step * dt, t0 = 8.0, c = 1.1;
        s = 1.0 * exp(-pow(t - t0, 2) / pow(2 * c, 2)) * cos(4.0 * (t - t0)) * cos(5.0 * (t - t0)) *
            cos(6.0 * (t - t0)) * cos(7.0 * (t - t0));
    } else {
        s = (step <= p->pulse_steps) ? sin(M_PI * (double)step / (double)p->pulse_steps) : 0.0;
    }
    return s / (double)p->nely; /* LoadIntensity = signal/(nn[1]-1) (:611) */
}

void or_region(const or_params* p, double* P) {
    long nel = or_nel(p);
    double xc[6], dx = ddx(p), dy = ddy(p), dz = ddz(p);
    phys_box(p, xc);
    for (long e = 0; e < nel; e++) {
        long i = e % p->nelx, j = (e / p->nelx) % p->nely, k = e / ((long)p->nelx * p->nely);
        double ex = (i + 1) * dx - dx / 2, ey = (j + 1) * dy - dy / 2, ez = (k + 1) * dz - dz / 2;
        int in = 0;
        switch (p->region) {
        case 0: in = ez > xc[5] - 1 && ex < xc[0] + 1; break;
        case 1: in = ez > xc[5] / 2.0 - 0.5 && ez < xc[5] / 2.0 + 0.5 && ex < xc[1] - 1; break;
        case 2: in = 1; break;
        case 3:
            in = ex > 0.75 * xc[1] - 2.5 && ex < 0.75 * xc[1] + 2.5 && ey > 0.50 * xc[3] - 2.5 &&
                 ey < 0.50 * xc[3] + 2.5 && ez > 0.50 * xc[5] - 2.5 && ez < 0.50 * xc[5] + 2.5;
            break;
        }
        P[e] = in ? 1.0 : 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* context and matrix-free operator                                     */
/* ------------------------------------------------------------------ */
typedef struct {
    const or_params* p;
    long nel, nn, nd;
    double ke[576], me[576];
    double *kf, *dkf, *mf, *dmf, *P;
    double a[6], dt;
} ctx_t;

static void ctx_init(ctx_t* c, const or_params* p, const double* xphys) {
    c->p = p;
    c->nel = or_nel(p);
    c->nn = or_nn(p);
    c->nd = 3 * c->nn;
    or_element_matrices(p, c->ke, c->me);
    c->kf = malloc(sizeof(double) * c->nel);
    c->dkf = malloc(sizeof(double) * c->nel);
    c->mf = malloc(sizeof(double) * c->nel);
    c->dmf = malloc(sizeof(double) * c->nel);
    c->P = malloc(sizeof(double) * c->nel);
    for (long e = 0; e < c->nel; e++) {
        double x = xphys ? xphys[e] : 1.0;
        or_interp(p->sInt, p->Emin, p->Emax, p->penalK, x, &c->kf[e], &c->dkf[e]);
        or_interp(p->mInt, p->Rmin, p->Rmax, p->penalM, x, &c->mf[e], &c->dmf[e]);
    }
    or_region(p, c->P);
    c->dt = p->t1 / p->nsteps;
    or_newmark_consts(p->ti, c->dt, c->a);
}

static void ctx_free(ctx_t* c) {
    free(c->kf); free(c->dkf); free(c->mf); free(c->dmf); free(c->P);
}

static void edofs(const or_params* p, long e, long ed[24]) {
    long nodes[8];
    or_element_nodes(p, e, nodes);
    for (int a = 0; a < 8; a++)
        for (int c = 0; c < 3; c++) ed[3 * a + c] = 3 * nodes[a] + c;
}

/* y = sum_e (cK*kf_e*KE + cM*mf_e*ME) u_e; 8-colour element loop (deterministic) */
static void ctx_apply(const ctx_t* c, double cK, double cM, const double* u, double* y) {
    const or_params* p = c->p;
    memset(y, 0, sizeof(double) * c->nd);
    for (int col = 0; col < 8; col++) {
        int ci = col & 1, cj = (col >> 1) & 1, ck = (col >> 2) & 1;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = ck; k < p->nelz; k += 2)
            for (int j = cj; j < p->nely; j += 2)
                for (int i = ci; i < p->nelx; i += 2) {
                    long e = i + (long)p->nelx * (j + (long)p->nely * k);
                    long ed[24];
                    edofs(p, e, ed);
                    double ue[24];
                    for (int h = 0; h < 24; h++) ue[h] = u[ed[h]];
                    double sk = cK * c->kf[e], sm = cM * c->mf[e];
                    for (int r = 0; r < 24; r++) {
                        double tk = 0.0, tm = 0.0;
                        for (int h = 0; h < 24; h++) {
                            tk += c->ke[24 * r + h] * ue[h];
                            tm += c->me[24 * r + h] * ue[h];
                        }
                        y[ed[r]] += sk * tk + sm * tm;
                    }
                }
    }
}

static void ctx_apply_bc(const ctx_t* c, double cK, double cM, const double* N, const double* u, double* y,
                         double* tmp) {
    for (long i = 0; i < c->nd; i++) tmp[i] = N[i] * u[i];
    ctx_apply(c, cK, cM, tmp, y);
    for (long i = 0; i < c->nd; i++) y[i] = N[i] * y[i] + (1.0 - N[i]) * u[i];
}

void or_apply(const or_params* p, const double* xphys, double cK, double cM, const double* u, double* y) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    ctx_apply(&c, cK, cM, u, y);
    ctx_free(&c);
}

void or_apply_bc(const or_params* p, const double* xphys, double cK, double cM, const double* N, const double* u,
                 double* y) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    double* tmp = malloc(sizeof(double) * c.nd);
    ctx_apply_bc(&c, cK, cM, N, u, y, tmp);
    free(tmp);
    ctx_free(&c);
}

static double vdot(long n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Jacobi-PCG on N.A.N + (I-N); returns iteration count, negative if maxit hit */
static int ctx_pcg(const ctx_t* c, double cK, double cM, const double* N, const double* b, double* x, double rtol,
                   int maxit, double* relres) {
    long nd = c->nd;
    double* r = malloc(sizeof(double) * nd);
    double* z = malloc(sizeof(double) * nd);
    double* pp = malloc(sizeof(double) * nd);
    double* q = malloc(sizeof(double) * nd);
    double* tmp = malloc(sizeof(double) * nd);
    double* dinv = malloc(sizeof(double) * nd);
    /* diagonal of the BC operator */
    memset(dinv, 0, sizeof(double) * nd);
    for (long e = 0; e < c->nel; e++) {
        long ed[24];
        edofs(c->p, e, ed);
        for (int h = 0; h < 24; h++)
            dinv[ed[h]] += cK * c->kf[e] * c->ke[25 * h] + cM * c->mf[e] * c->me[25 * h];
    }
    for (long i = 0; i < nd; i++) dinv[i] = 1.0 / (N[i] * dinv[i] + (1.0 - N[i]));
    double bn = sqrt(vdot(nd, b, b));
    int it = 0, conv = 0;
    double rel = 0.0;
    if (bn == 0.0) {
        memset(x, 0, sizeof(double) * nd);
        if (relres) *relres = 0.0;
        conv = 1;
        goto done;
    }
    while (it < maxit && !conv) {
        ctx_apply_bc(c, cK, cM, N, x, q, tmp);
        for (long i = 0; i < nd; i++) r[i] = b[i] - q[i];
        for (long i = 0; i < nd; i++) z[i] = dinv[i] * r[i];
        memcpy(pp, z, sizeof(double) * nd);
        double rz = vdot(nd, r, z);
        rel = sqrt(vdot(nd, r, r)) / bn;
        if (rel <= rtol) { conv = 1; break; }
        while (it < maxit) {
            it++;
            ctx_apply_bc(c, cK, cM, N, pp, q, tmp);
            double al = rz / vdot(nd, pp, q);
            for (long i = 0; i < nd; i++) { x[i] += al * pp[i]; r[i] -= al * q[i]; }
            rel = sqrt(vdot(nd, r, r)) / bn;
            if (rel <= rtol) break;
            for (long i = 0; i < nd; i++) z[i] = dinv[i] * r[i];
            double rzn = vdot(nd, r, z);
            double be = rzn / rz;
            rz = rzn;
            for (long i = 0; i < nd; i++) pp[i] = z[i] + be * pp[i];
        }
        /* confirm with the true residual; restart if the recursion drifted */
        ctx_apply_bc(c, cK, cM, N, x, q, tmp);
        for (long i = 0; i < nd; i++) r[i] = b[i] - q[i];
        rel = sqrt(vdot(nd, r, r)) / bn;
        if (rel <= rtol) conv = 1;
    }
    if (relres) *relres = rel;
done:
    free(r); free(z); free(pp); free(q); free(tmp); free(dinv);
    return conv ? it : -it - 1;
}

int or_pcg(const or_params* p, const double* xphys, double cK, double cM, const double* N, const double* b,
           double* x, double rtol, int maxit, double* relres) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    int it = ctx_pcg(&c, cK, cM, N, b, x, rtol, maxit, relres);
    ctx_free(&c);
    return it;
}

/* ------------------------------------------------------------------ */
/* forward Newmark: CompleteInitialState (:956-1206) + SolveStateBlock  */
/* (:1298-1397)                                                         */
/* ------------------------------------------------------------------ */
/* y = N.(X(N.w)) with X = sM*M + sK*K, added to acc */
static void add_B(const ctx_t* c, double sK, double sM, const double* N, const double* w, double* acc, double* t1,
                  double* t2) {
    for (long i = 0; i < c->nd; i++) t1[i] = N[i] * w[i];
    ctx_apply(c, sK, sM, t1, t2);
    for (long i = 0; i < c->nd; i++) acc[i] += N[i] * t2[i];
}

/* u0_in: initial displacement (NULL = 0), masked to N.u0 as LoadSolutionAsU0 + :584 leave it */
int or_forward_u0(const or_params* p, const double* xphys, const double* u0_in, double* states) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd;
    const double* a = c.a;
    double al = p->dalpha, be = p->dbeta;
    double* N = malloc(sizeof(double) * nd);
    double* F = malloc(sizeof(double) * nd);
    double* R = malloc(sizeof(double) * nd);
    double* t1 = malloc(sizeof(double) * nd);
    double* t2 = malloc(sizeof(double) * nd);
    or_bc_mask(p, N);
    or_load_location(p, F);
    int status = 0;
    double* u0 = states;
    double* v0 = states + nd;
    double* a0 = states + 2 * nd;
    memset(states, 0, sizeof(double) * 3 * nd); /* U0 = DU0 = 0 (:521-525) */
    if (u0_in)
        for (long i = 0; i < nd; i++) u0[i] = N[i] * u0_in[i]; /* -solutionAsU0 (:523, 584, 2621-2676) */
    /* R = N.(f(0) - N.C.N v0 - N.K.N u0); M~ a0 = R  (:1050-1078) */
    double s0 = or_load_intensity(p, 0);
    for (long i = 0; i < nd; i++) R[i] = F[i] * s0;
    for (long i = 0; i < nd; i++) t1[i] = N[i] * v0[i];
    ctx_apply(&c, be, al, t1, t2); /* C v0 */
    for (long i = 0; i < nd; i++) R[i] -= N[i] * t2[i];
    for (long i = 0; i < nd; i++) t1[i] = N[i] * u0[i];
    ctx_apply(&c, 1.0, 0.0, t1, t2); /* K u0 */
    for (long i = 0; i < nd; i++) R[i] -= N[i] * t2[i];
    for (long i = 0; i < nd; i++) R[i] *= N[i];
    if (ctx_pcg(&c, 0.0, 1.0, N, R, a0, p->rtol, p->maxit, NULL) < 0) status = -1;
    /* time stepping; system (K + a2 C + a5 M) on free dofs */
    double cK = 1.0 + a[2] * be, cM = a[5] + a[2] * al;
    for (int t = 1; t <= p->nsteps; t++) {
        double* up = states + (size_t)(t - 1) * 3 * nd;
        double* vp = up + nd;
        double* ap = up + 2 * nd;
        double* un = states + (size_t)t * 3 * nd;
        double* vn = un + nd;
        double* an = un + 2 * nd;
        double s = or_load_intensity(p, t);
        for (long i = 0; i < nd; i++) R[i] = F[i] * s;
        add_B(&c, a[2] * be, a[5] + a[2] * al, N, up, R, t1, t2);   /* B1 = a5 M + a2 C */
        add_B(&c, -a[0] * be, a[3] - a[0] * al, N, vp, R, t1, t2);  /* B2 = a3 M - a0 C */
        add_B(&c, -a[1] * be, a[4] - a[1] * al, N, ap, R, t1, t2);  /* B3 = a4 M - a1 C */
        for (long i = 0; i < nd; i++) R[i] *= N[i];
        memcpy(un, up, sizeof(double) * nd); /* initial guess u_n (:1358) */
        if (ctx_pcg(&c, cK, cM, N, R, un, p->rtol, p->maxit, NULL) < 0) status = -1;
        for (long i = 0; i < nd; i++) {
            double du = un[i] - up[i];
            double v = 0.0;
            v += a[0] * vp[i];
            v += a[1] * ap[i];
            v += a[2] * du;
            double acc = 0.0;
            acc += -a[3] * vp[i];
            acc += -a[4] * ap[i];
            acc += a[5] * du;
            vn[i] = v;
            an[i] = acc;
        }
    }
    free(N); free(F); free(R); free(t1); free(t2);
    ctx_free(&c);
    return status;
}

int or_forward(const or_params* p, const double* xphys, double* states) {
    return or_forward_u0(p, xphys, NULL, states);
}

/* ------------------------------------------------------------------ */
/* objective, objective gradient wrt state, sensitivities               */
/* ------------------------------------------------------------------ */
typedef struct {
    double uKu, duMdu, duCdu_interp, DxduCdu, uu, dudu, dduddu;
} elem_terms;

static void element_terms(const ctx_t* c, long e, const long* ed, const double* u, const double* du,
                          const double* ddu, elem_terms* t) {
    const or_params* p = c->p;
    memset(t, 0, sizeof(*t));
    double kf = c->kf[e], mf = c->mf[e], dkf = c->dkf[e], dmf = c->dmf[e];
    for (int k = 0; k < 24; k++) {
        for (int h = 0; h < 24; h++) {
            double KE = c->ke[24 * k + h], ME = c->me[24 * k + h];
            t->uKu += u[ed[k]] * KE * u[ed[h]];
            t->duMdu += du[ed[k]] * ME * du[ed[h]];
            t->duCdu_interp += du[ed[k]] * (p->dalpha * mf * ME + p->dbeta * kf * KE) * du[ed[h]];
            t->DxduCdu += du[ed[k]] * (p->dalpha * dmf * ME + p->dbeta * dkf * KE) * du[ed[h]];
        }
        t->uu += u[ed[k]] * u[ed[k]];
        t->dudu += du[ed[k]] * du[ed[k]];
        t->dduddu += ddu[ed[k]] * ddu[ed[k]];
    }
}

/* EvaluateObjective (:1416-1465) */
static double h0_of(const ctx_t* c, long e, const elem_terms* t) {
    double dt = c->dt, kf = c->kf[e], mf = c->mf[e];
    switch (c->p->objective) {
    case 0: return dt * 0.5 * t->uu;
    case 1: return dt * 0.5 * t->dudu;
    case 2: return dt * 0.5 * t->dduddu;
    case 3: return dt * 0.5 * (kf * t->uKu);
    case 4: return dt * 0.5 * (mf * t->duMdu);
    case 5: return dt * 0.5 * (kf * t->uKu + mf * t->duMdu);
    default: return dt * 0.5 * t->duCdu_interp;
    }
}

/* EvaluateObjectiveGradient (:1469-1514) */
static double dfx0_of(const ctx_t* c, long e, const elem_terms* t) {
    double dt = c->dt, dkf = c->dkf[e], dmf = c->dmf[e];
    switch (c->p->objective) {
    case 3: return dt * 0.5 * (dkf * t->uKu);
    case 4: return dt * 0.5 * (dmf * t->duMdu);
    case 5: return dt * 0.5 * (dkf * t->uKu + dmf * t->duMdu);
    case 6: return dt * 0.5 * t->DxduCdu;
    default: return 0.0;
    }
}

/* ComputeObjectiveElementLoop (:2359-2452) for one state */
static double objective_state(const ctx_t* c, const double* u, const double* du, const double* ddu) {
    double fx = 0.0;
    for (long e = 0; e < c->nel; e++) {
        long ed[24];
        edofs(c->p, e, ed);
        elem_terms t;
        element_terms(c, e, ed, u, du, ddu, &t);
        fx += h0_of(c, e, &t) * 1.0 * c->P[e]; /* s = 1 (:1400-1414) */
    }
    return fx;
}

/* ComputeGradientObjectiveElementLoop (:2206-2356) + GradObjectiveWrtState (:1516-1576) */
static void grad_obj_state(const ctx_t* c, const double* u, const double* du, const double* ddu, double* hdu,
                           double* hddu, double* hdddu) {
    const or_params* p = c->p;
    double dt = c->dt;
    memset(hdu, 0, sizeof(double) * c->nd);
    memset(hddu, 0, sizeof(double) * c->nd);
    memset(hdddu, 0, sizeof(double) * c->nd);
    for (long e = 0; e < c->nel; e++) {
        long ed[24];
        edofs(p, e, ed);
        double kf = c->kf[e], mf = c->mf[e], ps = c->P[e] * 1.0;
        for (int k = 0; k < 24; k++) {
            double h1 = 0.0, h2 = 0.0, h3 = 0.0;
            switch (p->objective) {
            case 0: h1 = dt * u[ed[k]]; break;
            case 1: h2 = dt * du[ed[k]]; break;
            case 2: h3 = dt * ddu[ed[k]]; break;
            case 3:
                for (int h = 0; h < 24; h++) h1 += dt * kf * c->ke[k * 24 + h] * u[ed[h]];
                break;
            case 4:
                for (int h = 0; h < 24; h++) h2 += dt * mf * c->me[k * 24 + h] * du[ed[h]];
                break;
            case 5:
                for (int h = 0; h < 24; h++) {
                    h1 += dt * kf * c->ke[k * 24 + h] * u[ed[h]];
                    h2 += dt * mf * c->me[k * 24 + h] * du[ed[h]];
                }
                break;
            default:
                for (int h = 0; h < 24; h++)
                    h2 += dt * (p->dalpha * mf * c->me[k * 24 + h] + p->dbeta * kf * c->ke[k * 24 + h]) * du[ed[h]];
                break;
            }
            hdu[ed[k]] += ps * h1;
            hddu[ed[k]] += ps * h2;
            hdddu[ed[k]] += ps * h3;
        }
    }
}

/* ComputeSensitivitiesElementLoopn (:1911-2077) */
static void sens_n(const ctx_t* c, const double* lam, const double* u, const double* du, const double* ddu,
                   const double* ub, const double* dub, const double* ddub, double* df) {
    const or_params* p = c->p;
    const double* a = c->a;
    for (long e = 0; e < c->nel; e++) {
        long ed[24];
        edofs(p, e, ed);
        double dkf = c->dkf[e], dmf = c->dmf[e];
        double t3 = 0.0;
        for (int k = 0; k < 24; k++)
            for (int h = 0; h < 24; h++) {
                double dme = dmf * c->me[k * 24 + h];
                double dke = dkf * c->ke[k * 24 + h];
                double dce = p->dalpha * dme + p->dbeta * dke;
                double dAdx = -dke - a[5] * dme - a[2] * dce;
                double dB1 = a[5] * dme + a[2] * dce;
                double dB2 = a[3] * dme - a[0] * dce;
                double dB3 = a[4] * dme - a[1] * dce;
                double dBu = dB1 * ub[ed[h]] + dB2 * dub[ed[h]] + dB3 * ddub[ed[h]];
                t3 += lam[ed[k]] * (dAdx * u[ed[h]] + dBu);
            }
        elem_terms t;
        element_terms(c, e, ed, u, du, ddu, &t);
        double dfx = c->P[e] * (h0_of(c, e, &t) * 0.0 + dfx0_of(c, e, &t) * 1.0);
        df[e] += dfx + 0.0 + t3;
    }
}

/* ComputeSensitivitiesElementLoop0 (:2080-2204) */
static void sens_0(const ctx_t* c, const double* nu0, const double* u, const double* du, const double* ddu,
                   double* df) {
    const or_params* p = c->p;
    for (long e = 0; e < c->nel; e++) {
        long ed[24];
        edofs(p, e, ed);
        double dkf = c->dkf[e], dmf = c->dmf[e];
        double t2 = 0.0;
        for (int k = 0; k < 24; k++)
            for (int h = 0; h < 24; h++) {
                double dme = dmf * c->me[k * 24 + h];
                double dke = dkf * c->ke[k * 24 + h];
                double dce = p->dalpha * dme + p->dbeta * dke;
                t2 += nu0[ed[k]] * dke * u[ed[h]] + nu0[ed[k]] * dce * du[ed[h]] + nu0[ed[k]] * dme * ddu[ed[h]];
            }
        elem_terms t;
        element_terms(c, e, ed, u, du, ddu, &t);
        double dfx = c->P[e] * (h0_of(c, e, &t) * 0.0 + dfx0_of(c, e, &t) * 1.0);
        df[e] += dfx + t2 + 0.0;
    }
}

double or_objective(const or_params* p, const double* xphys, const double* states) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd;
    double fx = 0.0;
    for (int t = 1; t <= p->nsteps; t++) {
        const double* s = states + (size_t)t * 3 * nd;
        fx += objective_state(&c, s, s + nd, s + 2 * nd);
    }
    fx += objective_state(&c, states, states + nd, states + 2 * nd);
    ctx_free(&c);
    return fx;
}

int or_objective_gradient(const or_params* p, const double* xphys, const double* states, double* fx, double* dfdx) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd, nel = c.nel;
    const double* a = c.a;
    double al = p->dalpha, be = p->dbeta;
    int T = p->nsteps, status = 0;
    double* N = malloc(sizeof(double) * nd);
    or_bc_mask(p, N);
    double* lam = calloc((size_t)(T + 1) * nd, sizeof(double)); /* lambda_t, t = 1..T */
    double* mu = malloc(sizeof(double) * nd);
    double* nu = malloc(sizeof(double) * nd);
    double* mun = calloc(nd, sizeof(double));
    double* nun = calloc(nd, sizeof(double));
    double* hdu = malloc(sizeof(double) * nd);
    double* hddu = malloc(sizeof(double) * nd);
    double* hdddu = malloc(sizeof(double) * nd);
    double* R1 = malloc(sizeof(double) * nd);
    double* B1l = malloc(sizeof(double) * nd);
    double* B2l = malloc(sizeof(double) * nd);
    double* B3l = malloc(sizeof(double) * nd);
    double* t1 = malloc(sizeof(double) * nd);
    double* t2 = malloc(sizeof(double) * nd);
    double cK = 1.0 + a[2] * be, cM = a[5] + a[2] * al;
    /* SolveLagrangen (:808-954), t = T .. 1 */
    for (int t = T; t >= 1; t--) {
        const double* s = states + (size_t)t * 3 * nd;
        double* lt = lam + (size_t)t * nd;
        grad_obj_state(&c, s, s + nd, s + 2 * nd, hdu, hddu, hdddu);
        if (t == T) {
            for (long i = 0; i < nd; i++) {
                mu[i] = -hddu[i];
                nu[i] = -hdddu[i];
                double r = hdu[i];
                r += -a[2] * mu[i];
                r += -a[5] * nu[i];
                R1[i] = -r;
            }
        } else {
            const double* ln = lam + (size_t)(t + 1) * nd;
            memset(B1l, 0, sizeof(double) * nd);
            memset(B2l, 0, sizeof(double) * nd);
            memset(B3l, 0, sizeof(double) * nd);
            add_B(&c, a[2] * be, a[5] + a[2] * al, N, ln, B1l, t1, t2);
            add_B(&c, -a[0] * be, a[3] - a[0] * al, N, ln, B2l, t1, t2);
            add_B(&c, -a[1] * be, a[4] - a[1] * al, N, ln, B3l, t1, t2);
            for (long i = 0; i < nd; i++) {
                double m = hddu[i];
                m += B2l[i];
                m += -a[0] * mun[i];
                m += a[3] * nun[i];
                double n_ = hdddu[i];
                n_ += B3l[i];
                n_ += a[4] * nun[i];
                double r = hdu[i];
                r += B1l[i];
                r += a[2] * mun[i];
                r += a[5] * nun[i];
                r = -r;
                mu[i] = -m;
                nu[i] = -n_;
                r += a[2] * mu[i];
                r += a[5] * nu[i];
                R1[i] = r;
            }
            memcpy(lt, ln, sizeof(double) * nd); /* initial guess lambda_{t+1} (:917) */
        }
        /* KSP operator N A N + I - N with A = -(K + a2 C + a5 M): solve the SPD form with -R1 */
        for (long i = 0; i < nd; i++) R1[i] = -N[i] * R1[i];
        if (ctx_pcg(&c, cK, cM, N, R1, lt, p->rtol, p->maxit, NULL) < 0) status = -1;
        memcpy(mun, mu, sizeof(double) * nd);
        memcpy(nun, nu, sizeof(double) * nd);
    }
    /* SolveLagrange0 (:678-804) */
    double* nu0 = calloc(nd, sizeof(double));
    {
        const double* l1 = lam + (size_t)1 * nd;
        grad_obj_state(&c, states, states + nd, states + 2 * nd, hdu, hddu, hdddu);
        memset(B3l, 0, sizeof(double) * nd);
        add_B(&c, -a[1] * be, a[4] - a[1] * al, N, l1, B3l, t1, t2);
        for (long i = 0; i < nd; i++) {
            double r = hdddu[i];
            r += B3l[i];
            r += -a[1] * mun[i];
            r += a[4] * nun[i];
            R1[i] = N[i] * (-r);
        }
        if (ctx_pcg(&c, 0.0, 1.0, N, R1, nu0, p->rtol, p->maxit, NULL) < 0) status = -1;
    }
    /* objective and sensitivities in the reference's accumulation order (:1778-1886, ncp = 1) */
    double f = 0.0;
    for (int t = 1; t <= T; t++) {
        const double* s = states + (size_t)t * 3 * nd;
        f += objective_state(&c, s, s + nd, s + 2 * nd);
    }
    f += objective_state(&c, states, states + nd, states + 2 * nd);
    memset(dfdx, 0, sizeof(double) * nel);
    for (int t = 1; t <= T; t++) {
        const double* s = states + (size_t)t * 3 * nd;
        const double* sb = states + (size_t)(t - 1) * 3 * nd;
        sens_n(&c, lam + (size_t)t * nd, s, s + nd, s + 2 * nd, sb, sb + nd, sb + 2 * nd, dfdx);
    }
    sens_0(&c, nu0, states, states + nd, states + 2 * nd, dfdx);
    *fx = f;
    free(N); free(lam); free(mu); free(nu); free(mun); free(nun); free(hdu); free(hddu); free(hdddu);
    free(R1); free(B1l); free(B2l); free(B3l); free(t1); free(t2); free(nu0);
    ctx_free(&c);
    return status;
}

/* ------------------------------------------------------------------ */
/* PDE filter and projection: PDEFilter.cc:28-224, Filter.cc, Filter.h  */
/* ------------------------------------------------------------------ */
int or_pde_filter(const or_params* p, double rmin, const double* in, double* out, double rtol) {
    long nn = or_nn(p), nel = or_nel(p);
    double dx = ddx(p), dy = ddy(p), dz = ddz(p);
    double R = rmin / 2.0 / sqrt(3.0); /* PDEFilter.cc:32 */
    double kf[64];
    or_pdefilter_ke(dx, dy, dz, R, kf);
    double vol = dx * dy * dz;
    double* rhs = calloc(nn, sizeof(double));
    double* U = malloc(sizeof(double) * nn);
    double* r = malloc(sizeof(double) * nn);
    double* z = malloc(sizeof(double) * nn);
    double* pp = malloc(sizeof(double) * nn);
    double* q = malloc(sizeof(double) * nn);
    double* dinv = calloc(nn, sizeof(double));
    long nodes[8];
    for (long e = 0; e < nel; e++) {
        or_element_nodes(p, e, nodes);
        for (int a = 0; a < 8; a++) {
            rhs[nodes[a]] += 0.125 * in[e]; /* T x, TF = 1/8 (:580-588) */
            dinv[nodes[a]] += kf[9 * a];
        }
    }
    memcpy(U, rhs, sizeof(double) * nn); /* warm start U = T x (:206) */
    for (long i = 0; i < nn; i++) { rhs[i] *= vol; dinv[i] = 1.0 / dinv[i]; }
    double bn = 0.0;
    for (long i = 0; i < nn; i++) bn += rhs[i] * rhs[i];
    bn = sqrt(bn);
    int it = 0, conv = 0, maxit = 100000;
#define KF_APPLY(src, dst)                                                         \
    do {                                                                           \
        memset(dst, 0, sizeof(double) * nn);                                       \
        for (long e = 0; e < nel; e++) {                                           \
            or_element_nodes(p, e, nodes);                                         \
            for (int a = 0; a < 8; a++) {                                          \
                double s = 0.0;                                                    \
                for (int b = 0; b < 8; b++) s += kf[8 * a + b] * src[nodes[b]];    \
                dst[nodes[a]] += s;                                                \
            }                                                                      \
        }                                                                          \
    } while (0)
    if (bn == 0.0) {
        memset(U, 0, sizeof(double) * nn);
        conv = 1;
    }
    while (!conv && it < maxit) {
        KF_APPLY(U, q);
        double rz = 0.0, rr = 0.0;
        for (long i = 0; i < nn; i++) { r[i] = rhs[i] - q[i]; z[i] = dinv[i] * r[i]; pp[i] = z[i]; rz += r[i] * z[i]; rr += r[i] * r[i]; }
        if (sqrt(rr) / bn <= rtol) { conv = 1; break; }
        while (it < maxit) {
            it++;
            KF_APPLY(pp, q);
            double pq = 0.0;
            for (long i = 0; i < nn; i++) pq += pp[i] * q[i];
            double al = rz / pq;
            rr = 0.0;
            for (long i = 0; i < nn; i++) { U[i] += al * pp[i]; r[i] -= al * q[i]; rr += r[i] * r[i]; }
            if (sqrt(rr) / bn <= rtol) break;
            double rzn = 0.0;
            for (long i = 0; i < nn; i++) { z[i] = dinv[i] * r[i]; rzn += r[i] * z[i]; }
            double be = rzn / rz;
            rz = rzn;
            for (long i = 0; i < nn; i++) pp[i] = z[i] + be * pp[i];
        }
        KF_APPLY(U, q);
        rr = 0.0;
        for (long i = 0; i < nn; i++) { r[i] = rhs[i] - q[i]; rr += r[i] * r[i]; }
        if (sqrt(rr) / bn <= rtol) conv = 1;
    }
#undef KF_APPLY
    for (long e = 0; e < nel; e++) { /* T' U (:216) */
        or_element_nodes(p, e, nodes);
        double s = 0.0;
        for (int a = 0; a < 8; a++) s += 0.125 * U[nodes[a]];
        out[e] = s;
    }
    free(rhs); free(U); free(r); free(z); free(pp); free(q); free(dinv);
    return conv ? it : -1;
}

double or_smooth_projection(double x, double beta, double eta) {
    return (tanh(beta * eta) + tanh(beta * (x - eta))) / (tanh(beta * eta) + tanh(beta * (1.0 - eta)));
}
double or_chainrule_projection(double x, double beta, double eta) {
    return beta * (1.0 - pow(tanh(beta * (x - eta)), 2.0)) / (tanh(beta * eta) + tanh(beta * (1.0 - eta)));
}

int or_filter_project(const or_params* p, double rmin, const double* x, double* xtilde, double* xphys,
                      int projection, double beta, double eta, double rtol) {
    long nel = or_nel(p);
    int it = or_pde_filter(p, rmin, x, xtilde, rtol);
    for (long i = 0; i < nel; i++) { /* Filter.cc:83-102 */
        if (xtilde[i] < 0.0) xtilde[i] = 0.0;
        if (xtilde[i] > 1.0) xtilde[i] = 1.0;
    }
    for (long i = 0; i < nel; i++) xphys[i] = projection ? or_smooth_projection(xtilde[i], beta, eta) : xtilde[i];
    return it < 0 ? -1 : 0;
}

int or_filter_gradients(const or_params* p, double rmin, const double* xtilde, double* dfdx, double* dgdx,
                        int projection, double beta, double eta, double rtol) {
    long nel = or_nel(p);
    if (projection) {
        for (long i = 0; i < nel; i++) {
            double d = or_chainrule_projection(xtilde[i], beta, eta);
            dfdx[i] *= d;
            dgdx[i] *= d;
        }
    }
    double* tmp = malloc(sizeof(double) * nel);
    int s1 = or_pde_filter(p, rmin, dfdx, tmp, rtol);
    memcpy(dfdx, tmp, sizeof(double) * nel);
    int s2 = or_pde_filter(p, rmin, dgdx, tmp, rtol);
    memcpy(dgdx, tmp, sizeof(double) * nel);
    free(tmp);
    return (s1 < 0 || s2 < 0) ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* MMA: MMA.cc:374-969 (single process: all reductions are local sums)  */
/* ------------------------------------------------------------------ */
struct or_mma {
    long n;
    int m, k;
    double asyminit, asymdec, asyminc;
    double *a, *c, *d, *y, *lam, *mu, *s, *b, *grad, *Hess;
    double z;
    double *L, *U, *alpha, *beta, *p0, *q0, *pij, *qij, *xo1, *xo2;
};

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

or_mma* or_mma_create(long n, int m, const double* x, const double* a, const double* c, const double* d) {
    (void)x;
    or_mma* o = calloc(1, sizeof(or_mma));
    o->n = n; o->m = m; o->k = 0;
    o->asyminit = 0.5; o->asymdec = 0.7; o->asyminc = 1.2;
    o->a = malloc(sizeof(double) * m); o->c = malloc(sizeof(double) * m); o->d = malloc(sizeof(double) * m);
    memcpy(o->a, a, sizeof(double) * m); memcpy(o->c, c, sizeof(double) * m); memcpy(o->d, d, sizeof(double) * m);
    o->y = calloc(m, sizeof(double)); o->lam = calloc(m, sizeof(double)); o->mu = calloc(m, sizeof(double));
    o->s = calloc(2 * m, sizeof(double)); o->b = calloc(m, sizeof(double)); o->grad = calloc(m, sizeof(double));
    o->Hess = calloc((size_t)m * m, sizeof(double));
    o->L = calloc(n, sizeof(double)); o->U = calloc(n, sizeof(double));
    o->alpha = calloc(n, sizeof(double)); o->beta = calloc(n, sizeof(double));
    o->p0 = calloc(n, sizeof(double)); o->q0 = calloc(n, sizeof(double));
    o->pij = calloc((size_t)n * m, sizeof(double)); o->qij = calloc((size_t)n * m, sizeof(double));
    o->xo1 = calloc(n, sizeof(double)); o->xo2 = calloc(n, sizeof(double));
    return o;
}

void or_mma_destroy(or_mma* o) {
    if (!o) return;
    free(o->a); free(o->c); free(o->d); free(o->y); free(o->lam); free(o->mu); free(o->s); free(o->b);
    free(o->grad); free(o->Hess); free(o->L); free(o->U); free(o->alpha); free(o->beta); free(o->p0);
    free(o->q0); free(o->pij); free(o->qij); free(o->xo1); free(o->xo2);
    free(o);
}

void or_mma_set_asymptotes(or_mma* o, double init, double dec, double inc) {
    o->asyminit = init; o->asymdec = dec; o->asyminc = inc;
}

void or_mma_set_outer_movelimit(long n, double Xmin, double Xmax, double movlim, const double* x, double* xmin,
                                double* xmax) {
    for (long i = 0; i < n; i++) { /* MMA.cc:374-393 */
        xmax[i] = dmin(Xmax, x[i] + movlim);
        xmin[i] = dmax(Xmin, x[i] - movlim);
    }
}

double or_mma_design_change(long n, const double* x, double* xold) {
    double ch = 0.0; /* MMA.cc:395-414 */
    for (long i = 0; i < n; i++) {
        ch = dmax(ch, fabs(x[i] - xold[i]));
        xold[i] = x[i];
    }
    return ch;
}

static void mma_gensub(or_mma* o, const double* xv, const double* df, const double* gx, const double* dg,
                       const double* xmin, const double* xmax) {
    long n = o->n;
    int m = o->m;
    o->k++;
    if (o->k < 3) { /* MMA.cc:500-505 */
        for (long i = 0; i < n; i++) {
            o->L[i] = xv[i] - o->asyminit * xmax[i];
            o->L[i] += o->asyminit * xmin[i];
            o->U[i] = xv[i] + o->asyminit * xmax[i];
            o->U[i] += -o->asyminit * xmin[i];
        }
    }
    if (o->k > 2) { /* MMA.cc:523-560, RobustAsymptotesType 0 */
        for (long i = 0; i < n; i++) {
            double hv = (xv[i] - o->xo1[i]) * (o->xo1[i] - o->xo2[i]);
            double gam = hv < 0.0 ? o->asymdec : (hv > 0.0 ? o->asyminc : 1.0);
            o->L[i] = xv[i] - gam * (o->xo1[i] - o->L[i]);
            o->U[i] = xv[i] + gam * (o->U[i] - o->xo1[i]);
            double xmi = dmax(1.0e-5, xmax[i] - xmin[i]);
            o->L[i] = dmax(o->L[i], xv[i] - 10.0 * xmi);
            o->L[i] = dmin(o->L[i], xv[i] - 0.01 * xmi);
            o->U[i] = dmax(o->U[i], xv[i] + 0.01 * xmi);
            o->U[i] = dmin(o->U[i], xv[i] + 10.0 * xmi);
        }
    }
    double feps = 1.0e-6; /* MMA.cc:562-580 */
    for (long i = 0; i < n; i++) {
        double L = o->L[i], U = o->U[i];
        o->alpha[i] = dmax(xmin[i], 0.9 * L + 0.1 * xv[i]);
        o->beta[i] = dmin(xmax[i], 0.9 * U + 0.1 * xv[i]);
        double dfp = dmax(0.0, df[i]), dfm = dmax(0.0, -1.0 * df[i]);
        o->p0[i] = pow(U - xv[i], 2.0) * (dfp + 0.001 * fabs(df[i]) + 0.5 * feps / (U - L));
        o->q0[i] = pow(xv[i] - L, 2.0) * (dfm + 0.001 * fabs(df[i]) + 0.5 * feps / (U - L));
        for (int j = 0; j < m; j++) {
            double g = dg[(size_t)j * n + i];
            double gp = dmax(0.0, g), gm = dmax(0.0, -1.0 * g);
            o->pij[(size_t)j * n + i] = pow(U - xv[i], 2.0) * gp;
            o->qij[(size_t)j * n + i] = pow(xv[i] - L, 2.0) * gm;
        }
    }
    for (int j = 0; j < m; j++) {
        double b = 0.0;
        for (long i = 0; i < n; i++)
            b += o->pij[(size_t)j * n + i] / (o->U[i] - xv[i]) + o->qij[(size_t)j * n + i] / (xv[i] - o->L[i]);
        o->b[j] = b - gx[j];
    }
}

static void mma_xyz(or_mma* o, double* x) {
    int m = o->m; /* MMA.cc:684-727 */
    double lamai = 0.0;
    for (int i = 0; i < m; i++) {
        if (o->lam[i] < 0.0) o->lam[i] = 0;
        o->y[i] = dmax(0.0, o->lam[i] - o->c[i]);
        lamai += o->lam[i] * o->a[i];
    }
    o->z = dmax(0.0, 10.0 * (lamai - 1.0));
    for (long i = 0; i < o->n; i++) {
        double pj = o->p0[i], qj = o->q0[i];
        for (int j = 0; j < m; j++) {
            pj += o->pij[(size_t)j * o->n + i] * o->lam[j];
            qj += o->qij[(size_t)j * o->n + i] * o->lam[j];
        }
        x[i] = (sqrt(pj) * o->L[i] + sqrt(qj) * o->U[i]) / (sqrt(pj) + sqrt(qj));
        if (x[i] < o->alpha[i]) x[i] = o->alpha[i];
        if (x[i] > o->beta[i]) x[i] = o->beta[i];
    }
}

static void mma_dualgrad(or_mma* o, const double* x) {
    for (int j = 0; j < o->m; j++) { /* MMA.cc:729-765 */
        double g = 0.0;
        for (long i = 0; i < o->n; i++)
            g += o->pij[(size_t)j * o->n + i] / (o->U[i] - x[i]) + o->qij[(size_t)j * o->n + i] / (x[i] - o->L[i]);
        o->grad[j] = g + (-o->b[j] - o->a[j] * o->z - o->y[j]);
    }
}

static void mma_dualhess(or_mma* o, const double* x) {
    long n = o->n; /* MMA.cc:767-871 */
    int m = o->m;
    double* df2 = malloc(sizeof(double) * n);
    double* PQ = malloc(sizeof(double) * n * m);
    for (long i = 0; i < n; i++) {
        double pj = o->p0[i], qj = o->q0[i];
        for (int j = 0; j < m; j++) {
            pj += o->pij[(size_t)j * n + i] * o->lam[j];
            qj += o->qij[(size_t)j * n + i] * o->lam[j];
            PQ[i * m + j] = o->pij[(size_t)j * n + i] / pow(o->U[i] - x[i], 2.0) -
                            o->qij[(size_t)j * n + i] / pow(x[i] - o->L[i], 2.0);
        }
        df2[i] = -1.0 / (2.0 * pj / pow(o->U[i] - x[i], 3.0) + 2.0 * qj / pow(x[i] - o->L[i], 3.0));
        double xp = (sqrt(pj) * o->L[i] + sqrt(qj) * o->U[i]) / (sqrt(pj) + sqrt(qj));
        if (xp < o->alpha[i]) df2[i] = 0.0;
        if (xp > o->beta[i]) df2[i] = 0.0;
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
            double h = 0.0;
            for (long k = 0; k < n; k++) h += (PQ[k * m + i] * df2[k]) * PQ[k * m + j];
            o->Hess[i * m + j] = h;
        }
    double lamai = 0.0;
    for (int j = 0; j < m; j++) {
        if (o->lam[j] < 0.0) o->lam[j] = 0.0;
        lamai += o->lam[j] * o->a[j];
        if (o->lam[j] > o->c[j]) o->Hess[j * m + j] += -1.0;
        o->Hess[j * m + j] += -o->mu[j] / o->lam[j];
    }
    if (lamai > 0.0)
        for (int j = 0; j < m; j++)
            for (int k = 0; k < m; k++) o->Hess[j * m + k] += -10.0 * o->a[j] * o->a[k];
    double tr = 0.0;
    for (int i = 0; i < m; i++) tr += o->Hess[i * m + i];
    double corr = 1e-4 * tr / m;
    if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
    for (int i = 0; i < m; i++) o->Hess[i * m + i] += corr;
    free(df2);
    free(PQ);
}

static void mma_factor_solve(double* K, double* x, int nn) {
    for (int ss = 0; ss < nn - 1; ss++) /* MMA.cc:936-947 */
        for (int i = ss + 1; i < nn; i++) {
            K[i * nn + ss] = K[i * nn + ss] / K[ss * nn + ss];
            for (int j = ss + 1; j < nn; j++) K[i * nn + j] = K[i * nn + j] - K[i * nn + ss] * K[ss * nn + j];
        }
    for (int i = 1; i < nn; i++) { /* MMA.cc:949-969 */
        double a = 0.0;
        for (int j = 0; j < i; j++) a = a - K[i * nn + j] * x[j];
        x[i] = x[i] + a;
    }
    x[nn - 1] = x[nn - 1] / K[(nn - 1) * nn + (nn - 1)];
    for (int i = nn - 2; i >= 0; i--) {
        double a = x[i];
        for (int j = i + 1; j < nn; j++) a = a - K[i * nn + j] * x[j];
        x[i] = a / K[i * nn + i];
    }
}

static void mma_linesearch(or_mma* o) {
    int m = o->m; /* MMA.cc:873-891 */
    double theta = 1.005;
    for (int i = 0; i < m; i++) {
        if (theta < -1.01 * o->s[i] / o->lam[i]) theta = -1.01 * o->s[i] / o->lam[i];
        if (theta < -1.01 * o->s[i + m] / o->mu[i]) theta = -1.01 * o->s[i + m] / o->mu[i];
    }
    theta = 1.0 / theta;
    for (int i = 0; i < m; i++) {
        o->lam[i] = o->lam[i] + theta * o->s[i];
        o->mu[i] = o->mu[i] + theta * o->s[i + m];
    }
}

static double mma_dualres(or_mma* o, const double* x, double epsi) {
    int m = o->m; /* MMA.cc:893-934 */
    double nrI = 0.0;
    for (int j = 0; j < m; j++) {
        double r = 0.0;
        for (long i = 0; i < o->n; i++)
            r += o->pij[(size_t)j * o->n + i] / (o->U[i] - x[i]) + o->qij[(size_t)j * o->n + i] / (x[i] - o->L[i]);
        r += -o->b[j] - o->a[j] * o->z - o->y[j] + o->mu[j];
        double r2 = 0.0 + o->mu[j] * o->lam[j] - epsi;
        if (nrI < fabs(r)) nrI = fabs(r);
        if (nrI < fabs(r2)) nrI = fabs(r2);
    }
    return nrI;
}

static void mma_solvedip(or_mma* o, double* x) {
    int m = o->m; /* MMA.cc:645-682 */
    for (int j = 0; j < m; j++) { o->lam[j] = o->c[j] / 2.0; o->mu[j] = 1.0; }
    double tol = 1.0e-9 * sqrt(m + o->n), epsi = 1.0, err = 1.0;
    while (epsi > tol) {
        int loop = 0;
        while (err > 0.9 * epsi && loop < 100) {
            loop++;
            mma_xyz(o, x);
            mma_dualgrad(o, x);
            for (int j = 0; j < m; j++) o->grad[j] = -1.0 * o->grad[j] - epsi / o->lam[j];
            mma_dualhess(o, x);
            mma_factor_solve(o->Hess, o->grad, m);
            for (int j = 0; j < m; j++) o->s[j] = o->grad[j];
            for (int i = 0; i < m; i++) o->s[m + i] = -o->mu[i] + epsi / o->lam[i] - o->s[i] * o->mu[i] / o->lam[i];
            mma_linesearch(o);
            mma_xyz(o, x);
            err = mma_dualres(o, x, epsi);
        }
        epsi = epsi * 0.1;
    }
}

void or_mma_update(or_mma* o, double* x, const double* dfdx, const double* gx, const double* dgdx,
                   const double* xmin, const double* xmax) {
    mma_gensub(o, x, dfdx, gx, dgdx, xmin, xmax); /* MMA.cc:487-505 */
    memcpy(o->xo2, o->xo1, sizeof(double) * o->n);
    memcpy(o->xo1, x, sizeof(double) * o->n);
    mma_solvedip(o, x);
}

void or_mma_kkt_residual(or_mma* o, const double* x, const double* dfdx, const double* gx, const double* dgdx,
                         const double* xmin, const double* xmax, double* norm2, double* normInf) {
    double n2 = 0.0, nI = 0.0; /* MMA.cc:416-484 */
    for (long i = 0; i < o->n; i++) {
        double ri = dfdx[i];
        for (int j = 0; j < o->m; j++) ri += o->lam[j] * dgdx[(size_t)j * o->n + i];
        double mmin = 0.0, mmax = 0.0;
        if (x[i] < xmin[i] + 1.0e-5 && ri > 0.0) mmin = ri;
        if (x[i] > xmax[i] - 1.0e-5 && ri < 0.0) mmax = -ri;
        ri += -mmin + mmax;
        n2 += pow(ri, 2.0);
        nI = dmax(fabs(ri), nI);
        double resi = mmin * (x[i] - xmin[i]);
        n2 += pow(resi, 2.0);
        nI = dmax(fabs(resi), nI);
        resi = mmax * (xmax[i] - x[i]);
        n2 += pow(resi, 2.0);
        nI = dmax(fabs(resi), nI);
    }
    double ri = 0.0;
    for (int j = 0; j < o->m; j++) ri += o->lam[j] * (o->a[j] * o->z + o->y[j] - gx[j]);
    n2 += pow(ri, 2.0);
    nI = dmax(fabs(ri), nI);
    *norm2 = sqrt(n2);
    *normInf = nI;
}

void or_mma_get_state(or_mma* o, double* L, double* U, double* lam) {
    if (L) memcpy(L, o->L, sizeof(double) * o->n);
    if (U) memcpy(U, o->U, sizeof(double) * o->n);
    if (lam) memcpy(lam, o->lam, sizeof(double) * o->m);
}

/* ------------------------------------------------------------------ */
/* main.cc:52-161 optimisation loop (filter = 2, m = 1, minmax = -1)     */
/* ------------------------------------------------------------------ */
int or_design_loop(const or_params* p, double rmin, double volfrac, double V0, int maxit, int projection,
                   double beta, double eta, double movlim, double* x, double* xphys, double* hist) {
    long n = or_nel(p), nd = 3 * or_nn(p);
    double ftol = 1e-12;
    double* xt = malloc(sizeof(double) * n);
    double* xold = malloc(sizeof(double) * n);
    double* df = malloc(sizeof(double) * n);
    double* dg = malloc(sizeof(double) * n);
    double* xmin = malloc(sizeof(double) * n);
    double* xmax = malloc(sizeof(double) * n);
    double* states = malloc(sizeof(double) * (size_t)(p->nsteps + 1) * 3 * nd);
    for (long i = 0; i < n; i++) { x[i] = V0; xt[i] = V0; xphys[i] = V0; xold[i] = V0; } /* TopOpt.cc:500-511 */
    double am = 0.0, cm = 1000.0, dm = 0.0; /* TopOpt.cc:527-531 */
    or_mma* mma = or_mma_create(n, 1, x, &am, &cm, &dm);
    or_mma_set_asymptotes(mma, 0.1, 0.6, 1.05); /* main.cc:58 */
    or_filter_project(p, rmin, x, xt, xphys, projection, beta, eta, ftol);
    int itr = 0, status = 0;
    double ch = 1.0, fscale = 1.0, betaFinal = 48.0;
    while (itr < maxit && ch > 0.01) {
        itr++;
        double fx, gx;
        if (or_forward(p, xphys, states) < 0) status = -1;
        if (or_objective_gradient(p, xphys, states, &fx, df) < 0) status = -1;
        gx = 0.0; /* LinearElasticity.cc:1283-1296 */
        for (long i = 0; i < n; i++) gx += xphys[i];
        gx = gx / (double)n - volfrac;
        for (long i = 0; i < n; i++) dg[i] = 1.0 / (double)n;
        or_filter_gradients(p, rmin, xt, df, dg, projection, beta, eta, ftol);
        double mnd = 0.0; /* Filter.cc:208-227 */
        for (long i = 0; i < n; i++) mnd += 4 * xphys[i] * (1.0 - xphys[i]);
        mnd = mnd / (double)n;
        if (itr == 1) fscale = 1.0 / fx;
        fx = fx * fscale;
        for (long i = 0; i < n; i++) df[i] *= fscale;
        or_mma_set_outer_movelimit(n, 0.0, 1.0, movlim, x, xmin, xmax);
        or_mma_update(mma, x, df, &gx, dg, xmin, xmax);
        ch = or_mma_design_change(n, x, xold);
        if (hist) {
            hist[5 * (itr - 1) + 0] = fx / fscale;
            hist[5 * (itr - 1) + 1] = fx;
            hist[5 * (itr - 1) + 2] = gx;
            hist[5 * (itr - 1) + 3] = ch;
            hist[5 * (itr - 1) + 4] = mnd;
        }
        if (projection && (itr >= 200 || ch < 0.01)) { /* Filter.cc:274-300 */
            if ((ch < 0.01 || itr % 10 == 0) && beta < betaFinal && gx <= 0.0) {
                int change = 1;
                beta = beta < 7 ? beta + 1 : beta * 1.2;
                if (beta > betaFinal) { beta = betaFinal; change = 0; }
                if (change) ch = 1.0;
            }
        }
        if (gx > 0) ch = 1;
        or_filter_project(p, rmin, x, xt, xphys, projection, beta, eta, ftol);
    }
    or_mma_destroy(mma);
    free(xt); free(xold); free(df); free(dg); free(xmin); free(xmax); free(states);
    return status < 0 ? status : itr;
}

/* ------------------------------------------------------------------ */
/* SOMMG reduced-order model (-reduction): the basis of                 */
/* SOMMG::ConstructBasis (SOMMG.cc:40-265) and the reduced branches of  */
/* LinearElasticity.cc (:404-461, 994-1003, 1015-1019, 1089-1107,       */
/* 1683-1764, 2318-2335), where K, M, C, the load and the state-side    */
/* objective gradients become their projections on the basis, the      */
/* reduced systems are solved by LU (SetUpSolver2, :3007-3014), and the */
/* objective and sensitivities use the states and multipliers projected */
/* back to the full mesh.                                               */
/* ------------------------------------------------------------------ */
/* dense Gaussian elimination with partial pivoting: x = A^-1 b (A, n x n row-major, not modified) */
static int dense_solve(int n, const double* A, const double* b, double* x) {
    double* w = malloc(sizeof(double) * (size_t)n * (n + 1));
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) w[i * (n + 1) + j] = A[i * n + j];
        w[i * (n + 1) + n] = b[i];
    }
    int ok = 1;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(w[i * (n + 1) + k]) > fabs(w[piv * (n + 1) + k])) piv = i;
        if (w[piv * (n + 1) + k] == 0.0) { ok = 0; break; }
        if (piv != k)
            for (int j = 0; j <= n; j++) {
                double t = w[k * (n + 1) + j];
                w[k * (n + 1) + j] = w[piv * (n + 1) + j];
                w[piv * (n + 1) + j] = t;
            }
        for (int i = k + 1; i < n; i++) {
            double f = w[i * (n + 1) + k] / w[k * (n + 1) + k];
            for (int j = k; j <= n; j++) w[i * (n + 1) + j] -= f * w[k * (n + 1) + j];
        }
    }
    if (ok)
        for (int i = n - 1; i >= 0; i--) {
            double s = w[i * (n + 1) + n];
            for (int j = i + 1; j < n; j++) s -= w[i * (n + 1) + j] * x[j];
            x[i] = s / w[i * (n + 1) + i];
        }
    free(w);
    return ok ? 0 : -1;
}

/* y = X w for a dense n x n X */
static void dense_mv(int n, const double* X, const double* w, double* y) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += X[i * n + j] * w[j];
        y[i] = s;
    }
}

/* full = sum_k r[k] Q_k */
static void expand(long nd, int nb, const double* Q, const double* r, double* full) {
    memset(full, 0, sizeof(double) * nd);
    for (int k = 0; k < nb; k++)
        for (long i = 0; i < nd; i++) full[i] += r[k] * Q[(size_t)k * nd + i];
}

/* SOMMG::ConstructBasis (SOMMG.cc:40-230): Q_0 ~ P0^-1 N.F with P0 = K - om^2 M on the free dofs,
   then Qb_j = P0^-1 N.(2 om M Q_{j-1}) orthonormalised against Q_0..Q_{j-1} by modified Gram-Schmidt
   (:206-218).  The moment-matching correction of :136-187 enters as fvec += (-1.0)*(-1,0)*massVec
   (:176): the comma expression (-1,0) is 0, so fvec is unchanged and the correction is not formed
   here.  Q: nb vectors of nd; Ut: the nb x nb factor Utemp (row-major, upper triangle). */
int or_sommg_basis(const or_params* p, const double* xphys, int nb, double om, double* Q, double* Ut) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd;
    double* N = malloc(sizeof(double) * nd);
    double* f = malloc(sizeof(double) * nd);
    double* qb = malloc(sizeof(double) * nd);
    or_bc_mask(p, N);
    or_load_location(p, f);
    for (long i = 0; i < nd; i++) f[i] *= N[i]; /* :77 */
    memset(Ut, 0, sizeof(double) * nb * nb);
    int status = 0;
    for (int j = 0; j < nb; j++) {
        if (j > 0) {
            ctx_apply(&c, 0.0, 1.0, Q + (size_t)(j - 1) * nd, f); /* fvec = -(-2 om) M Q_{j-1} (:129-131) */
            for (long i = 0; i < nd; i++) f[i] = N[i] * (f[i] * (-1.0 * (-2.0 * om))); /* N.fvec (:192) */
        }
        memset(qb, 0, sizeof(double) * nd);
        if (ctx_pcg(&c, 1.0, -om * om, N, f, qb, p->rtol, p->maxit, NULL) < 0) status = -1;
        for (int i = 0; i < j; i++) { /* :206-213 */
            double d = vdot(nd, Q + (size_t)i * nd, qb);
            Ut[i * nb + j] = d;
            for (long h = 0; h < nd; h++) qb[h] -= d * Q[(size_t)i * nd + h];
        }
        double nrm = sqrt(vdot(nd, qb, qb)); /* :216-218 */
        Ut[j * nb + j] = nrm;
        for (long h = 0; h < nd; h++) Q[(size_t)j * nd + h] = qb[h] * (1.0 / nrm);
    }
    free(N); free(f); free(qb);
    ctx_free(&c);
    return status;
}

/* reduced matrices and load (SOMMG.cc:236-262): Kr_ij = Q_i.(K Q_j), Mr, Cr with C = dalpha M + dbeta K
   (AssembleDampingMatrix, LinearElasticity.cc:2605-2619), Pr_i = (N.F).Q_i */
void or_sommg_reduce(const or_params* p, const double* xphys, int nb, const double* Q, double* Kr, double* Mr,
                     double* Cr, double* Pr) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd;
    double* N = malloc(sizeof(double) * nd);
    double* f = malloc(sizeof(double) * nd);
    double* t = malloc(sizeof(double) * nd);
    or_bc_mask(p, N);
    or_load_location(p, f);
    for (long i = 0; i < nd; i++) f[i] *= N[i];
    for (int i = 0; i < nb; i++) Pr[i] = vdot(nd, f, Q + (size_t)i * nd);
    for (int j = 0; j < nb; j++) {
        const double* qj = Q + (size_t)j * nd;
        ctx_apply(&c, 1.0, 0.0, qj, t);
        for (int i = 0; i < nb; i++) Kr[i * nb + j] = vdot(nd, t, Q + (size_t)i * nd);
        ctx_apply(&c, 0.0, 1.0, qj, t);
        for (int i = 0; i < nb; i++) Mr[i * nb + j] = vdot(nd, t, Q + (size_t)i * nd);
        ctx_apply(&c, p->dbeta, p->dalpha, qj, t);
        for (int i = 0; i < nb; i++) Cr[i * nb + j] = vdot(nd, t, Q + (size_t)i * nd);
    }
    free(N); free(f); free(t);
    ctx_free(&c);
}

/* reduced Newmark (CompleteInitialState :1015-1078 and SolveStateBlock :1298-1397 on the reduced
   matrices, N = 1): rstates[(t*3 + {0,1,2})*nb + k] */
int or_reduced_forward(const or_params* p, int nb, const double* Kr, const double* Mr, const double* Cr,
                       const double* Pr, double* rstates) {
    double a[6];
    double dt = p->t1 / p->nsteps;
    or_newmark_consts(p->ti, dt, a);
    size_t n2 = (size_t)nb * nb;
    double* S = malloc(sizeof(double) * n2);
    double* B1 = malloc(sizeof(double) * n2);
    double* B2 = malloc(sizeof(double) * n2);
    double* B3 = malloc(sizeof(double) * n2);
    double* R = malloc(sizeof(double) * nb);
    double* w = malloc(sizeof(double) * nb);
    for (size_t i = 0; i < n2; i++) {
        S[i] = a[5] * Mr[i] + a[2] * Cr[i] + Kr[i]; /* -A (:1021-1024) */
        B1[i] = a[5] * Mr[i] + a[2] * Cr[i];
        B2[i] = a[3] * Mr[i] - a[0] * Cr[i];
        B3[i] = a[4] * Mr[i] - a[1] * Cr[i];
    }
    int status = 0;
    memset(rstates, 0, sizeof(double) * 3 * nb); /* U0 = DU0 = 0 */
    double s0 = or_load_intensity(p, 0);
    for (int i = 0; i < nb; i++) R[i] = Pr[i] * s0; /* R = RHS - C DU0 - K U0 with U0 = DU0 = 0 */
    if (dense_solve(nb, Mr, R, rstates + 2 * nb)) status = -1;
    for (int t = 1; t <= p->nsteps; t++) {
        const double* up = rstates + (size_t)(t - 1) * 3 * nb;
        const double* vp = up + nb;
        const double* ap = up + 2 * nb;
        double* un = rstates + (size_t)t * 3 * nb;
        double* vn = un + nb;
        double* an = un + 2 * nb;
        double s = or_load_intensity(p, t);
        for (int i = 0; i < nb; i++) R[i] = Pr[i] * s;
        dense_mv(nb, B1, up, w);
        for (int i = 0; i < nb; i++) R[i] += w[i];
        dense_mv(nb, B2, vp, w);
        for (int i = 0; i < nb; i++) R[i] += w[i];
        dense_mv(nb, B3, ap, w);
        for (int i = 0; i < nb; i++) R[i] += w[i];
        if (dense_solve(nb, S, R, un)) status = -1;
        for (int i = 0; i < nb; i++) {
            double du = un[i] - up[i];
            vn[i] = a[0] * vp[i] + a[1] * ap[i] + a[2] * du;
            an[i] = -a[3] * vp[i] - a[4] * ap[i] + a[5] * du;
        }
    }
    free(S); free(B1); free(B2); free(B3); free(R); free(w);
    return status;
}

/* full states Q r for every step: full[(t*3 + c)*nd + i] */
void or_reduced_expand(long nd, int nb, int nsteps, const double* Q, const double* rstates, double* full) {
    for (int t = 0; t <= nsteps; t++)
        for (int c = 0; c < 3; c++)
            expand(nd, nb, Q, rstates + ((size_t)t * 3 + c) * nb, full + ((size_t)t * 3 + c) * nd);
}

/* objective and sensitivities of the reduced model: the adjoint recurrences of SolveLagrangen /
   SolveLagrange0 (:678-954) on the reduced matrices with the projected state gradients
   Q^T hdu, Q^T hddu, Q^T hdddu (:2318-2335); the multipliers are expanded as Q lambda (:1750-1764,
   1850-1863) and enter the full-mesh sensitivity loops with the expanded states (the basis is held
   fixed: its dependence on the design is not differentiated, as in the reference). */
int or_reduced_objective_gradient(const or_params* p, const double* xphys, int nb, const double* Q,
                                  const double* Kr, const double* Mr, const double* Cr, const double* rstates,
                                  double* fx, double* dfdx) {
    ctx_t c;
    ctx_init(&c, p, xphys);
    long nd = c.nd, nel = c.nel;
    const double* a = c.a;
    int T = p->nsteps, status = 0;
    size_t n2 = (size_t)nb * nb;
    double* full = malloc(sizeof(double) * (size_t)(T + 1) * 3 * nd);
    or_reduced_expand(nd, nb, T, Q, rstates, full);
    double* S = malloc(sizeof(double) * n2);
    double* B1 = malloc(sizeof(double) * n2);
    double* B2 = malloc(sizeof(double) * n2);
    double* B3 = malloc(sizeof(double) * n2);
    for (size_t i = 0; i < n2; i++) {
        S[i] = a[5] * Mr[i] + a[2] * Cr[i] + Kr[i];
        B1[i] = a[5] * Mr[i] + a[2] * Cr[i];
        B2[i] = a[3] * Mr[i] - a[0] * Cr[i];
        B3[i] = a[4] * Mr[i] - a[1] * Cr[i];
    }
    double* hdu = malloc(sizeof(double) * nd);
    double* hddu = malloc(sizeof(double) * nd);
    double* hdddu = malloc(sizeof(double) * nd);
    double* lamf = malloc(sizeof(double) * nd);
    double* h1 = malloc(sizeof(double) * nb);
    double* h2 = malloc(sizeof(double) * nb);
    double* h3 = malloc(sizeof(double) * nb);
    double* lam = calloc((size_t)(T + 2) * nb, sizeof(double));
    double* mu = malloc(sizeof(double) * nb);
    double* nu = malloc(sizeof(double) * nb);
    double* mun = calloc(nb, sizeof(double));
    double* nun = calloc(nb, sizeof(double));
    double* b1 = malloc(sizeof(double) * nb);
    double* b2 = malloc(sizeof(double) * nb);
    double* b3 = malloc(sizeof(double) * nb);
    double* R1 = malloc(sizeof(double) * nb);
    double* nu0 = malloc(sizeof(double) * nb);
#define PROJECT(s)                                                                        \
    do {                                                                                  \
        grad_obj_state(&c, (s), (s) + nd, (s) + 2 * nd, hdu, hddu, hdddu);                \
        for (int k = 0; k < nb; k++) {                                                    \
            h1[k] = vdot(nd, hdu, Q + (size_t)k * nd);                                    \
            h2[k] = vdot(nd, hddu, Q + (size_t)k * nd);                                   \
            h3[k] = vdot(nd, hdddu, Q + (size_t)k * nd);                                  \
        }                                                                                 \
    } while (0)
    for (int t = T; t >= 1; t--) {
        PROJECT(full + (size_t)t * 3 * nd);
        double* lt = lam + (size_t)t * nb;
        if (t == T) {
            for (int i = 0; i < nb; i++) {
                mu[i] = -h2[i];
                nu[i] = -h3[i];
                R1[i] = -(h1[i] - a[2] * mu[i] - a[5] * nu[i]);
            }
        } else {
            const double* ln = lam + (size_t)(t + 1) * nb;
            dense_mv(nb, B1, ln, b1);
            dense_mv(nb, B2, ln, b2);
            dense_mv(nb, B3, ln, b3);
            for (int i = 0; i < nb; i++) {
                double m = h2[i] + b2[i] - a[0] * mun[i] + a[3] * nun[i];
                double n_ = h3[i] + b3[i] + a[4] * nun[i];
                double r = -(h1[i] + b1[i] + a[2] * mun[i] + a[5] * nun[i]);
                mu[i] = -m;
                nu[i] = -n_;
                R1[i] = r + a[2] * mu[i] + a[5] * nu[i];
            }
        }
        /* A lambda = R1 with A = -S (:1118-1125) */
        for (int i = 0; i < nb; i++) R1[i] = -R1[i];
        if (dense_solve(nb, S, R1, lt)) status = -1;
        memcpy(mun, mu, sizeof(double) * nb);
        memcpy(nun, nu, sizeof(double) * nb);
    }
    {
        PROJECT(full);
        dense_mv(nb, B3, lam + (size_t)1 * nb, b3);
        for (int i = 0; i < nb; i++) R1[i] = -(h3[i] + b3[i] - a[1] * mun[i] + a[4] * nun[i]);
        if (dense_solve(nb, Mr, R1, nu0)) status = -1; /* dRdS = M (:763-769) */
    }
#undef PROJECT
    double f = 0.0;
    for (int t = 1; t <= T; t++) {
        const double* s = full + (size_t)t * 3 * nd;
        f += objective_state(&c, s, s + nd, s + 2 * nd);
    }
    f += objective_state(&c, full, full + nd, full + 2 * nd);
    memset(dfdx, 0, sizeof(double) * nel);
    for (int t = 1; t <= T; t++) {
        const double* s = full + (size_t)t * 3 * nd;
        const double* sb = full + (size_t)(t - 1) * 3 * nd;
        expand(nd, nb, Q, lam + (size_t)t * nb, lamf);
        sens_n(&c, lamf, s, s + nd, s + 2 * nd, sb, sb + nd, sb + 2 * nd, dfdx);
    }
    expand(nd, nb, Q, nu0, lamf);
    sens_0(&c, lamf, full, full + nd, full + 2 * nd, dfdx);
    *fx = f;
    free(full); free(S); free(B1); free(B2); free(B3); free(hdu); free(hddu); free(hdddu); free(lamf);
    free(h1); free(h2); free(h3); free(lam); free(mu); free(nu); free(mun); free(nun); free(b1); free(b2);
    free(b3); free(R1); free(nu0);
    ctx_free(&c);
    return status;
}
