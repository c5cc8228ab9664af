// Element matrices of the trilinear hex8 on an axis-aligned box element.
//
// Same integrals as the reference's Hex8IsoparametricKE/ME (LinearElasticity.cc:
// 3089-3333) and closed-form PDE-filter matrix (PDEFilter.cc:485-578): 2x2x2
// Gauss quadrature of B^T C B (E = 1), of N^T N (density 1), and of
// R^2 grad N . grad N + N N.  For a box element the Jacobian is diagonal,
// so shape-function gradients are the parametric ones scaled by 2/h.
#include <cmath>
#include <cstring>

#include "tt_internal.h"

namespace tt {

// Gauss abscissa as written in the reference (LinearElasticity.cc:3150)
static const double kGP = 0.577350269189626;

static const int kSx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const int kSy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const int kSz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

struct GaussPoint {
    double N[8], dx[8], dy[8], dz[8], w;
};

static GaussPoint gauss_point(double xi, double eta, double zeta, double hx, double hy, double hz) {
    GaussPoint g;
    for (int a = 0; a < 8; a++) {
        const double fx = 1.0 + kSx[a] * xi, fy = 1.0 + kSy[a] * eta, fz = 1.0 + kSz[a] * zeta;
        g.N[a] = 0.125 * fx * fy * fz;
        g.dx[a] = 0.125 * kSx[a] * fy * fz * (2.0 / hx);
        g.dy[a] = 0.125 * kSy[a] * fx * fz * (2.0 / hy);
        g.dz[a] = 0.125 * kSz[a] * fx * fy * (2.0 / hz);
    }
    g.w = 0.125 * hx * hy * hz;  // unit weights times det J
    return g;
}

void hex8_stiffness(double hx, double hy, double hz, double nu, double* ke) {
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    double D[6][6] = {};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) D[i][j] = lam;
        D[i][i] = lam + 2.0 * mu;
        D[3 + i][3 + i] = mu;
    }
    std::memset(ke, 0, sizeof(double) * 576);
    const double gp[2] = {-kGP, kGP};
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                GaussPoint g = gauss_point(gp[ix], gp[iy], gp[iz], hx, hy, hz);
                // strain rows: xx, yy, zz, xy, yz, zx
                double B[6][24] = {};
                for (int a = 0; a < 8; a++) {
                    const int u = 3 * a, v = 3 * a + 1, w = 3 * a + 2;
                    B[0][u] = g.dx[a];
                    B[1][v] = g.dy[a];
                    B[2][w] = g.dz[a];
                    B[3][u] = g.dy[a];
                    B[3][v] = g.dx[a];
                    B[4][v] = g.dz[a];
                    B[4][w] = g.dy[a];
                    B[5][u] = g.dz[a];
                    B[5][w] = g.dx[a];
                }
                double DB[6][24];
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 24; c++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += D[r][k] * B[k][c];
                        DB[r][c] = s;
                    }
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += B[k][i] * DB[k][j];
                        ke[24 * i + j] += g.w * s;
                    }
            }
}

void hex8_mass(double hx, double hy, double hz, double* m8) {
    std::memset(m8, 0, sizeof(double) * 64);
    const double gp[2] = {-kGP, kGP};
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                GaussPoint g = gauss_point(gp[ix], gp[iy], gp[iz], hx, hy, hz);
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++) m8[8 * a + b] += g.w * g.N[a] * g.N[b];
            }
}

void hex8_helmholtz(double hx, double hy, double hz, double R, double* kf) {
    std::memset(kf, 0, sizeof(double) * 64);
    const double q = 1.0 / std::sqrt(3.0);
    const double gp[2] = {-q, q};
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                GaussPoint g = gauss_point(gp[ix], gp[iy], gp[iz], hx, hy, hz);
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++)
                        kf[8 * a + b] += g.w * (R * R * (g.dx[a] * g.dx[b] + g.dy[a] * g.dy[b] + g.dz[a] * g.dz[b]) +
                                                g.N[a] * g.N[b]);
            }
}

}  // namespace tt
