// Host-side element matrices.
//
// FP64 mode: the hex8 brick stiffness in closed form.  For a brick the
// bilinear form  int C_ikjl dN_a/dx_k dN_b/dx_l  factorises into products of
// 1-D 2x2 integrals (mass h/6 [2 1; 1 2], stiffness 1/h [1 -1; -1 1], mixed
// [-1/2 -1/2; 1/2 1/2]), and the reference's 2x2x2 Gauss rule
// (R/src/Core/Element.jl:73-111) integrates it exactly, so the two agree to
// rounding.
//
// F32-widened mode: the reference evaluates the same Gauss rule in Float32
// and widens (GPUSolver.jl:248-249); this mode repeats that evaluation in
// float (the reference's BLAS summation order is not reproducible, so the
// two agree to float rounding).
#include <cmath>
#include <cstring>

#include "ke.h"

namespace hexa {
namespace {

const int CN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// A_kl[a][b] = int dN_a/dx_k dN_b/dx_l over the brick h.
void grad_pair(const double h[3], int k, int l, double A[8][8]) {
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double v = 1.0;
            for (int d = 0; d < 3; ++d) {
                const int ia = CN[a][d], ib = CN[b][d];
                double f;
                if (d == k && d == l) f = (ia == ib ? 1.0 : -1.0) / h[d];
                else if (d == k) f = ia ? 0.5 : -0.5;             // int phi'_ia phi_ib
                else if (d == l) f = ib ? 0.5 : -0.5;             // int phi_ia phi'_ib
                else f = (ia == ib ? 2.0 : 1.0) * h[d] / 6.0;
                v *= f;
            }
            A[a][b] = v;
        }
}

void ke_closed_form(double dx, double dy, double dz, double nu, double* Ke) {
    const double h[3] = {dx, dy, dz};
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    static thread_local double A[3][3][8][8];
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) grad_pair(h, k, l, A[k][l]);
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            const double lap = A[0][0][a][b] + A[1][1][a][b] + A[2][2][a][b];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    Ke[(3 * a + i) * 24 + 3 * b + j] =
                        lam * A[i][j][a][b] + mu * A[j][i][a][b] + (i == j ? mu * lap : 0.0);
        }
}

void ke_gauss_f32(double dxd, double dyd, double dzd, double nud, double* Ke) {
    const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd, nu = (float)nud;
    float X[8][3];
    for (int i = 0; i < 8; ++i) {
        X[i][0] = (CN[i][0] ? dx : 0.0f) - dx / 2;
        X[i][1] = (CN[i][1] ? dy : 0.0f) - dy / 2;
        X[i][2] = (CN[i][2] ? dz : 0.0f) - dz / 2;
    }
    const float fac = 1.0f / ((1.0f + nu) * (1.0f - 2.0f * nu));
    const float c1 = (1.0f - nu) * fac, c2 = nu * fac, c3 = ((1.0f - 2.0f * nu) / 2.0f) * fac;
    float D[6][6] = {};
    D[0][0] = D[1][1] = D[2][2] = c1;
    D[0][1] = D[0][2] = D[1][0] = D[1][2] = D[2][0] = D[2][1] = c2;
    D[3][3] = D[4][4] = D[5][5] = c3;
    const float gp = 1.0f / std::sqrt(3.0f);
    const float G[2] = {-gp, gp};
    float K[24][24] = {};
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q)
            for (int r = 0; r < 2; ++r) {
                const float xi = G[p], eta = G[q], zeta = G[r];
                float dN[8][3];
                for (int i = 0; i < 8; ++i) {
                    const float sx = CN[i][0] ? 1.0f : -1.0f, sy = CN[i][1] ? 1.0f : -1.0f,
                                sz = CN[i][2] ? 1.0f : -1.0f;
                    const float tx = 1.0f + xi * sx, ty = 1.0f + eta * sy, tz = 1.0f + zeta * sz;
                    dN[i][0] = 0.125f * sx * ty * tz;
                    dN[i][1] = 0.125f * tx * sy * tz;
                    dN[i][2] = 0.125f * tx * ty * sz;
                }
                float J[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        float s = 0;
                        for (int i = 0; i < 8; ++i) s += dN[i][a] * X[i][b];
                        J[a][b] = s;
                    }
                const float det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                float Ji[3][3];
                Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
                Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
                Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
                Ji[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
                Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
                Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
                Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
                Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
                Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
                float B[6][24] = {};
                for (int i = 0; i < 8; ++i) {
                    float g[3];
                    for (int a = 0; a < 3; ++a) g[a] = dN[i][0] * Ji[a][0] + dN[i][1] * Ji[a][1] + dN[i][2] * Ji[a][2];
                    const int o = 3 * i;
                    B[0][o] = g[0]; B[1][o + 1] = g[1]; B[2][o + 2] = g[2];
                    B[3][o] = g[1]; B[3][o + 1] = g[0];
                    B[4][o + 1] = g[2]; B[4][o + 2] = g[1];
                    B[5][o] = g[2]; B[5][o + 2] = g[0];
                }
                float BtD[24][6];
                for (int a = 0; a < 24; ++a)
                    for (int c = 0; c < 6; ++c) {
                        float s = 0;
                        for (int t = 0; t < 6; ++t) s += B[t][a] * D[t][c];
                        BtD[a][c] = s;
                    }
                for (int a = 0; a < 24; ++a)
                    for (int b = 0; b < 24; ++b) {
                        float s = 0;
                        for (int t = 0; t < 6; ++t) s += BtD[a][t] * B[t][b];
                        K[a][b] += s * det;
                    }
            }
    for (int a = 0; a < 24; ++a)
        for (int b = 0; b < 24; ++b) Ke[a * 24 + b] = (double)K[a][b];
}

}  // namespace

bool element_stiffness(double dx, double dy, double dz, double nu, int mode, double* Ke) {
    if (!(dx > 0 && dy > 0 && dz > 0) || !(nu > -1.0 && nu < 0.5)) return false;
    if (mode == HEXA_KE_F32_WIDENED) ke_gauss_f32(dx, dy, dz, nu, Ke);
    else ke_closed_form(dx, dy, dz, nu, Ke);
    return true;
}

void helmholtz_element(double dx, double dy, double dz, int bug_compat, double* K8, double* M8) {
    // The reference's node table puts the top face at z=+dz over a bottom
    // face at -dz/2 (Element.jl:115-116): a brick of height 1.5*dz.
    const double h[3] = {dx, dy, bug_compat ? 1.5 * dz : dz};
    double A[8][8];
    std::memset(K8, 0, sizeof(double) * 64);
    for (int k = 0; k < 3; ++k) {
        grad_pair(h, k, k, A);
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) K8[a * 8 + b] += A[a][b];
    }
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double v = 1.0;
            for (int d = 0; d < 3; ++d) v *= (CN[a][d] == CN[b][d] ? 2.0 : 1.0) * h[d] / 6.0;
            M8[a * 8 + b] = v;
        }
}

}  // namespace hexa
