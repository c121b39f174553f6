// Host-side construction of the tetrahedral basis operators; see tet_ops.h.
#include "tet_ops.h"

#include <cmath>
#include <stdexcept>

#include "tri_ops.h"

namespace hpg {

namespace {
const int kVS[6][2] = {{2, 3}, {3, 1}, {2, 1}, {1, 0}, {2, 0}, {3, 0}};

// derivative matrix of the 1-D Lagrange interpolant on nodes x: L[i][j] = l_j'(x_i)
std::vector<double> lagrange_derivative(const std::vector<double>& x) {
    const int n = (int)x.size();
    std::vector<double> w(n, 1.0), L(n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) w[j] /= (x[j] - x[k]);
    for (int i = 0; i < n; ++i) {
        double diag = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            L[i * n + j] = (w[j] / w[i]) / (x[i] - x[j]);
            diag += 1.0 / (x[i] - x[j]);
        }
        L[i * n + i] = diag;
    }
    return L;
}

// barycentrics (l0, l1, l2, l3) and their constant (r, s, t) gradients
void bary(double r, double s, double t, double l[4]) {
    l[0] = 0.5 * (1 + t);
    l[1] = 0.5 * (1 + s);
    l[3] = 0.5 * (1 + r);
    l[2] = 1.0 - l[0] - l[1] - l[3];
}
const double kGrad[4][3] = {{0, 0, .5}, {0, .5, 0}, {-.5, -.5, -.5}, {.5, 0, 0}};

void modes(int tm, const double l[4], double* val, double (*grad)[3]) {
    for (int v = 0; v < 4; ++v) {
        val[v] = l[v];
        for (int a = 0; a < 3; ++a) grad[v][a] = kGrad[v][a];
    }
    if (tm > 4)
        for (int e = 0; e < 6; ++e) {
            const int A = kVS[e][0], B = kVS[e][1];
            val[4 + e] = l[A] * l[B];
            for (int a = 0; a < 3; ++a) grad[4 + e][a] = l[A] * kGrad[B][a] + l[B] * kGrad[A][a];
        }
}
}  // namespace

TetOps::TetOps(int p_) : p(p_) {
    if (p != 1 && p != 2) throw std::runtime_error("tet_ops: p must be 1 or 2");
    em = p - 1;
    tm = (p + 1) * (p + 2) * (p + 3) / 6;
    gp = p + 1;
    G = gp * gp * gp;
    tm2d = 3 + 3 * em;
    GF = gp * gp;
    std::vector<double> x(gp), wx(gp), y(gp), wy(gp), z(gp), wz(gp);
    gauss_nodes(gp, 0.0, 0.0, x.data(), wx.data());
    gauss_nodes(gp, 1.0, 0.0, y.data(), wy.data());
    gauss_nodes(gp, 2.0, 0.0, z.data(), wz.data());
    for (auto& w : wy) w *= 0.5;
    for (auto& w : wz) w *= 0.25;
    P.assign(G * tm, 0.0); Pr = Ps = Pt = P;
    Ir.assign(tm * G, 0.0); Is = It = Ir;
    std::vector<double> W(G);
    double val[10], grad[10][3], l[4];
    for (int i = 0; i < gp; ++i)
        for (int j = 0; j < gp; ++j)
            for (int k = 0; k < gp; ++k) {
                const int g = (i * gp + j) * gp + k;
                const double t = z[k], s = 0.5 * (1 + y[j]) * (1 - t) - 1, r = 0.5 * (1 + x[i]) * (-s - t) - 1;
                bary(r, s, t, l);
                modes(tm, l, val, grad);
                W[g] = wx[i] * wy[j] * wz[k];
                for (int m = 0; m < tm; ++m) {
                    P[g * tm + m] = val[m];
                    Pr[g * tm + m] = grad[m][0];
                    Ps[g * tm + m] = grad[m][1];
                    Pt[g * tm + m] = grad[m][2];
                    Ir[m * G + g] = -W[g] * grad[m][0];
                    Is[m * G + g] = -W[g] * grad[m][1];
                    It[m * G + g] = -W[g] * grad[m][2];
                }
            }
    // Gauss-grid derivatives: d/dr = y0 z0 d/dx, d/ds = x0 y0 z0 d/dx + z0 d/dy,
    // d/dt = x0 y0 z0 d/dx + y1 z0 d/dy + d/dz (x0 = (1+x)/2, y0 = 2/(1-y),
    // y1 = (1+y)/2, z0 = 2/(1-z))
    const std::vector<double> Lx = lagrange_derivative(x), Ly = lagrange_derivative(y), Lz = lagrange_derivative(z);
    Dr.assign(G * G, 0.0); Ds = Dt = Dr;
    for (int i = 0; i < gp; ++i)
        for (int j = 0; j < gp; ++j)
            for (int k = 0; k < gp; ++k) {
                const int g = (i * gp + j) * gp + k;
                const double x0 = 0.5 * (1 + x[i]), y0 = 2.0 / (1 - y[j]), y1 = 0.5 * (1 + y[j]), z0 = 2.0 / (1 - z[k]);
                for (int n = 0; n < gp; ++n) {
                    const int gx = (n * gp + j) * gp + k, gy = (i * gp + n) * gp + k, gz = (i * gp + j) * gp + n;
                    const double dx = Lx[i * gp + n], dy = Ly[j * gp + n], dz = Lz[k * gp + n];
                    Dr[g * G + gx] += y0 * z0 * dx;
                    Ds[g * G + gx] += x0 * y0 * z0 * dx;
                    Ds[g * G + gy] += z0 * dy;
                    Dt[g * G + gx] += x0 * y0 * z0 * dx;
                    Dt[g * G + gy] += y1 * z0 * dy;
                    Dt[g * G + gz] += dz;
                }
            }
    // face 0 grid (x, y): V1 = (1+s)/2, V3 = (1+r)/2, V2 = 1 - V1 - V3
    P2.assign(GF * tm2d, 0.0);
    W2.assign(GF, 0.0);
    for (int i = 0; i < gp; ++i)
        for (int j = 0; j < gp; ++j) {
            const int g = i * gp + j;
            const double s = y[j], r = 0.5 * (1 + x[i]) * (1 - s) - 1;
            const double a1 = 0.5 * (1 + s), a3 = 0.5 * (1 + r), a2 = 1.0 - a1 - a3;
            const double fv[6] = {a1, a2, a3, a2 * a3, a1 * a3, a1 * a2};
            for (int m = 0; m < tm2d; ++m) P2[g * tm2d + m] = fv[m];
            W2[g] = wx[i] * wy[j];
        }
    // mass matrix and lumped factors
    mass.assign(tm * tm, 0.0);
    for (int a = 0; a < tm; ++a)
        for (int b = 0; b < tm; ++b) {
            double sm = 0.0;
            for (int g = 0; g < G; ++g) sm += W[g] * P[g * tm + a] * P[g * tm + b];
            mass[a * tm + b] = sm;
        }
    if (p == 1) {
        vdiag = 1.0 / (mass[0] + mass[1] + mass[2] + mass[3]);
        return;
    }
    // vertex 3 made mass-orthogonal to vertices 0, 1, 2 by its edges (2,3), (3,1), (3,0)
    const int dof[4] = {3, 4, 5, 9};
    std::vector<double> A(16, 0.0);
    double rhs[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) A[j * 4 + k] = mass[j * tm + dof[k]];
    for (int k = 0; k < 4; ++k) A[3 * 4 + k] = 0.0;
    A[3 * 4 + 0] = 1.0;
    rhs[3] = 1.0;
    // plain elimination with partial pivoting
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(A[r * 4 + c]) > std::fabs(A[piv * 4 + c])) piv = r;
        for (int k = 0; k < 4; ++k) std::swap(A[c * 4 + k], A[piv * 4 + k]);
        std::swap(rhs[c], rhs[piv]);
        for (int r = c + 1; r < 4; ++r) {
            const double f = A[r * 4 + c] / A[c * 4 + c];
            for (int k = c; k < 4; ++k) A[r * 4 + k] -= f * A[c * 4 + k];
            rhs[r] -= f * rhs[c];
        }
    }
    double xs[4];
    for (int r = 3; r >= 0; --r) {
        double sm = rhs[r];
        for (int k = r + 1; k < 4; ++k) sm -= A[r * 4 + k] * xs[k];
        xs[r] = sm / A[r * 4 + r];
    }
    sfmv = {-xs[2], -xs[1]};   // (first end, second end) of an edge
    vdiag = 1.0 / (mass[0] - sfmv[1] * (mass[7 * tm] + mass[8 * tm] + mass[9 * tm]));
    vfms.assign(4 * tm, 0.0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < tm; ++j) vfms[i * tm + j] = mass[i * tm + j];
    ediag = {100.0};
}

}  // namespace hpg
