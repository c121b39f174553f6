// Oracle (test infrastructure only): tet_hp_cns restatement (straight tetrahedra,
// steady, p <= 2, Weiss-Smith preconditioner, inflow faces).  See otet_hp.h.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "olinalg.h"
#include "otet_hp.h"

namespace oracle {

namespace {
// faces opposite vertex i (tet_mesh/setup.cpp:305-308)
const int kVF[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
// MAX / MIN of sys/param.h, as the reference uses them
inline double pmax(double a, double b) { return a > b ? a : b; }
inline double pmin(double a, double b) { return a < b ? a : b; }

void matmul5(const double A[25], const double B[25], double C[25]) {
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            double t = 0.0;
            for (int k = 0; k < 5; ++k) t += A[i * 5 + k] * B[k * 5 + j];
            C[i * 5 + j] = t;
        }
}
}  // namespace

TetHpCns::TetHpCns(const TetMesh& mesh, const TetParams& p) : TetHp<5>(mesh, p) {
    tau.assign(m.ntet * 25, 0.0);
    vpc.assign(m.npnt * 25, 0.0);
    epc.assign(m.nseg * 25, 0.0);
    tstep.assign(m.ntet, 0.0);
}

// tet_hp/cns/update.cpp:349-375
void TetHpCns::switch_variables(const double pvu[5], double a[5]) const {
    const double gm1 = prm.gamma - 1.0;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], w = pvu[3], rt = pvu[4];
    const double rho = (pr + prm.atm_pressure) / rt;
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double dpdc[25] = {ke * gm1,          -u * gm1,      -v * gm1,      -w * gm1,      gm1,
                             -u / rho,          1.0 / rho,     0.0,           0.0,           0.0,
                             -v / rho,          0.0,           1.0 / rho,     0.0,           0.0,
                             -w / rho,          0.0,           0.0,           1.0 / rho,     0.0,
                             (gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, -w * gm1 / rho, gm1 / rho};
    double t[5];
    for (int i = 0; i < 5; ++i) {
        t[i] = 0.0;
        for (int j = 0; j < 5; ++j) t[i] += dpdc[i * 5 + j] * a[j];
    }
    for (int i = 0; i < 5; ++i) a[i] = t[i];
}

// |P dpdc A|, |P dpdc B|, |P dpdc C| of setup_preconditioner.cpp:289-400 from their
// eigen-decompositions V diag(|eig|) VINV
void TetHpCns::abs_flux_jacobians(const double pvu[5], double b2, double A[25], double Bm[25], double C[25]) const {
    const double gam = prm.gamma, gm1 = gam - 1.0;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], w = pvu[3], rt = pvu[4];
    const double rho = pr / rt;
    const double c2 = gam * rt;
    const double bb = 1.0 - 2.0 * b2 + b2 * b2;
    const double t1 = std::sqrt(u * u * bb + 4.0 * b2 * c2);
    const double t2 = std::sqrt(v * v * bb + 4.0 * b2 * c2);
    const double t3 = std::sqrt(w * w * bb + 4.0 * b2 * c2);
    {
        const double V[25] = {0.5 * (u * (b2 - 1.0) + t1) * rho, 0.5 * (u * (b2 - 1.0) - t1) * rho, 0, 0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              0.5 * (u * (b2 - 1.0) * gm1 + gm1 * t1) / gam, 0.5 * (u * (b2 - 1.0) * gm1 - gm1 * t1) / gam, 0, 0, 1.0};
        const double eig[5] = {std::fabs(0.5 * (u + u * b2 + t1)), std::fabs(0.5 * (u + u * b2 - t1)), std::fabs(u),
                               std::fabs(u), std::fabs(u)};
        double VI[25] = {1.0 / (rho * t1), 0.5 * (t1 - u * (b2 - 1.0)) / t1, 0, 0, 0,
                         -1.0 / (rho * t1), 0.5 * (u * (b2 - 1.0) + t1) / t1, 0, 0, 0,
                         0, 0, 1.0, 0, 0,
                         0, 0, 0, 1.0, 0,
                         -gm1 / (gam * rho), 0, 0, 0, 1.0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) VI[i * 5 + j] *= eig[i];
        matmul5(V, VI, A);
    }
    {
        const double V[25] = {0.5 * (v * (b2 - 1.0) + t2) * rho, 0.5 * (v * (b2 - 1.0) - t2) * rho, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              0.5 * (gm1 * v * (b2 - 1.0) + gm1 * t2) / gam, 0.5 * (gm1 * v * (b2 - 1.0) - gm1 * t2) / gam, 0, 0, 1.0};
        const double eig[5] = {std::fabs(0.5 * (v + v * b2 + t2)), std::fabs(0.5 * (v + v * b2 - t2)), std::fabs(v),
                               std::fabs(v), std::fabs(v)};
        double VI[25] = {1.0 / (rho * t2), 0, 0.5 * (t2 - v * (b2 - 1.0)) / t2, 0, 0,
                         -1.0 / (rho * t2), 0, 0.5 * (v * (b2 - 1.0) + t2) / t2, 0, 0,
                         0, 1.0, 0, 0, 0,
                         0, 0, 0, 1.0, 0,
                         -gm1 / (gam * rho), 0, 0, 0, 1.0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) VI[i * 5 + j] *= eig[i];
        matmul5(V, VI, Bm);
    }
    {
        // entry (4, 1) uses temp2, not temp3, exactly as tet_hp/cns/setup_preconditioner.cpp:391
        const double V[25] = {0.5 * (w * (b2 - 1.0) + t3) * rho, 0.5 * (w * (b2 - 1.0) - t3) * rho, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0.5 * (gm1 * w * (b2 - 1.0) + gm1 * t3) / gam, 0.5 * (gm1 * w * (b2 - 1.0) - gm1 * t2) / gam, 0, 0, 1.0};
        const double eig[5] = {std::fabs(0.5 * (w + w * b2 + t3)), std::fabs(0.5 * (w + w * b2 - t3)), std::fabs(w),
                               std::fabs(w), std::fabs(w)};
        double VI[25] = {1.0 / (rho * t3), 0, 0, 0.5 * (t3 - w * (b2 - 1.0)) / t3, 0,
                         -1.0 / (rho * t3), 0, 0, 0.5 * (w * (b2 - 1.0) + t3) / t3, 0,
                         0, 1.0, 0, 0, 0,
                         0, 0, 1.0, 0, 0,
                         -gm1 / (rho * gam), 0, 0, 0, 1.0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) VI[i * 5 + j] *= eig[i];
        matmul5(V, VI, C);
    }
}

// tet_hp/cns/setup_preconditioner.cpp:206-489 (steady: bd(0) = 0)
void TetHpCns::preconditioner_tau_timestep(const double pvu[5], double h, double hmax, double Pinv[25],
                                           double Tau[25], double& timestep) const {
    const double gam = prm.gamma, gm1 = gam - 1.0, gogm1 = gam / gm1;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], w = pvu[3], rt = pvu[4];
    const double rho = pr / rt;
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double c2 = gam * rt, c = std::sqrt(c2);
    const double nu = prm.mu / rho;
    const double kcond = prm.R * prm.mu / prm.prandtl * prm.gamma / (prm.gamma - 1.0);
    const double cp = gogm1 * prm.R;
    const double alpha = kcond / (rho * cp);
    const double bd0 = 0.0;
    const double hdt = 0.5 * hmax * bd0 / c;
    const double umag = std::sqrt(u * u + v * v + w * w);
    const double M = pmax(umag / c, 1.0e-5);
    const double nuh = 3.0 * pmax(4.0 * nu / (3.0 * hmax * c), alpha / (hmax * c));
    double b2 = 1.0;
    if (M < 0.8 && prm.preconditioner > 0) b2 = pmin(3.0 * M * M + (hdt + nuh) * (hdt + nuh), 1.0);

    double P[25] = {b2, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, (b2 - 1.0) / (gogm1 * rho), 0, 0, 0, 1};
    const double pinv[25] = {1.0 / b2, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0,
                             -(b2 - 1.0) / (gogm1 * rho * b2), 0, 0, 0, 1};
    for (int k = 0; k < 25; ++k) Pinv[k] = pinv[k];
    const double dpdc[25] = {ke * gm1,          -u * gm1,      -v * gm1,      -w * gm1,      gm1,
                             -u / rho,          1.0 / rho,     0.0,           0.0,           0.0,
                             -v / rho,          0.0,           1.0 / rho,     0.0,           0.0,
                             -w / rho,          0.0,           0.0,           1.0 / rho,     0.0,
                             (gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, -w * gm1 / rho, gm1 / rho};
    const double dcdp[25] = {1.0 / rt, 0, 0, 0, -rho / rt,
                             u / rt, rho, 0, 0, -rho * u / rt,
                             v / rt, 0, rho, 0, -rho * v / rt,
                             w / rt, 0, 0, rho, -rho * w / rt,
                             (rt + gm1 * ke) / (gm1 * rt), rho * u, rho * v, rho * w, -rho * ke / rt};
    double tmp[25];
    matmul5(P, dpdc, tmp);   // P <- P dpdc
    for (int k = 0; k < 25; ++k) P[k] = tmp[k];
    if (prm.preconditioner == 2) {
        matmul5(dcdp, pinv, tmp);
        for (int k = 0; k < 25; ++k) Pinv[k] = tmp[k];
    }

    const double bb = 1.0 - 2.0 * b2 + b2 * b2;
    double A[25], Bm[25], C[25];
    abs_flux_jacobians(pvu, b2, A, Bm, C);
    double S[5] = {0.0, 4.0 * nu / (3.0 * h * h), 4.0 * nu / (3.0 * h * h), 4.0 * nu / (3.0 * h * h), alpha / (h * h)};
    for (int i = 0; i < 5; ++i) S[i] = 3.0 * S[i] + 0.5 * bd0;
    const double maxeig = 0.5 * (umag + umag * b2 + std::sqrt(umag * umag * bb + 4.0 * b2 * c2));
    std::vector<double> Tinv(25);
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            Tinv[i * 5 + j] = 2.0 * (A[i * 5 + j] + Bm[i * 5 + j] + C[i * 5 + j] + (i == j ? hmax * S[i] : 0.0)) / h;
    timestep = pmax(3.0 * maxeig / h, pmax(0.5 * bd0, 3.0 * pmax(4.0 * nu / (3.0 * h * h), alpha / (h * h))));
    // Tau = Tinv^-1 P (the reference factors Tinv with dgetrf and solves the transposed
    // system for P^T, :446-486); here each column of P is solved against Tinv directly
    for (int col = 0; col < 5; ++col) {
        double x[5];
        for (int i = 0; i < 5; ++i) x[i] = P[i * 5 + col];
        lu_solve_inplace(Tinv, 5, 5, x);
        for (int i = 0; i < 5; ++i) Tau[i * 5 + col] = x[i];
    }
}

// flux Jacobians dF/dw, dG/dw, dH/dw with respect to the primitive unknowns
// (p, u, v, w, RT) at one point (tet_hp/cns/rsdl.cpp:799-826); u[n * stride]
void TetHpCns::flux_jacobians(const double* u, int stride, double A[25], double Bq[25], double Cq[25]) const {
    const double gogm1 = prm.gamma / (prm.gamma - 1.0);
    const double pr = u[0], uv = u[stride], vv = u[2 * stride], wv = u[3 * stride], rt = u[4 * stride];
    const double ke = 0.5 * (uv * uv + vv * vv + wv * wv);
    const double rho = (pr + prm.atm_pressure) / rt;
    const double a[25] = {uv / rt, rho, 0.0, 0.0, -rho * uv / rt,
                          uv * uv / rt + 1.0, 2.0 * rho * uv, 0.0, 0.0, -rho * uv * uv / rt,
                          uv * vv / rt, rho * vv, rho * uv, 0.0, -rho * uv * vv / rt,
                          uv * wv / rt, rho * wv, 0.0, rho * uv, -rho * uv * wv / rt,
                          uv * (gogm1 * rt + ke) / rt, rho * (gogm1 * rt + ke) + rho * uv * uv, rho * uv * vv, rho * uv * wv, -rho * uv * ke / rt};
    const double b[25] = {vv / rt, 0.0, rho, 0.0, -rho * vv / rt,
                          uv * vv / rt, rho * vv, rho * uv, 0.0, -rho * uv * vv / rt,
                           vv * vv / rt + 1.0, 0.0, 2.0 * rho * vv, 0.0, -rho * vv * vv / rt,
                           vv * wv / rt, 0.0, rho * wv, rho * vv, -rho * vv * wv / rt,
                           vv * (gogm1 * rt + ke) / rt, rho * uv * vv, rho * (gogm1 * rt + ke) + rho * vv * vv, rho * vv * wv, -rho * vv * ke / rt};
    const double c[25] = {wv / rt, 0.0, 0.0, rho, -rho * wv / rt,
                           uv * wv / rt, rho * wv, 0.0, rho * uv, -rho * uv * wv / rt,
                           vv * wv / rt, 0.0, rho * wv, rho * vv, -rho * vv * wv / rt,
                           wv * wv / rt + 1.0, 0.0, 0.0, 2.0 * rho * wv, -rho * wv * wv / rt,
                           wv * (gogm1 * rt + ke) / rt, rho * uv * wv, rho * vv * wv, rho * (gogm1 * rt + ke) + rho * wv * wv, -rho * wv * ke / rt};
    for (int k = 0; k < 25; ++k) { A[k] = a[k]; Bq[k] = b[k]; Cq[k] = c[k]; }
}

// tet_hp/cns/rsdl.cpp:16-87 and the straight branch :472-843; steady (mvel = 0,
// dugdt = 0, bd(0) = 0); body force and reference density at their defaults (0)
void TetHpCns::element_rsdl(int t, int stage, const double uht[NV][10], double lf_re[NV][10],
                            double lf_im[NV][10]) const {
    const TetBasis& B = b();
    const int G = B.G, T = B.tm;
    const double beta = stage < prm.nstage ? prm.beta[stage] : 1.0;
    const double gam = prm.gamma, gm1 = gam - 1.0, gogm1 = gam / gm1;
    const double lmu = prm.mu, lkcond = prm.R * prm.mu / prm.prandtl * prm.gamma / (prm.gamma - 1.0);
    std::vector<double> u(NV * G), du(NV * 3 * G, 0.0), cv(NV * 3 * G), df(NV * 3 * G), res(NV * G);
    auto U = [&](int n, int g) -> double& { return u[n * G + g]; };
    auto DU = [&](int n, int a, int g) -> double& { return du[(n * 3 + a) * G + g]; };
    auto CV = [&](int n, int a, int g) -> double& { return cv[(n * 3 + a) * G + g]; };
    auto DF = [&](int n, int a, int g) -> double& { return df[(n * 3 + a) * G + g]; };
    const double* P[4];
    for (int i = 0; i < 4; ++i) P[i] = &m.pnts[3 * m.tet_pnt[4 * t + i]];
    double ld[3][3];
    for (int n = 0; n < 3; ++n) {
        ld[n][0] = 0.5 * (P[3][n] - P[2][n]);
        ld[n][1] = 0.5 * (P[1][n] - P[2][n]);
        ld[n][2] = 0.5 * (P[0][n] - P[2][n]);
    }
    if (beta > 0.0) {
        B.proj(uht[0], &U(0, 0));
        for (int n = 1; n < NV; ++n) B.proj(uht[n], &U(n, 0), &DU(n, 0, 0), &DU(n, 1, 0), &DU(n, 2, 0));
    } else {
        for (int n = 0; n < NV; ++n) B.proj(uht[n], &U(n, 0));
    }
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < T; ++i) { lf_re[n][i] = 0.0; lf_im[n][i] = 0.0; }
    double d[3][3];
    d[0][0] = ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1];
    d[0][1] = -ld[0][1] * ld[2][2] + ld[0][2] * ld[2][1];
    d[0][2] = ld[0][1] * ld[1][2] - ld[0][2] * ld[1][1];
    d[1][0] = -ld[1][0] * ld[2][2] + ld[1][2] * ld[2][0];
    d[1][1] = ld[0][0] * ld[2][2] - ld[0][2] * ld[2][0];
    d[1][2] = -ld[0][0] * ld[1][2] + ld[0][2] * ld[1][0];
    d[2][0] = ld[1][0] * ld[2][1] - ld[1][1] * ld[2][0];
    d[2][1] = -ld[0][0] * ld[2][1] + ld[0][1] * ld[2][0];
    d[2][2] = ld[0][0] * ld[1][1] - ld[0][1] * ld[1][0];
    const double cjcb = ld[0][0] * (ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1]) -
                        ld[0][1] * (ld[1][0] * ld[2][2] - ld[2][0] * ld[1][2]) +
                        ld[0][2] * (ld[1][0] * ld[2][1] - ld[2][0] * ld[1][1]);
    // convective, pressure and energy fluxes (:494-540)
    for (int g = 0; g < G; ++g) {
        const double rho = (U(0, g) + prm.atm_pressure) / U(4, g);
        const double fx = rho * U(1, g), fy = rho * U(2, g), fz = rho * U(3, g);
        for (int a = 0; a < 3; ++a) CV(0, a, g) = d[a][0] * fx + d[a][1] * fy + d[a][2] * fz;
        for (int n = 1; n < NV - 1; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) = U(n, g) * CV(0, a, g);
        for (int n = 1; n < 4; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) += d[a][n - 1] * U(0, g);
        const double hh = gogm1 * U(4, g) + 0.5 * (U(1, g) * U(1, g) + U(2, g) * U(2, g) + U(3, g) * U(3, g));
        for (int a = 0; a < 3; ++a) CV(4, a, g) = hh * CV(0, a, g);
    }
    for (int n = 0; n < NV; ++n) B.intgrtrst(lf_im[n], &CV(n, 0, 0), &CV(n, 1, 0), &CV(n, 2, 0));
    if (!(beta > 0.0)) return;

    // viscous tensor visc(n,m)(a,b) and heat-diffusion tensor (:584-689)
    const double mujcbi = lmu / cjcb, kcjcbi = lkcond / cjcb / prm.R;
    double visc[3][3][3][3];
    for (int n = 0; n < 3; ++n)
        for (int mm = 0; mm < 3; ++mm)
            for (int a = 0; a < 3; ++a)
                for (int bb = 0; bb < 3; ++bb) {
                    double s;
                    if (n == mm) {
                        s = 4. / 3. * d[a][n] * d[bb][n];
                        for (int c = 0; c < 3; ++c)
                            if (c != n) s += d[a][c] * d[bb][c];
                    } else {
                        s = d[a][mm] * d[bb][n] - 2. / 3. * d[a][n] * d[bb][mm];
                    }
                    visc[n][mm][a][bb] = -mujcbi * s;
                }
    double kc[3][3];
    for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 3; ++bb) kc[a][bb] = -kcjcbi * (d[a][0] * d[bb][0] + d[a][1] * d[bb][1] + d[a][2] * d[bb][2]);
    for (int g = 0; g < G; ++g) {
        for (int n = 0; n < NV; ++n) res[n * G + g] = 0.0;  // rhorbd0 = 0, dugdt = 0, body = 0
        for (int n = 0; n < 3; ++n)
            for (int a = 0; a < 3; ++a) {
                double s = 0.0;
                for (int bb = 0; bb < 3; ++bb)
                    for (int mm = 0; mm < 3; ++mm) s += visc[n][mm][a][bb] * DU(mm + 1, bb, g);
                DF(n + 1, a, g) = s;
            }
        for (int a = 0; a < 3; ++a) {
            DF(4, a, g) = kc[a][0] * DU(4, 0, g) + kc[a][1] * DU(4, 1, g) + kc[a][2] * DU(4, 2, g);
            DF(4, a, g) += DF(1, a, g) * U(1, g) + DF(2, a, g) * U(2, g) + DF(3, a, g) * U(3, g);
        }
        for (int n = 1; n < NV; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) += DF(n, a, g);
    }
    for (int n = 0; n < NV; ++n) B.intgrt(lf_re[n], &res[n * G]);
    for (int n = 0; n < NV; ++n) {
        B.derivr(&CV(n, 0, 0), &res[n * G]);
        B.derivs(&CV(n, 1, 0), &res[n * G]);
        B.derivt(&CV(n, 2, 0), &res[n * G]);
    }
    // stabilisation with the element tau (:776-827)
    const double* tt = &tau[25 * t];
    for (int g = 0; g < G; ++g) {
        for (int a = 0; a < 3; ++a) DF(0, a, g) = 0.0;
        double tres[5] = {0, 0, 0, 0, 0};
        for (int mm = 0; mm < 5; ++mm)
            for (int n = 0; n < 5; ++n) tres[mm] += tt[mm * 5 + n] * res[n * G + g];
        double A[25], Bq[25], Cq[25];
        flux_jacobians(&U(0, g), G, A, Bq, Cq);
        for (int mm = 0; mm < 5; ++mm)
            for (int n = 0; n < 5; ++n) {
                const int k = mm * 5 + n;
                DF(mm, 0, g) -= (d[0][0] * A[k] + d[0][1] * Bq[k] + d[0][2] * Cq[k]) * tres[n];
                DF(mm, 1, g) -= (d[1][0] * A[k] + d[1][1] * Bq[k] + d[1][2] * Cq[k]) * tres[n];
                DF(mm, 2, g) -= (d[2][0] * A[k] + d[2][1] * Bq[k] + d[2][2] * Cq[k]) * tres[n];
            }
    }
    for (int n = 0; n < NV; ++n) B.intgrtrst(lf_re[n], &DF(n, 0, 0), &DF(n, 1, 0), &DF(n, 2, 0));
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < T; ++i) lf_re[n][i] *= beta;
}

// inflow face (tet_hp/cns/bdry.cpp:58-104, 129-150): pressure from the solution, velocity
// and RT from the boundary values; n = dx/dr x dx/ds at every face Gauss point
void TetHpCns::face_rsdl(int f, double lf[NV][6]) const {
    const TetBasis& B = b();
    const int GF = B.gpx * B.gpy, TF = B.tm2d;
    std::vector<double> crd(3 * GF), dcr(3 * GF), dcs(3 * GF), pr(GF), flx(NV * GF);
    double cht[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 3; ++n) {
        for (int j = 0; j < 3; ++j) cht[j] = m.pnts[3 * m.bf_pnt[3 * f + j] + n];
        B.proj2d(cht, &crd[n * GF], &dcr[n * GF], &dcs[n * GF]);
    }
    double uh[6];
    for (int j = 0; j < 3; ++j) uh[j] = ugv[m.bf_pnt[3 * f + j] * NV + 0];
    for (int j = 0; j < 3; ++j) {
        const int e = m.bf_seg[3 * f + j], sg = m.bf_sgn[3 * f + j];
        int msgn = 1;
        for (int k = 0; k < B.em; ++k) {
            uh[3 + j * B.em + k] = msgn * uge[(e * em0 + k) * NV + 0];
            msgn *= sg;
        }
    }
    B.proj2d(uh, pr.data());
    const double vb[3] = {prm.bc[1], prm.bc[2], prm.bc[3]}, RT = prm.bc[4];
    const double hh = prm.gamma / (prm.gamma - 1.0) * RT + 0.5 * (vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
    for (int g = 0; g < GF; ++g) {
        const double v1[3] = {dcr[0 * GF + g], dcr[1 * GF + g], dcr[2 * GF + g]};
        const double v2[3] = {dcs[0 * GF + g], dcs[1 * GF + g], dcs[2 * GF + g]};
        const double nrm[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
        const double rho = (pr[g] + prm.atm_pressure) / RT;
        const double f0 = rho * (vb[0] * nrm[0] + vb[1] * nrm[1] + vb[2] * nrm[2]);
        flx[0 * GF + g] = f0;
        for (int n = 1; n < 4; ++n) flx[n * GF + g] = f0 * vb[n - 1] + pr[g] * nrm[n - 1];
        flx[4 * GF + g] = hh * f0;
    }
    for (int n = 0; n < NV; ++n) {
        double r[6];
        B.intgrt2d(r, &flx[n * GF]);
        for (int i = 0; i < TF; ++i) lf[n][i] = r[i];
    }
}

// tet_hp/cns/setup_preconditioner.cpp:9-203 (non-TIMEACCURATE) and
// tet_hp/setup_preconditioner.cpp:428-496 (diagonal part inverted).  The reference
// aborts on a non-positive size or a NaN; here they count as errors.
int TetHpCns::setup_preconditioner() {
    const TetBasis& B = b();
    const int G = B.G;
    int err = 0;
    std::fill(vprcn.begin(), vprcn.end(), 0.0);
    std::fill(eprcn.begin(), eprcn.end(), 0.0);
    std::fill(vpc.begin(), vpc.end(), 0.0);
    std::fill(epc.begin(), epc.end(), 0.0);
    std::vector<double> u(NV * G);
    double uht[NV][10];
    for (int t = 0; t < m.ntet; ++t) {
        const double jcb = 0.125 * vol[t];
        ugtouht(t, uht);
        for (int n = 0; n < NV; ++n) B.proj(uht[n], &u[n * G]);
        double umax[5] = {0, 0, 0, 0, 0}, qmax = 0.0;
        for (int g = 0; g < G; ++g) {
            const double q = std::pow(u[1 * G + g], 2.0) + std::pow(u[2 * G + g], 2.0) + std::pow(u[3 * G + g], 2.0);
            qmax = pmax(qmax, q);
            umax[0] = pmax(umax[0], std::fabs(u[0 * G + g] + prm.atm_pressure));
            for (int n = 1; n < 5; ++n) umax[n] = pmax(umax[n], std::fabs(u[n * G + g]));
        }
        double amin = 1.0e99, amax = 0.0;
        for (int j = 0; j < 4; ++j) {
            const double* A = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][0]]];
            const double* Bp = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][1]]];
            const double* C = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][2]]];
            const double v1[3] = {A[0] - Bp[0], A[1] - Bp[1], A[2] - Bp[2]};
            const double v2[3] = {A[0] - C[0], A[1] - C[1], A[2] - C[2]};
            const double c[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
            const double a = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            amax = (a > amax ? a : amax);
            amin = (a < amin ? a : amin);
        }
        const double pp = (B.p + 1) * (B.p + 1);
        const double hmin = 4.0 * jcb / (0.25 * pp * amax);
        const double hmax = 4.0 * jcb / (0.25 * pp * amin);
        if (!(hmin > 0.0) || std::isnan(qmax)) {
            err = 1;
            continue;
        }
        double Pinv[25], Tau[25];
        preconditioner_tau_timestep(umax, hmin, hmax, Pinv, Tau, tstep[t]);
        for (int k = 0; k < 25; ++k) tau[25 * t + k] = prm.adis * Tau[k] / jcb;
        for (int n = 0; n < NV; ++n) iprcn[NV * t + n] = jcb;
        for (int i = 0; i < 4; ++i) {
            const int v = m.tet_pnt[4 * t + i];
            for (int n = 0; n < NV; ++n) vprcn[v * NV + n] += jcb;
            for (int k = 0; k < 25; ++k) vpc[25 * v + k] += Pinv[k] * tstep[t] * jcb;
        }
        if (B.em > 0)
            for (int i = 0; i < 6; ++i) {
                const int s = m.tet_seg[6 * t + i];
                for (int n = 0; n < NV; ++n) eprcn[s * NV + n] += jcb;
                for (int k = 0; k < 25; ++k) epc[25 * s + k] += Pinv[k] * tstep[t] * jcb;
            }
    }
    for (int v = 0; v < m.npnt; ++v)
        for (int k = 0; k < 25; ++k) vpc[25 * v + k] /= vprcn[v * NV + 0];
    if (B.em > 0)
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < 25; ++k) epc[25 * s + k] /= eprcn[s * NV + 0];
    for (auto& x : vprcn) x = 1.0 / x;
    if (B.em > 0)
        for (auto& x : eprcn) x = 1.0 / x;
    return err;
}

// inflow: u, v, w, RT held (bdry_cns.h:70-98, bdry.cpp:205-222)
void TetHpCns::vdirichlet() {
    for (int v = 0; v < m.npnt; ++v)
        if (vess[v])
            for (int n = 1; n < NV; ++n) resv[v * NV + n] = 0.0;
}

void TetHpCns::edirichlet() {
    for (int s = 0; s < m.nseg; ++s)
        if (eess[s])
            for (int k = 0; k < em0; ++k)
                for (int n = 1; n < NV; ++n) rese[(s * em0 + k) * NV + n] = 0.0;
}

// tet_hp/cns/update.cpp:146-347, Weiss-Smith branch (preconditioner 0 or 1)
void TetHpCns::minvrt() {
    const TetBasis& B = b();
    if (prm.preconditioner == 2) throw std::runtime_error("oracle TetHpCns: the squared preconditioner is not restated");
    if (B.em > 0) {
        for (int e = 0; e < m.nseg; ++e)
            for (int k = 0; k < B.em; ++k)
                for (int i = 0; i < 2; ++i) {
                    const int v0 = m.seg_pnt[2 * e + i];
                    for (int n = 0; n < NV; ++n) resv[v0 * NV + n] -= B.sfmv[i * B.em + k] * rese[(e * em0 + k) * NV + n];
                }
    }
    for (int v = 0; v < m.npnt; ++v)
        for (int n = 0; n < NV; ++n) resv[v * NV + n] *= vprcn[v * NV + n] * B.vdiag;
    auto lower_solve = [](const double* Pm, double* r) {
        for (int j = 0; j < 5; ++j) {
            double l0 = r[j];
            for (int k = 0; k < j; ++k) l0 -= Pm[j * 5 + k] * r[k];
            r[j] = l0 / Pm[j * 5 + j];
        }
    };
    for (int v = 0; v < m.npnt; ++v) {
        double r[5], ug[5];
        for (int n = 0; n < NV; ++n) { r[n] = resv[v * NV + n]; ug[n] = ugv[v * NV + n]; }
        switch_variables(ug, r);
        lower_solve(&vpc[25 * v], r);
        for (int n = 0; n < NV; ++n) resv[v * NV + n] = r[n];
    }
    vdirichlet();
    if (B.em == 0) return;
    for (int s = 0; s < m.nseg; ++s) {
        double r[5], ua[5];
        const int p0 = m.seg_pnt[2 * s], p1 = m.seg_pnt[2 * s + 1];
        for (int n = 0; n < NV; ++n) {
            r[n] = rese[(s * em0 + 0) * NV + n];
            ua[n] = 0.5 * (ugv[p0 * NV + n] + ugv[p1 * NV + n]);
        }
        switch_variables(ua, r);
        lower_solve(&epc[25 * s], r);
        for (int n = 0; n < NV; ++n) rese[(s * em0 + 0) * NV + n] = r[n];
    }
    for (int t = 0; t < m.ntet; ++t) {
        double uh[4][NV];
        for (int i = 0; i < 4; ++i)
            for (int n = 0; n < NV; ++n) uh[i][n] = resv[m.tet_pnt[4 * t + i] * NV + n] * iprcn[NV * t + n];
        for (int i = 0; i < 6; ++i) {
            const int e = m.tet_seg[6 * t + i], sg = m.tet_sgn[6 * t + i];
            for (int j = 0; j < 4; ++j) {
                int msgn = 1;
                for (int k = 0; k < B.em; ++k) {
                    for (int n = 0; n < NV; ++n)
                        rese[(e * em0 + k) * NV + n] -= msgn * B.vfms[j * B.tm + 4 + k + i * B.em] * uh[j][n];
                    msgn *= sg;
                }
            }
        }
    }
    // ediag(0) = 100 (:320-322)
    for (int e = 0; e < m.nseg; ++e)
        for (int n = 0; n < NV; ++n) rese[(e * em0 + 0) * NV + n] *= eprcn[e * NV + n] * 100.0;
    edirichlet();
}

// tet_hp/cns/update.cpp:18-143 (no Dirichlet pass before minvrt, no variable projection)
void TetHpCns::update() {
    const TetBasis& B = b();
    ug0v = ugv;
    ug0e = uge;
    for (int stage = 0; stage < prm.nstage; ++stage) {
        rsdl(stage);
        minvrt();
        const double cfa = prm.alpha[stage] * prm.cfl[log2p];
        for (int i = 0; i < m.npnt * NV; ++i) ugv[i] = ug0v[i] - cfa * resv[i];
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.em; ++k)
                for (int n = 0; n < NV; ++n) {
                    const int ix = (s * em0 + k) * NV + n;
                    uge[ix] = ug0e[ix] - cfa * rese[ix];
                }
    }
}

}  // namespace oracle
