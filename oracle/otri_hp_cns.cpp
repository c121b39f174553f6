// Oracle (test infrastructure only).  See otri_hp_cns.h for the reference map.
#include "otri_hp_cns.h"
#include "olinalg.h"
#include <cmath>
#include <algorithm>
#include <stdexcept>

namespace oracle {

static const int S = 6;  // MXGP stride of the reference work arrays (tri_hp/tri_hp.h:31-33)
static inline double MAXM(double a, double b) { return a > b ? a : b; }  // sys/param.h MAX
static inline double MINM(double a, double b) { return a < b ? a : b; }  // sys/param.h MIN

TriHpCns::TriHpCns(const Mesh& mesh, const Params& p) : TriHp<4>(mesh, p) {
    // tri_hp/cns/allocate.cpp:45: Pr = mu/(k/cp), cp = gamma/(gamma-1) R
    kcond = prm.mu / prm.prandtl * prm.Rgas * prm.gamma / (prm.gamma - 1.0);
    vpc.assign(m.npnt * 16, 0.0); spc.assign(m.nseg * 16, 0.0); tpc.assign(m.ntri * 16, 0.0);
    res_tv = resv; res_ts = ress; res_ti = resi;
}

// utilities/matrix.cpp:16-37: 20 power iterations from x = 1, then the Rayleigh quotient
double spectral_radius4(const double A[4][4]) {
    double x[4] = {1, 1, 1, 1}, temp[4] = {0, 0, 0, 0};
    for (int it = 0; it < 20; ++it) {
        for (int i = 0; i < 4; ++i) {
            temp[i] = 0.0;
            for (int j = 0; j < 4; ++j) temp[i] += A[i][j] * x[j];
        }
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += temp[i] * temp[i];
        const double nrm = std::pow(s, 0.5);
        for (int i = 0; i < 4; ++i) x[i] = temp[i] / nrm;
    }
    double num = 0.0, den = 0.0;
    for (int i = 0; i < 4; ++i) num += x[i] * temp[i];
    for (int i = 0; i < 4; ++i) den += x[i] * x[i];
    return std::fabs(num / den);
}

static void matmul4(const double A[4][4], const double B[4][4], double C[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double t = 0.0;
            for (int k = 0; k < 4; ++k) t += A[i][k] * B[k][j];
            C[i][j] = t;
        }
}

// tri_hp/cns/setup_preconditioner.cpp:241-515 (steady: hdt = 0; alph forced to 0 at :353)
void TriHpCns::calculate_tau(const double pvu[4], double h, double Pinv[4][4], double Tau[4][4], double& timestep) const {
    const double gam = prm.gamma, gm1 = gam - 1.0, gogm1 = gam / gm1;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], rt = pvu[3];
    const double rho = pr / rt, ke = 0.5 * (u * u + v * v), c2 = gam * rt, c = std::sqrt(c2);
    const double nu = prm.mu / rho, cp = gogm1 * prm.Rgas, alpha = kcond / (rho * cp);
    const double hdt = 0.0;
    const double umag = std::sqrt(u * u + v * v);
    const double M = MAXM(umag / c, 1.0e-5);
    const double nuh = 4.0 * nu / (h * c);
    const double alh = 2.0 * alpha / (h * c);
    double b2;
    if (M < 0.8)
        b2 = MINM(M * M / (1.0 - M * M) + hdt * hdt + nuh * nuh + alh * alh, 1.0);
    else
        b2 = 1.0;
    const double alph = 0.0;

    double P[4][4] = {{b2, 0, 0, 0},
                      {-alph * u / (pr * gam), 1, 0, 0},
                      {-alph * v / (pr * gam), 0, 1, 0},
                      {(b2 - 1.0) / (gogm1 * rho), 0, 0, 1}};
    const double Pi[4][4] = {{1.0 / b2, 0, 0, 0},
                             {alph * u / (pr * gam * b2), 1, 0, 0},
                             {alph * v / (pr * gam * b2), 0, 1, 0},
                             {-(b2 - 1.0) / (gogm1 * rho * b2), 0, 0, 1}};
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) Pinv[i][j] = Pi[i][j];
    const double dpdc[4][4] = {{ke * gm1, -u * gm1, -v * gm1, gm1},
                               {-u / rho, 1.0 / rho, 0, 0},
                               {-v / rho, 0, 1.0 / rho, 0},
                               {(gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, gm1 / rho}};
    double tmp[4][4];
    matmul4(P, dpdc, tmp);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) P[i][j] = tmp[i][j];

    // |A| = V |Lambda| V^-1 of the preconditioned x-flux Jacobian
    const double temp1 = std::sqrt(u * u * (1.0 - 2.0 * b2 + b2 * b2) + 4.0 * b2 * c2);
    const double V1[4][4] = {{0.5 * (u * (b2 - 1.0) + temp1) * rho, 0.5 * (u * (b2 - 1.0) - temp1) * rho, 0, 0},
                             {1, 1, 0, 0},
                             {0, 0, 1, 0},
                             {0.5 * (u * (b2 - 1.0) * gm1 + gm1 * temp1) / gam, 0.5 * (u * (b2 - 1.0) * gm1 - gm1 * temp1) / gam, 0, 1}};
    const double Ae[4] = {std::fabs(0.5 * (u + u * b2 + temp1)), std::fabs(0.5 * (u + u * b2 - temp1)), std::fabs(u), std::fabs(u)};
    double VI1[4][4] = {{1.0 / (temp1 * rho), -0.5 * (u * (b2 - 1.0) - temp1) / temp1, 0, 0},
                        {-1.0 / (temp1 * rho), 0.5 * (u * (b2 - 1.0) + temp1) / temp1, 0, 0},
                        {0, 0, 1, 0},
                        {-gm1 / (gam * rho), 0, 0, 1}};
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) VI1[i][j] = Ae[i] * VI1[i][j];
    double A[4][4];
    matmul4(V1, VI1, A);

    const double temp2 = std::sqrt(v * v * (1.0 - 2.0 * b2 + b2 * b2) + 4.0 * b2 * c2);
    const double V2[4][4] = {{0, 0, 0.5 * (v * (b2 - 1.0) + temp2) * rho, 0.5 * (v * (b2 - 1.0) - temp2) * rho},
                             {0, 1, 0, 0},
                             {0, 0, 1, 1},
                             {1, 0, 0.5 * (gm1 * v * (b2 - 1.0) + gm1 * temp2) / gam, 0.5 * (gm1 * v * (b2 - 1.0) - gm1 * temp2) / gam}};
    const double Be[4] = {std::fabs(v), std::fabs(v), std::fabs(0.5 * (v + v * b2 + temp2)), std::fabs(0.5 * (v + v * b2 - temp2))};
    double VI2[4][4] = {{-gm1 / (gam * rho), 0, 0, 1},
                        {0, 1, 0, 0},
                        {1.0 / (temp2 * rho), 0, 0.5 * (-v * (b2 - 1.0) + temp2) / temp2, 0},
                        {-1.0 / (rho * temp2), 0, 0.5 * (v * (b2 - 1.0) + temp2) / temp2, 0}};
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) VI2[i][j] = Be[i] * VI2[i][j];
    double B[4][4];
    matmul4(V2, VI2, B);

    const double Sd[4][4] = {{0, 0, 0, 0}, {0, nu / (h * h), 0, 0}, {0, 0, nu / (h * h), 0}, {0, 0, 0, alpha / (h * h)}};
    double Sm[4][4];
    matmul4(P, Sd, Sm);

    double Tinv[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Tinv[i][j] = 2.0 * (A[i][j] + B[i][j] + h * Sm[i][j]) / h;
    timestep = 1.0 / spectral_radius4(Tinv);

    // Tau = Tinv^-1 P: the reference factors Tinv with dgetrf and solves the transposed
    // system for P^T (:487-514); here each column of P is solved against Tinv directly
    std::vector<double> a(16);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) a[i * 4 + j] = Tinv[i][j];
    for (int col = 0; col < 4; ++col) {
        double b[4] = {P[0][col], P[1][col], P[2][col], P[3][col]};
        lu_solve_inplace(a, 4, 4, b);
        for (int i = 0; i < 4; ++i) Tau[i][col] = b[i];
    }
}

// tri_hp/cns/rsdl.cpp:16-90 and the straight branch :312-530
void TriHpCns::element_rsdl(int tind, int stage, const double uh[NV][15], double lf_re[NV][15], double lf_im[NV][15]) {
    const TriBasis& B = b();
    const int lgpx = B.gpx, lgpn = B.gpn;
    const double gam = prm.gamma, gm1 = gam - 1.0, ogm1 = 1.0 / gm1, gogm1 = gam * ogm1;
    const int v0 = m.tri_pnt[3 * tind], v1 = m.tri_pnt[3 * tind + 1], v2 = m.tri_pnt[3 * tind + 2];
    double ld[2][2];
    for (int n = 0; n < 2; ++n) {
        ld[n][0] = 0.5 * (m.pnts[2 * v2 + n] - m.pnts[2 * v1 + n]);
        ld[n][1] = 0.5 * (m.pnts[2 * v0 + n] - m.pnts[2 * v1 + n]);
    }
    double u[NV][S * S], du[NV][2][S * S], cv[NV][2][S * S], df[NV][2][S * S], res[NV][S * S];
    if (prm.beta[stage] > 0.0) {
        B.proj(uh[0], u[0], S);
        for (int n = 1; n < NV; ++n) B.proj(uh[n], u[n], du[n][0], du[n][1], S);
    } else {
        for (int n = 0; n < NV; ++n) B.proj(uh[n], u[n], S);
    }
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < B.tm; ++i) { lf_re[n][i] = 0.0; lf_im[n][i] = 0.0; }

    // convective terms
    for (int i = 0; i < lgpx; ++i)
        for (int j = 0; j < lgpn; ++j) {
            const int g = i * S + j;
            const double rho = u[0][g] / u[NV - 1][g];
            const double fluxx = rho * (u[1][g] - 0.0);
            const double fluxy = rho * (u[2][g] - 0.0);
            cv[0][0][g] = +ld[1][1] * fluxx - ld[0][1] * fluxy;
            cv[0][1][g] = -ld[1][0] * fluxx + ld[0][0] * fluxy;
            for (int n = 1; n < NV - 1; ++n) {
                cv[n][0][g] = u[n][g] * cv[0][0][g];
                cv[n][1][g] = u[n][g] * cv[0][1][g];
            }
            cv[1][0][g] += ld[1][1] * u[0][g];
            cv[1][1][g] -= ld[1][0] * u[0][g];
            cv[2][0][g] -= ld[0][1] * u[0][g];
            cv[2][1][g] += ld[0][0] * u[0][g];
            const double E = u[3][g] / gm1 + 0.5 * (u[1][g] * u[1][g] + u[2][g] * u[2][g]);
            const double nfx = E * fluxx + u[0][g] * u[1][g];
            const double nfy = E * fluxy + u[0][g] * u[2][g];
            cv[3][0][g] = +ld[1][1] * nfx - ld[0][1] * nfy;
            cv[3][1][g] = -ld[1][0] * nfx + ld[0][0] * nfy;
        }
    for (int n = 0; n < NV; ++n) B.intgrtrs(lf_im[n], cv[n][0], cv[n][1], S);
    if (!(prm.beta[stage] > 0.0)) return;

    const double cjcb = ld[0][0] * ld[1][1] - ld[1][0] * ld[0][1];
    for (int i = 0; i < lgpx; ++i)
        for (int j = 0; j < lgpn; ++j) {
            const int g = i * S + j;
            const double lmu = prm.mu, lkcond = kcond;
            const double rhorbd0 = 0.0;
            const double mujcbi = lmu / cjcb;
            const double kcjcbi = lkcond / cjcb / prm.Rgas;
            res[0][g] = rhorbd0 + 0.0;
            for (int n = 1; n < NV - 1; ++n) res[n][g] = rhorbd0 * u[n][g] + 0.0;
            const double e = ogm1 * u[NV - 1][g] + 0.5 * (u[1][g] * u[1][g] + u[2][g] * u[2][g]);
            res[NV - 1][g] = rhorbd0 * e + 0.0;
            // source "zero": res(n) -= cjcb * 0
            for (int n = 0; n < NV; ++n) res[n][g] -= cjcb * 0.0;
            double visc[2][2][2][2];
            visc[0][0][0][0] = -mujcbi * (4. / 3. * ld[1][1] * ld[1][1] + ld[0][1] * ld[0][1]);
            visc[0][0][1][1] = -mujcbi * (4. / 3. * ld[1][0] * ld[1][0] + ld[0][0] * ld[0][0]);
            visc[0][0][0][1] = mujcbi * (4. / 3. * ld[1][1] * ld[1][0] + ld[0][1] * ld[0][0]);
            visc[1][1][0][0] = -mujcbi * (ld[1][1] * ld[1][1] + 4. / 3. * ld[0][1] * ld[0][1]);
            visc[1][1][1][1] = -mujcbi * (ld[1][0] * ld[1][0] + 4. / 3. * ld[0][0] * ld[0][0]);
            visc[1][1][0][1] = mujcbi * (ld[1][1] * ld[1][0] + 4. / 3. * ld[0][1] * ld[0][0]);
            visc[0][1][0][0] = mujcbi * 1. / 3. * ld[0][1] * ld[1][1];
            visc[0][1][1][1] = mujcbi * 1. / 3. * ld[0][0] * ld[1][0];
            visc[0][1][0][1] = -mujcbi * (ld[0][1] * ld[1][0] - 2. / 3. * ld[0][0] * ld[1][1]);
            visc[0][1][1][0] = -mujcbi * (ld[0][0] * ld[1][1] - 2. / 3. * ld[0][1] * ld[1][0]);
            double kc[2][2];
            kc[0][0] = -kcjcbi * (ld[1][1] * ld[1][1] + ld[0][1] * ld[0][1]);
            kc[1][1] = -kcjcbi * (ld[1][0] * ld[1][0] + ld[0][0] * ld[0][0]);
            kc[0][1] = kcjcbi * (ld[1][1] * ld[1][0] + ld[0][1] * ld[0][0]);
            df[1][0][g] = +visc[0][0][0][0] * du[1][0][g] + visc[0][1][0][0] * du[2][0][g] +
                          visc[0][0][0][1] * du[1][1][g] + visc[0][1][0][1] * du[2][1][g];
            df[1][1][g] = +visc[0][0][0][1] * du[1][0][g] + visc[0][1][1][0] * du[2][0][g] +
                          visc[0][0][1][1] * du[1][1][g] + visc[0][1][1][1] * du[2][1][g];
            df[2][0][g] = +visc[0][1][0][0] * du[1][0][g] + visc[1][1][0][0] * du[2][0][g] +
                          visc[0][1][1][0] * du[1][1][g] + visc[1][1][0][1] * du[2][1][g];
            df[2][1][g] = +visc[0][1][0][1] * du[1][0][g] + visc[1][1][0][1] * du[2][0][g] +
                          visc[0][1][1][1] * du[1][1][g] + visc[1][1][1][1] * du[2][1][g];
            df[3][0][g] = +kc[0][0] * du[NV - 1][0][g] + kc[0][1] * du[NV - 1][1][g];
            df[3][1][g] = +kc[0][1] * du[NV - 1][0][g] + kc[1][1] * du[NV - 1][1][g];
            df[3][0][g] += df[1][0][g] * u[1][g] + df[2][0][g] * u[2][g];
            df[3][1][g] += df[1][1][g] * u[1][g] + df[2][1][g] * u[2][g];
            for (int n = 1; n < NV; ++n) {
                cv[n][0][g] += df[n][0][g];
                cv[n][1][g] += df[n][1][g];
            }
        }
    for (int n = 0; n < NV; ++n) B.intgrt(lf_re[n], res[n], S);
    for (int n = 0; n < NV; ++n) {
        B.derivr(cv[n][0], res[n], S);
        B.derivs(cv[n][1], res[n], S);
    }
    const double hmax = 0.25 * area(tind) / (inscribedradius(tind) * (0.25 * (B.p + 1) * (B.p + 1)));
    for (int i = 0; i < lgpx; ++i)
        for (int j = 0; j < lgpn; ++j) {
            const int g = i * S + j;
            df[0][0][g] = 0.0;
            df[0][1][g] = 0.0;
            const double pr = u[0][g], uv = u[1][g], vv = u[2][g], rt = u[3][g];
            const double pvu[4] = {pr, uv, vv, rt};
            double P[4][4], tau[4][4], tstep;
            calculate_tau(pvu, hmax, P, tau, tstep);
            for (int a = 0; a < 4; ++a) for (int c = 0; c < 4; ++c) tau[a][c] *= prm.adis / cjcb;
            double tres[4] = {0, 0, 0, 0};
            for (int a = 0; a < NV; ++a)
                for (int c = 0; c < NV; ++c) tres[a] += tau[a][c] * res[c][g];
            const double rho = pr / rt, ke = 0.5 * (uv * uv + vv * vv);
            const double A[4][4] = {
                {uv / rt, rho, 0.0, -rho * uv / rt},
                {uv * uv / rt + 1.0, 2.0 * rho * uv, 0.0, -rho * uv * uv / rt},
                {uv * vv / rt, rho * vv, rho * uv, -rho * uv * vv / rt},
                {uv * (gogm1 * rt + ke) / rt, rho * (gogm1 * rt + ke) + rho * uv * uv, rho * uv * vv,
                 -rho * uv * (gogm1 * rt + ke) / rt + rho * uv * gogm1}};
            const double Bm[4][4] = {
                {vv / rt, 0.0, rho, -rho * vv / rt},
                {uv * vv / rt, rho * vv, rho * uv, -rho * uv * vv / rt},
                {vv * vv / rt + 1.0, 0.0, 2.0 * rho * vv, -rho * vv * vv / rt},
                {vv * (gogm1 * rt + ke) / rt, rho * uv * vv, rho * (gogm1 * rt + ke) + rho * vv * vv,
                 -rho * vv * (gogm1 * rt + ke) / rt + rho * vv * gogm1}};
            for (int a = 0; a < NV; ++a)
                for (int c = 0; c < NV; ++c) {
                    df[a][0][g] -= (+ld[1][1] * A[a][c] - ld[0][1] * Bm[a][c]) * tres[c];
                    df[a][1][g] -= (-ld[1][0] * A[a][c] + ld[0][0] * Bm[a][c]) * tres[c];
                }
        }
    for (int n = 0; n < NV; ++n) B.intgrtrs(lf_re[n], df[n][0], df[n][1], S);
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < B.tm; ++i) lf_re[n][i] *= prm.beta[stage];
}

// tri_hp/hp_boundary.cpp:973-1049 with the bdry_cns::generic flux (tri_hp/cns/bdry.cpp:192-245);
// ibc f(0) is the boundary pressure Params::bc[0]
void TriHpCns::edge_rsdl() {
    const TriBasis& B = b();
    const double gm1 = prm.gamma - 1.0, pbc = prm.bc[0];
    const int nebd = int(m.ebdry_start.size()) - 1;
    for (int e = 0; e < nebd; ++e)
        for (int j = m.ebdry_start[e]; j < m.ebdry_start[e + 1]; ++j) {
            const int sind = m.ebdry_seg[j];
            const int v0 = m.seg_pnt[2 * sind], v1 = m.seg_pnt[2 * sind + 1];
            double u1d[NV][8], cht[2][8], crd[2][8], dcrd[2][8], uu[NV][8], r[NV][8], lfl[NV][8];
            for (int n = 0; n < NV; ++n) {
                u1d[n][0] = ugv[v0 * NV + n];
                u1d[n][1] = ugv[v1 * NV + n];
                for (int mm = 0; mm < B.sm; ++mm) u1d[n][mm + 2] = ugs[(sind * sm0 + mm) * NV + n];
            }
            for (int n = 0; n < 2; ++n) {
                cht[n][0] = m.pnts[2 * v0 + n];
                cht[n][1] = m.pnts[2 * v1 + n];
                for (int mm = 0; mm < B.sm; ++mm) cht[n][mm + 2] = 0.0;
                B.proj1d(cht[n], crd[n], dcrd[n]);
            }
            for (int n = 0; n < NV; ++n) B.proj1d(u1d[n], uu[n]);
            for (int k = 0; k < B.gpx; ++k) {
                double nrm0 = dcrd[1][k], nrm1 = -dcrd[0][k];
                const double jcb = std::sqrt(nrm0 * nrm0 + nrm1 * nrm1);
                nrm0 /= jcb; nrm1 /= jcb;
                const double mvx = 0.0, mvy = 0.0;
                double flx[NV];
                flx[0] = pbc / uu[NV - 1][k] * ((uu[1][k] - mvx) * nrm0 + (uu[2][k] - mvy) * nrm1);
                flx[1] = flx[0] * uu[1][k] + pbc * nrm0;
                flx[2] = flx[0] * uu[2][k] + pbc * nrm1;
                const double rho = uu[0][k] / uu[NV - 1][k];
                const double E = pbc / (rho * gm1) + 0.5 * (uu[1][k] * uu[1][k] + uu[2][k] * uu[2][k]);
                flx[NV - 1] = rho * E * ((uu[1][k] - mvx) * nrm0 + (uu[2][k] - mvy) * nrm1) +
                              pbc * (uu[1][k] * nrm0 + uu[2][k] * nrm1);
                for (int n = 0; n < NV; ++n) r[n][k] = flx[n] * jcb;
            }
            for (int n = 0; n < NV; ++n) B.intgrt1d(lfl[n], r[n]);
            for (int n = 0; n < NV; ++n) resv[v0 * NV + n] += lfl[n][0];
            for (int n = 0; n < NV; ++n) resv[v1 * NV + n] += lfl[n][1];
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) ress[(sind * smf + k) * NV + n] += lfl[n][k + 2];
        }
}

// tri_hp/cns/setup_preconditioner.cpp:7-239 (diagonal branch, straight elements),
// then the scalar assembly and inversion of tri_hp/setup_preconditioner.cpp:68-76
int TriHpCns::setup_preconditioner() {
    const TriBasis& B = b();
    int err = 0;
    std::fill(tprcn.begin(), tprcn.end(), 0.0);
    std::fill(vpc.begin(), vpc.end(), 0.0);
    std::fill(spc.begin(), spc.end(), 0.0);
    std::fill(tpc.begin(), tpc.end(), 0.0);
    double u[NV][S * S];
    int done = 0;
    for (int tind = 0; tind < m.ntri; ++tind) {
        const double jcb = 0.25 * area(tind);
        const int* v = &m.tri_pnt[3 * tind];
        ugtouht(tind);
        for (int n = 0; n < NV; ++n) B.proj(uht[n], u[n], S);
        double hmax = 0.0, qmax = 0.0, umax[4] = {0, 0, 0, 0};
        for (int i = 0; i < B.gpx; ++i)
            for (int j = 0; j < B.gpn; ++j) {
                const int g = i * S + j;
                const double q = std::pow(u[1][g] - 0.5 * 0.0, 2.0) + std::pow(u[2][g] - 0.5 * 0.0, 2.0);
                qmax = MAXM(qmax, q);
                umax[0] = MAXM(umax[0], std::fabs(u[0][g]));
                umax[1] = MAXM(umax[1], std::fabs(u[1][g] - 0.5 * 0.0));
                umax[2] = MAXM(umax[2], std::fabs(u[2][g] - 0.5 * 0.0));
                umax[3] = MAXM(umax[3], std::fabs(u[3][g]));
            }
        for (int j = 0; j < 3; ++j) {
            const int a = v[j], c = v[(j + 1) % 3];
            const double hh = std::pow(m.pnts[2 * a] - m.pnts[2 * c], 2.0) + std::pow(m.pnts[2 * a + 1] - m.pnts[2 * c + 1], 2.0);
            hmax = (hh > hmax ? hh : hmax);
        }
        hmax = std::sqrt(hmax);
        const double pf = 0.25 * (B.p + 1) * (B.p + 1);
        const double h = 4. * jcb / (pf * hmax);
        hmax = hmax / pf;
        if (!(h > 0.0)) { err = 1; break; }
        if (std::isnan(qmax)) { err = 1; break; }
        double Pinv[4][4], tau[4][4], tstep;
        calculate_tau(umax, hmax, Pinv, tau, tstep);
        const double dtstari = 1.0 / tstep;
        for (int n = 0; n < NV; ++n) tprcn[tind * NV + n] = jcb;
        for (int a = 0; a < 16; ++a) tpc[tind * 16 + a] = Pinv[a / 4][a % 4] * dtstari;
        ++done;
    }
    // the reference assembles inside the element loop, in element order; the sums here
    // visit the same elements in the same order
    for (int tind = 0; tind < done; ++tind) {
        const double jcb = 0.25 * area(tind);
        for (int i = 0; i < 3; ++i) {
            const int vv = m.tri_pnt[3 * tind + i];
            for (int a = 0; a < 16; ++a) vpc[vv * 16 + a] += tpc[tind * 16 + a] * jcb;
            if (B.sm > 0) {
                const int side = m.tri_seg[3 * tind + i];
                for (int a = 0; a < 16; ++a) spc[side * 16 + a] += tpc[tind * 16 + a];
            }
        }
    }
    assemble_prcn(done);
    return err;
}

// tri_hp/cns/update.cpp:247-271: a <- dpdc(pvu) a (conservative to primitive change)
void TriHpCns::switch_variables(const double pvu[4], double a[4]) const {
    const double gm1 = prm.gamma - 1.0;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], rt = pvu[3];
    const double rho = pr / rt, ke = 0.5 * (u * u + v * v);
    const double dpdc[4][4] = {{ke * gm1, -u * gm1, -v * gm1, gm1},
                               {-u / rho, 1.0 / rho, 0.0, 0.0},
                               {-v / rho, 0.0, 1.0 / rho, 0.0},
                               {(gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, gm1 / rho}};
    double t[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t[i] += dpdc[i][j] * a[j];
    for (int i = 0; i < 4; ++i) a[i] = t[i];
}

static inline void lower_solve(const double* P, double* r, int stride_row, int stride_col) {
    // forward substitution with the lower triangle of a 4x4 matrix (diagonal included)
    for (int j = 0; j < 4; ++j) {
        double l0 = r[j];
        for (int k = 0; k < j; ++k) l0 -= P[j * stride_row + k * stride_col] * r[k];
        r[j] = l0 / P[j * stride_row + j * stride_col];
    }
}

// tri_hp/cns/bdry.cpp:247-340 (inflow): boundary values prescribed for u, v, RT, so the
// conservative residual of momentum and energy follows from the mass residual
void TriHpCns::modify_boundary_residual() {
    const TriBasis& B = b();
    const double ogm1 = 1.0 / (prm.gamma - 1.0);
    const double bu = prm.bc[1], bv = prm.bc[2], brt = prm.bc[3];
    const double KE = 0.5 * (bu * bu + bv * bv);
    const int nebd = int(m.ebdry_start.size()) - 1;
    for (int e = 0; e < nebd; ++e) {
        if (m.ebdry_essential[e] != 0b1110) continue;
        const int j0 = m.ebdry_start[e], j1 = m.ebdry_start[e + 1];
        auto fix_vertex = [&](int v0) {
            resv[v0 * NV + 1] = resv[v0 * NV + 0] * bu;
            resv[v0 * NV + 2] = resv[v0 * NV + 0] * bv;
            resv[v0 * NV + 3] = resv[v0 * NV + 0] * (brt * ogm1 + KE);
        };
        int sind = -1;
        for (int j = j0; j < j1; ++j) {
            sind = m.ebdry_seg[j];
            fix_vertex(m.seg_pnt[2 * sind]);
        }
        fix_vertex(m.seg_pnt[2 * sind + 1]);
        if (!B.sm) continue;
        for (int j = j0; j < j1; ++j) {
            sind = m.ebdry_seg[j];
            const int v0 = m.seg_pnt[2 * sind], v1 = m.seg_pnt[2 * sind + 1];
            double rescoef[8], res1d[8], rr[NV][8], lfl[NV][8];
            rescoef[0] = resv[v0 * NV + 0];
            rescoef[1] = resv[v1 * NV + 0];
            for (int mm = 0; mm < B.sm; ++mm) rescoef[mm + 2] = ress[(sind * smf + mm) * NV + 0];
            B.proj1d(rescoef, res1d);
            for (int n = 1; n < NV; ++n) B.proj1d(resv[v0 * NV + n], resv[v1 * NV + n], rr[n]);
            for (int k = 0; k < B.gpx; ++k) {
                rr[1][k] -= res1d[k] * bu;
                rr[2][k] -= res1d[k] * bv;
                rr[3][k] -= res1d[k] * (brt * ogm1 + KE);
            }
            for (int n = 1; n < NV; ++n) {
                B.intgrt1d(lfl[n], rr[n]);
                band_chol_solve(B.sdiag1d.data(), B.sm, B.sbwth, &lfl[n][2]);
                for (int mm = 0; mm < B.sm; ++mm) ress[(sind * smf + mm) * NV + n] = -lfl[n][2 + mm];
            }
        }
    }
}

// tri_hp/cns/update.cpp:5-245
void TriHpCns::project_new_variables() {
    const TriBasis& B = b();
    res_tv = resv; res_ts = ress; res_ti = resi;
    for (int i = 0; i < m.npnt; ++i) {
        double lug[4], lres[4];
        for (int n = 0; n < NV; ++n) { lug[n] = ugv[i * NV + n]; lres[n] = resv[i * NV + n]; }
        switch_variables(lug, lres);
        lower_solve(&vpc[i * 16], lres, 4, 1);
        for (int n = 0; n < NV; ++n) resv[i * NV + n] = lres[n];
    }
    vdirichlet();
    if (!B.sm) return;

    for (int sind = 0; sind < m.nseg; ++sind) {
        const int p0 = m.seg_pnt[2 * sind], p1 = m.seg_pnt[2 * sind + 1];
        double temp1d[NV][8], P[4][4][8], ucoef[NV][8], rcoef[NV][8], u1d[NV][8], res1d[NV][8];
        for (int n = 0; n < NV; ++n) B.proj1d(resv[p0 * NV + n], resv[p1 * NV + n], temp1d[n]);
        for (int a = 0; a < NV; ++a)
            for (int c = 0; c < NV; ++c) B.proj1d(vpc[p0 * 16 + a * 4 + c], vpc[p1 * 16 + a * 4 + c], P[a][c]);
        for (int n = 0; n < NV; ++n) {
            ucoef[n][0] = ugv[p0 * NV + n]; ucoef[n][1] = ugv[p1 * NV + n];
            rcoef[n][0] = res_tv[p0 * NV + n]; rcoef[n][1] = res_tv[p1 * NV + n];
            for (int mm = 0; mm < B.sm; ++mm) {
                ucoef[n][mm + 2] = ugs[(sind * sm0 + mm) * NV + n];
                rcoef[n][mm + 2] = res_ts[(sind * smf + mm) * NV + n];
            }
        }
        for (int n = 0; n < NV; ++n) { B.proj1d(ucoef[n], u1d[n]); B.proj1d(rcoef[n], res1d[n]); }
        for (int i = 0; i < B.gpx; ++i) {
            double lug[4], lres[4], Pl[16];
            for (int n = 0; n < NV; ++n) { lug[n] = u1d[n][i]; lres[n] = res1d[n][i]; }
            switch_variables(lug, lres);
            for (int a = 0; a < 16; ++a) Pl[a] = P[a / 4][a % 4][i];
            lower_solve(Pl, lres, 4, 1);
            for (int n = 0; n < NV; ++n) temp1d[n][i] -= lres[n];
        }
        for (int n = 0; n < NV; ++n) {
            B.intgrt1d(rcoef[n], temp1d[n]);
            band_chol_solve(B.sdiag1d.data(), B.sm, B.sbwth, &rcoef[n][2]);
            for (int mm = 0; mm < B.sm; ++mm) ress[(sind * smf + mm) * NV + n] = -rcoef[n][mm + 2];
        }
    }
    for (int mode = 0; mode < B.sm; ++mode) sdirichlet(mode);
    if (!B.im) return;

    for (int tind = 0; tind < m.ntri; ++tind) {
        double ucoef[NV][15], rcoef[NV][15], tcoef[NV][15];
        for (int i = 0; i < 3; ++i) {
            const int vr = m.tri_pnt[3 * tind + i];
            for (int n = 0; n < NV; ++n) {
                ucoef[n][i] = ugv[vr * NV + n];
                rcoef[n][i] = res_tv[vr * NV + n];
                tcoef[n][i] = resv[vr * NV + n];
            }
        }
        int cnt = 3;
        for (int i = 0; i < 3; ++i) {
            const int sind = m.tri_seg[3 * tind + i], sign = m.tri_sgn[3 * tind + i];
            int msgn = 1;
            for (int mm = 0; mm < B.sm; ++mm) {
                for (int n = 0; n < NV; ++n) {
                    ucoef[n][cnt] = msgn * ugs[(sind * sm0 + mm) * NV + n];
                    rcoef[n][cnt] = msgn * res_ts[(sind * smf + mm) * NV + n];
                    tcoef[n][cnt] = msgn * ress[(sind * smf + mm) * NV + n];
                }
                msgn *= sign;
                ++cnt;
            }
        }
        int indx = 0;
        for (int mm = 1; mm < B.sm; ++mm) {
            for (int k = 0; k < B.sm - mm; ++k) {
                for (int n = 0; n < NV; ++n) {
                    ucoef[n][cnt] = ugi[(tind * im0 + indx) * NV + n];
                    rcoef[n][cnt] = res_ti[(tind * imf + indx) * NV + n];
                    tcoef[n][cnt] = 0.0;
                }
                ++cnt; ++indx;
            }
            indx += sm0 - B.sm;
        }
        double u2d[NV][S * S], res2d[NV][S * S], temp2d[NV][S * S], P2d[4][4][S * S];
        for (int n = 0; n < NV; ++n) {
            B.proj(ucoef[n], u2d[n], S);
            B.proj(rcoef[n], res2d[n], S);
            // proj_bdry (tri_basis/proj.cpp:283-337) of the boundary modes; with the
            // interior coefficients zero, proj adds exact zeros to the same sums
            B.proj(tcoef[n], temp2d[n], S);
        }
        const int q0 = m.tri_pnt[3 * tind], q1 = m.tri_pnt[3 * tind + 1], q2 = m.tri_pnt[3 * tind + 2];
        for (int a = 0; a < NV; ++a)
            for (int c = 0; c < NV; ++c)
                B.proj(vpc[q0 * 16 + a * 4 + c], vpc[q1 * 16 + a * 4 + c], vpc[q2 * 16 + a * 4 + c], P2d[a][c], S);
        for (int i = 0; i < B.gpx; ++i)
            for (int j = 0; j < B.gpn; ++j) {
                const int g = i * S + j;
                double lug[4], lres[4], Pl[16];
                for (int n = 0; n < NV; ++n) { lug[n] = u2d[n][g]; lres[n] = res2d[n][g]; }
                switch_variables(lug, lres);
                for (int a = 0; a < 16; ++a) Pl[a] = P2d[a / 4][a % 4][g];
                lower_solve(Pl, lres, 4, 1);
                for (int n = 0; n < NV; ++n) temp2d[n][g] -= lres[n];
            }
        for (int n = 0; n < NV; ++n) {
            B.intgrt(rcoef[n], temp2d[n], S);
            band_chol_solve(B.idiag.data(), B.im, B.ibwth, &rcoef[n][B.bm]);
            for (int k = 0; k < B.im; ++k) resi[(tind * imf + k) * NV + n] = -rcoef[n][B.bm + k];
        }
    }
}

// tri_hp/cns/update.cpp:276-357
void TriHpCns::update() {
    ug0v = ugv; ug0s = ugs; ug0i = ugi;
    const TriBasis& B = b();
    for (int stage = 0; stage < prm.nstage; ++stage) {
        rsdl(stage);
        minvrt();
        modify_boundary_residual();
        project_new_variables();
        const double cflalpha = prm.alpha[stage] * prm.cfl[log2p];
        for (int i = 0; i < m.npnt * NV; ++i) ugv[i] = ug0v[i] - cflalpha * resv[i];
        if (B.sm > 0) {
            for (int s = 0; s < m.nseg; ++s)
                for (int k = 0; k < B.sm; ++k)
                    for (int n = 0; n < NV; ++n)
                        ugs[(s * sm0 + k) * NV + n] = ug0s[(s * sm0 + k) * NV + n] - cflalpha * ress[(s * smf + k) * NV + n];
            if (B.im > 0)
                for (int t = 0; t < m.ntri; ++t) {
                    int indx = 0, indx1 = 0;
                    for (int mm = 1; mm < B.sm; ++mm) {
                        for (int k = 0; k < B.sm - mm; ++k) {
                            for (int n = 0; n < NV; ++n)
                                ugi[(t * im0 + indx1) * NV + n] = ug0i[(t * im0 + indx1) * NV + n] - cflalpha * resi[(t * imf + indx) * NV + n];
                            ++indx; ++indx1;
                        }
                        indx1 += sm0 - B.sm;
                    }
                }
        }
    }
}

}  // namespace oracle
