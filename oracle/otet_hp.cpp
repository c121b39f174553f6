// Oracle (test infrastructure only): the generic tet_hp part and tet_hp_ins.  See otet_hp.h.
#include "otet_hp.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace oracle {

namespace {
const int kVS[6][2] = {{2, 3}, {3, 1}, {2, 1}, {1, 0}, {2, 0}, {3, 0}};  // tet_mesh/setup.cpp:230-235
// faces opposite vertex i, ordered as tet_mesh::create_tri_from_tet orders them
// (tet_mesh/setup.cpp:305-308)
const int kVF[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
}  // namespace

template <int NV_>
TetHp<NV_>::TetHp(const TetMesh& mesh, const TetParams& p) : m(mesh), prm(p) {
    if (prm.log2pmax < 0 || prm.log2pmax > 1) throw std::runtime_error("oracle TetHp: log2p must be 0 or 1");
    for (int l = 0; l <= prm.log2pmax; ++l) bases.emplace_back(1 << l);
    log2p = prm.log2pmax;
    em0 = bases[log2p].em;
    const int nv = m.npnt * NV, ne = m.nseg * em0 * NV;
    for (auto* a : {&ugv, &ug0v, &resv, &res_rv, &res0v}) a->assign(nv, 0.0);
    for (auto* a : {&uge, &ug0e, &rese, &res_re, &res0e}) a->assign(ne, 0.0);
    dresv.assign(prm.log2pmax + 1, std::vector<double>(nv, 0.0));
    drese.assign(prm.log2pmax + 1, std::vector<double>(ne, 0.0));
    vprcn.assign(nv, 0.0);
    eprcn.assign(m.nseg * NV, 0.0);
    iprcn.assign(m.ntet * NV, 0.0);
    vol.assign(m.ntet, 0.0);
    // tet(tind).vol: |triple product| (tet_mesh/setup.cpp:58-90, tet_prims.cpp:180-191)
    for (int t = 0; t < m.ntet; ++t) {
        const double* P[4];
        for (int i = 0; i < 4; ++i) P[i] = &m.pnts[3 * m.tet_pnt[4 * t + i]];
        double d1[3], d2[3], d3[3];
        for (int n = 0; n < 3; ++n) { d1[n] = P[3][n] - P[2][n]; d2[n] = P[1][n] - P[2][n]; d3[n] = P[0][n] - P[2][n]; }
        const double c[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
        const double v = c[0] * d3[0] + c[1] * d3[1] + c[2] * d3[2];
        if (!(v > 0.0)) throw std::runtime_error("oracle TetHp: tet not positively oriented");
        vol[t] = std::fabs(v);
    }
    vess.assign(m.npnt, 0);
    eess.assign(m.nseg, 0);
    for (int f = 0; f < m.nbf; ++f)
        for (int j = 0; j < 3; ++j) {
            vess[m.bf_pnt[3 * f + j]] = 1;
            eess[m.bf_seg[3 * f + j]] = 1;
        }
}

TetHpIns::TetHpIns(const TetMesh& mesh, const TetParams& p) : TetHp<4>(mesh, p) { tau.assign(m.ntet * NV, 0.0); }

// tet_hp/gtol.cpp:17-85 (vertices, edges with sgn^m)
template <int NV_>
void TetHp<NV_>::ugtouht(int t, double uht[NV][10]) const {
    const TetBasis& B = b();
    for (int i = 0; i < 4; ++i)
        for (int n = 0; n < NV; ++n) uht[n][i] = ugv[m.tet_pnt[4 * t + i] * NV + n];
    int cnt = 4;
    for (int i = 0; i < 6; ++i) {
        const int e = m.tet_seg[6 * t + i], sg = m.tet_sgn[6 * t + i];
        int msgn = 1;
        for (int k = 0; k < B.em; ++k) {
            for (int n = 0; n < NV; ++n) uht[n][cnt] = msgn * uge[(e * em0 + k) * NV + n];
            msgn *= sg;
            ++cnt;
        }
    }
}

// tet_hp/ins/rsdl.cpp:15-87, 442-785: straight element, steady (mvel = 0, dugdt = 0)
void TetHpIns::element_rsdl(int t, int stage, const double uht[NV][10], double lf_re[NV][10],
                            double lf_im[NV][10]) const {
    const TetBasis& B = b();
    const int G = B.G, T = B.tm;
    const double beta = stage < prm.nstage ? prm.beta[stage] : 1.0;
    std::vector<double> u(NV * G), du(NV * 3 * G, 0.0), cv(3 * 3 * G), df(3 * 3 * G), res(NV * G);
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
        for (int n = 0; n < NV - 1; ++n) B.proj(uht[n], &U(n, 0), &DU(n, 0, 0), &DU(n, 1, 0), &DU(n, 2, 0));
        B.proj(uht[NV - 1], &U(NV - 1, 0));
    } else {
        for (int n = 0; n < NV; ++n) B.proj(uht[n], &U(n, 0));
    }
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < T; ++i) { lf_re[n][i] = 0.0; lf_im[n][i] = 0.0; }
    // cofactors d(a)(c) = d r_a / d x_c times cjcb (rsdl.cpp:449-458)
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
    const double rho = prm.rho;
    // convective and pressure fluxes (rsdl.cpp:461-508)
    for (int g = 0; g < G; ++g) {
        const double fx = rho * U(0, g), fy = rho * U(1, g), fz = rho * U(2, g);
        for (int a = 0; a < 3; ++a) DU(NV - 1, a, g) = d[a][0] * fx + d[a][1] * fy + d[a][2] * fz;
        for (int n = 0; n < NV - 1; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) = U(n, g) * DU(NV - 1, a, g);
        for (int n = 0; n < 3; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) += d[a][n] * U(NV - 1, g);
    }
    for (int n = 0; n < NV - 1; ++n) B.intgrtrst(lf_im[n], &CV(n, 0, 0), &CV(n, 1, 0), &CV(n, 2, 0));
    B.intgrtrst(lf_im[NV - 1], &DU(NV - 1, 0, 0), &DU(NV - 1, 1, 0), &DU(NV - 1, 2, 0));
    if (!(beta > 0.0)) return;

    // viscous tensor (rsdl.cpp:518-614): visc(n,m)(a,b)
    const double cjcbi = prm.mu / cjcb;
    double visc[3][3][3][3];
    for (int n = 0; n < 3; ++n)
        for (int mm = 0; mm < 3; ++mm)
            for (int a = 0; a < 3; ++a)
                for (int bb = 0; bb < 3; ++bb) {
                    if (n == mm) {
                        double s = 2. * d[a][n] * d[bb][n];
                        for (int c = 0; c < 3; ++c)
                            if (c != n) s += d[a][c] * d[bb][c];
                        visc[n][mm][a][bb] = -cjcbi * s;
                    } else {
                        visc[n][mm][a][bb] = -cjcbi * d[a][mm] * d[bb][n];
                    }
                }
    for (int g = 0; g < G; ++g) {
        for (int n = 0; n < NV; ++n) res[n * G + g] = 0.0;  // rhorbd0 = 0, dugdt = 0
        for (int n = 0; n < 3; ++n)
            for (int a = 0; a < 3; ++a) {
                double s = 0.0;
                for (int bb = 0; bb < 3; ++bb)
                    for (int mm = 0; mm < 3; ++mm) s += visc[n][mm][a][bb] * DU(mm, bb, g);
                DF(n, a, g) = s;
            }
        for (int n = 0; n < NV - 1; ++n)
            for (int a = 0; a < 3; ++a) CV(n, a, g) += DF(n, a, g);
    }
    for (int n = 0; n < NV; ++n) B.intgrt(lf_re[n], &res[n * G]);
    for (int n = 0; n < NV - 1; ++n) {
        B.derivr(&CV(n, 0, 0), &res[n * G]);
        B.derivs(&CV(n, 1, 0), &res[n * G]);
        B.derivt(&CV(n, 2, 0), &res[n * G]);
    }
    B.derivr(&DU(NV - 1, 0, 0), &res[(NV - 1) * G]);
    B.derivs(&DU(NV - 1, 1, 0), &res[(NV - 1) * G]);
    B.derivt(&DU(NV - 1, 2, 0), &res[(NV - 1) * G]);
    // stabilisation with the tau frozen by setup_preconditioner (rsdl.cpp:699-767)
    for (int g = 0; g < G; ++g) {
        const double t0 = tau[NV * t + 0] * res[0 * G + g], t1 = tau[NV * t + 0] * res[1 * G + g];
        const double t2 = tau[NV * t + 0] * res[2 * G + g], t3 = tau[NV * t + NV - 1] * res[3 * G + g];
        const double u0 = U(0, g), u1 = U(1, g), u2 = U(2, g);
        for (int a = 0; a < 3; ++a) {
            DF(0, a, g) -= (d[a][0] * (2. * u0) + d[a][1] * u1 + d[a][2] * u2) * t0 + d[a][1] * u0 * t1 +
                           d[a][2] * u0 * t2 + d[a][0] * t3;
            DF(1, a, g) -= d[a][0] * u1 * t0 + (d[a][0] * u0 + d[a][1] * (2. * u1) + d[a][2] * u2) * t1 +
                           d[a][2] * u1 * t2 + d[a][1] * t3;
            DF(2, a, g) -= d[a][0] * u2 * t0 + d[a][1] * u2 * t1 +
                           (d[a][0] * u0 + d[a][1] * u1 + d[a][2] * (2. * u2)) * t2 + d[a][2] * t3;
            DU(NV - 1, a, g) = -d[a][0] * t0 - d[a][1] * t1 - d[a][2] * t2;
        }
    }
    for (int n = 0; n < NV - 1; ++n) B.intgrtrst(lf_re[n], &DF(n, 0, 0), &DF(n, 1, 0), &DF(n, 2, 0));
    B.intgrtrst(lf_re[NV - 1], &DU(NV - 1, 0, 0), &DU(NV - 1, 1, 0), &DU(NV - 1, 2, 0));
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < T; ++i) lf_re[n][i] *= beta;
}

// inflow face: continuity flux rho (u.n) through the face, momentum flux 0
// (tet_hp/ins/bdry.cpp:59-102, bdry_ins.h:122-134); n = dx/dr x dx/ds (not normalised)
void TetHpIns::face_rsdl(int f, double lf[NV][6]) const {
    const TetBasis& B = b();
    const int GF = B.gpx * B.gpy, TF = B.tm2d;
    std::vector<double> crd(3 * GF), dcr(3 * GF), dcs(3 * GF), u(NV * GF), flx(NV * GF);
    double cht[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 3; ++n) {
        for (int j = 0; j < 3; ++j) cht[j] = m.pnts[3 * m.bf_pnt[3 * f + j] + n];
        B.proj2d(cht, &crd[n * GF], &dcr[n * GF], &dcs[n * GF]);
    }
    double uh[6];
    for (int n = 0; n < NV; ++n) {
        for (int j = 0; j < 3; ++j) uh[j] = ugv[m.bf_pnt[3 * f + j] * NV + n];
        for (int j = 0; j < 3; ++j) {
            const int e = m.bf_seg[3 * f + j], sg = m.bf_sgn[3 * f + j];
            int msgn = 1;
            for (int k = 0; k < B.em; ++k) {
                uh[3 + j * B.em + k] = msgn * uge[(e * em0 + k) * NV + n];
                msgn *= sg;
            }
        }
        B.proj2d(uh, &u[n * GF]);
    }
    for (int g = 0; g < GF; ++g) {
        const double v1[3] = {dcr[0 * GF + g], dcr[1 * GF + g], dcr[2 * GF + g]};
        const double v2[3] = {dcs[0 * GF + g], dcs[1 * GF + g], dcs[2 * GF + g]};
        const double nrm[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                               v1[0] * v2[1] - v1[1] * v2[0]};
        flx[(NV - 1) * GF + g] =
            prm.rho * (u[0 * GF + g] * nrm[0] + u[1 * GF + g] * nrm[1] + u[2 * GF + g] * nrm[2]);
        for (int n = 0; n < NV - 1; ++n) flx[n * GF + g] = 0.0;
    }
    for (int n = 0; n < NV; ++n) {
        double r[6];
        B.intgrt2d(r, &flx[n * GF]);
        for (int i = 0; i < TF; ++i) lf[n][i] = r[i];
    }
}

// tet_hp/update.cpp:9-131
template <int NV_>
void TetHp<NV_>::rsdl(int stage) {
    const TetBasis& B = b();
    const double omb = 1.0 - (stage < prm.nstage ? prm.beta[stage] : 1.0);
    for (int i = 0; i < m.npnt * NV; ++i) { resv[i] = 0.0; res_rv[i] *= omb; }
    for (int s = 0; s < m.nseg; ++s)
        for (int k = 0; k < B.em; ++k)
            for (int n = 0; n < NV; ++n) {
                rese[(s * em0 + k) * NV + n] = 0.0;
                res_re[(s * em0 + k) * NV + n] *= omb;
            }
    // boundary faces first (tet_hp/ins/bdry.cpp:16-57)
    for (int f = 0; f < m.nbf; ++f) {
        double lf[NV][6];
        face_rsdl(f, lf);
        for (int j = 0; j < 3; ++j)
            for (int n = 0; n < NV; ++n) resv[m.bf_pnt[3 * f + j] * NV + n] += lf[n][j];
        int indx = 3;
        for (int j = 0; j < 3; ++j) {
            const int e = m.bf_seg[3 * f + j];
            double msgn = 1.0;
            for (int k = 0; k < B.em; ++k) {
                for (int n = 0; n < NV; ++n) rese[(e * em0 + k) * NV + n] += msgn * lf[n][indx];
                msgn *= m.bf_sgn[3 * f + j];
                ++indx;
            }
        }
    }
    double uht[NV][10], lre[NV][10], lim[NV][10];
    for (int t = 0; t < m.ntet; ++t) {
        ugtouht(t, uht);
        element_rsdl(t, stage, uht, lre, lim);
        // lftog (tet_hp/gtol.cpp:815-876) into res (imaginary) and res_r (real)
        for (int pass = 0; pass < 2; ++pass) {
            double (*lf)[10] = pass == 0 ? lim : lre;
            std::vector<double>& gv = pass == 0 ? resv : res_rv;
            std::vector<double>& ge = pass == 0 ? rese : res_re;
            for (int i = 0; i < 4; ++i)
                for (int n = 0; n < NV; ++n) gv[m.tet_pnt[4 * t + i] * NV + n] += lf[n][i];
            int cnt = 4;
            for (int i = 0; i < 6; ++i) {
                const int e = m.tet_seg[6 * t + i], sg = m.tet_sgn[6 * t + i];
                int msgn = 1;
                for (int k = 0; k < B.em; ++k) {
                    for (int n = 0; n < NV; ++n) ge[(e * em0 + k) * NV + n] += msgn * lf[n][cnt];
                    msgn *= sg;
                    ++cnt;
                }
            }
        }
    }
    for (int i = 0; i < m.npnt * NV; ++i) resv[i] += res_rv[i];
    for (int s = 0; s < m.nseg; ++s)
        for (int k = 0; k < B.em; ++k)
            for (int n = 0; n < NV; ++n) rese[(s * em0 + k) * NV + n] += res_re[(s * em0 + k) * NV + n];
    if (coarse_flag) {
        std::vector<double>& dv = dresv[log2p];
        std::vector<double>& de = drese[log2p];
        if (isfrst) {
            for (int i = 0; i < m.npnt * NV; ++i) dv[i] = prm.fadd * res0v[i] - resv[i];
            for (int s = 0; s < m.nseg; ++s)
                for (int k = 0; k < B.em; ++k)
                    for (int n = 0; n < NV; ++n) {
                        const int ix = (s * em0 + k) * NV + n;
                        de[ix] = prm.fadd * res0e[ix] - rese[ix];
                    }
            isfrst = false;
        }
        for (int i = 0; i < m.npnt * NV; ++i) resv[i] += dv[i];
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.em; ++k)
                for (int n = 0; n < NV; ++n) rese[(s * em0 + k) * NV + n] += de[(s * em0 + k) * NV + n];
    }
}

// tet_hp/ins/setup_preconditioner.cpp:7-135 and tet_hp/setup_preconditioner.cpp:428-496
int TetHpIns::setup_preconditioner() {
    const TetBasis& B = b();
    int err = 0;
    const double nu = prm.mu / prm.rho;
    std::fill(vprcn.begin(), vprcn.end(), 0.0);
    std::fill(eprcn.begin(), eprcn.end(), 0.0);
    for (int t = 0; t < m.ntet; ++t) {
        double jcb = 0.125 * vol[t];
        double amin = 1.0e99, amax = 0.0;
        for (int j = 0; j < 4; ++j) {
            const double* A = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][0]]];
            const double* Bp = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][1]]];
            const double* C = &m.pnts[3 * m.tet_pnt[4 * t + kVF[j][2]]];
            const double v1[3] = {A[0] - Bp[0], A[1] - Bp[1], A[2] - Bp[2]};
            const double v2[3] = {A[0] - C[0], A[1] - C[1], A[2] - C[2]};
            const double c[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                                 v1[0] * v2[1] - v1[1] * v2[0]};
            const double a = .5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            amax = a > amax ? a : amax;
            amin = a < amin ? a : amin;
        }
        const double pp = (B.p + 1) * (B.p + 1);
        const double h = 1.0 * 4.0 * jcb / (0.25 * pp * amax);
        const double hmax = 1.0 * 4.0 * jcb / (0.25 * pp * amin);
        double qmax = 0.0;
        for (int j = 0; j < 4; ++j) {
            const int v = m.tet_pnt[4 * t + j];
            const double q = std::pow(ugv[v * NV + 0], 2.0) + std::pow(ugv[v * NV + 1], 2.0) +
                             std::pow(ugv[v * NV + 2], 2.0);
            qmax = (qmax > q) ? qmax : q;  // MAX of sys/param.h (a NaN q propagates)
        }
        if (!(jcb > 0.0)) err = 1;
        if (!(qmax >= 0.0)) err = 1;
        double gam = 3.0 * qmax + (2. * nu / hmax) * (2. * nu / hmax);
        if (prm.mu == 0.0) gam = std::max(gam, 0.1);
        const double q = std::sqrt(qmax);
        const double lam1 = q + std::sqrt(qmax + gam);
        tau[NV * t + 0] = prm.adis * h / (jcb * std::sqrt(gam));
        tau[NV * t + NV - 1] = qmax * tau[NV * t + 0];
        jcb *= 2. * nu * (1. / (h * h) + 1. / (h * h) + 1. / (h * h)) + 3 * lam1 / h;
        iprcn[NV * t + 0] = prm.rho * jcb;
        iprcn[NV * t + 1] = prm.rho * jcb;
        iprcn[NV * t + 2] = prm.rho * jcb;
        iprcn[NV * t + 3] = jcb / gam;
        for (int i = 0; i < 4; ++i)
            for (int n = 0; n < NV; ++n) vprcn[m.tet_pnt[4 * t + i] * NV + n] += iprcn[NV * t + n];
        if (B.em > 0)
            for (int i = 0; i < 6; ++i)
                for (int n = 0; n < NV; ++n) eprcn[m.tet_seg[6 * t + i] * NV + n] += iprcn[NV * t + n];
    }
    for (auto& x : vprcn) x = 1.0 / x;
    if (B.em > 0)
        for (auto& x : eprcn) x = 1.0 / x;
    return err;
}

// inflow faces: velocity components of the face vertices / edges (bdry_ins.h:145-172)
void TetHpIns::vdirichlet() {
    for (int v = 0; v < m.npnt; ++v)
        if (vess[v])
            for (int n = 0; n < NV - 1; ++n) resv[v * NV + n] = 0.0;
}

void TetHpIns::edirichlet() {
    const int em = b().em;
    for (int s = 0; s < m.nseg; ++s)
        if (eess[s])
            for (int k = 0; k < em; ++k)
                for (int n = 0; n < NV - 1; ++n) rese[(s * em0 + k) * NV + n] = 0.0;
}

// tet_hp/setup_preconditioner.cpp:11-210 for em <= 1 (diagonal preconditioner)
void TetHpIns::minvrt() {
    const TetBasis& B = b();
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
    vdirichlet();
    if (B.em == 0) return;
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
    for (int e = 0; e < m.nseg; ++e)
        for (int n = 0; n < NV; ++n) rese[(e * em0 + 0) * NV + n] *= eprcn[e * NV + n] * B.ediag[0];
    edirichlet();
}

// tet_hp/update.cpp:181-417
void TetHpIns::update() {
    const TetBasis& B = b();
    ug0v = ugv;
    ug0e = uge;
    for (int stage = 0; stage < prm.nstage; ++stage) {
        rsdl(stage);
        vdirichlet();
        edirichlet();
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

// tet_hp/allocate.cpp:530-559
template <int NV_>
double TetHp<NV_>::maxres(double* per_var) const {
    double mx[NV];
    for (int n = 0; n < NV; ++n) mx[n] = 0.0;
    for (int i = 0; i < m.npnt; ++i)
        for (int n = 0; n < NV; ++n) mx[n] = std::max(mx[n], std::fabs(resv[i * NV + n]));
    double e = 0.0;
    for (int n = 0; n < NV; ++n) { e = std::max(e, mx[n]); if (per_var) per_var[n] = mx[n]; }
    return e;
}

// tet_hp/mg_restrict.cpp:4-25 (p-levels)
template <int NV_>
void TetHp<NV_>::mg_restrict() {
    isfrst = true;
    coarse_flag = true;
    if (log2p == 0) throw std::runtime_error("oracle TetHp: no coarser p-level");
    --log2p;
    res0v = resv;
    const TetBasis& B = b();
    for (int s = 0; s < m.nseg; ++s)
        for (int k = 0; k < B.em; ++k)
            for (int n = 0; n < NV; ++n) res0e[(s * em0 + k) * NV + n] = rese[(s * em0 + k) * NV + n];
}

// tet_hp/mg_prolongate.cpp:7-14
template <int NV_>
void TetHp<NV_>::mg_prolongate() {
    ++log2p;
    if (log2p == prm.log2pmax) coarse_flag = false;
}

// tet_hp/update.cpp:134-178.  dw is one value for every variable: the reference adds
// each |uht(n)(i)| of the four vertices to the whole dw array (a Blitz array plus a
// scalar), then scales by eps_r and adds eps_a.
template <int NV_>
void TetHp<NV_>::element_jacobian(int tind, double eps_r, double eps_a, double* K) {
    const int T = b().tm, N = T * NV;
    double uht[NV][10], lre[NV][10], lim[NV][10], Rbar[NV][10];
    ugtouht(tind, uht);
    double dw = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int n = 0; n < NV; ++n) dw = dw + std::fabs(uht[n][i]);
    dw = dw * eps_r;
    dw = dw + eps_a;
    element_rsdl(tind, 0, uht, lre, lim);
    for (int i = 0; i < T; ++i)
        for (int n = 0; n < NV; ++n) Rbar[n][i] = lre[n][i] + lim[n][i];
    int kcol = 0;
    for (int mode = 0; mode < T; ++mode)
        for (int var = 0; var < NV; ++var) {
            uht[var][mode] += dw;
            element_rsdl(tind, 0, uht, lre, lim);
            int krow = 0;
            for (int i = 0; i < T; ++i)
                for (int n = 0; n < NV; ++n) K[(krow++) * N + kcol] = (lre[n][i] + lim[n][i] - Rbar[n][i]) / dw;
            ++kcol;
            uht[var][mode] -= dw;
        }
}

template class TetHp<4>;
template class TetHp<5>;

}  // namespace oracle
