// Oracle (test infrastructure only).  See otri_hp_ins.h for the reference map.
#include "otri_hp_ins.h"
#include "olinalg.h"
#include <cmath>
#include <algorithm>
#include <stdexcept>

namespace oracle {

static const int S = 6;  // MXGP stride of the reference work arrays (tri_hp/tri_hp.h:31-33)

TriHpIns::TriHpIns(const Mesh& mesh, const Params& p) : TriHp<3>(mesh, p) {
    tau.assign(m.ntri * NV, 0.0);
}

// tri_hp/ins/rsdl.cpp:17-82 and the straight branch :259-433 (steady: mvel = dugdt = 0)
void TriHpIns::element_rsdl(int tind, int stage, const double uh[NV][15], double lf_re[NV][15], double lf_im[NV][15]) {
    if (curved(tind)) { element_rsdl_curved(tind, stage, uh, lf_re, lf_im); return; }
    const TriBasis& B = b();
    const int lgpx = B.gpx, lgpn = B.gpn;
    const double bd0 = 0.0, rhobd0 = prm.rho * bd0, lmu = prm.mu;
    const int v0 = m.tri_pnt[3 * tind], v1 = m.tri_pnt[3 * tind + 1], v2 = m.tri_pnt[3 * tind + 2];
    double ld[2][2];
    for (int n = 0; n < 2; ++n) {
        ld[n][0] = 0.5 * (m.pnts[2 * v2 + n] - m.pnts[2 * v1 + n]);
        ld[n][1] = 0.5 * (m.pnts[2 * v0 + n] - m.pnts[2 * v1 + n]);
    }
    double u[NV][S * S], du[NV][2][S * S], cv[2][2][S * S], df[2][2][S * S], res[NV][S * S];
    const double mv0 = 0.0, mv1 = 0.0;
    if (prm.beta[stage] > 0.0) {
        for (int n = 0; n < NV - 1; ++n) B.proj(uh[n], u[n], du[n][0], du[n][1], S);
        B.proj(uh[NV - 1], u[NV - 1], S);
    } else {
        for (int n = 0; n < NV; ++n) B.proj(uh[n], u[n], S);
    }
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < B.tm; ++i) { lf_re[n][i] = 0.0; lf_im[n][i] = 0.0; }

    for (int i = 0; i < lgpx; ++i)
        for (int j = 0; j < lgpn; ++j) {
            const int g = i * S + j;
            const double fluxx = prm.rho * (u[0][g] - mv0);
            const double fluxy = prm.rho * (u[1][g] - mv1);
            du[2][0][g] = +ld[1][1] * fluxx - ld[0][1] * fluxy;
            du[2][1][g] = -ld[1][0] * fluxx + ld[0][0] * fluxy;
            for (int n = 0; n < NV - 1; ++n) {
                cv[n][0][g] = u[n][g] * du[2][0][g];
                cv[n][1][g] = u[n][g] * du[2][1][g];
            }
            cv[0][0][g] += ld[1][1] * u[2][g];
            cv[0][1][g] -= ld[1][0] * u[2][g];
            cv[1][0][g] -= ld[0][1] * u[2][g];
            cv[1][1][g] += ld[0][0] * u[2][g];
        }
    for (int n = 0; n < NV - 1; ++n) B.intgrtrs(lf_im[n], cv[n][0], cv[n][1], S);
    B.intgrtrs(lf_im[NV - 1], du[2][0], du[2][1], S);

    if (prm.beta[stage] > 0.0) {
        const double cjcb = ld[0][0] * ld[1][1] - ld[1][0] * ld[0][1];
        const double cjcbi = lmu / cjcb;
        const double lrhorbd0 = rhobd0 * cjcb;
        double visc[2][2][2][2];
        visc[0][0][0][0] = -cjcbi * (2. * ld[1][1] * ld[1][1] + ld[0][1] * ld[0][1]);
        visc[0][0][1][1] = -cjcbi * (2. * ld[1][0] * ld[1][0] + ld[0][0] * ld[0][0]);
        visc[0][0][0][1] = cjcbi * (2. * ld[1][1] * ld[1][0] + ld[0][1] * ld[0][0]);
        visc[1][1][0][0] = -cjcbi * (ld[1][1] * ld[1][1] + 2. * ld[0][1] * ld[0][1]);
        visc[1][1][1][1] = -cjcbi * (ld[1][0] * ld[1][0] + 2. * ld[0][0] * ld[0][0]);
        visc[1][1][0][1] = cjcbi * (ld[1][1] * ld[1][0] + 2. * ld[0][1] * ld[0][0]);
        visc[0][1][0][0] = cjcbi * ld[0][1] * ld[1][1];
        visc[0][1][1][1] = cjcbi * ld[0][0] * ld[1][0];
        visc[0][1][0][1] = -cjcbi * ld[0][1] * ld[1][0];
        visc[0][1][1][0] = -cjcbi * ld[0][0] * ld[1][1];
        for (int i = 0; i < lgpx; ++i)
            for (int j = 0; j < lgpn; ++j) {
                const int g = i * S + j;
                const double rhorbd0 = lrhorbd0;
                for (int n = 0; n < NV - 1; ++n) res[n][g] = rhorbd0 * u[n][g];
                res[NV - 1][g] = rhorbd0;
                df[0][0][g] = visc[0][0][0][0] * du[0][0][g] + visc[0][1][0][0] * du[1][0][g] +
                              visc[0][0][0][1] * du[0][1][g] + visc[0][1][0][1] * du[1][1][g];
                df[0][1][g] = visc[0][0][0][1] * du[0][0][g] + visc[0][1][1][0] * du[1][0][g] +
                              visc[0][0][1][1] * du[0][1][g] + visc[0][1][1][1] * du[1][1][g];
                df[1][0][g] = visc[0][1][0][0] * du[0][0][g] + visc[1][1][0][0] * du[1][0][g] +
                              visc[0][1][1][0] * du[0][1][g] + visc[1][1][0][1] * du[1][1][g];
                df[1][1][g] = visc[0][1][0][1] * du[0][0][g] + visc[1][1][0][1] * du[1][0][g] +
                              visc[0][1][1][1] * du[0][1][g] + visc[1][1][1][1] * du[1][1][g];
                for (int n = 0; n < NV - 1; ++n) { cv[n][0][g] += df[n][0][g]; cv[n][1][g] += df[n][1][g]; }
            }
        for (int n = 0; n < NV; ++n) B.intgrt(lf_re[n], res[n], S);
        for (int n = 0; n < NV - 1; ++n) { B.derivr(cv[n][0], res[n], S); B.derivs(cv[n][1], res[n], S); }
        B.derivr(du[2][0], res[2], S);
        B.derivs(du[2][1], res[2], S);

        const double h = inscribedradius(tind) / (0.25 * (B.p + 1) * (B.p + 1));
        for (int i = 0; i < lgpx; ++i)
            for (int j = 0; j < lgpn; ++j) {
                const int g = i * S + j;
                const double q = std::pow(u[0][g] - 0.5 * mv0, 2.0) + std::pow(u[1][g] - 0.5 * mv1, 2.0);
                const double nu = prm.mu / prm.rho;
                double gam = 3.0 * q + (0.5 * h * bd0 + 2. * nu / h) * (0.5 * h * bd0 + 2. * nu / h);
                if (prm.mu + bd0 == 0.0) gam = std::max(gam, 0.1);
                tau[tind * NV + 0] = prm.adis * h / (cjcb * std::sqrt(gam));
                tau[tind * NV + NV - 1] = std::sqrt(q) * tau[tind * NV + 0];
                const double t0 = tau[tind * NV] * res[0][g], t1 = tau[tind * NV] * res[1][g];
                const double t2 = tau[tind * NV + NV - 1] * res[NV - 1][g];
                df[0][0][g] -= (ld[1][1] * (2 * u[0][g] - mv0) - ld[0][1] * (u[1][g] - mv1)) * t0 -
                               ld[0][1] * u[0][g] * t1 + ld[1][1] * t2;
                df[0][1][g] -= (-ld[1][0] * (2 * u[0][g] - mv0) + ld[0][0] * (u[1][g] - mv1)) * t0 +
                               ld[0][0] * u[0][g] * t1 - ld[1][0] * t2;
                df[1][0][g] -= +ld[1][1] * u[1][g] * t0 +
                               (ld[1][1] * (u[0][g] - mv0) - ld[0][1] * (2. * u[1][g] - mv1)) * t1 - ld[0][1] * t2;
                df[1][1][g] -= -ld[1][0] * u[1][g] * t0 +
                               (-ld[1][0] * (u[0][g] - mv0) + ld[0][0] * (2. * u[1][g] - mv1)) * t1 + ld[0][0] * t2;
                du[2][0][g] = -(ld[1][1] * t0 - ld[0][1] * t1);
                du[2][1][g] = -(-ld[1][0] * t0 + ld[0][0] * t1);
            }
        for (int n = 0; n < NV - 1; ++n) B.intgrtrs(lf_re[n], df[n][0], df[n][1], S);
        B.intgrtrs(lf_re[NV - 1], du[2][0], du[2][1], S);
        for (int n = 0; n < NV; ++n)
            for (int i = 0; i < B.tm; ++i) lf_re[n][i] *= prm.beta[stage];
    }
}

// tri_hp/ins/rsdl.cpp:33-36 (tri_hp::metric::calc_metrics, tri_hp/metric.cpp:12-21:
// crdtocht + proj_bdry) and the curved branch :83-257, steady (mvel = dugdt = 0)
void TriHpIns::element_rsdl_curved(int tind, int stage, const double uh[NV][15], double lf_re[NV][15],
                                   double lf_im[NV][15]) {
    const TriBasis& B = b();
    const int lgpx = B.gpx, lgpn = B.gpn;
    const double bd0 = 0.0, rhobd0 = prm.rho * bd0, lmu = prm.mu;
    double cht[2][15], crd[2][S * S], dcrd[2][2][S * S];
    crdtocht(tind, cht);
    for (int n = 0; n < 2; ++n) B.proj(cht[n], crd[n], dcrd[n][0], dcrd[n][1], S);
    double u[NV][S * S], du[NV][2][S * S], cv[2][2][S * S], df[2][2][S * S], res[NV][S * S];
    const double mv0 = 0.0, mv1 = 0.0;
    if (prm.beta[stage] > 0.0) {
        for (int n = 0; n < NV - 1; ++n) B.proj(uh[n], u[n], du[n][0], du[n][1], S);
        B.proj(uh[NV - 1], u[NV - 1], S);
    } else {
        for (int n = 0; n < NV; ++n) B.proj(uh[n], u[n], S);
    }
    for (int n = 0; n < NV; ++n)
        for (int i = 0; i < B.tm; ++i) { lf_re[n][i] = 0.0; lf_im[n][i] = 0.0; }

    for (int i = 0; i < lgpx; ++i)
        for (int j = 0; j < lgpn; ++j) {
            const int g = i * S + j;
            const double d00 = dcrd[0][0][g], d01 = dcrd[0][1][g], d10 = dcrd[1][0][g], d11 = dcrd[1][1][g];
            const double fluxx = prm.rho * (u[0][g] - mv0);
            const double fluxy = prm.rho * (u[1][g] - mv1);
            du[NV - 1][0][g] = +d11 * fluxx - d01 * fluxy;
            du[NV - 1][1][g] = -d10 * fluxx + d00 * fluxy;
            for (int n = 0; n < NV - 1; ++n) {
                cv[n][0][g] = u[n][g] * du[NV - 1][0][g];
                cv[n][1][g] = u[n][g] * du[NV - 1][1][g];
            }
            cv[0][0][g] += d11 * u[NV - 1][g];
            cv[0][1][g] -= d10 * u[NV - 1][g];
            cv[1][0][g] -= d01 * u[NV - 1][g];
            cv[1][1][g] += d00 * u[NV - 1][g];
        }
    for (int n = 0; n < NV - 1; ++n) B.intgrtrs(lf_im[n], cv[n][0], cv[n][1], S);
    B.intgrtrs(lf_im[NV - 1], du[NV - 1][0], du[NV - 1][1], S);

    if (prm.beta[stage] > 0.0) {
        for (int i = 0; i < lgpx; ++i)
            for (int j = 0; j < lgpn; ++j) {
                const int g = i * S + j;
                const double d00 = dcrd[0][0][g], d01 = dcrd[0][1][g], d10 = dcrd[1][0][g], d11 = dcrd[1][1][g];
                const double cjcb = d00 * d11 - d10 * d01;
                const double rhorbd0 = rhobd0 * cjcb;
                const double cjcbi = lmu / cjcb;
                for (int n = 0; n < NV - 1; ++n) res[n][g] = rhorbd0 * u[n][g];
                res[NV - 1][g] = rhorbd0;
                double visc[2][2][2][2];
                visc[0][0][0][0] = -cjcbi * (2. * d11 * d11 + d01 * d01);
                visc[0][0][1][1] = -cjcbi * (2. * d10 * d10 + d00 * d00);
                visc[0][0][0][1] = cjcbi * (2. * d11 * d10 + d01 * d00);
                visc[1][1][0][0] = -cjcbi * (d11 * d11 + 2. * d01 * d01);
                visc[1][1][1][1] = -cjcbi * (d10 * d10 + 2. * d00 * d00);
                visc[1][1][0][1] = cjcbi * (d11 * d10 + 2. * d01 * d00);
                visc[0][1][0][0] = cjcbi * d01 * d11;
                visc[0][1][1][1] = cjcbi * d00 * d10;
                visc[0][1][0][1] = -cjcbi * d01 * d10;
                visc[0][1][1][0] = -cjcbi * d00 * d11;
                df[0][0][g] = visc[0][0][0][0] * du[0][0][g] + visc[0][1][0][0] * du[1][0][g] +
                              visc[0][0][0][1] * du[0][1][g] + visc[0][1][0][1] * du[1][1][g];
                df[0][1][g] = visc[0][0][0][1] * du[0][0][g] + visc[0][1][1][0] * du[1][0][g] +
                              visc[0][0][1][1] * du[0][1][g] + visc[0][1][1][1] * du[1][1][g];
                df[1][0][g] = visc[0][1][0][0] * du[0][0][g] + visc[1][1][0][0] * du[1][0][g] +
                              visc[0][1][1][0] * du[0][1][g] + visc[1][1][0][1] * du[1][1][g];
                df[1][1][g] = visc[0][1][0][1] * du[0][0][g] + visc[1][1][0][1] * du[1][0][g] +
                              visc[0][1][1][1] * du[0][1][g] + visc[1][1][1][1] * du[1][1][g];
                for (int n = 0; n < NV - 1; ++n) { cv[n][0][g] += df[n][0][g]; cv[n][1][g] += df[n][1][g]; }
            }
        for (int n = 0; n < NV; ++n) B.intgrt(lf_re[n], res[n], S);
        for (int n = 0; n < NV - 1; ++n) { B.derivr(cv[n][0], res[n], S); B.derivs(cv[n][1], res[n], S); }
        B.derivr(du[NV - 1][0], res[NV - 1], S);
        B.derivs(du[NV - 1][1], res[NV - 1], S);

        const double h = inscribedradius(tind) / (0.25 * (B.p + 1) * (B.p + 1));
        for (int i = 0; i < lgpx; ++i)
            for (int j = 0; j < lgpn; ++j) {
                const int g = i * S + j;
                const double d00 = dcrd[0][0][g], d01 = dcrd[0][1][g], d10 = dcrd[1][0][g], d11 = dcrd[1][1][g];
                const double q = std::pow(u[0][g] - 0.5 * mv0, 2.0) + std::pow(u[1][g] - 0.5 * mv1, 2.0);
                const double nu = prm.mu / prm.rho;
                const double cjcb = d00 * d11 - d10 * d01;
                double gam = 3.0 * q + (0.5 * h * bd0 + 2. * nu / h) * (0.5 * h * bd0 + 2. * nu / h);
                if (prm.mu + bd0 == 0.0) gam = std::max(gam, 0.1);
                tau[tind * NV + 0] = prm.adis * h / (cjcb * std::sqrt(gam));
                tau[tind * NV + NV - 1] = std::sqrt(q) * tau[tind * NV + 0];
                const double t0 = tau[tind * NV] * res[0][g], t1 = tau[tind * NV] * res[1][g];
                const double t2 = tau[tind * NV + NV - 1] * res[NV - 1][g];
                df[0][0][g] -= (d11 * (2 * u[0][g] - mv0) - d01 * (u[1][g] - mv1)) * t0 - d01 * u[0][g] * t1 + d11 * t2;
                df[0][1][g] -= (-d10 * (2 * u[0][g] - mv0) + d00 * (u[1][g] - mv1)) * t0 + d00 * u[0][g] * t1 - d10 * t2;
                df[1][0][g] -= +d11 * u[1][g] * t0 + (d11 * (u[0][g] - mv0) - d01 * (2. * u[1][g] - mv1)) * t1 - d01 * t2;
                df[1][1][g] -= -d10 * u[1][g] * t0 + (-d10 * (u[0][g] - mv0) + d00 * (2. * u[1][g] - mv1)) * t1 + d00 * t2;
                du[NV - 1][0][g] = -(d11 * t0 - d01 * t1);
                du[NV - 1][1][g] = -(-d10 * t0 + d00 * t1);
            }
        for (int n = 0; n < NV - 1; ++n) B.intgrtrs(lf_re[n], df[n][0], df[n][1], S);
        B.intgrtrs(lf_re[NV - 1], du[NV - 1][0], du[NV - 1][1], S);
        for (int n = 0; n < NV; ++n)
            for (int i = 0; i < B.tm; ++i) lf_re[n][i] *= prm.beta[stage];
    }
}

void ins_edge_flux(int type, const double u[3], const double nrm[2], double side_length, int p, double rho,
                   double mu, const double ub_in[3], double flx[3]) {
    const double mv[2] = {0.0, 0.0};
    if (type == EDGE_GENERIC) {
        // bdry_ins::generic::flux: the ibc pressure on the momentum
        flx[2] = rho * ((u[0] - mv[0]) * nrm[0] + (u[1] - mv[1]) * nrm[1]);
        for (int n = 0; n < 2; ++n) flx[n] = flx[2] * u[n] + ub_in[2] * nrm[n];
        return;
    }
    if (type == EDGE_EULER) {
        // bdry_ins::euler::flux: ibc velocity, interior pressure
        flx[2] = rho * ((ub_in[0] - mv[0]) * nrm[0] + (ub_in[1] - mv[1]) * nrm[1]);
        for (int n = 0; n < 2; ++n) flx[n] = flx[2] * ub_in[n] + u[2] * nrm[n];
        return;
    }
    // bdry_ins::characteristic::flux (natural fluxes zero, bd(0) = 0)
    for (int n = 0; n < 3; ++n) flx[n] = 0.0;
    const double bd0 = 0.0;
    const double rhoi = 1. / rho, nu = mu / rho;
    const double hmax = side_length * 2.0 / (0.25 * (p + 1) * (p + 1));
    const double qmax = std::pow(u[0] - 0.5 * mv[0], 2.0) + std::pow(u[1] - 0.5 * mv[1], 2.0);
    const double gam = 3.0 * qmax + (0.5 * hmax * bd0 + 2. * nu / hmax) * (0.5 * hmax * bd0 + 2. * nu / hmax);
    const double ul = u[0] * nrm[0] + u[1] * nrm[1];
    const double vl = -u[0] * nrm[1] + u[1] * nrm[0];
    const double pl = u[2];
    double ub[3] = {ub_in[0], ub_in[1], ub_in[2]}, uvp[3];
    const double ur = ub[0] * nrm[0] + ub[1] * nrm[1];
    const double vr = -ub[0] * nrm[1] + ub[1] * nrm[0];
    const double pr = ub[2];
    const double um = mv[0] * nrm[0] + mv[1] * nrm[1];
    const double cl = std::sqrt((ul - .5 * um) * (ul - .5 * um) + gam);
    const double cr = std::sqrt((ur - .5 * um) * (ur - .5 * um) + gam);
    const double c = 0.5 * (cl + cr), s = 0.5 * (ul + ur), v = 0.5 * (vl + vr);
    const double den = 1. / (2 * c);
    const double lam0 = s - um, lam1 = s - .5 * um + c, lam2 = s - .5 * um - c;
    uvp[0] = ((pl - pr) * rhoi + (ul * lam1 - ur * lam2)) * den;
    if (lam0 > 0.0)
        uvp[1] = v * ((pr - pl) * rhoi + lam2 * (ur - ul)) * den / (lam0 - lam2) + vl;
    else
        uvp[1] = v * ((pr - pl) * rhoi + lam1 * (ur - ul)) * den / (lam0 - lam1) + vr;
    uvp[2] = (rho * (ul - ur) * gam - lam2 * pl + lam1 * pr) * den;
    ub[0] = uvp[0] * nrm[0] - uvp[1] * nrm[1];
    ub[1] = uvp[0] * nrm[1] + uvp[1] * nrm[0];
    ub[2] = uvp[2];
    const double f2 = rho * ((ub[0] - mv[0]) * nrm[0] + (ub[1] - mv[1]) * nrm[1]);
    for (int n = 0; n < 2; ++n) flx[n] += f2 * ub[n] + ub[2] * nrm[n];
    flx[2] = f2;
}

// tri_hp/hp_boundary.cpp:973-1049 with the boundary's flux (ins_edge_flux)
void TriHpIns::edge_rsdl() {
    const TriBasis& B = b();
    const int nebd = int(m.ebdry_start.size()) - 1;
    for (int e = 0; e < nebd; ++e) {
        // a periodic boundary is a comm boundary: no flux of its own
        if (!prdc_bdry.empty() && prdc_bdry[e]) continue;
        for (int j = m.ebdry_start[e]; j < m.ebdry_start[e + 1]; ++j) {
            const int sind = m.ebdry_seg[j];
            const int v0 = m.seg_pnt[2 * sind], v1 = m.seg_pnt[2 * sind + 1];
            double u1d[NV][8], cht[2][8], crd[2][8], dcrd[2][8], uu[NV][8], r[NV][8], lfl[NV][8];
            for (int n = 0; n < NV; ++n) {
                u1d[n][0] = ugv[v0 * NV + n];
                u1d[n][1] = ugv[v1 * NV + n];
                for (int mm = 0; mm < B.sm; ++mm) u1d[n][mm + 2] = ugs[(sind * sm0 + mm) * NV + n];
            }
            crdtocht1d(sind, cht);
            for (int n = 0; n < 2; ++n) B.proj1d(cht[n], crd[n], dcrd[n]);
            for (int n = 0; n < NV; ++n) B.proj1d(u1d[n], uu[n]);
            for (int k = 0; k < B.gpx; ++k) {
                double nrm0 = dcrd[1][k], nrm1 = -dcrd[0][k];
                const double jcb = std::sqrt(nrm0 * nrm0 + nrm1 * nrm1);
                nrm0 /= jcb; nrm1 /= jcb;
                double flx[NV];
                const double uk[3] = {uu[0][k], uu[1][k], uu[2][k]}, nk[2] = {nrm0, nrm1};
                const double ub[3] = {ibc_u[0], ibc_u[1], prm.bc_pressure};
                ins_edge_flux(ebdry_flux.empty() ? EDGE_GENERIC : ebdry_flux[e], uk, nk, jcb, B.p, prm.rho, prm.mu,
                              ub, flx);
                for (int n = 0; n < NV; ++n) r[n][k] = flx[n] * jcb;
            }
            for (int n = 0; n < NV; ++n) B.intgrt1d(lfl[n], r[n]);
            for (int n = 0; n < NV; ++n) resv[v0 * NV + n] += lfl[n][0];
            for (int n = 0; n < NV; ++n) resv[v1 * NV + n] += lfl[n][1];
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) ress[(sind * smf + k) * NV + n] += lfl[n][k + 2];
        }
    }
}

// tri_hp/ins/setup_preconditioner.cpp:9-221 (diagonal, REFINED_WAY, curved and straight)
// followed by tri_hp/setup_preconditioner.cpp:9-96.  The reference's diagonal branch
// declares a second `err` that shadows the returned one (:15), so it breaks out of
// the element loop but reports 0; this restatement reports the error count (1).
int TriHpIns::setup_preconditioner() {
    const TriBasis& B = b();
    const double bd0 = 0.0;
    const double nu = prm.mu / prm.rho;
    int err = 0;
    double u[2][S * S];
    for (int tind = 0; tind < m.ntri; ++tind) {
        double jcb = 0.25 * area(tind);
        const int* v = &m.tri_pnt[3 * tind];
        ugtouht(tind);
        for (int n = 0; n < 2; ++n) B.proj(uht[n], u[n], S);
        double qmax = 0.0, hmax = 0.0, h;
        const double pf = 0.25 * (B.p + 1) * (B.p + 1);
        if (curved(tind)) {
            // setup_preconditioner.cpp:44-90
            double cht[2][15], crd[2][S * S], dcrd[2][2][S * S];
            crdtocht(tind, cht);
            for (int n = 0; n < 2; ++n) B.proj(cht[n], crd[n], dcrd[n][0], dcrd[n][1], S);
            double jcbmin = jcb;
            for (int i = 0; i < B.gpx; ++i)
                for (int j = 0; j < B.gpn; ++j) {
                    const int g = i * S + j;
                    jcbmin = std::min(jcbmin, dcrd[0][0][g] * dcrd[1][1][g] - dcrd[1][0][g] * dcrd[0][1][g]);
                    double hh = 0.0;
                    for (int n = 0; n < 2; ++n) hh += dcrd[n][0][g] * dcrd[n][0][g];
                    hmax = std::max(hh, hmax);
                    hh = 0.0;
                    for (int n = 0; n < 2; ++n) hh += dcrd[n][1][g] * dcrd[n][1][g];
                    hmax = std::max(hh, hmax);
                    hh = 0.0;
                    for (int n = 0; n < 2; ++n) hh += (dcrd[n][1][g] - dcrd[n][0][g]) * (dcrd[n][1][g] - dcrd[n][0][g]);
                    hmax = std::max(hh, hmax);
                    const double q = std::pow(u[0][g] - 0.5 * 0.0, 2.0) + std::pow(u[1][g] - 0.5 * 0.0, 2.0);
                    qmax = std::max(qmax, q);
                }
            hmax = 2. * std::sqrt(hmax);
            h = 4. * jcbmin / (pf * hmax);
            hmax = hmax / pf;
        } else {
            for (int i = 0; i < B.gpx; ++i)
                for (int j = 0; j < B.gpn; ++j) {
                    const double q = std::pow(u[0][i * S + j] - 0.5 * 0.0, 2.0) + std::pow(u[1][i * S + j] - 0.5 * 0.0, 2.0);
                    qmax = (qmax > q) ? qmax : q;  // MAX of sys/param.h
                }
            for (int j = 0; j < 3; ++j) {
                const int a = v[j], c = v[(j + 1) % 3];
                const double hh = std::pow(m.pnts[2 * a] - m.pnts[2 * c], 2.0) + std::pow(m.pnts[2 * a + 1] - m.pnts[2 * c + 1], 2.0);
                hmax = (hh > hmax ? hh : hmax);
            }
            hmax = std::sqrt(hmax);
            h = 4. * jcb / (pf * hmax);
            hmax = hmax / pf;
        }
        if (!(h > 0.0)) { err = 1; break; }
        if (std::isnan(qmax)) { err = 1; break; }
        double gam = 3.0 * qmax + (0.5 * hmax * bd0 + 2. * nu / hmax) * (0.5 * hmax * bd0 + 2. * nu / hmax);
        if (prm.mu + bd0 == 0.0) gam = std::max(gam, 0.1);
        const double q = std::sqrt(qmax);
        const double lam1 = q + std::sqrt(qmax + gam);
        tau[tind * NV + 0] = prm.adis * h / (jcb * std::sqrt(gam));
        tau[tind * NV + NV - 1] = qmax * tau[tind * NV + 0];
        jcb *= 2. * nu * (1. / (hmax * hmax) + 1. / (h * h)) + 3 * lam1 / h;
        tprcn[tind * NV + 0] = prm.rho * jcb;
        tprcn[tind * NV + 1] = prm.rho * jcb;
        tprcn[tind * NV + 2] = jcb / gam;
    }
    assemble_prcn(err ? 0 : m.ntri);
    return err;
}

}  // namespace oracle
