// Oracle (test infrastructure only): tetrahedral basis restatement, p in {1, 2}.
// See otet_basis.h for the reference files followed.
#include "otet_basis.h"

#include <cmath>
#include <stdexcept>

#include "olinalg.h"
#include "otri_basis.h"

namespace oracle {

namespace {

// Mode m = gx(ix) * gy(iy) * gz(iz) for the vertex and edge modes, read off the
// sum-factorisation of tet_basis::proj (tet_basis/proj.cpp:52-239) and
// tet_basis::intgrt (tet_basis/intgrt.cpp:34-200): modes V0..V3, E1..E6 (edge i of
// the tet joins local vertices (2,3), (3,1), (2,1), (1,0), (2,0), (3,0),
// tet_mesh/setup.cpp:230-235), edge mode k of every edge uses polynomial k.
struct ModeFactors {
    int ix, iy, iz;
};

ModeFactors mode_factors(int m, int em) {
    if (m == 0) return {0, 0, 0};
    if (m == 1) return {0, 1, 1};
    if (m == 2) return {1, 2, 1};
    if (m == 3) return {2, 2, 1};
    const int e = (m - 4) / em, k = (m - 4) - e * em;
    switch (e) {
        case 0: return {3 + k, 3 + k, 2 + k};            // E1: bubble in x
        case 1: return {2, 3 + em + k, 2 + k};           // E2
        case 2: return {1, 3 + em + k, 2 + k};           // E3
        case 3: return {0, 1, 2 + em + k};               // E4: bubble in z
        case 4: return {1, 2, 2 + em + k};               // E5
        default: return {2, 2, 2 + em + k};              // E6
    }
}

// Lagrange factors of tet_basis/initialize.cpp:151-175 (x), :225-244 (y), :286-304 (z)
void lagrange_factors(const std::vector<double>& x, std::vector<double>& dlt, std::vector<double>& dlt2,
                      std::vector<double>* dlt3) {
    const int n = (int)x.size();
    dlt.assign(n, 1.0);
    dlt2.assign(n * n, 0.0);
    if (dlt3) dlt3->assign(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            if (k != i) dlt[i] *= (x[i] - x[k]);
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < n; ++m) {
            if (m == i) continue;
            dlt2[i * n + m] = dlt[i] / (x[i] - x[m]);
            if (dlt3) (*dlt3)[i * n + m] = dlt2[i * n + m] * (1.0 + x[i]) / 2.0;
        }
    for (int i = 0; i < n; ++i) dlt[i] = 1.0 / dlt[i];
}

}  // namespace

TetBasis::TetBasis(int p_) : p(p_) {
    if (p < 1 || p > 2) throw std::runtime_error("oracle TetBasis: p must be 1 or 2");
    em = p - 1;
    fm = 0;
    im = 0;
    tm = (p + 1) * (p + 2) * (p + 3) / 6;
    bm = 4 + 6 * em;
    nmodx = em + 3;
    nmody = 3 + 2 * em + fm;
    nmodz = 2 + 2 * em + fm + im;
    gpx = gpy = gpz = p + 1;
    G = gpx * gpy * gpz;
    tm2d = 3 + 3 * em;

    const int nr = std::max(gpx + 1, em + 1) + 1;
    std::vector<double> ra(nr), rb(nr);
    // x: Gauss-Legendre (initialize.cpp:136-150)
    recur_coeffs(gpx + 1, 1, 0.0, 0.0, ra.data(), rb.data());
    xp.resize(gpx);
    wtx.resize(gpx);
    gauss_rule(gpx, ra.data(), rb.data(), xp.data(), wtx.data());
    lagrange_factors(xp, dltx, dltx2, &dltx3);
    // edge functions: monic Jacobi(2,2) recurrence (initialize.cpp:180-189)
    a0.assign(em + 1, 0.0);
    b0.assign(em + 1, 0.0);
    recur_coeffs(em + 1, 6, 2.0, 2.0, a0.data(), b0.data());
    // y: Gauss-Jacobi(1,0), weights halved (initialize.cpp:215-233)
    recur_coeffs(gpy + 1, 6, 1.0, 0.0, ra.data(), rb.data());
    yp.resize(gpy);
    wty.resize(gpy);
    gauss_rule(gpy, ra.data(), rb.data(), yp.data(), wty.data());
    for (auto& w : wty) w *= 0.5;
    lagrange_factors(yp, dlty, dlty2, &dlty3);
    // z: Gauss-Jacobi(2,0), weights quartered (initialize.cpp:277-295)
    recur_coeffs(gpz + 1, 6, 2.0, 0.0, ra.data(), rb.data());
    zp.resize(gpz);
    wtz.resize(gpz);
    gauss_rule(gpz, ra.data(), rb.data(), zp.data(), wtz.data());
    for (auto& w : wtz) w *= 0.25;
    lagrange_factors(zp, dltz, dltz2, nullptr);

    x0.resize(gpx);
    y0.resize(gpy);
    y1.resize(gpy);
    z0.resize(gpz);
    gx.assign(gpx * nmodx, 0.0); dgx.assign(gpx * nmodx, 0.0);
    gy.assign(gpy * nmody, 0.0); dgy.assign(gpy * nmody, 0.0);
    gz.assign(gpz * nmodz, 0.0); dgz.assign(gpz * nmodz, 0.0);
    std::vector<double> px(nmodx), dpx(nmodx), py(nmody), dpy(nmody), pz(nmodz), dpz(nmodz);
    for (int i = 0; i < gpx; ++i) {
        x0[i] = 0.5 * (1.0 + xp[i]);
        ptvalues_deriv(xp[i], 0.0, 0.0, px.data(), dpx.data(), py.data(), dpy.data(), pz.data(), dpz.data());
        for (int m = 0; m < nmodx; ++m) { gx[i * nmodx + m] = px[m]; dgx[i * nmodx + m] = dpx[m]; }
    }
    for (int j = 0; j < gpy; ++j) {
        y0[j] = 2.0 / (1.0 - yp[j]);
        y1[j] = 0.5 * (1.0 + yp[j]);
        ptvalues_deriv(0.0, yp[j], 0.0, px.data(), dpx.data(), py.data(), dpy.data(), pz.data(), dpz.data());
        for (int m = 0; m < nmody; ++m) { gy[j * nmody + m] = py[m]; dgy[j * nmody + m] = dpy[m]; }
    }
    for (int k = 0; k < gpz; ++k) {
        z0[k] = 2.0 / (1.0 - zp[k]);
        ptvalues_deriv(0.0, 0.0, zp[k], px.data(), dpx.data(), py.data(), dpy.data(), pz.data(), dpz.data());
        for (int m = 0; m < nmodz; ++m) { gz[k * nmodz + m] = pz[m]; dgz[k * nmodz + m] = dpz[m]; }
    }
    // weighted tables (initialize.cpp:371-401)
    gxwtx.assign(nmodx * gpx, 0.0); dgxwtx.assign(nmodx * gpx, 0.0);
    for (int m = 0; m < nmodx; ++m)
        for (int i = 0; i < gpx; ++i) {
            gxwtx[m * gpx + i] = gx[i * nmodx + m] * wtx[i];
            dgxwtx[m * gpx + i] = dgx[i * nmodx + m] * wtx[i];
        }
    gywty.assign(nmody * gpy, 0.0); dgywty.assign(nmody * gpy, 0.0);
    for (int m = 0; m < nmody; ++m)
        for (int j = 0; j < gpy; ++j) {
            gywty[m * gpy + j] = gy[j * nmody + m] * wty[j];
            dgywty[m * gpy + j] = dgy[j * nmody + m] * wty[j];
        }
    gzwtz.assign(nmodz * gpz, 0.0); dgzwtz.assign(nmodz * gpz, 0.0);
    for (int m = 0; m < nmodz; ++m)
        for (int k = 0; k < gpz; ++k) {
            gzwtz[m * gpz + k] = gz[k * nmodz + m] * wtz[k];
            dgzwtz[m * gpz + k] = dgz[k * nmodz + m] * wtz[k];
        }
    lumpinv();
}

// tet_basis/ptvalues.cpp:154-336, vertex and edge functions (face/interior loops are
// empty for em <= 1)
void TetBasis::ptvalues_deriv(double x, double y, double z, double* pgx, double* dpgx, double* pgy, double* dpgy,
                              double* pgz, double* dpgz) const {
    int ind = 0;
    pgx[ind] = 1.0; dpgx[ind] = 0.0;
    pgx[++ind] = 0.5 * (1 - x); dpgx[ind] = -0.5;
    pgx[++ind] = 0.5 * (1 + x); dpgx[ind] = 0.5;
    double pk = 1.0, pkm = 0.0, dpk = 0.0, dpkm = 0.0;
    for (int m = 0; m < em; ++m) {
        pgx[++ind] = 0.25 * (1.0 + x) * (1.0 - x) * pk;
        dpgx[ind] = -x * 0.5 * pk + (1.0 + x) * (1.0 - x) * 0.25 * dpk;
        const double pkp = (x - a0[m]) * pk - b0[m] * pkm;
        const double dpkp = pk + (x - a0[m]) * dpk - b0[m] * dpkm;
        dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
    }
    ind = 0;
    pgy[ind] = 1.0; dpgy[ind] = 0.0;
    pgy[++ind] = 0.5 * (1 + y); dpgy[ind] = 0.5;
    pgy[++ind] = 0.5 * (1 - y); dpgy[ind] = -0.5;
    for (int m = 2; m <= em + 1; ++m) {
        pgy[++ind] = std::pow(0.5 * (1 - y), m);
        dpgy[ind] = -0.5 * m * std::pow(0.5 * (1.0 - y), m - 1);
    }
    pk = 1.0; pkm = 0.0; dpk = 0.0; dpkm = 0.0;
    for (int n = 0; n < em; ++n) {
        pgy[++ind] = (1.0 - y) * (1.0 + y) * 0.25 * pk;
        dpgy[ind] = -0.5 * y * pk + (1.0 - y) * (1.0 + y) * 0.25 * dpk;
        const double pkp = (y - a0[n]) * pk - b0[n] * pkm;
        const double dpkp = pk + (y - a0[n]) * dpk - b0[n] * dpkm;
        dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
    }
    ind = 0;
    pgz[ind] = 0.5 * (1 + z); dpgz[ind] = 0.5;
    pgz[++ind] = 0.5 * (1 - z); dpgz[ind] = -0.5;
    for (int m = 2; m <= em + 1; ++m) {
        pgz[++ind] = std::pow(0.5 * (1 - z), m);
        dpgz[ind] = -0.5 * m * std::pow(0.5 * (1.0 - z), m - 1);
    }
    pk = 1.0; pkm = 0.0; dpk = 0.0; dpkm = 0.0;
    for (int n = 0; n < em; ++n) {
        pgz[++ind] = (1.0 - z) * (1.0 + z) * 0.25 * pk;
        dpgz[ind] = -0.5 * z * pk + (1.0 - z) * (1.0 + z) * 0.25 * dpk;
        const double pkp = (z - a0[n]) * pk - b0[n] * pkm;
        const double dpkp = pk + (z - a0[n]) * dpk - b0[n] * dpkm;
        dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
    }
}

// tet_basis/proj.cpp:22-264: z pass, then y pass, then x pass; the r, s, t
// derivatives through the collapsed-coordinate chain rule (c0 = z0, b0 = y0,
// b1 = y1, a0 = x0): dx = y0 z0 d/dx, dy = z0 d/dy + x0 dx, dz = d/dz + y1 z0 d/dy + x0 dx
void TetBasis::proj(const double* lin, double* f, double* dx, double* dy, double* dz) const {
    for (int g = 0; g < G; ++g) { f[g] = 0.0; dx[g] = 0.0; dy[g] = 0.0; dz[g] = 0.0; }
    for (int m = 0; m < tm; ++m) {
        const ModeFactors mf = mode_factors(m, em);
        for (int k = 0; k < gpz; ++k) {
            const double cz = lin[m] * gz[k * nmodz + mf.iz], cdz = lin[m] * dgz[k * nmodz + mf.iz];
            for (int j = 0; j < gpy; ++j) {
                const double gyv = gy[j * nmody + mf.iy], dgyv = dgy[j * nmody + mf.iy];
                const double v = cz * gyv;                 // wk3 part
                const double vr = y0[j] * z0[k] * cz * gyv;  // wk4 part
                const double vs = z0[k] * cz * dgyv;        // wk6 part
                const double vt = cdz * gyv + y1[j] * vs;   // wk5 + wk7 part
                for (int i = 0; i < gpx; ++i) {
                    const int g = (i * gpy + j) * gpz + k;
                    f[g] += v * gx[i * nmodx + mf.ix];
                    dx[g] += vr * dgx[i * nmodx + mf.ix];
                    dy[g] += vs * gx[i * nmodx + mf.ix];
                    dz[g] += vt * gx[i * nmodx + mf.ix];
                }
            }
        }
    }
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpy; ++j)
            for (int k = 0; k < gpz; ++k) {
                const int g = (i * gpy + j) * gpz + k;
                dy[g] += x0[i] * dx[g];
                dz[g] += x0[i] * dx[g];
            }
}

void TetBasis::proj(const double* lin, double* f) const {
    for (int g = 0; g < G; ++g) f[g] = 0.0;
    for (int m = 0; m < tm; ++m) {
        const ModeFactors mf = mode_factors(m, em);
        for (int i = 0; i < gpx; ++i)
            for (int j = 0; j < gpy; ++j)
                for (int k = 0; k < gpz; ++k)
                    f[(i * gpy + j) * gpz + k] +=
                        lin[m] * gz[k * nmodz + mf.iz] * gy[j * nmody + mf.iy] * gx[i * nmodx + mf.ix];
    }
}

// tet_basis/proj.cpp:401-416: linear field from the 4 vertex values
void TetBasis::proj(double u0, double u1, double u2, double u3, double* f) const {
    std::vector<double> lin(tm, 0.0);
    lin[0] = u0; lin[1] = u1; lin[2] = u2; lin[3] = u3;
    proj(lin.data(), f);
}

// tet_basis/intgrt.cpp:23-290: rslt_m = sum_g w_g phi_m(g) f(g)
void TetBasis::intgrt(double* rslt, const double* f) const {
    for (int m = 0; m < tm; ++m) {
        const ModeFactors mf = mode_factors(m, em);
        double s = 0.0;
        for (int k = 0; k < gpz; ++k) {
            double sy = 0.0;
            for (int j = 0; j < gpy; ++j) {
                double sx = 0.0;
                for (int i = 0; i < gpx; ++i) sx += f[(i * gpy + j) * gpz + k] * gxwtx[mf.ix * gpx + i];
                sy += sx * gywty[mf.iy * gpy + j];
            }
            s += gzwtz[mf.iz * gpz + k] * sy;
        }
        rslt[m] = s;
    }
}

// tet_basis/intgrt.cpp:292-609: the transpose of proj's derivative part, subtracted
void TetBasis::intgrtrst(double* rslt, const double* dx, const double* dy, const double* dz) const {
    for (int m = 0; m < tm; ++m) {
        const ModeFactors mf = mode_factors(m, em);
        double s = 0.0;
        for (int k = 0; k < gpz; ++k) {
            double sk = 0.0;
            for (int j = 0; j < gpy; ++j) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int i = 0; i < gpx; ++i) {
                    const int g = (i * gpy + j) * gpz + k;
                    const double w3 = y0[j] * z0[k] * (dx[g] + x0[i] * (dy[g] + dz[g]));
                    const double w4 = z0[k] * (dy[g] + y1[j] * dz[g]);
                    s0 += w3 * dgxwtx[mf.ix * gpx + i];
                    s1 += w4 * gxwtx[mf.ix * gpx + i];
                    s2 += dz[g] * gxwtx[mf.ix * gpx + i];
                }
                sk += gzwtz[mf.iz * gpz + k] * (s0 * gywty[mf.iy * gpy + j] + s1 * dgywty[mf.iy * gpy + j]) +
                      dgzwtz[mf.iz * gpz + k] * s2 * gywty[mf.iy * gpy + j];
            }
            s += sk;
        }
        rslt[m] -= s;
    }
}

// tet_basis/proj.cpp:781-808
void TetBasis::derivr(const double* f, double* d) const {
    std::vector<double> wk(gpx);
    for (int j = 0; j < gpy; ++j)
        for (int k = 0; k < gpz; ++k) {
            const double l0 = y0[j] * z0[k];
            for (int i = 0; i < gpx; ++i) wk[i] = f[(i * gpy + j) * gpz + k] * dltx[i] * l0;
            for (int i = 0; i < gpx; ++i) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpx; ++n)
                    if (n != i) s += (wk[i] + wk[n]) * dltx2[i * gpx + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
}

// tet_basis/proj.cpp:810-852
void TetBasis::derivs(const double* f, double* d) const {
    std::vector<double> wk(std::max(gpx, gpy));
    for (int j = 0; j < gpy; ++j)
        for (int k = 0; k < gpz; ++k) {
            const double l0 = y0[j] * z0[k];
            for (int i = 0; i < gpx; ++i) wk[i] = f[(i * gpy + j) * gpz + k] * dltx[i] * l0;
            for (int i = 0; i < gpx; ++i) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpx; ++n)
                    if (n != i) s += (wk[i] + wk[n]) * dltx3[i * gpx + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
    for (int i = 0; i < gpx; ++i)
        for (int k = 0; k < gpz; ++k) {
            const double l0 = z0[k];
            for (int j = 0; j < gpy; ++j) wk[j] = f[(i * gpy + j) * gpz + k] * dlty[j] * l0;
            for (int j = 0; j < gpy; ++j) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpy; ++n)
                    if (n != j) s += (wk[j] + wk[n]) * dlty2[j * gpy + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
}

// tet_basis/proj.cpp:854-911
void TetBasis::derivt(const double* f, double* d) const {
    std::vector<double> wk(std::max(gpx, std::max(gpy, gpz)));
    for (int j = 0; j < gpy; ++j)
        for (int k = 0; k < gpz; ++k) {
            const double l0 = y0[j] * z0[k];
            for (int i = 0; i < gpx; ++i) wk[i] = f[(i * gpy + j) * gpz + k] * dltx[i] * l0;
            for (int i = 0; i < gpx; ++i) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpx; ++n)
                    if (n != i) s += (wk[i] + wk[n]) * dltx3[i * gpx + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
    for (int i = 0; i < gpx; ++i)
        for (int k = 0; k < gpz; ++k) {
            const double l0 = z0[k];
            for (int j = 0; j < gpy; ++j) wk[j] = f[(i * gpy + j) * gpz + k] * dlty[j] * l0;
            for (int j = 0; j < gpy; ++j) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpy; ++n)
                    if (n != j) s += (wk[j] + wk[n]) * dlty3[j * gpy + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpy; ++j) {
            for (int k = 0; k < gpz; ++k) wk[k] = f[(i * gpy + j) * gpz + k] * dltz[k];
            for (int k = 0; k < gpz; ++k) {
                double s = d[(i * gpy + j) * gpz + k];
                for (int n = 0; n < gpz; ++n)
                    if (n != k) s += (wk[k] + wk[n]) * dltz2[k * gpz + n];
                d[(i * gpy + j) * gpz + k] = s;
            }
        }
}

// Face 0 (z = -1) modes of tet_basis/proj2d.cpp:20-120: V1 = gy1 gx0, V2 = gy2 gx1,
// V3 = gy2 gx2, E1 = gy(3+k) gx(3+k), E2 = gy(3+em+k) gx2, E3 = (-1)^k gy(3+em+k) gx1.
// dx = d/dr = y0 d/dx, dy = d/ds = d/dy + x0 dx.
static void face_factors(int m, int em, int& ix, int& iy, double& sg) {
    sg = 1.0;
    if (m == 0) { ix = 0; iy = 1; return; }
    if (m == 1) { ix = 1; iy = 2; return; }
    if (m == 2) { ix = 2; iy = 2; return; }
    const int e = (m - 3) / em, k = (m - 3) - e * em;
    if (e == 0) { ix = 3 + k; iy = 3 + k; return; }
    if (e == 1) { ix = 2; iy = 3 + em + k; return; }
    ix = 1; iy = 3 + em + k; sg = (k & 1) ? -1.0 : 1.0;
}

void TetBasis::proj2d(const double* lin, double* f, double* dx, double* dy) const {
    for (int g = 0; g < gpx * gpy; ++g) { f[g] = 0.0; dx[g] = 0.0; dy[g] = 0.0; }
    for (int m = 0; m < tm2d; ++m) {
        int ix, iy;
        double sg;
        face_factors(m, em, ix, iy, sg);
        for (int j = 0; j < gpy; ++j) {
            const double v = sg * lin[m] * gy[j * nmody + iy];
            const double dv = sg * lin[m] * dgy[j * nmody + iy];
            for (int i = 0; i < gpx; ++i) {
                const int g = i * gpy + j;
                f[g] += v * gx[i * nmodx + ix];
                dx[g] += v * y0[j] * dgx[i * nmodx + ix];
                dy[g] += dv * gx[i * nmodx + ix];
            }
        }
    }
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpy; ++j) dy[i * gpy + j] += x0[i] * dx[i * gpy + j];
}

void TetBasis::proj2d(const double* lin, double* f) const {
    std::vector<double> a(gpx * gpy), b(gpx * gpy);
    proj2d(lin, f, a.data(), b.data());
}

// tet_basis/intgrt2d.cpp:22-113
void TetBasis::intgrt2d(double* rslt, const double* f) const {
    for (int m = 0; m < tm2d; ++m) {
        int ix, iy;
        double sg;
        face_factors(m, em, ix, iy, sg);
        double s = 0.0;
        for (int j = 0; j < gpy; ++j) {
            double sx = 0.0;
            for (int i = 0; i < gpx; ++i) sx += f[i * gpy + j] * gxwtx[ix * gpx + i];
            s += sx * gywty[iy * gpy + j];
        }
        rslt[m] = sg * s;
    }
}

// tet_basis/initialize.cpp:521-800 (p <= 2): mass matrix, vdiag, sfmv, vfms, ediag
void TetBasis::lumpinv() {
    mm.assign(tm * tm, 0.0);
    std::vector<double> u(tm), f(G), l(tm);
    for (int m = 0; m < tm; ++m) {
        std::fill(u.begin(), u.end(), 0.0);
        u[m] = 1.0;
        proj(u.data(), f.data());
        intgrt(l.data(), f.data());
        for (int i = 0; i < tm; ++i) mm[m * tm + i] = l[i];
    }
    if (p < 2) {
        vdiag = 1.0 / (mm[0] + mm[1] + mm[2] + mm[3]);  // lumped (initialize.cpp:634-637)
        return;
    }
    // vertex modes made mass-orthogonal to the other vertices' linear functions:
    // for vertex i, solve M_i x = e_i with M_i(j,k) = mm(j, dind(i,k)) and row i
    // replaced by e_0 (initialize.cpp:642-757; the LAPACK transposed solve on the
    // row-major array is the plain solve of M_i).  Only vertex 3's solution is kept
    // for sfmv (initialize.cpp:781-784).
    const int vdof = 1 + 3 * em;
    const int dind[4][4] = {{0, 4 + 3 * em, 5 + 3 * em, 6 + 3 * em},
                            {1, 4 + em, 5 + em, 6 + em},
                            {2, 4, 4 + 2 * em, 4 + 4 * em},
                            {3, 4, 4 + em, 4 + 5 * em}};
    if (em != 1 || vdof != 4) throw std::runtime_error("oracle TetBasis: lumpinv restated for p = 2");
    std::vector<double> vwk(4);
    for (int i = 0; i < 4; ++i) {
        std::vector<double> mw(16);
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) mw[j * 4 + k] = mm[j * tm + dind[i][k]];
        for (int k = 0; k < 4; ++k) mw[i * 4 + k] = 0.0;
        mw[i * 4 + 0] = 1.0;
        std::fill(vwk.begin(), vwk.end(), 0.0);
        vwk[i] = 1.0;
        lu_solve_inplace(mw, 4, 4, vwk.data());
    }
    sfmv.assign(2 * em, 0.0);
    for (int k = 0; k < em; ++k) {
        sfmv[1 * em + k] = -vwk[k + 1];
        sfmv[0 * em + k] = -vwk[k + 1 + em];
    }
    double vd = mm[0];
    for (int k = 0; k < em; ++k) {
        vd -= sfmv[1 * em + k] * mm[(4 + k + 3 * em) * tm + 0];
        vd -= sfmv[1 * em + k] * mm[(4 + k + 4 * em) * tm + 0];
        vd -= sfmv[1 * em + k] * mm[(4 + k + 5 * em) * tm + 0];
    }
    vdiag = 1.0 / vd;
    vfms.assign(4 * tm, 0.0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < tm; ++j) vfms[i * tm + j] = mm[i * tm + j];
    // tet_hp::minvrt overwrites ediag(0) with 100 at p = 2 "for fast convergence"
    // (tet_hp/setup_preconditioner.cpp:199-200); the mass-matrix value of
    // initialize.cpp:806-816 is never used there
    ediag.assign(em, 0.0);
    ediag[0] = 100.0;
}

}  // namespace oracle
