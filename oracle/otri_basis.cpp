// Oracle (test infrastructure only).  See otri_basis.h for the reference map.
#include "otri_basis.h"
#include "olinalg.h"
#include <cmath>
#include <algorithm>
#include <stdexcept>

namespace oracle {

void recur_coeffs(int n, int ipoly, double al, double be, double* a, double* b) {
    // Monic three-term recurrence p_{k+1} = (x - a_k) p_k - b_k p_{k-1}; b_0 = total mass.
    for (int k = 0; k < n; ++k) { a[k] = 0.0; b[k] = 0.0; }
    if (ipoly == 1) {
        b[0] = 2.0;
        for (int k = 1; k < n; ++k) b[k] = 1.0 / (4.0 - 1.0 / (double(k) * k));
        return;
    }
    if (ipoly != 6) throw std::runtime_error("recur_coeffs: unsupported family");
    const double s = al + be;
    a[0] = (be - al) / (s + 2.0);
    b[0] = std::exp((s + 1.0) * std::log(2.0) + std::lgamma(al + 1.0) + std::lgamma(be + 1.0) -
                    std::lgamma(s + 2.0));
    if (n == 1) return;
    a[1] = (be * be - al * al) / ((s + 2.0) * (s + 4.0));
    b[1] = 4.0 * (al + 1.0) * (be + 1.0) / ((s + 3.0) * (s + 2.0) * (s + 2.0));
    for (int k = 2; k < n; ++k) {
        const double t = 2.0 * k + s;
        a[k] = (be * be - al * al) / (t * (t + 2.0));
        b[k] = 4.0 * k * (k + al) * (k + be) * (k + s) / (t * t * (t + 1.0) * (t - 1.0));
    }
}

TriBasis::TriBasis(int p_) : p(p_) {
    sm = p - 1;
    im = (p - 1) * (p - 2) / 2;
    tm = (p + 1) * (p + 2) / 2;
    bm = 3 * p;
    nmodx = p + 2;
    gpx = p + 1;
    gpn = p + 1;
    arow = gpx + 1;
    a0.assign((sm + 2) * arow, 0.0);
    b0.assign((sm + 2) * arow, 0.0);
    gx.assign(gpx * nmodx, 0.0); dgx = gx;
    wtx.assign(gpx, 0.0); xp = wtx; x0 = wtx; dltx = wtx;
    gxwtx.assign(nmodx * gpx, 0.0); dgxwtx = gxwtx;
    dltx1.assign(gpx * gpx, 0.0); dltn1 = dltx1;
    gn.assign(gpn * tm, 0.0); dgn = gn;
    wtn.assign(gpn, 0.0); np = wtn; n0 = wtn; dltn = wtn;
    gnwtn.assign(tm * gpn, 0.0); gnwtnn0 = gnwtn; dgnwtn = gnwtn;
    dltn2.assign(gpn * gpn, 0.0);
    norm.assign(tm, 1.0);
    pgx.assign(nmodx, 0.0); dpgx = pgx;
    pgn.assign(tm, 0.0); dpgn = pgn;
    initialize_values();
    lumpinv();
    lumpinv1d();
}

// tri_basis/initialize.cpp:1085-1130: the 1-D mass matrix of the edge modes
// (columns 1..p+1 of gx, Gauss-Legendre weights) and the banded Cholesky factor
// of its side-mode block, upper band storage with bandwidth sbwth.
void TriBasis::lumpinv1d() {
    const int n1 = p + 1;
    std::vector<double> mm(n1 * n1, 0.0);
    for (int a = 1; a < p + 2; ++a)
        for (int c = 1; c < p + 2; ++c) {
            double s = 0.0;
            for (int i = 0; i < gpx; ++i) s += wtx[i] * gx[i * nmodx + a] * gx[i * nmodx + c];
            mm[(a - 1) * n1 + (c - 1)] = s;
        }
    sdiag1d.assign(std::max(sm, 1) * (sbwth + 1), 0.0);
    if (p > 1) {
        for (int j = 0; j < sm; ++j)
            for (int i = std::max(0, j - sbwth); i <= j; ++i) sdiag1d[j * (sbwth + 1) + i - (j - sbwth)] = mm[(i + 2) * n1 + j + 2];
        band_chol_upper(sdiag1d.data(), sm, sbwth);
    }
}

// tri_basis/ptvalues.cpp:84-195
void TriBasis::ptvalues_deriv(double x, double eta) const {
    pgx[0] = 1.0; dpgx[0] = 0.0;
    pgx[1] = 0.5 * (1 - x); dpgx[1] = -0.5;
    pgx[2] = 0.5 * (1 + x); dpgx[2] = 0.5;
    double pk = 1.0, pkm = 0.0, dpk = 0.0, dpkm = 0.0;
    for (int m = 0; m < sm; ++m) {
        pgx[m + 3] = (1. + x) * (1. - x) * .25 * pk * norm[m + 3];
        dpgx[m + 3] = (-x * .5 * pk + (1. + x) * (1. - x) * .25 * dpk) * norm[m + 3];
        double pkp = (x - a0[0 * arow + m]) * pk - b0[0 * arow + m] * pkm;
        double dpkp = pk + (x - a0[0 * arow + m]) * dpk - b0[0 * arow + m] * dpkm;
        dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
    }
    int ind = 0;
    pgn[ind] = (1 + eta) * .5; dpgn[ind++] = 0.5;
    pgn[ind] = (1 - eta) * .5; dpgn[ind++] = -0.5;
    pgn[ind] = (1 - eta) * .5; dpgn[ind++] = -0.5;
    for (int m = 2; m <= sm + 1; ++m) {
        pgn[ind] = std::pow(.5 * (1 - eta), m);
        dpgn[ind] = -.5 * m * std::pow(.5 * (1. - eta), m - 1);
        ++ind;
    }
    for (int side = 1; side <= 2; ++side) {
        pk = 1.0; pkm = 0.0; dpk = 0.0; dpkm = 0.0;
        for (int n = 0; n < sm; ++n) {
            double sg = (side == 2 && (n % 2)) ? -1.0 : 1.0;
            pgn[ind] = sg * (1. - eta) * (1. + eta) * .25 * pk * norm[n + 3];
            dpgn[ind] = sg * (-.5 * eta * pk + (1. - eta) * (1. + eta) * .25 * dpk) * norm[n + 3];
            double pkp = (eta - a0[1 * arow + n]) * pk - b0[1 * arow + n] * pkm;
            double dpkp = pk + (eta - a0[1 * arow + n]) * dpk - b0[1 * arow + n] * dpkm;
            dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
            ++ind;
        }
    }
    for (int m = 2; m < sm + 1; ++m) {
        pk = 1.0; pkm = 0.0; dpk = 0.0; dpkm = 0.0;
        for (int k = 0; k < sm + 1 - m; ++k) {
            pgn[ind] = (std::pow(.5 * (1. - eta), m) * .5 * (1. + eta) * pk) * norm[ind];
            dpgn[ind] = (-.25 * m * std::pow(.5 * (1. - eta), m - 1) * (1. + eta) * pk +
                         std::pow(.5 * (1. - eta), m) * .5 * (pk + (1. + eta) * dpk)) * norm[ind];
            double pkp = (eta - a0[m * arow + k]) * pk - b0[m * arow + k] * pkm;
            double dpkp = pk + (eta - a0[m * arow + k]) * dpk - b0[m * arow + k] * dpkm;
            dpkm = dpk; dpk = dpkp; pkm = pk; pk = pkp;
            ++ind;
        }
    }
}

// tri_basis/initialize.cpp:161-455
void TriBasis::initialize_values() {
    std::vector<double> a(arow + 8), b(arow + 8);
    // Gauss-Legendre in xi
    recur_coeffs(gpx + 1, 1, 0.0, 0.0, a.data(), b.data());
    for (int k = 0; k < gpx + 1; ++k) { a0[k] = a[k]; b0[k] = b[k]; }
    gauss_rule(gpx, a.data(), b.data(), x0.data(), wtx.data());
    for (int i = 0; i < gpx; ++i) {
        dltx[i] = 1.0;
        for (int k = 0; k < gpx; ++k) if (k != i) dltx[i] *= (x0[i] - x0[k]);
    }
    for (int i = 0; i < gpx; ++i)
        for (int n = 0; n < gpx; ++n) if (n != i) {
            dltx1[i * gpx + n] = dltx[i] / (x0[i] - x0[n]);
            dltn1[i * gpx + n] = dltx[i] / (x0[i] - x0[n]) * .5 * (1 + x0[i]);
        }
    for (int i = 0; i < gpx; ++i) dltx[i] = 1.0 / dltx[i];

    // side-mode recursion in xi: Jacobi(2,2) (MORTHOGONAL)
    recur_coeffs(sm + 1, 6, 2.0, 2.0, a.data(), b.data());
    for (int k = 0; k < sm + 1; ++k) { a0[k] = a[k]; b0[k] = b[k]; }
    for (int i = 0; i < tm; ++i) norm[i] = 1.0;
    for (int i = 0; i < gpx; ++i) {
        double x = x0[i];
        xp[i] = x;
        x0[i] = 0.5 * (1 + x);
        ptvalues_deriv(x, 0.0);
        for (int m = 0; m < nmodx; ++m) { gx[i * nmodx + m] = pgx[m]; dgx[i * nmodx + m] = dpgx[m]; }
    }

    // Gauss-Jacobi(1,0) in eta, weights halved
    recur_coeffs(gpn + 1, 6, 1.0, 0.0, a.data(), b.data());
    for (int k = 0; k < gpn + 1; ++k) { a0[arow + k] = a[k]; b0[arow + k] = b[k]; }
    gauss_rule(gpn, a.data(), b.data(), n0.data(), wtn.data());
    for (int i = 0; i < gpn; ++i) wtn[i] *= 0.5;
    for (int i = 0; i < gpn; ++i) {
        dltn[i] = 1.0;
        for (int k = 0; k < gpn; ++k) if (k != i) dltn[i] *= (n0[i] - n0[k]);
    }
    for (int j = 0; j < gpn; ++j)
        for (int n = 0; n < gpn; ++n) if (n != j) dltn2[j * gpn + n] = dltn[j] / (n0[j] - n0[n]);
    for (int j = 0; j < gpn; ++j) dltn[j] = 1.0 / dltn[j];

    // eta recursions: side modes Jacobi(2,2), interior group m Jacobi(2m+1,2)
    recur_coeffs(sm + 1, 6, 2.0, 2.0, a.data(), b.data());
    for (int k = 0; k < sm + 1; ++k) { a0[arow + k] = a[k]; b0[arow + k] = b[k]; }
    for (int m = 2; m < sm + 1; ++m) {
        recur_coeffs(sm + 2 - m, 6, 2. * m + 1, 2.0, a.data(), b.data());
        for (int k = 0; k < sm + 2 - m; ++k) { a0[m * arow + k] = a[k]; b0[m * arow + k] = b[k]; }
    }
    for (int j = 0; j < gpn; ++j) {
        double eta = n0[j];
        np[j] = eta;
        n0[j] = 2.0 / (1 - eta);
        ptvalues_deriv(0.0, eta);
        for (int m = 0; m < tm; ++m) { gn[j * tm + m] = pgn[m]; dgn[j * tm + m] = dpgn[m]; }
    }

    // normalisation constants
    for (int n = 0; n < 3; ++n) norm[n] = 1.0;
    for (int n = 3; n < sm + 3; ++n) {
        double s = 0.0;
        for (int i = 0; i < gpx; ++i) s += wtx[i] * gx[i * nmodx + n] * gx[i * nmodx + n];
        norm[n] = 1. / std::sqrt(s);
        norm[n + sm] = norm[n];
        norm[n + 2 * sm] = norm[n];
    }
    for (int m = bm; m < tm; ++m) {
        double s = 0.0;
        for (int j = 0; j < gpn; ++j) s += wtn[j] * gn[j * tm + m] * gn[j * tm + m];
        norm[m] = 1. / std::sqrt(s);
    }
    for (int n = 3; n < nmodx; ++n) {
        for (int i = 0; i < gpx; ++i) { gx[i * nmodx + n] *= norm[n]; dgx[i * nmodx + n] *= norm[n]; }
        for (int j = 0; j < gpn; ++j) {
            gn[j * tm + n + sm] *= norm[n]; dgn[j * tm + n + sm] *= norm[n];
            gn[j * tm + n + 2 * sm] *= norm[n]; dgn[j * tm + n + 2 * sm] *= norm[n];
        }
    }
    for (int m = bm; m < tm; ++m)
        for (int j = 0; j < gpn; ++j) { gn[j * tm + m] *= norm[m]; dgn[j * tm + m] *= norm[m]; }

    for (int m = 0; m < nmodx; ++m)
        for (int i = 0; i < gpx; ++i) {
            gxwtx[m * gpx + i] = gx[i * nmodx + m] * wtx[i];
            dgxwtx[m * gpx + i] = dgx[i * nmodx + m] * wtx[i];
        }
    for (int m = 0; m < tm; ++m)
        for (int j = 0; j < gpn; ++j) {
            gnwtn[m * gpn + j] = gn[j * tm + m] * wtn[j];
            gnwtnn0[m * gpn + j] = gn[j * tm + m] * wtn[j] * n0[j];
            dgnwtn[m * gpn + j] = dgn[j * tm + m] * wtn[j];
        }
}

// tri_basis/proj.cpp:22-119
void TriBasis::proj(const double* lin, double* f, double* dx, double* dy, int stride) const {
    const int bs1 = sm + 3, bs2 = 2 * sm + 3, bint = bm;
    std::vector<double> wk0(gpn * nmodx), wk1(gpn * nmodx), wk2(gpn * nmodx);
    for (int j = 0; j < gpn; ++j) {
        const double oeta = n0[j];
        const double* g = &gn[j * tm];
        const double* dg = &dgn[j * tm];
        wk0[j * nmodx + 0] = lin[0] * g[0];
        wk1[j * nmodx + 0] = lin[0] * dg[0];
        wk2[j * nmodx + 0] = wk0[j * nmodx + 0] * oeta;
        double l0 = lin[1] * g[1], l1 = lin[1] * dg[1];
        for (int m = bs2; m < bint; ++m) { l0 += lin[m] * g[m]; l1 += lin[m] * dg[m]; }
        wk0[j * nmodx + 1] = l0; wk1[j * nmodx + 1] = l1; wk2[j * nmodx + 1] = l0 * oeta;
        l0 = lin[2] * g[2]; l1 = lin[2] * dg[2];
        for (int m = bs1; m < bs2; ++m) { l0 += lin[m] * g[m]; l1 += lin[m] * dg[m]; }
        wk0[j * nmodx + 2] = l0; wk1[j * nmodx + 2] = l1; wk2[j * nmodx + 2] = l0 * oeta;
        int ind = bint, cnt = sm - 1;
        for (int m = 3; m < bs1; ++m) {
            l0 = lin[m] * g[m]; l1 = lin[m] * dg[m];
            for (int n = 0; n < cnt; ++n) { l0 += lin[ind + n] * g[ind + n]; l1 += lin[ind + n] * dg[ind + n]; }
            ind += cnt--;
            wk0[j * nmodx + m] = l0; wk1[j * nmodx + m] = l1; wk2[j * nmodx + m] = l0 * oeta;
        }
    }
    for (int i = 0; i < gpx; ++i) {
        const double xp1 = x0[i];
        for (int j = 0; j < gpn; ++j) {
            double l0 = wk0[j * nmodx] * gx[i * nmodx];
            double l1 = wk1[j * nmodx] * gx[i * nmodx];
            double l2 = wk2[j * nmodx] * dgx[i * nmodx];
            for (int n = 1; n < nmodx; ++n) {
                l0 += wk0[j * nmodx + n] * gx[i * nmodx + n];
                l1 += wk1[j * nmodx + n] * gx[i * nmodx + n];
                l2 += wk2[j * nmodx + n] * dgx[i * nmodx + n];
            }
            f[i * stride + j] = l0;
            dy[i * stride + j] = l1 + xp1 * l2;
            dx[i * stride + j] = l2;
        }
    }
}

// tri_basis/proj.cpp:121-186
void TriBasis::proj(const double* lin, double* f, int stride) const {
    const int bs1 = sm + 3, bs2 = 2 * sm + 3, bint = bm;
    std::vector<double> wk0(gpn * nmodx);
    for (int j = 0; j < gpn; ++j) {
        const double* g = &gn[j * tm];
        wk0[j * nmodx + 0] = lin[0] * g[0];
        double l0 = lin[1] * g[1];
        for (int m = bs2; m < bint; ++m) l0 += lin[m] * g[m];
        wk0[j * nmodx + 1] = l0;
        l0 = lin[2] * g[2];
        for (int m = bs1; m < bs2; ++m) l0 += lin[m] * g[m];
        wk0[j * nmodx + 2] = l0;
        int ind = bint, cnt = sm - 1;
        for (int m = 3; m < bs1; ++m) {
            l0 = lin[m] * g[m];
            for (int n = 0; n < cnt; ++n) l0 += lin[ind + n] * g[ind + n];
            ind += cnt--;
            wk0[j * nmodx + m] = l0;
        }
    }
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpn; ++j) {
            double l0 = 0.0;
            for (int n = 0; n < nmodx; ++n) l0 += wk0[j * nmodx + n] * gx[i * nmodx + n];
            f[i * stride + j] = l0;
        }
}

// tri_basis/proj.cpp:188-199
void TriBasis::proj(double u1, double u2, double u3, double* f, int stride) const {
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpn; ++j)
            f[i * stride + j] = u1 * gx[i * nmodx + 0] * gn[j * tm + 0] + u2 * gx[i * nmodx + 1] * gn[j * tm + 1] +
                                u3 * gx[i * nmodx + 2] * gn[j * tm + 2];
}

// tri_basis/proj.cpp:339-363
void TriBasis::derivr(const double* f, double* dr, int stride) const {
    std::vector<double> wk(gpx);
    for (int j = 0; j < gpn; ++j) {
        const double ln0 = n0[j];
        for (int i = 0; i < gpx; ++i) wk[i] = f[i * stride + j] * dltx[i] * ln0;
        for (int i = 0; i < gpx; ++i) {
            double l0 = dr[i * stride + j];
            for (int n = 0; n < gpx; ++n) if (n != i) l0 += (wk[i] + wk[n]) * dltx1[i * gpx + n];
            dr[i * stride + j] = l0;
        }
    }
}

// tri_basis/proj.cpp:365-402
void TriBasis::derivs(const double* f, double* ds, int stride) const {
    std::vector<double> wk(std::max(gpx, gpn));
    for (int j = 0; j < gpn; ++j) {
        const double ln0 = n0[j];
        for (int i = 0; i < gpx; ++i) wk[i] = f[i * stride + j] * dltx[i] * ln0;
        for (int i = 0; i < gpx; ++i) {
            double l0 = ds[i * stride + j];
            for (int n = 0; n < gpx; ++n) if (n != i) l0 += (wk[i] + wk[n]) * dltn1[i * gpx + n];
            ds[i * stride + j] = l0;
        }
    }
    for (int i = 0; i < gpx; ++i) {
        for (int j = 0; j < gpn; ++j) wk[j] = f[i * stride + j] * dltn[j];
        for (int j = 0; j < gpn; ++j) {
            double l0 = ds[i * stride + j];
            for (int n = 0; n < gpn; ++n) if (n != j) l0 += (wk[j] + wk[n]) * dltn2[j * gpn + n];
            ds[i * stride + j] = l0;
        }
    }
}

// tri_basis/intgrt.cpp:22-78
void TriBasis::intgrt(double* rslt, const double* f, int stride) const {
    const int bs1 = sm + 3, bs2 = 2 * sm + 3, bint = bm, lsm2 = sm + 2;
    std::vector<double> wk0(nmodx * gpn);
    for (int n = 0; n < nmodx; ++n)
        for (int j = 0; j < gpn; ++j) {
            double l0 = f[0 * stride + j] * gxwtx[n * gpx + 0];
            for (int i = 1; i < gpx; ++i) l0 += f[i * stride + j] * gxwtx[n * gpx + i];
            wk0[n * gpn + j] = l0;
        }
    for (int m = 0; m < bs1; ++m) {
        double l0 = 0.0;
        for (int j = 0; j < gpn; ++j) l0 += gnwtn[m * gpn + j] * wk0[m * gpn + j];
        rslt[m] = l0;
    }
    for (int m = bs1; m < bs2; ++m) {
        double l0 = 0.0;
        for (int j = 0; j < gpn; ++j) l0 += gnwtn[m * gpn + j] * wk0[2 * gpn + j];
        rslt[m] = l0;
    }
    for (int m = bs2; m < bint; ++m) {
        double l0 = 0.0;
        for (int j = 0; j < gpn; ++j) l0 += gnwtn[m * gpn + j] * wk0[1 * gpn + j];
        rslt[m] = l0;
    }
    int indx = bint;
    for (int m = 3; m < lsm2; ++m)
        for (int n = 0; n < lsm2 - m; ++n) {
            double l0 = 0.0;
            for (int j = 0; j < gpn; ++j) l0 += gnwtn[indx * gpn + j] * wk0[m * gpn + j];
            rslt[indx] = l0;
            ++indx;
        }
}

// tri_basis/intgrt.cpp:81-158 (accumulates into rslt)
void TriBasis::intgrtrs(double* rslt, const double* dx, const double* dy, int stride) const {
    const int bs1 = sm + 3, bs2 = 2 * sm + 3, bint = bm, lsm2 = sm + 2;
    std::vector<double> wk0(nmodx * gpn), wk1(nmodx * gpn), wk2(gpx * gpn);
    for (int j = 0; j < gpn; ++j)
        for (int i = 0; i < gpx; ++i) wk2[i * gpn + j] = dx[i * stride + j] + dy[i * stride + j] * x0[i];
    for (int n = 0; n < nmodx; ++n)
        for (int j = 0; j < gpn; ++j) {
            double l0 = wk2[0 * gpn + j] * dgxwtx[n * gpx + 0];
            double l1 = dy[0 * stride + j] * gxwtx[n * gpx + 0];
            for (int i = 1; i < gpx; ++i) {
                l0 += wk2[i * gpn + j] * dgxwtx[n * gpx + i];
                l1 += dy[i * stride + j] * gxwtx[n * gpx + i];
            }
            wk0[n * gpn + j] = l0; wk1[n * gpn + j] = l1;
        }
    for (int m = 0; m < bs1; ++m) {
        double l0 = rslt[m];
        for (int j = 0; j < gpn; ++j) l0 -= gnwtnn0[m * gpn + j] * wk0[m * gpn + j] + dgnwtn[m * gpn + j] * wk1[m * gpn + j];
        rslt[m] = l0;
    }
    for (int j = 0; j < gpn; ++j) {
        double l0 = wk0[2 * gpn + j], l1 = wk1[2 * gpn + j];
        for (int m = bs1; m < bs2; ++m) rslt[m] -= gnwtnn0[m * gpn + j] * l0 + dgnwtn[m * gpn + j] * l1;
        l0 = wk0[1 * gpn + j]; l1 = wk1[1 * gpn + j];
        for (int m = bs2; m < bint; ++m) rslt[m] -= gnwtnn0[m * gpn + j] * l0 + dgnwtn[m * gpn + j] * l1;
    }
    int indx = bint;
    for (int m = 3; m < lsm2; ++m) {
        const int nmax = lsm2 - m;
        for (int j = 0; j < gpn; ++j) {
            const double l0 = wk0[m * gpn + j], l1 = wk1[m * gpn + j];
            for (int n = 0; n < nmax; ++n) {
                rslt[indx] -= gnwtnn0[indx * gpn + j] * l0 + dgnwtn[indx * gpn + j] * l1;
                ++indx;
            }
            indx -= nmax;
        }
        indx += nmax;
    }
}

// tri_basis/proj1d.cpp:61-71 (linear edge field from its two end values)
void TriBasis::proj1d(double u1, double u2, double* f) const {
    for (int i = 0; i < gpx; ++i) f[i] = u1 * gx[i * nmodx + 1] + u2 * gx[i * nmodx + 2];
}

// tri_basis/proj1d.cpp:19-59
void TriBasis::proj1d(const double* lin, double* f, double* dx) const {
    for (int i = 0; i < gpx; ++i) {
        double l0 = lin[0] * gx[i * nmodx + 1] + lin[1] * gx[i * nmodx + 2];
        double l1 = lin[0] * dgx[i * nmodx + 1] + lin[1] * dgx[i * nmodx + 2];
        for (int n = 2; n < sm + 2; ++n) { l0 += lin[n] * gx[i * nmodx + n + 1]; l1 += lin[n] * dgx[i * nmodx + n + 1]; }
        f[i] = l0; dx[i] = l1;
    }
}
void TriBasis::proj1d(const double* lin, double* f) const {
    for (int i = 0; i < gpx; ++i) {
        double l0 = lin[0] * gx[i * nmodx + 1] + lin[1] * gx[i * nmodx + 2];
        for (int n = 2; n < sm + 2; ++n) l0 += lin[n] * gx[i * nmodx + n + 1];
        f[i] = l0;
    }
}

// tri_basis/intgrt1d.cpp:17-43
void TriBasis::intgrt1d(double* rslt, const double* f) const {
    double l0 = 0.0;
    for (int i = 0; i < gpx; ++i) l0 += gxwtx[1 * gpx + i] * f[i];
    rslt[0] = l0;
    l0 = 0.0;
    for (int i = 0; i < gpx; ++i) l0 += gxwtx[2 * gpx + i] * f[i];
    rslt[1] = l0;
    for (int n = 3; n < nmodx; ++n) {
        l0 = 0.0;
        for (int i = 0; i < gpx; ++i) l0 += gxwtx[n * gpx + i] * f[i];
        rslt[n - 1] = l0;
    }
}

// tri_basis/initialize.cpp:590-958
void TriBasis::lumpinv() {
    const int T = tm;
    std::vector<double> mm(T * T), mwk(T * T), u(T), l(T), vwk(T), wk1(T * gpx);
    if (sm > 0) { vfms.assign(3 * sm, 0.0); sfmv.assign(2 * sm, 0.0); sdiag.assign(sm, 0.0); }
    if (sm > 1) sfms.assign((sm - 1) * sm * 3, 0.0);
    if (im > 0) { ifmb.assign(bm * im, 0.0); bfmi.assign(bm * im, 0.0); idiag.assign(im * (ibwth + 1), 0.0); }
    auto MM = [&](int r, int c) -> double& { return mm[r * T + c]; };
    auto MW = [&](int r, int c) -> double& { return mwk[r * T + c]; };

    for (int m = 0; m < T; ++m) {
        std::fill(u.begin(), u.end(), 0.0);
        u[m] = 1.0;
        proj(u.data(), wk1.data(), gpn);
        intgrt(l.data(), wk1.data(), gpn);
        for (int i = 0; i < T; ++i) MM(m, i) = l[i];
    }

    if (p > 1) {
        for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            for (int k = 0; k < 2; ++k) {
                const int ind = (i + 1 + k) % 3;
                vwk[k] = MM(ind, i);
                for (int j = 0; j < sm; ++j) { MW(k, j) = MM(ind, 3 + j + i1 * sm); MW(k, j + sm) = MM(ind, 3 + j + i2 * sm); }
                for (int j = 0; j < im; ++j) MW(k, 2 * sm + j) = MM(ind, j + bm);
            }
            for (int k = 0; k < 3; ++k)
                for (int m = 0; m < sm - 1; ++m) {
                    const int row = m + 2 + k * (sm - 1);
                    vwk[row] = MM(3 + m + k * sm, i);
                    for (int j = 0; j < sm; ++j) {
                        MW(row, j) = MM(3 + m + k * sm, 3 + j + i1 * sm);
                        MW(row, j + sm) = MM(3 + m + k * sm, 3 + j + i2 * sm);
                    }
                    for (int j = 0; j < im; ++j) MW(row, 2 * sm + j) = MM(3 + m + k * sm, j + bm);
                }
            int ind = 3 * sm - 1, ind1 = 0;
            for (int m = 2; m < sm + 1; ++m) {
                for (int n = 1; n < sm + 1 - m; ++n) {
                    vwk[ind] = MM(ind1 + bm, i);
                    for (int j = 0; j < sm; ++j) {
                        MW(ind, j) = MM(ind1 + bm, 3 + j + i1 * sm);
                        MW(ind, j + sm) = MM(ind1 + bm, 3 + j + i2 * sm);
                    }
                    for (int j = 0; j < im; ++j) MW(ind, 2 * sm + j) = MM(ind1 + bm, j + bm);
                    ++ind; ++ind1;
                }
                ++ind1;
            }
            const int neq = (sm + 1) * (sm + 2) / 2 - 1;
            lu_solve_inplace(mwk, neq, T, vwk.data());
            for (int k = 0; k < im; ++k) ifmb[i * im + k] = vwk[k + 2 * sm];
        }
        for (int k = 0; k < sm; ++k) { sfmv[0 * sm + k] = vwk[k + sm]; sfmv[1 * sm + k] = vwk[k]; }
        vdiag = MM(2, 2);
        for (int k = 0; k < sm; ++k) {
            vdiag -= sfmv[0 * sm + k] * MM(2, 3 + k + 1 * sm);
            vdiag -= sfmv[1 * sm + k] * MM(2, 3 + k);
        }
        for (int k = 0; k < im; ++k) vdiag -= ifmb[2 * im + k] * MM(2, k + bm);
    } else {
        vdiag = MM(0, 0) + MM(0, 1) + MM(0, 2);
    }

    if (im > 0) {
        for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            for (int k = 0; k < sm - 1; ++k) {
                int ind = 0;
                for (int j = k; j < sm - 1; ++j) {
                    vwk[ind] = MM(i * sm + 3 + k, i1 * sm + 3 + j);
                    for (int m = 0; m < im; ++m) MW(ind, m) = MM(bm + m, i1 * sm + 3 + j);
                    ++ind;
                }
                for (int j = k; j < std::min(2 * k, sm - 1); ++j) {
                    vwk[ind] = MM(i * sm + 3 + k, i2 * sm + 3 + j);
                    for (int m = 0; m < im; ++m) MW(ind, m) = MM(bm + m, i2 * sm + 3 + j);
                    ++ind;
                }
                int ind1 = 0;
                for (int m = 2; m < sm + 1; ++m) {
                    for (int n = 1; n < sm + 1 - m; ++n) {
                        vwk[ind] = MM(i * sm + 3 + k, bm + ind1);
                        for (int j = 0; j < im; ++j) MW(ind, j) = MM(bm + j, bm + ind1);
                        ++ind; ++ind1;
                    }
                    ++ind1;
                }
                if (ind != im) throw std::runtime_error("lumpinv: interior constraint count");
                lu_solve_inplace(mwk, im, T, vwk.data());
                for (int j = 0; j < im; ++j) ifmb[(i * sm + 3 + k) * im + j] = vwk[j];
            }
            for (int j = 0; j < im; ++j) {
                vwk[j] = MM(i * sm + sm + 2, j + bm);
                for (int k = 0; k < im; ++k) MW(j, k) = MM(k + bm, j + bm);
            }
            lu_solve_inplace(mwk, im, T, vwk.data());
            for (int j = 0; j < im; ++j) ifmb[(i * sm + sm + 2) * im + j] = vwk[j];
        }
    }

    for (int k = 0; k < sm; ++k) {
        std::fill(u.begin(), u.end(), 0.0);
        u[k + 3] = 1.0;
        for (int j = 0; j < im; ++j) u[j + bm] = -ifmb[(k + 3) * im + j];
        proj(u.data(), wk1.data(), gpn);
        intgrt(l.data(), wk1.data(), gpn);
        for (int j = 0; j < 3; ++j) vfms[j * sm + k] = l[j];
        sdiag[k] = 1.0 / l[3 + k];
        for (int j = 0; j < k; ++j) {
            sfms[(j * sm + k) * 3 + 0] = l[3 + j];
            sfms[(j * sm + k) * 3 + 1] = l[3 + j + sm];
            sfms[(j * sm + k) * 3 + 2] = l[3 + j + 2 * sm];
        }
    }

    if (im > 0) {
        for (int j = 0; j < im; ++j) {
            const int i1 = std::max(0, j - ibwth);
            for (int i = i1; i <= j; ++i) idiag[j * (ibwth + 1) + (i - (j - ibwth))] = MM(i + bm, j + bm);
        }
        band_chol_upper(idiag.data(), im, ibwth);
        for (int i = 0; i < bm; ++i) {
            for (int j = 0; j < im; ++j) bfmi[i * im + j] = MM(i, j + bm);
            band_chol_solve(idiag.data(), im, ibwth, &bfmi[i * im]);
        }
    }
}

}  // namespace oracle
