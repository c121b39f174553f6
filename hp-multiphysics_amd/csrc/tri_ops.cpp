// Host-side construction of the triangle basis operators; see tri_ops.h.
#include "tri_ops.h"
#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace hpg {

void monic_jacobi(int n, double al, double be, double* a, double* b) {
    const double s = al + be;
    for (int k = 0; k < n; ++k) {
        if (k == 0) {
            a[0] = (be - al) / (s + 2.0);
            b[0] = std::exp((s + 1.0) * std::log(2.0) + std::lgamma(al + 1.0) + std::lgamma(be + 1.0) - std::lgamma(s + 2.0));
            continue;
        }
        const double t = 2.0 * k + s;
        a[k] = (s == 0.0 || be * be == al * al) ? 0.0 : (be * be - al * al) / (t * (t + 2.0));
        if (k == 1)
            b[1] = 4.0 * (al + 1.0) * (be + 1.0) / ((s + 3.0) * (s + 2.0) * (s + 2.0));
        else
            b[k] = 4.0 * k * (k + al) * (k + be) * (k + s) / (t * t * (t + 1.0) * (t - 1.0));
    }
}

// p_n and p_n' at x through the monic recurrence
static void eval_monic(int n, const double* a, const double* b, double x, double& pn, double& dpn) {
    double pkm = 0.0, pk = 1.0, dpkm = 0.0, dpk = 0.0;
    for (int k = 0; k < n; ++k) {
        const double pkp = (x - a[k]) * pk - (k ? b[k] * pkm : 0.0);
        const double dpkp = pk + (x - a[k]) * dpk - (k ? b[k] * dpkm : 0.0);
        pkm = pk; pk = pkp; dpkm = dpk; dpk = dpkp;
    }
    pn = pk; dpn = dpk;
}

// Gauss-Jacobi rule for weight (1-x)^al (1+x)^be by Newton iteration with deflation
// on p_n, then Christoffel weights w_k = 1 / sum_j phat_j(x_k)^2.
void gauss_nodes(int n, double al, double be, double* x, double* w) {
    std::vector<double> a(n + 1), b(n + 1);
    monic_jacobi(n + 1, al, be, a.data(), b.data());
    std::vector<double> r;
    for (int k = 0; k < n; ++k) {
        double z = -std::cos(M_PI * (k + 0.75) / (n + 0.5));  // ascending first guesses
        for (int it = 0; it < 100; ++it) {
            double pn, dpn;
            eval_monic(n, a.data(), b.data(), z, pn, dpn);
            double defl = 0.0;
            for (double rr : r) defl += 1.0 / (z - rr);
            const double dz = pn / (dpn - pn * defl);
            z -= dz;
            if (std::fabs(dz) <= 1e-17 * std::max(1.0, std::fabs(z))) break;
        }
        // two polishing steps without deflation
        for (int it = 0; it < 2; ++it) {
            double pn, dpn;
            eval_monic(n, a.data(), b.data(), z, pn, dpn);
            z -= pn / dpn;
        }
        r.push_back(z);
    }
    std::sort(r.begin(), r.end());
    for (int k = 0; k < n; ++k) {
        x[k] = r[k];
        double pkm = 0.0, pk = 1.0 / std::sqrt(b[0]), s = pk * pk;
        for (int j = 0; j + 1 < n; ++j) {
            const double pkp = ((r[k] - a[j]) * pk - (j ? std::sqrt(b[j]) * pkm : 0.0)) / std::sqrt(b[j + 1]);
            pkm = pk; pk = pkp;
            s += pk * pk;
        }
        w[k] = 1.0 / s;
    }
}

// monic Jacobi polynomial values/derivatives of degrees 0..n-1 at x
static void jacobi_values(int n, double al, double be, double x, double* P, double* dP) {
    std::vector<double> a(std::max(n, 1)), b(std::max(n, 1));
    monic_jacobi(std::max(n, 1), al, be, a.data(), b.data());
    double pkm = 0.0, pk = 1.0, dpkm = 0.0, dpk = 0.0;
    for (int k = 0; k < n; ++k) {
        P[k] = pk; dP[k] = dpk;
        const double pkp = (x - a[k]) * pk - (k ? b[k] * pkm : 0.0);
        const double dpkp = pk + (x - a[k]) * dpk - (k ? b[k] * dpkm : 0.0);
        pkm = pk; pk = pkp; dpkm = dpk; dpk = dpkp;
    }
}

// small dense solve with partial pivoting (row-major n x n, leading dim ld)
static void solve_dense(std::vector<double> A, int n, int ld, double* x) {
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(A[r * ld + k]) > std::fabs(A[piv * ld + k])) piv = r;
        if (A[piv * ld + k] == 0.0) throw std::runtime_error("tri_ops: singular constraint system");
        if (piv != k) {
            for (int c = 0; c < n; ++c) std::swap(A[k * ld + c], A[piv * ld + c]);
            std::swap(x[k], x[piv]);
        }
        for (int r = k + 1; r < n; ++r) {
            const double f = A[r * ld + k] / A[k * ld + k];
            for (int c = k; c < n; ++c) A[r * ld + c] -= f * A[k * ld + c];
            x[r] -= f * x[k];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = x[r];
        for (int c = r + 1; c < n; ++c) s -= A[r * ld + c] * x[c];
        x[r] = s / A[r * ld + r];
    }
}

TriOps::TriOps(int p_) : p(p_) {
    sm = p - 1; im = (p - 1) * (p - 2) / 2; tm = (p + 1) * (p + 2) / 2; bm = 3 * p;
    gpx = p + 1; gpn = p + 1; G = gpx * gpn;
    xg.resize(gpx); wx.resize(gpx); eg.resize(gpn); we.resize(gpn);
    gauss_nodes(gpx, 0.0, 0.0, xg.data(), wx.data());
    gauss_nodes(gpn, 1.0, 0.0, eg.data(), we.data());
    for (double& v : we) v *= 0.5;

    // ---- 1D functions along xi: x-modes 0..sm+2 ------------------------------------
    const int nmodx = p + 2;
    std::vector<double> Jx(std::max(sm, 1)), dJx(std::max(sm, 1));
    std::vector<double> gxv(gpx * nmodx), dgxv(gpx * nmodx);
    std::vector<double> xnorm(nmodx, 1.0);
    auto fill_x = [&](double x, double* g, double* dg) {
        g[0] = 1.0; dg[0] = 0.0;
        g[1] = 0.5 * (1 - x); dg[1] = -0.5;
        g[2] = 0.5 * (1 + x); dg[2] = 0.5;
        jacobi_values(sm, 2.0, 2.0, x, Jx.data(), dJx.data());
        for (int k = 0; k < sm; ++k) {
            const double bub = 0.25 * (1 - x) * (1 + x);
            g[3 + k] = bub * Jx[k];
            dg[3 + k] = -0.5 * x * Jx[k] + bub * dJx[k];
        }
    };
    for (int i = 0; i < gpx; ++i) fill_x(xg[i], &gxv[i * nmodx], &dgxv[i * nmodx]);
    for (int k = 3; k < nmodx; ++k) {
        double s = 0.0;
        for (int i = 0; i < gpx; ++i) s += wx[i] * gxv[i * nmodx + k] * gxv[i * nmodx + k];
        xnorm[k] = 1.0 / std::sqrt(s);
        for (int i = 0; i < gpx; ++i) { gxv[i * nmodx + k] *= xnorm[k]; dgxv[i * nmodx + k] *= xnorm[k]; }
    }

    // ---- eta functions for every mode, with the x-mode each multiplies ------------------
    std::vector<int> xmode(tm);
    std::vector<double> gnv(gpn * tm), dgnv(gpn * tm);
    for (int j = 0; j < gpn; ++j) {
        const double e = eg[j];
        double* g = &gnv[j * tm];
        double* dg = &dgnv[j * tm];
        int ind = 0;
        g[ind] = 0.5 * (1 + e); dg[ind] = 0.5; xmode[ind++] = 0;
        g[ind] = 0.5 * (1 - e); dg[ind] = -0.5; xmode[ind++] = 1;
        g[ind] = 0.5 * (1 - e); dg[ind] = -0.5; xmode[ind++] = 2;
        for (int k = 0; k < sm; ++k) {  // side 0: ((1-eta)/2)^(k+2)
            const int q = k + 2;
            g[ind] = std::pow(0.5 * (1 - e), q);
            dg[ind] = -0.5 * q * std::pow(0.5 * (1 - e), q - 1);
            xmode[ind++] = 3 + k;
        }
        jacobi_values(sm, 2.0, 2.0, e, Jx.data(), dJx.data());
        for (int side = 1; side <= 2; ++side)
            for (int k = 0; k < sm; ++k) {
                const double sg = (side == 2 && (k & 1)) ? -1.0 : 1.0;
                const double bub = 0.25 * (1 - e) * (1 + e);
                g[ind] = sg * bub * Jx[k] * xnorm[3 + k];
                dg[ind] = sg * (-0.5 * e * Jx[k] + bub * dJx[k]) * xnorm[3 + k];
                xmode[ind++] = (side == 1) ? 2 : 1;
            }
        for (int m = 2; m <= sm; ++m) {
            const int nk = sm + 1 - m;
            std::vector<double> Jm(nk), dJm(nk);
            jacobi_values(nk, 2.0 * m + 1.0, 2.0, e, Jm.data(), dJm.data());
            const double c = std::pow(0.5 * (1 - e), m), dc = -0.5 * m * std::pow(0.5 * (1 - e), m - 1);
            for (int k = 0; k < nk; ++k) {
                g[ind] = c * 0.5 * (1 + e) * Jm[k];
                dg[ind] = dc * 0.5 * (1 + e) * Jm[k] + c * 0.5 * (Jm[k] + (1 + e) * dJm[k]);
                xmode[ind++] = m + 1;
            }
        }
        if (ind != tm) throw std::runtime_error("tri_ops: mode count");
    }
    for (int m = bm; m < tm; ++m) {
        double s = 0.0;
        for (int j = 0; j < gpn; ++j) s += we[j] * gnv[j * tm + m] * gnv[j * tm + m];
        const double nrm = 1.0 / std::sqrt(s);
        for (int j = 0; j < gpn; ++j) { gnv[j * tm + m] *= nrm; dgnv[j * tm + m] *= nrm; }
    }

    // ---- dense 2D operators ------------------------------------------------------------
    P.assign(G * tm, 0.0); Pr = P; Ps = P;
    I.assign(tm * G, 0.0); Ir = I; Is = I;
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpn; ++j) {
            const int g = i * gpn + j;
            const double n0 = 2.0 / (1.0 - eg[j]), x0 = 0.5 * (1.0 + xg[i]), w = wx[i] * we[j];
            for (int m = 0; m < tm; ++m) {
                const double gx = gxv[i * nmodx + xmode[m]], dgx = dgxv[i * nmodx + xmode[m]];
                const double gn = gnv[j * tm + m], dgn = dgnv[j * tm + m];
                const double v = gx * gn, vr = n0 * dgx * gn, vs = gx * dgn + x0 * n0 * dgx * gn;
                P[g * tm + m] = v; Pr[g * tm + m] = vr; Ps[g * tm + m] = vs;
                I[m * G + g] = w * v; Ir[m * G + g] = -w * vr; Is[m * G + g] = -w * vs;
            }
        }
    // Lagrange derivative matrices on the xi and eta Gauss nodes
    auto lagr_deriv = [](const std::vector<double>& z) {
        const int n = int(z.size());
        std::vector<double> D(n * n, 0.0), c(n, 1.0);
        for (int a = 0; a < n; ++a)
            for (int k = 0; k < n; ++k) if (k != a) c[a] *= (z[a] - z[k]);
        for (int a = 0; a < n; ++a) {
            double diag = 0.0;
            for (int k = 0; k < n; ++k) if (k != a) {
                D[a * n + k] = c[a] / (c[k] * (z[a] - z[k]));
                diag += 1.0 / (z[a] - z[k]);
            }
            D[a * n + a] = diag;
        }
        return D;
    };
    const std::vector<double> Dx = lagr_deriv(xg), De = lagr_deriv(eg);
    Dr.assign(G * G, 0.0); Ds = Dr;
    for (int i = 0; i < gpx; ++i)
        for (int j = 0; j < gpn; ++j) {
            const int g = i * gpn + j;
            const double n0 = 2.0 / (1.0 - eg[j]), x0 = 0.5 * (1.0 + xg[i]);
            for (int i2 = 0; i2 < gpx; ++i2) {
                Dr[g * G + i2 * gpn + j] += n0 * Dx[i * gpx + i2];
                Ds[g * G + i2 * gpn + j] += x0 * n0 * Dx[i * gpx + i2];
            }
            for (int j2 = 0; j2 < gpn; ++j2) Ds[g * G + i * gpn + j2] += De[j * gpn + j2];
        }

    // edge basis (tri_basis/proj1d.cpp:19-59): modes {x-mode 1, x-mode 2, x-modes 3..}
    const int n1 = sm + 2;
    g1.assign(gpx * n1, 0.0); dg1 = g1; w1 = wx;
    for (int i = 0; i < gpx; ++i)
        for (int k = 0; k < n1; ++k) {
            g1[i * n1 + k] = gxv[i * nmodx + k + 1];
            dg1[i * n1 + k] = dgxv[i * nmodx + k + 1];
        }

    // 1-D edge mass matrix and the banded Cholesky factor of its side block
    {
        std::vector<double> mm(n1 * n1, 0.0);
        for (int a = 0; a < n1; ++a)
            for (int c = 0; c < n1; ++c) {
                double acc = 0.0;
                for (int i = 0; i < gpx; ++i) acc += wx[i] * g1[i * n1 + a] * g1[i * n1 + c];
                mm[a * n1 + c] = acc;
            }
        const int kd = 2, ld = 3;
        sdiag1d.assign(std::max(sm, 1) * ld, 0.0);
        for (int j = 0; j < sm; ++j)
            for (int i = std::max(0, j - kd); i <= j; ++i) sdiag1d[j * ld + kd + i - j] = mm[(i + 2) * n1 + j + 2];
        for (int j = 0; j < sm; ++j) {
            const int i0 = std::max(0, j - kd);
            for (int i = i0; i <= j; ++i) {
                double acc = sdiag1d[j * ld + kd + i - j];
                for (int k = std::max(i0, std::max(0, i - kd)); k < i; ++k)
                    acc -= sdiag1d[i * ld + kd + k - i] * sdiag1d[j * ld + kd + k - j];
                if (i == j) {
                    if (!(acc > 0.0)) throw std::runtime_error("edge mass matrix not positive definite");
                    sdiag1d[j * ld + kd] = std::sqrt(acc);
                } else {
                    sdiag1d[j * ld + kd + i - j] = acc / sdiag1d[i * ld + kd];
                }
            }
        }
    }

    // ---- lumped mass inversion (tri_basis/initialize.cpp:590-958) ------------------------
    mass.assign(tm * tm, 0.0);
    for (int a = 0; a < tm; ++a)
        for (int c = 0; c < tm; ++c) {
            double s = 0.0;
            for (int g = 0; g < G; ++g) s += I[c * G + g] * P[g * tm + a];
            mass[a * tm + c] = s;  // mm(a, c) = integral(phi_c phi_a)
        }
    auto MM = [&](int r, int c) { return mass[r * tm + c]; };
    const int T = tm;
    std::vector<double> A(T * T), rhs(T);
    auto AW = [&](int r, int c) -> double& { return A[r * T + c]; };
    if (sm > 0) { vfms.assign(3 * sm, 0.0); sfmv.assign(2 * sm, 0.0); sdiag.assign(sm, 0.0); }
    if (sm > 1) sfms.assign((sm - 1) * sm * 3, 0.0);
    if (im > 0) { ifmb.assign(bm * im, 0.0); bfmi.assign(bm * im, 0.0); idiag.assign(im * 2, 0.0); }
    if (p > 1) {
        // vertex functions that are orthogonal to the low side/interior modes
        for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            int row = 0;
            auto put_row = [&](int src, double rv) {
                rhs[row] = rv;
                for (int j = 0; j < sm; ++j) { AW(row, j) = MM(src, 3 + j + i1 * sm); AW(row, j + sm) = MM(src, 3 + j + i2 * sm); }
                for (int j = 0; j < im; ++j) AW(row, 2 * sm + j) = MM(src, j + bm);
                ++row;
            };
            for (int k = 0; k < 2; ++k) { const int src = (i + 1 + k) % 3; put_row(src, MM(src, i)); }
            for (int k = 0; k < 3; ++k)
                for (int m = 0; m < sm - 1; ++m) { const int src = 3 + m + k * sm; put_row(src, MM(src, i)); }
            int q = 0;
            for (int m = 2; m < sm + 1; ++m) {
                for (int n = 1; n < sm + 1 - m; ++n) { put_row(q + bm, MM(q + bm, i)); ++q; }
                ++q;
            }
            const int neq = (sm + 1) * (sm + 2) / 2 - 1;
            if (row != neq) throw std::runtime_error("tri_ops: vertex constraint count");
            solve_dense(A, neq, T, rhs.data());
            for (int k = 0; k < im; ++k) ifmb[i * im + k] = rhs[k + 2 * sm];
        }
        for (int k = 0; k < sm; ++k) { sfmv[k] = rhs[k + sm]; sfmv[sm + k] = rhs[k]; }
        vdiag = MM(2, 2);
        for (int k = 0; k < sm; ++k) { vdiag -= sfmv[k] * MM(2, 3 + k + sm); vdiag -= sfmv[sm + k] * MM(2, 3 + k); }
        for (int k = 0; k < im; ++k) vdiag -= ifmb[2 * im + k] * MM(2, k + bm);
    } else {
        vdiag = MM(0, 0) + MM(0, 1) + MM(0, 2);
    }
    if (im > 0) {
        for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            for (int k = 0; k < sm - 1; ++k) {
                const int src = i * sm + 3 + k;
                int row = 0;
                for (int j = k; j < sm - 1; ++j) {
                    rhs[row] = MM(src, i1 * sm + 3 + j);
                    for (int m = 0; m < im; ++m) AW(row, m) = MM(bm + m, i1 * sm + 3 + j);
                    ++row;
                }
                for (int j = k; j < std::min(2 * k, sm - 1); ++j) {
                    rhs[row] = MM(src, i2 * sm + 3 + j);
                    for (int m = 0; m < im; ++m) AW(row, m) = MM(bm + m, i2 * sm + 3 + j);
                    ++row;
                }
                int q = 0;
                for (int m = 2; m < sm + 1; ++m) {
                    for (int n = 1; n < sm + 1 - m; ++n) {
                        rhs[row] = MM(src, bm + q);
                        for (int j = 0; j < im; ++j) AW(row, j) = MM(bm + j, bm + q);
                        ++row; ++q;
                    }
                    ++q;
                }
                if (row != im) throw std::runtime_error("tri_ops: interior constraint count");
                solve_dense(A, im, T, rhs.data());
                for (int j = 0; j < im; ++j) ifmb[src * im + j] = rhs[j];
            }
            const int src = i * sm + sm + 2;
            for (int j = 0; j < im; ++j) {
                rhs[j] = MM(src, j + bm);
                for (int k = 0; k < im; ++k) AW(j, k) = MM(k + bm, j + bm);
            }
            solve_dense(A, im, T, rhs.data());
            for (int j = 0; j < im; ++j) ifmb[src * im + j] = rhs[j];
        }
    }
    for (int k = 0; k < sm; ++k) {
        // l = mass . u with u = e_{3+k} - sum_j ifmb(3+k,j) e_{bm+j}
        std::vector<double> l(tm, 0.0);
        for (int c = 0; c < tm; ++c) {
            double s = MM(3 + k, c);
            for (int j = 0; j < im; ++j) s -= ifmb[(k + 3) * im + j] * MM(bm + j, c);
            l[c] = s;
        }
        for (int j = 0; j < 3; ++j) vfms[j * sm + k] = l[j];
        sdiag[k] = 1.0 / l[3 + k];
        for (int j = 0; j < k; ++j) {
            sfms[(j * sm + k) * 3 + 0] = l[3 + j];
            sfms[(j * sm + k) * 3 + 1] = l[3 + j + sm];
            sfms[(j * sm + k) * 3 + 2] = l[3 + j + 2 * sm];
        }
    }
    if (im > 0) {
        // band (bandwidth 1) Cholesky U^T U of the interior mass block, stored
        // idiag[j][0] = U(j-1,j), idiag[j][1] = U(j,j)
        for (int j = 0; j < im; ++j) {
            double off = 0.0;
            if (j > 0) { off = MM(bm + j - 1, bm + j) / idiag[(j - 1) * 2 + 1]; idiag[j * 2 + 0] = off; }
            const double d = MM(bm + j, bm + j) - off * off;
            if (!(d > 0.0)) throw std::runtime_error("tri_ops: interior mass not SPD");
            idiag[j * 2 + 1] = std::sqrt(d);
        }
        for (int i = 0; i < bm; ++i) {
            double* x = &bfmi[i * im];
            for (int j = 0; j < im; ++j) x[j] = MM(i, j + bm);
            for (int j = 0; j < im; ++j) x[j] = (x[j] - (j ? idiag[j * 2] * x[j - 1] : 0.0)) / idiag[j * 2 + 1];
            for (int j = im - 1; j >= 0; --j) {
                x[j] /= idiag[j * 2 + 1];
                if (j) x[j - 1] -= idiag[j * 2] * x[j];
            }
        }
    }
}

}  // namespace hpg
