// Oracle (test infrastructure only): small dense linear-algebra kernels that the
// reference obtains from LAPACK and from its own utilities/myblas.c.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// oracle; the product library never links it.
//
// Restated routines and the reference interfaces they replace:
//   lu_factor/lu_solve   LAPACK dgetrf_/dgetrs_ as called in tri_basis/initialize.cpp:699-881
//                        (partial pivoting, dgetf2 rule: first row of maximal |a|)
//   band_chol_upper      LAPACK dpbtrf_ 'U' (tri_basis/initialize.cpp:905-915)
//   band_chol_solve      LAPACK dpbtrs_ 'U' (tri_basis/initialize.cpp:925-932)
//   dpbtrsnu2_literal    utilities/myblas.c:91-122 (DPBTRSNU2), index arithmetic kept
//                        exactly as the reference executes it, including the use of
//                        b[j] rather than b[i-ofdg+j] once i >= ofdg
//   jacobi_matrix_eigs   utilities/myblas.c:766- (gauss: Golub-Welsch through implicit QL)
#pragma once
#include <cmath>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace oracle {

// Row-major n x n matrix a (leading dimension lda).  Solves a x = b in place.
inline void lu_solve_inplace(std::vector<double> a, int n, int lda, double* b) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(a[k * lda + k]);
        for (int r = k + 1; r < n; ++r) {
            double v = std::fabs(a[r * lda + k]);
            if (v > best) { best = v; p = r; }
        }
        if (best == 0.0) throw std::runtime_error("oracle LU: singular matrix");
        piv[k] = p;
        if (p != k)
            for (int c = 0; c < n; ++c) std::swap(a[k * lda + c], a[p * lda + c]);
        double inv = 1.0 / a[k * lda + k];
        for (int r = k + 1; r < n; ++r) {
            a[r * lda + k] *= inv;
            double f = a[r * lda + k];
            if (f != 0.0)
                for (int c = k + 1; c < n; ++c) a[r * lda + c] -= f * a[k * lda + c];
        }
    }
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int r = 1; r < n; ++r)
        for (int c = 0; c < r; ++c) b[r] -= a[r * lda + c] * b[c];
    for (int r = n - 1; r >= 0; --r) {
        for (int c = r + 1; c < n; ++c) b[r] -= a[r * lda + c] * b[c];
        b[r] /= a[r * lda + r];
    }
}

// Upper band Cholesky A = U^T U.  Band storage ab[j*(kd+1) + (kd + i - j)] = A(i,j),
// max(0,j-kd) <= i <= j (the C view of LAPACK's column-major upper band layout).
inline void band_chol_upper(double* ab, int n, int kd) {
    const int ld = kd + 1;
    for (int j = 0; j < n; ++j) {
        int i0 = std::max(0, j - kd);
        for (int i = i0; i <= j; ++i) {
            double s = ab[j * ld + kd + i - j];
            int k0 = std::max(i0, std::max(0, i - kd));
            for (int k = k0; k < i; ++k)
                s -= ab[i * ld + kd + k - i] * ab[j * ld + kd + k - j];
            if (i == j) {
                if (!(s > 0.0)) throw std::runtime_error("oracle band Cholesky: not SPD");
                ab[j * ld + kd] = std::sqrt(s);
            } else {
                ab[j * ld + kd + i - j] = s / ab[i * ld + kd];
            }
        }
    }
}

// Solve U^T U x = b with the factor from band_chol_upper (one right-hand side).
inline void band_chol_solve(const double* ab, int n, int kd, double* b) {
    const int ld = kd + 1;
    for (int j = 0; j < n; ++j) {
        double s = b[j];
        for (int k = std::max(0, j - kd); k < j; ++k) s -= ab[j * ld + kd + k - j] * b[k];
        b[j] = s / ab[j * ld + kd];
    }
    for (int j = n - 1; j >= 0; --j) {
        double s = b[j] / ab[j * ld + kd];
        b[j] = s;
        for (int k = std::max(0, j - kd); k < j; ++k) b[k] -= ab[j * ld + kd + k - j] * s;
    }
}

// utilities/myblas.c:91-122 with its literal indexing (stride = row length of abd).
inline void dpbtrsnu2_literal(const double* abd, int stride, int ordr, int ofdg, double* b, int rhs) {
    for (int i = 0; i < ordr; ++i) {
        int frst = (ofdg - i > 0 ? ofdg - i : 0);
        int ba = -frst;
        for (int j = frst; j < ofdg; ++j)
            for (int n = 0; n < rhs; ++n) b[rhs * i + n] -= abd[i * stride + j] * b[(ba + j) * rhs + n];
        for (int n = 0; n < rhs; ++n) b[i * rhs + n] /= abd[i * stride + ofdg];
    }
    for (int i = ordr - 1; i >= 0; --i) {
        for (int n = 0; n < rhs; ++n) b[i * rhs + n] /= abd[i * stride + ofdg];
        int frst = (ofdg - i > 0 ? ofdg - i : 0);
        int ba = -frst;
        for (int j = frst; j < ofdg; ++j)
            for (int n = 0; n < rhs; ++n) b[rhs * (ba + j) + n] -= abd[i * stride + j] * b[i * rhs + n];
    }
}

// Gauss rule from monic recurrence coefficients a[0..n-1], b[0..n-1] (b[0] = total mass):
// nodes are the eigenvalues of the symmetric Jacobi matrix, weights b[0] * (first
// eigenvector component)^2 (Golub & Welsch 1969).  Implicit QL with Wilkinson shifts.
inline void gauss_rule(int n, const double* a, const double* b, double* x, double* w) {
    std::vector<double> d(a, a + n), e(n, 0.0), z(n, 0.0);
    for (int k = 1; k < n; ++k) e[k - 1] = std::sqrt(b[k]);
    z[0] = 1.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) throw std::runtime_error("oracle gauss_rule: no convergence");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            for (; i >= l; --i) {
                double f = s * e[i], bb = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
                s = f / r; c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                double zf = z[i + 1];
                z[i + 1] = s * z[i] + c * zf;
                z[i] = c * z[i] - s * zf;
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p; e[l] = g; e[m] = 0.0;
        }
    }
    std::vector<int> idx(n);
    for (int k = 0; k < n; ++k) idx[k] = k;
    std::sort(idx.begin(), idx.end(), [&](int p, int q) { return d[p] < d[q]; });
    for (int k = 0; k < n; ++k) { x[k] = d[idx[k]]; w[k] = b[0] * z[idx[k]] * z[idx[k]]; }
}

}  // namespace oracle
