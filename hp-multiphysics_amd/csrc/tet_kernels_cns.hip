// HIP kernels of the tet_hp_cns residual/relaxation path (gfx950), p in {1, 2},
// NV = 5 (p, u, v, w, RT).  Same shapes as the tet_hp_ins kernels: at p = 2 the
// matrix-core element residual (k_elem_tetc_mfma, 16 elements per wave), at p = 1 one
// element per half-wave with a lane per Gauss point and the transform operators in
// LDS (k_elem_tetc); element residuals go to the element-major buffer lf[t][m][im/re][n] and are gathered per
// vertex / edge in ascending element order (k_gather_tet<5>).  The 5x5 algebra of the
// Weiss-Smith preconditioner and the element tau runs one element per lane.
#include <hip/hip_runtime.h>

namespace hpg {

constexpr int CNV = 5;

// MAX / MIN of sys/param.h, as the reference uses them
__device__ __forceinline__ double pmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double pmin(double a, double b) { return a < b ? a : b; }

// a <- dpdc(pvu) a: conservative residual to primitive (tet_hp/cns/update.cpp:349-375)
__device__ __forceinline__ void tetc_switch_variables(const double pvu[5], double a[5], double gm1, double atm) {
    const double pr = pvu[0], u = pvu[1], v = pvu[2], w = pvu[3], rt = pvu[4];
    const double rho = (pr + atm) / rt;
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double d[25] = {ke * gm1,          -u * gm1,      -v * gm1,      -w * gm1,      gm1,
                          -u / rho,          1.0 / rho,     0.0,           0.0,           0.0,
                          -v / rho,          0.0,           1.0 / rho,     0.0,           0.0,
                          -w / rho,          0.0,           0.0,           1.0 / rho,     0.0,
                          (gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, -w * gm1 / rho, gm1 / rho};
    double t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 5; ++j) s += d[i * 5 + j] * a[j];
        t[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) a[i] = t[i];
}

// forward substitution with the lower triangle of P (update.cpp:176-183, 284-290)
__device__ __forceinline__ void tetc_lower_solve(const double* __restrict__ P, double r[5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        double l0 = r[j];
#pragma unroll
        for (int k = 0; k < j; ++k) l0 -= P[j * 5 + k] * r[k];
        r[j] = l0 / P[j * 5 + j];
    }
}

__device__ __forceinline__ void matmul5(const double A[25], const double B[25], double C[25]) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) t += A[i * 5 + k] * B[k * 5 + j];
            C[i * 5 + j] = t;
        }
}

// acc += V diag(eig) VI (the |P dpdc A| construction of setup_preconditioner.cpp:289-400)
__device__ __forceinline__ void add_abs_jacobian(const double V[25], const double eig[5], const double VI[25],
                                                 double acc[25]) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) t += V[i * 5 + k] * (eig[k] * VI[k * 5 + j]);
            acc[i * 5 + j] += t;
        }
}

// X = A^-1 B for the five columns of B (overwritten); Gaussian elimination with partial
// pivoting (first row of largest |pivot|), A is destroyed.
__device__ __forceinline__ void solve5(double a[25], double b[25]) {
    int piv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        int p = k;
        double best = fabs(a[k * 5 + k]);
#pragma unroll
        for (int r = k + 1; r < 5; ++r) {
            const double v = fabs(a[r * 5 + k]);
            if (v > best) { best = v; p = r; }
        }
        piv[k] = p;
#pragma unroll
        for (int r = k + 1; r < 5; ++r)
            if (p == r) {
#pragma unroll
                for (int c = 0; c < 5; ++c) { const double t = a[k * 5 + c]; a[k * 5 + c] = a[r * 5 + c]; a[r * 5 + c] = t; }
            }
        const double inv = 1.0 / a[k * 5 + k];
#pragma unroll
        for (int r = k + 1; r < 5; ++r) {
            a[r * 5 + k] *= inv;
            const double f = a[r * 5 + k];
            if (f != 0.0) {
#pragma unroll
                for (int c = k + 1; c < 5; ++c) a[r * 5 + c] -= f * a[k * 5 + c];
            }
        }
    }
#pragma unroll
    for (int col = 0; col < 5; ++col) {
        double x[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) x[i] = b[i * 5 + col];
#pragma unroll
        for (int k = 0; k < 5; ++k)
#pragma unroll
            for (int r = k + 1; r < 5; ++r)
                if (piv[k] == r) { const double t = x[k]; x[k] = x[r]; x[r] = t; }
#pragma unroll
        for (int r = 1; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < r; ++c) x[r] -= a[r * 5 + c] * x[c];
#pragma unroll
        for (int r = 4; r >= 0; --r) {
#pragma unroll
            for (int c = r + 1; c < 5; ++c) x[r] -= a[r * 5 + c] * x[c];
            x[r] /= a[r * 5 + r];
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) b[i * 5 + col] = x[i];
    }
}

// calculate_preconditioner_tau_timestep (tet_hp/cns/setup_preconditioner.cpp:206-489),
// Weiss-Smith preconditioner (type 1), steady (bd(0) = 0): Pinv, Tau = Tinv^-1 P dpdc
// and the pseudo-time step factor
__device__ void tetc_tau_timestep(const double pvu[5], double h, double hmax, double gam, double mu, double kcond,
                                  double R, int prec, double Pinv[25], double Tau[25], double& timestep) {
    const double gm1 = gam - 1.0, gogm1 = gam / gm1;
    const double pr = pvu[0], u = pvu[1], v = pvu[2], w = pvu[3], rt = pvu[4];
    const double rho = pr / rt;
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double c2 = gam * rt, c = sqrt(c2);
    const double nu = mu / rho;
    const double cp = gogm1 * R;
    const double alpha = kcond / (rho * cp);
    const double hdt = 0.0;  // 0.5 hmax bd(0) / c
    const double umag = sqrt(u * u + v * v + w * w);
    const double M = pmax(umag / c, 1.0e-5);
    const double nuh = 3.0 * pmax(4.0 * nu / (3.0 * hmax * c), alpha / (hmax * c));
    double b2 = 1.0;
    if (M < 0.8 && prec > 0) b2 = pmin(3.0 * M * M + (hdt + nuh) * (hdt + nuh), 1.0);
#pragma unroll
    for (int k = 0; k < 25; ++k) Pinv[k] = 0.0;
    Pinv[0] = 1.0 / b2;
    Pinv[6] = Pinv[12] = Pinv[18] = Pinv[24] = 1.0;
    Pinv[20] = -(b2 - 1.0) / (gogm1 * rho * b2);
    // P <- P dpdc with P = diag(b2, 1, 1, 1, 1) + (b2 - 1)/(gogm1 rho) e4 e0^T
    double P[25];
    {
        const double dpdc[25] = {ke * gm1,          -u * gm1,      -v * gm1,      -w * gm1,      gm1,
                                 -u / rho,          1.0 / rho,     0.0,           0.0,           0.0,
                                 -v / rho,          0.0,           1.0 / rho,     0.0,           0.0,
                                 -w / rho,          0.0,           0.0,           1.0 / rho,     0.0,
                                 (gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, -w * gm1 / rho, gm1 / rho};
        const double Pws[25] = {b2, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0,
                                (b2 - 1.0) / (gogm1 * rho), 0, 0, 0, 1};
        matmul5(Pws, dpdc, P);
    }
    const double bb = 1.0 - 2.0 * b2 + b2 * b2;
    const double t1 = sqrt(u * u * bb + 4.0 * b2 * c2);
    const double t2 = sqrt(v * v * bb + 4.0 * b2 * c2);
    const double t3 = sqrt(w * w * bb + 4.0 * b2 * c2);
    double T[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) T[k] = 0.0;
    {
        const double V[25] = {0.5 * (u * (b2 - 1.0) + t1) * rho, 0.5 * (u * (b2 - 1.0) - t1) * rho, 0, 0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              0.5 * (u * (b2 - 1.0) * gm1 + gm1 * t1) / gam, 0.5 * (u * (b2 - 1.0) * gm1 - gm1 * t1) / gam, 0, 0, 1.0};
        const double eig[5] = {fabs(0.5 * (u + u * b2 + t1)), fabs(0.5 * (u + u * b2 - t1)), fabs(u), fabs(u), fabs(u)};
        const double VI[25] = {1.0 / (rho * t1), 0.5 * (t1 - u * (b2 - 1.0)) / t1, 0, 0, 0,
                               -1.0 / (rho * t1), 0.5 * (u * (b2 - 1.0) + t1) / t1, 0, 0, 0,
                               0, 0, 1.0, 0, 0,
                               0, 0, 0, 1.0, 0,
                               -gm1 / (gam * rho), 0, 0, 0, 1.0};
        add_abs_jacobian(V, eig, VI, T);
    }
    {
        const double V[25] = {0.5 * (v * (b2 - 1.0) + t2) * rho, 0.5 * (v * (b2 - 1.0) - t2) * rho, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              0.5 * (gm1 * v * (b2 - 1.0) + gm1 * t2) / gam, 0.5 * (gm1 * v * (b2 - 1.0) - gm1 * t2) / gam, 0, 0, 1.0};
        const double eig[5] = {fabs(0.5 * (v + v * b2 + t2)), fabs(0.5 * (v + v * b2 - t2)), fabs(v), fabs(v), fabs(v)};
        const double VI[25] = {1.0 / (rho * t2), 0, 0.5 * (t2 - v * (b2 - 1.0)) / t2, 0, 0,
                               -1.0 / (rho * t2), 0, 0.5 * (v * (b2 - 1.0) + t2) / t2, 0, 0,
                               0, 1.0, 0, 0, 0,
                               0, 0, 0, 1.0, 0,
                               -gm1 / (gam * rho), 0, 0, 0, 1.0};
        add_abs_jacobian(V, eig, VI, T);
    }
    {
        // entry (4, 1) with temp2, exactly as setup_preconditioner.cpp:391
        const double V[25] = {0.5 * (w * (b2 - 1.0) + t3) * rho, 0.5 * (w * (b2 - 1.0) - t3) * rho, 0, 0, 0,
                              0, 0, 1.0, 0, 0,
                              0, 0, 0, 1.0, 0,
                              1.0, 1.0, 0, 0, 0,
                              0.5 * (gm1 * w * (b2 - 1.0) + gm1 * t3) / gam, 0.5 * (gm1 * w * (b2 - 1.0) - gm1 * t2) / gam, 0, 0, 1.0};
        const double eig[5] = {fabs(0.5 * (w + w * b2 + t3)), fabs(0.5 * (w + w * b2 - t3)), fabs(w), fabs(w), fabs(w)};
        const double VI[25] = {1.0 / (rho * t3), 0, 0, 0.5 * (t3 - w * (b2 - 1.0)) / t3, 0,
                               -1.0 / (rho * t3), 0, 0, 0.5 * (w * (b2 - 1.0) + t3) / t3, 0,
                               0, 1.0, 0, 0, 0,
                               0, 0, 1.0, 0, 0,
                               -gm1 / (rho * gam), 0, 0, 0, 1.0};
        add_abs_jacobian(V, eig, VI, T);
    }
    const double s1 = 3.0 * (4.0 * nu / (3.0 * h * h)), s4 = 3.0 * (alpha / (h * h));
    const double maxeig = 0.5 * (umag + umag * b2 + sqrt(umag * umag * bb + 4.0 * b2 * c2));
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const double sd = i != j ? 0.0 : (i == 0 ? 0.0 : (i == 4 ? s4 : s1));
            T[i * 5 + j] = 2.0 * (T[i * 5 + j] + hmax * sd) / h;
        }
    timestep = pmax(3.0 * maxeig / h, pmax(0.0, 3.0 * pmax(4.0 * nu / (3.0 * h * h), alpha / (h * h))));
    solve5(T, P);
#pragma unroll
    for (int k = 0; k < 25; ++k) Tau[k] = P[k];
}

struct TetcSetupArgs {
    const double* pnts;
    const int* tet_pnt;
    const int* tet_seg;
    const double* ugv;
    const double* uge;
    const double* Pg;   // G x TM values at the Gauss points
    int ntet, em0, p, prec;
    double mu, adis, gamma, kcond, R, atm;
    double* tau;   // ntet x 25: adis Tau / jcb
    double* tpc;   // ntet x 25: Pinv * timestep
    double* iprcn; // ntet x 5: jcb
    int* err;
};

// tet_hp_cns::setup_preconditioner element loop (setup_preconditioner.cpp:46-131 and
// :148-150 without TIMEACCURATE); one element per lane
template <int P>
__global__ __launch_bounds__(256) void k_setup_tetc(TetcSetupArgs a) {
    using D = TDims<P>;
    constexpr int TM = D::TM, G = D::G;
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= a.ntet) return;
    int v[4];
    double X[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = a.tet_pnt[4 * t + k];
#pragma unroll
        for (int n = 0; n < 3; ++n) X[k][n] = a.pnts[3 * v[k] + n];
    }
    double d1[3], d2[3], d3[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) { d1[n] = X[3][n] - X[2][n]; d2[n] = X[1][n] - X[2][n]; d3[n] = X[0][n] - X[2][n]; }
    const double cx = d1[1] * d2[2] - d1[2] * d2[1], cy = d1[2] * d2[0] - d1[0] * d2[2], cz = d1[0] * d2[1] - d1[1] * d2[0];
    const double jcb = 0.125 * fabs(cx * d3[0] + cy * d3[1] + cz * d3[2]);
    const int vf[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
    double amin = 1.0e99, amax = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double e1[3], e2[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            e1[n] = X[vf[j][0]][n] - X[vf[j][1]][n];
            e2[n] = X[vf[j][0]][n] - X[vf[j][2]][n];
        }
        const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2], c2 = e1[0] * e2[1] - e1[1] * e2[0];
        const double ar = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        amax = (ar > amax ? ar : amax);
        amin = (ar < amin ? ar : amin);
    }
    const double pp = (a.p + 1) * (a.p + 1);
    const double hmin = 4.0 * jcb / (0.25 * pp * amax);
    const double hmax = 4.0 * jcb / (0.25 * pp * amin);
    // ugtouht (edge mode 0 carries sgn^0 = 1) and the Gauss-point maxima
    double uh[CNV][TM];
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
#pragma unroll
        for (int k = 0; k < 4; ++k) uh[n][k] = a.ugv[v[k] * CNV + n];
#pragma unroll
        for (int k = 4; k < TM; ++k) uh[n][k] = a.uge[(a.tet_seg[6 * t + (k - 4)] * a.em0) * CNV + n];
    }
    double umax[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, qmax = 0.0;
    for (int g = 0; g < G; ++g) {
        double u[CNV];
#pragma unroll
        for (int n = 0; n < CNV; ++n) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < TM; ++k) s += a.Pg[g * TM + k] * uh[n][k];
            u[n] = s;
        }
        const double q = pow(u[1], 2.0) + pow(u[2], 2.0) + pow(u[3], 2.0);
        qmax = pmax(qmax, q);
        umax[0] = pmax(umax[0], fabs(u[0] + a.atm));
#pragma unroll
        for (int n = 1; n < CNV; ++n) umax[n] = pmax(umax[n], fabs(u[n]));
    }
    double* tau = a.tau + 25L * t;
    double* tpc = a.tpc + 25L * t;
#pragma unroll
    for (int n = 0; n < CNV; ++n) a.iprcn[CNV * t + n] = jcb;
    if (!(hmin > 0.0) || qmax != qmax) {
        // the reference aborts here; the entity is flagged and given a harmless block
        atomicOr(a.err, 1);
#pragma unroll
        for (int k = 0; k < 25; ++k) { tau[k] = 0.0; tpc[k] = (k % 6 == 0) ? 1.0 : 0.0; }
        return;
    }
    double Pinv[25], Tau[25], ts;
    tetc_tau_timestep(umax, hmin, hmax, a.gamma, a.mu, a.kcond, a.R, a.prec, Pinv, Tau, ts);
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        tau[k] = a.adis * Tau[k] / jcb;
        tpc[k] = Pinv[k] * ts;
    }
}

// per vertex / edge: prcn = 1 / sum of the incident jcb, pc = sum(tpc jcb) / sum(jcb),
// incident elements ascending (setup_preconditioner.cpp:152-183, tet_hp/setup_preconditioner.cpp:455-476)
__global__ __launch_bounds__(256) void k_setup_gather_tetc(int nent, const int* __restrict__ ptr,
                                                           const int* __restrict__ code, const double* __restrict__ iprcn,
                                                           const double* __restrict__ tpc, double* __restrict__ prcn,
                                                           double* __restrict__ pc) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    double sj = 0.0, s[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) s[k] = 0.0;
    for (int k = ptr[e]; k < ptr[e + 1]; ++k) {
        const int t = code[k] >> 4;
        const double j = iprcn[CNV * t];
        sj += j;
#pragma unroll
        for (int q = 0; q < 25; ++q) s[q] += tpc[25L * t + q] * j;
    }
#pragma unroll
    for (int q = 0; q < 25; ++q) pc[25L * e + q] = s[q] / sj;
#pragma unroll
    for (int n = 0; n < CNV; ++n) prcn[CNV * e + n] = 1.0 / sj;
}

// T: arithmetic of the element kernel (double, or float in the FP32 inner solve of the
// defect correction); state, tau and the element records are FP64 in both
template <typename T>
struct TetcElemArgsT {
    const double* pnts;
    const int* tet_pnt;
    const int* tet_seg;
    int ntet, em0;      // ntet: number of elements to evaluate (of elist when set)
    const int* elist;   // optional list of elements; null: elements 0 .. ntet - 1
    const double* ugv;
    const double* uge;
    const double* tau;  // ntet x 25
    double* lf;         // ntet x TM x 2 x 5: [t][m][0 = imaginary, 1 = real][n]
    const T* ops;
    double mu, kcond, R, gamma, atm, beta;
};
typedef TetcElemArgsT<double> TetcElemArgs;
typedef TetcElemArgsT<float> TetcElemArgs32;

// df(m, a) -= sum_n (d(a)(0) A + d(a)(1) B + d(a)(2) C)(m, n) tres(n) with the
// conservative flux Jacobians of rsdl.cpp:799-826
template <typename T>
__device__ __forceinline__ void tetc_stabilise(const T u[5], const T tres[5], const T d[3][3],
                                               T gogm1, T atm, T df[5][3]) {
    const T pr = u[0], uv = u[1], vv = u[2], wv = u[3], rt = u[4];
    const T ke = T(0.5) * (uv * uv + vv * vv + wv * wv);
    const T rho = (pr + atm) / rt;
    const T A[25] = {uv / rt, rho, T(0.0), T(0.0), -rho * uv / rt,
                          uv * uv / rt + T(1.0), T(2.0) * rho * uv, T(0.0), T(0.0), -rho * uv * uv / rt,
                          uv * vv / rt, rho * vv, rho * uv, T(0.0), -rho * uv * vv / rt,
                          uv * wv / rt, rho * wv, T(0.0), rho * uv, -rho * uv * wv / rt,
                          uv * (gogm1 * rt + ke) / rt, rho * (gogm1 * rt + ke) + rho * uv * uv, rho * uv * vv, rho * uv * wv, -rho * uv * ke / rt};
    const T B[25] = {vv / rt, T(0.0), rho, T(0.0), -rho * vv / rt,
                          uv * vv / rt, rho * vv, rho * uv, T(0.0), -rho * uv * vv / rt,
                          vv * vv / rt + T(1.0), T(0.0), T(2.0) * rho * vv, T(0.0), -rho * vv * vv / rt,
                          vv * wv / rt, T(0.0), rho * wv, rho * vv, -rho * vv * wv / rt,
                          vv * (gogm1 * rt + ke) / rt, rho * uv * vv, rho * (gogm1 * rt + ke) + rho * vv * vv, rho * vv * wv, -rho * vv * ke / rt};
    const T C[25] = {wv / rt, T(0.0), T(0.0), rho, -rho * wv / rt,
                          uv * wv / rt, rho * wv, T(0.0), rho * uv, -rho * uv * wv / rt,
                          vv * wv / rt, T(0.0), rho * wv, rho * vv, -rho * vv * wv / rt,
                          wv * wv / rt + T(1.0), T(0.0), T(0.0), T(2.0) * rho * wv, -rho * wv * wv / rt,
                          wv * (gogm1 * rt + ke) / rt, rho * uv * wv, rho * vv * wv, rho * (gogm1 * rt + ke) + rho * wv * wv, -rho * wv * ke / rt};
#pragma unroll
    for (int m = 0; m < 5; ++m)
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const int k = m * 5 + n;
#pragma unroll
            for (int aa = 0; aa < 3; ++aa) df[m][aa] -= (d[aa][0] * A[k] + d[aa][1] * B[k] + d[aa][2] * C[k]) * tres[n];
        }
}

// tet_hp_cns::element_rsdl, straight branch (tet_hp/cns/rsdl.cpp:16-87, 472-843), steady,
// tau frozen by setup_preconditioner, body force and reference density zero
template <typename T, int P, bool FULL>
__device__ __forceinline__ void elem_tetc_body(const TetcElemArgsT<T>& a, int noct) {
    using D = TDims<P>;
    constexpr int TM = D::TM, G = D::G;
    __shared__ __attribute__((aligned(16))) T sops[D::SIZE];
    __shared__ T suh[8][CNV][17];  // padded: the two half-waves' rows fall on different LDS banks
    __shared__ __attribute__((aligned(16))) T sf[8][15][D::FROW];  // point flux rows, padded
    __shared__ T stau[8][25];
    for (int k = threadIdx.x; k < D::SIZE; k += blockDim.x) sops[k] = a.ops[k];
    const int hw = threadIdx.x >> 5, g = threadIdx.x & 31;
    const T gm1 = T(a.gamma) - T(1.0), gogm1 = T(a.gamma) / gm1;
    __syncthreads();  // operators staged
    HPG_OCTET_LOOP(noct, oc) {
        wave_lds_sync();  // previous octet done with this half-wave's LDS rows
        const int slot = oc * 8 + hw;
        const bool valid = slot < a.ntet;
        const int tt = valid ? (a.elist ? a.elist[slot] : slot) : 0;
        int vtx[4], seg[6];
#pragma unroll
        for (int k = 0; k < 4; ++k) vtx[k] = a.tet_pnt[4 * tt + k];
#pragma unroll
        for (int k = 0; k < 6; ++k) seg[k] = a.tet_seg[6 * tt + k];
        // ugtouht (tet_hp/gtol.cpp:17-44); padding half-waves carry p = RT = 1
        if (g < TM) {
#pragma unroll
            for (int n = 0; n < CNV; ++n) {
                double v;
                if (g < 4) {
                    const int ix = (g == 0 ? vtx[0] : g == 1 ? vtx[1] : g == 2 ? vtx[2] : vtx[3]) * CNV + n;
                    v = a.ugv[ix];
                } else {
                    const int e = g - 4;
                    const int s = e == 0 ? seg[0] : e == 1 ? seg[1] : e == 2 ? seg[2] : e == 3 ? seg[3] : e == 4 ? seg[4] : seg[5];
                    v = a.uge[(s * a.em0) * CNV + n];
                }
                suh[hw][n][g] = valid ? T(v) : ((n == 0 || n == 4) && g < 4 ? T(1.0) : T(0.0));
            }
        }
        if (FULL && g < 25) stau[hw][g] = valid ? T(a.tau[25L * tt + g]) : T(0.0);
        T X[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int n = 0; n < 3; ++n) X[k][n] = valid ? T(a.pnts[3 * vtx[k] + n]) : ((k == 0 && n == 2) || (k == 1 && n == 1) || (k == 3 && n == 0) ? T(1.0) : -T(1.0));
        T ld[3][3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            ld[n][0] = T(0.5) * (X[3][n] - X[2][n]);
            ld[n][1] = T(0.5) * (X[1][n] - X[2][n]);
            ld[n][2] = T(0.5) * (X[0][n] - X[2][n]);
        }
        T d[3][3];
        d[0][0] = ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1];
        d[0][1] = -ld[0][1] * ld[2][2] + ld[0][2] * ld[2][1];
        d[0][2] = ld[0][1] * ld[1][2] - ld[0][2] * ld[1][1];
        d[1][0] = -ld[1][0] * ld[2][2] + ld[1][2] * ld[2][0];
        d[1][1] = ld[0][0] * ld[2][2] - ld[0][2] * ld[2][0];
        d[1][2] = -ld[0][0] * ld[1][2] + ld[0][2] * ld[1][0];
        d[2][0] = ld[1][0] * ld[2][1] - ld[1][1] * ld[2][0];
        d[2][1] = -ld[0][0] * ld[2][1] + ld[0][1] * ld[2][0];
        d[2][2] = ld[0][0] * ld[1][1] - ld[0][1] * ld[1][0];
        const T cjcb = ld[0][0] * (ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1]) -
                            ld[0][1] * (ld[1][0] * ld[2][2] - ld[2][0] * ld[1][2]) +
                            ld[0][2] * (ld[1][0] * ld[2][1] - ld[2][0] * ld[1][1]);
        wave_lds_sync();
        // values (and r, s, t derivatives of u, v, w, RT) at this lane's Gauss point
        T u[CNV], du[4][3];
#pragma unroll
        for (int n = 0; n < CNV; ++n) u[n] = T(0.0);
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) du[n][c] = T(0.0);
        if (g < G) {
#pragma unroll
            for (int k = 0; k < TM; ++k) {
                const T pk = sops[D::OP_P + k * 32 + g];
#pragma unroll
                for (int n = 0; n < CNV; ++n) u[n] += pk * suh[hw][n][k];
                if constexpr (FULL) {
                    const T pr = sops[D::OP_PR + k * 32 + g], ps = sops[D::OP_PS + k * 32 + g],
                                 pt = sops[D::OP_PT + k * 32 + g];
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        du[n][0] += pr * suh[hw][n + 1][k];
                        du[n][1] += ps * suh[hw][n + 1][k];
                        du[n][2] += pt * suh[hw][n + 1][k];
                    }
                }
            }
        } else {
            u[0] = T(1.0);
            u[4] = T(1.0);  // keeps the lanes past the last Gauss point finite
        }
        // convective, pressure and energy fluxes (rsdl.cpp:494-540)
        T cv[CNV][3];
        {
            const T rho = (u[0] + T(a.atm)) / u[4];
            const T fx = rho * u[1], fy = rho * u[2], fz = rho * u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) cv[0][c] = d[c][0] * fx + d[c][1] * fy + d[c][2] * fz;
#pragma unroll
            for (int n = 1; n < 4; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) cv[n][c] = u[n] * cv[0][c] + d[c][n - 1] * u[0];
            const T hh = gogm1 * u[4] + T(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
#pragma unroll
            for (int c = 0; c < 3; ++c) cv[4][c] = hh * cv[0][c];
        }
#pragma unroll
        for (int n = 0; n < CNV; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = g < G ? cv[n][c] : T(0.0);
        wave_lds_sync();
        // lf_im = intgrtrst of the fluxes (rsdl.cpp:542-543); mode g / 3 lands in lane g
        double* lfo = a.lf + (size_t)tt * (2 * CNV * TM);
        const bool owner = valid && g < 3 * TM && g % 3 == 0;
        {
            T li[CNV];
            tet_intgrtrst<D, CNV, D::FROW, T>(sops, sf[hw], g, li);
            if (owner)
#pragma unroll
                for (int n = 0; n < CNV; ++n) lfo[(g / 3) * (2 * CNV) + n] = li[n];
        }
        if constexpr (FULL) {
            // viscous stress and heat flux (rsdl.cpp:584-755): visc(n,m)(a,b) applied to
            // du(m+1, b); energy: kcond tensor on d(RT) plus the viscous dissipation
            const T mujcbi = T(a.mu) / cjcb, kcjcbi = T(a.kcond) / cjcb / T(a.R);
            T df[CNV][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) df[0][c] = T(0.0);
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
                    T s = T(0.0);
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb)
#pragma unroll
                        for (int m = 0; m < 3; ++m) {
                            T vs;
                            if (n == m) {
                                vs = T(4.) / T(3.) * d[aa][n] * d[bb][n];
#pragma unroll
                                for (int c = 0; c < 3; ++c)
                                    if (c != n) vs += d[aa][c] * d[bb][c];
                            } else {
                                vs = d[aa][m] * d[bb][n] - T(2.) / T(3.) * d[aa][n] * d[bb][m];
                            }
                            s += (-mujcbi * vs) * du[m][bb];
                        }
                    df[n + 1][aa] = s;
                }
#pragma unroll
            for (int aa = 0; aa < 3; ++aa) {
                const T k0 = -kcjcbi * (d[aa][0] * d[0][0] + d[aa][1] * d[0][1] + d[aa][2] * d[0][2]);
                const T k1 = -kcjcbi * (d[aa][0] * d[1][0] + d[aa][1] * d[1][1] + d[aa][2] * d[1][2]);
                const T k2 = -kcjcbi * (d[aa][0] * d[2][0] + d[aa][1] * d[2][1] + d[aa][2] * d[2][2]);
                df[4][aa] = k0 * du[3][0] + k1 * du[3][1] + k2 * du[3][2];
                df[4][aa] += df[1][aa] * u[1] + df[2][aa] * u[2] + df[3][aa] * u[3];
            }
            wave_lds_sync();  // the half-wave is done reading the convective fluxes
#pragma unroll
            for (int n = 0; n < CNV; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = g < G ? cv[n][c] + df[n][c] : T(0.0);
            wave_lds_sync();
            // strong residual: derivr/derivs/derivt of the total fluxes (rsdl.cpp:768-773)
            T res[CNV] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
            if (g < G) tet_strong_residual<D, CNV, T>(sops, sf[hw], g, res);
            // stabilisation with the element tau (rsdl.cpp:776-827)
            T tres[CNV];
#pragma unroll
            for (int m = 0; m < CNV; ++m) {
                T s = T(0.0);
#pragma unroll
                for (int n = 0; n < CNV; ++n) s += stau[hw][5 * m + n] * res[n];
                tres[m] = s;
            }
            tetc_stabilise<T>(u, tres, d, gogm1, T(a.atm), df);
            wave_lds_sync();  // the half-wave is done reading the total fluxes
#pragma unroll
            for (int n = 0; n < CNV; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = g < G ? df[n][c] : T(0.0);
            wave_lds_sync();
            // lf_re = beta * intgrtrst (rsdl.cpp:829-835)
            T lr[CNV];
            tet_intgrtrst<D, CNV, D::FROW, T>(sops, sf[hw], g, lr);
            if (owner)
#pragma unroll
                for (int n = 0; n < CNV; ++n) lfo[(g / 3) * (2 * CNV) + CNV + n] = lr[n] * T(a.beta);
        } else {
            if (owner)
#pragma unroll
                for (int n = 0; n < CNV; ++n) lfo[(g / 3) * (2 * CNV) + CNV + n] = T(0.0);
        }
    }
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tetc(TetcElemArgs a, int noct) {
    elem_tetc_body<double, P, FULL>(a, noct);
}

// FP32 arithmetic (the inner solve of the defect correction); FP64 state and records
template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tetc32(TetcElemArgs32 a, int noct) {
    elem_tetc_body<float, P, FULL>(a, noct);
}

// Matrix-core form of the cns element residual (p = 2), the layout of
// elem_tet_mfma_body with NV = 5 (p, u, v, w, RT): 16 tetrahedra per wave, Gauss-point
// rows in two 16-row tiles, accumulators fed back as B operands, point fluxes formed in
// the lane that holds the point.  The element tau (5 x 5) sits in LDS per wave.
// Padding points and padding elements carry p = RT = 1 so that rho = (p + atm) / RT
// stays finite (a zero operator column does not cancel an inf in a matrix product).
template <typename T, int P, bool FULL>
__device__ __forceinline__ void elem_tetc_mfma_body(const TetcElemArgsT<T>& a, int ngrp) {
    using D = TDims<P>;
    using M = TMDims<P>;
    using V4 = typename Vec4<T>::type;
    constexpr int GT = M::GT, KT = M::KT, QK = M::QK, TM = D::TM, G = D::G;
    constexpr bool F64 = sizeof(T) == 8;
    // first Gauss point of K-step (tile kt, register i): its four points are
    // 16 kt + 4 i + l / 16 (f64) or 16 kt + 4 (l / 16) + i (f32)
    auto kfirst = [](int kt, int i) { return F64 ? 16 * kt + 4 * i : 16 * kt + i; };
    __shared__ T sops[M::SIZE];
    __shared__ T stau[4][16][25];
    for (int k = threadIdx.x; k < M::SIZE; k += blockDim.x) sops[k] = a.ops[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = lane & 15, q = lane >> 4;
    const int WX = ((int)gridDim.x >> 3) * 4;
    const int rank = ((int)blockIdx.x >> 3) * 4 + w;
    const int xcd = (int)blockIdx.x & 7;
    const int gpx = (ngrp + 7) >> 3;
    const int gend = min(ngrp, (xcd + 1) * gpx);
    const V4 z = {T(0), T(0), T(0), T(0)};
    const T gm1 = T(a.gamma) - T(1), gogm1 = T(a.gamma) / gm1, atm = T(a.atm);
    // Gauss point of accumulator register i of tile kt in this lane
    auto point = [&](int kt, int i) { return F64 ? 16 * kt + 4 * i + q : 16 * kt + 4 * q + i; };
    for (int grp = xcd * gpx + rank; grp < gend; grp += WX) {
        const int slot = grp * 16 + e;
        const bool valid = slot < a.ntet;
        const int tt = valid ? (a.elist ? a.elist[slot] : slot) : 0;
        T b[CNV][KT];
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            const int k = 4 * s + q;
            const double* src = nullptr;
            if (valid && k < TM)
                src = k < 4 ? a.ugv + (size_t)a.tet_pnt[4 * tt + k] * CNV
                            : a.uge + (size_t)a.tet_seg[6 * tt + (k - 4)] * a.em0 * CNV;
#pragma unroll
            for (int n = 0; n < CNV; ++n)
                b[n][s] = src ? T(src[n]) : (!valid && k < 4 && (n == 0 || n == 4) ? T(1) : T(0));
        }
        if constexpr (FULL) {
            wave_lds_sync();  // the previous group is done with this wave's tau rows
            for (int k = q; k < 25; k += 4) stau[w][e][k] = valid ? T(a.tau[25L * tt + k]) : T(0);
            wave_lds_sync();
        }
        T X[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int v = a.tet_pnt[4 * tt + k];
#pragma unroll
            for (int n = 0; n < 3; ++n)
                X[k][n] = valid ? T(a.pnts[3 * v + n]) : ((k == 0 && n == 2) || (k == 1 && n == 1) || (k == 3 && n == 0) ? T(1) : T(-1));
        }
        T ld[3][3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            ld[n][0] = T(0.5) * (X[3][n] - X[2][n]);
            ld[n][1] = T(0.5) * (X[1][n] - X[2][n]);
            ld[n][2] = T(0.5) * (X[0][n] - X[2][n]);
        }
        T d[3][3];
        d[0][0] = ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1];
        d[0][1] = -ld[0][1] * ld[2][2] + ld[0][2] * ld[2][1];
        d[0][2] = ld[0][1] * ld[1][2] - ld[0][2] * ld[1][1];
        d[1][0] = -ld[1][0] * ld[2][2] + ld[1][2] * ld[2][0];
        d[1][1] = ld[0][0] * ld[2][2] - ld[0][2] * ld[2][0];
        d[1][2] = -ld[0][0] * ld[1][2] + ld[0][2] * ld[1][0];
        d[2][0] = ld[1][0] * ld[2][1] - ld[1][1] * ld[2][0];
        d[2][1] = -ld[0][0] * ld[2][1] + ld[0][1] * ld[2][0];
        d[2][2] = ld[0][0] * ld[1][1] - ld[0][1] * ld[1][0];
        const T cjcb = ld[0][0] * (ld[1][1] * ld[2][2] - ld[1][2] * ld[2][1]) -
                       ld[0][1] * (ld[1][0] * ld[2][2] - ld[2][0] * ld[1][2]) +
                       ld[0][2] * (ld[1][0] * ld[2][1] - ld[2][0] * ld[1][1]);
        // values on Gauss-point tile mt (proj); each pass projects the tile it walks,
        // which keeps a single tile of values live
        auto values = [&](int mt, V4 UT[CNV]) {
#pragma unroll
            for (int n = 0; n < CNV; ++n) {
                V4 c = z;
#pragma unroll
                for (int s = 0; s < KT; ++s) c = mfma(sops[M::OFF_PROJ + (mt * KT + s) * 64 + lane], b[n][s], c);
                UT[n] = c;
            }
        };
        // the state at point (kt, i) of the tile values UT; p = RT = 1 past the last
        // Gauss point
        auto state = [&](const V4 UT[CNV], int kt, int i, T u[CNV]) {
            const bool pad = point(kt, i) >= G;
#pragma unroll
            for (int n = 0; n < CNV; ++n) u[n] = pad ? ((n == 0 || n == 4) ? T(1) : T(0)) : UT[n][i];
        };
        // convective, pressure and energy fluxes (rsdl.cpp:494-540)
        auto conv = [&](const T u[CNV], T cv[CNV][3]) {
            const T rho = (u[0] + atm) / u[4];
            const T fx = rho * u[1], fy = rho * u[2], fz = rho * u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) cv[0][c] = d[c][0] * fx + d[c][1] * fy + d[c][2] * fz;
#pragma unroll
            for (int n = 1; n < 4; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) cv[n][c] = u[n] * cv[0][c] + d[c][n - 1] * u[0];
            const T hh = gogm1 * u[4] + T(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
#pragma unroll
            for (int c = 0; c < 3; ++c) cv[4][c] = hh * cv[0][c];
        };
        // lf_im = intgrtrst of the fluxes (rsdl.cpp:542-543)
        {
            V4 L[CNV];
#pragma unroll
            for (int n = 0; n < CNV; ++n) L[n] = z;
#pragma unroll
            for (int kt = 0; kt < GT; ++kt) {
                V4 UT[CNV];
                values(kt, UT);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T u[CNV], cv[CNV][3];
                    state(UT, kt, i, u);
                    conv(u, cv);
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const T ai = sops[M::OFF_INT + (x * QK + kk) * 64 + lane];
#pragma unroll
                        for (int n = 0; n < CNV; ++n) L[n] = mfma(ai, cv[n][x], L[n]);
                    }
                }
            }
            double* lfo = a.lf + (size_t)tt * (2 * CNV * TM);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = F64 ? 4 * i + q : 4 * q + i;
                if (valid && m < TM) {
                    double* o = lfo + m * (2 * CNV);
#pragma unroll
                    for (int n = 0; n < CNV; ++n) o[n] = double(L[n][i]);
                    if constexpr (!FULL)
#pragma unroll
                        for (int n = 0; n < CNV; ++n) o[CNV + n] = 0.0;
                }
            }
        }
        if constexpr (FULL) {
            // viscous stress and heat flux (rsdl.cpp:584-755): the stress of velocity n
            // in direction a, -mu/cjcb (sum_b dd(a,b) du(n,b) + sum_m d(a)(m) gm(m,n)
            // - 2/3 d(a)(n) sum_m gm(m,m)) with gm(m,n) = sum_b d(b)(n) du(m,b); the
            // energy row kcond/(R cjcb) dd(a,.) d(RT) plus the viscous dissipation
            const T mujcbi = T(a.mu) / cjcb, kcjcbi = T(a.kcond) / cjcb / T(a.R);
            T dd[3][3];
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) dd[x][y] = d[x][0] * d[y][0] + d[x][1] * d[y][1] + d[x][2] * d[y][2];
            V4 R[CNV][GT];
#pragma unroll
            for (int n = 0; n < CNV; ++n)
#pragma unroll
                for (int mt = 0; mt < GT; ++mt) R[n][mt] = z;
            V4 L[CNV];
#pragma unroll
            for (int n = 0; n < CNV; ++n) L[n] = z;
#pragma unroll
            for (int kt = 0; kt < GT; ++kt) {
                // values and r, s, t derivatives of u, v, w, RT on this tile
                V4 UT[CNV];
                values(kt, UT);
                V4 DU[4][3];
#pragma unroll
                for (int bb = 0; bb < 3; ++bb)
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        V4 c = z;
#pragma unroll
                        for (int s = 0; s < KT; ++s)
                            c = mfma(sops[M::OFF_PROJ + (((1 + bb) * GT + kt) * KT + s) * 64 + lane], b[n + 1][s], c);
                        DU[n][bb] = c;
                    }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T u[CNV], cv[CNV][3], df[CNV][3], gm[3][3];
                    state(UT, kt, i, u);
                    conv(u, cv);
#pragma unroll
                    for (int m = 0; m < 3; ++m)
#pragma unroll
                        for (int n = 0; n < 3; ++n)
                            gm[m][n] = d[0][n] * DU[m][0][i] + d[1][n] * DU[m][1][i] + d[2][n] * DU[m][2][i];
                    const T dv = T(2) / T(3) * (gm[0][0] + gm[1][1] + gm[2][2]);
#pragma unroll
                    for (int x = 0; x < 3; ++x) df[0][x] = T(0);
#pragma unroll
                    for (int n = 0; n < 3; ++n)
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const T h = dd[x][0] * DU[n][0][i] + dd[x][1] * DU[n][1][i] + dd[x][2] * DU[n][2][i];
                            df[n + 1][x] = -mujcbi * (h + d[x][0] * gm[0][n] + d[x][1] * gm[1][n] + d[x][2] * gm[2][n] - d[x][n] * dv);
                        }
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        df[4][x] = -kcjcbi * (dd[x][0] * DU[3][0][i] + dd[x][1] * DU[3][1][i] + dd[x][2] * DU[3][2][i]);
                        df[4][x] += df[1][x] * u[1] + df[2][x] * u[2] + df[3][x] * u[3];
                    }
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const T ai = sops[M::OFF_INT + (x * QK + kk) * 64 + lane];
#pragma unroll
                        for (int n = 1; n < CNV; ++n) L[n] = mfma(ai, df[n][x], L[n]);
                    }
                    // strong residual: derivr/derivs/derivt of the total fluxes (rsdl.cpp:768-773)
#pragma unroll
                    for (int n = 0; n < CNV; ++n)
#pragma unroll
                        for (int x = 0; x < 3; ++x) cv[n][x] += df[n][x];
#pragma unroll
                    for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const T ad = sops[M::OFF_DER + ((x * GT + mt) * QK + kk) * 64 + lane];
#pragma unroll
                            for (int n = 0; n < CNV; ++n) R[n][mt] = mfma(ad, cv[n][x], R[n][mt]);
                        }
                }
            }
            // stabilisation with the element tau (rsdl.cpp:776-827) into lf_re
#pragma unroll
            for (int kt = 0; kt < GT; ++kt) {
                V4 UT[CNV];
                values(kt, UT);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T u[CNV], tres[CNV], sf[CNV][3];
                    state(UT, kt, i, u);
#pragma unroll
                    for (int m = 0; m < CNV; ++m) {
                        T s = T(0);
#pragma unroll
                        for (int n = 0; n < CNV; ++n) s += stau[w][e][5 * m + n] * R[n][kt][i];
                        tres[m] = s;
                    }
#pragma unroll
                    for (int n = 0; n < CNV; ++n)
#pragma unroll
                        for (int x = 0; x < 3; ++x) sf[n][x] = T(0);
                    tetc_stabilise<T>(u, tres, d, gogm1, atm, sf);
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const T ai = sops[M::OFF_INT + (x * QK + kk) * 64 + lane];
#pragma unroll
                        for (int n = 0; n < CNV; ++n) L[n] = mfma(ai, sf[n][x], L[n]);
                    }
                }
            }
            double* lfo = a.lf + (size_t)tt * (2 * CNV * TM);
            const T bt = T(a.beta);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = F64 ? 4 * i + q : 4 * q + i;
                if (valid && m < TM) {
                    double* o = lfo + m * (2 * CNV) + CNV;
#pragma unroll
                    for (int n = 0; n < CNV; ++n) o[n] = double(L[n][i] * bt);
                }
            }
        }
    }
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tetc_mfma(TetcElemArgs a, int ngrp) {
    elem_tetc_mfma_body<double, P, FULL>(a, ngrp);
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tetc32_mfma(TetcElemArgs32 a, int ngrp) {
    elem_tetc_mfma_body<float, P, FULL>(a, ngrp);
}

// inflow face flux (tet_hp/cns/bdry.cpp:58-104, 129-150): pressure from the solution,
// u, v, w, RT from the boundary values; n = (p2 - p1) x (p0 - p1) / 4 for a straight face.
// flf[f][n][m]
template <int P>
__global__ __launch_bounds__(256) void k_face_tetc(int nbf, const int* __restrict__ bf_pnt, const int* __restrict__ bf_seg,
                                                    const double* __restrict__ pnts, const double* __restrict__ ugv,
                                                    const double* __restrict__ uge, int em0, const double* __restrict__ P2,
                                                    const double* __restrict__ W2, double gamma, double atm, double bu,
                                                    double bv, double bw, double brt, double* __restrict__ flf) {
    using D = TDims<P>;
    const int f = xcd_block() * blockDim.x + threadIdx.x;
    if (f >= nbf) return;
    int pv[3], se[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) { pv[j] = bf_pnt[3 * f + j]; se[j] = bf_seg[3 * f + j]; }
    double x[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int n = 0; n < 3; ++n) x[j][n] = pnts[3 * pv[j] + n];
    double v1[3], v2[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        v1[n] = 0.5 * (x[2][n] - x[1][n]);
        v2[n] = 0.5 * (x[0][n] - x[1][n]);
    }
    const double nrm[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
    double uh[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) uh[j] = ugv[pv[j] * CNV];
#pragma unroll
    for (int j = 0; j < 3 * D::EM; ++j) uh[3 + j] = uge[(se[j] * em0) * CNV];
    const double hh = gamma / (gamma - 1.0) * brt + 0.5 * (bu * bu + bv * bv + bw * bw);
    const double un = bu * nrm[0] + bv * nrm[1] + bw * nrm[2];
    double lf[CNV][6];
#pragma unroll
    for (int n = 0; n < CNV; ++n)
#pragma unroll
        for (int m = 0; m < 6; ++m) lf[n][m] = 0.0;
#pragma unroll
    for (int q = 0; q < D::GF; ++q) {
        double pr = 0.0;
#pragma unroll
        for (int m = 0; m < D::TM2; ++m) pr += P2[q * D::TM2 + m] * uh[m];
        const double f0 = (pr + atm) / brt * un;
        const double fl[CNV] = {f0, f0 * bu + pr * nrm[0], f0 * bv + pr * nrm[1], f0 * bw + pr * nrm[2], hh * f0};
#pragma unroll
        for (int n = 0; n < CNV; ++n) {
            const double wq = fl[n] * W2[q];
#pragma unroll
            for (int m = 0; m < D::TM2; ++m) lf[n][m] += P2[q * D::TM2 + m] * wq;
        }
    }
#pragma unroll
    for (int n = 0; n < CNV; ++n)
#pragma unroll
        for (int m = 0; m < D::TM2; ++m) flf[(f * CNV + n) * D::TM2 + m] = lf[n][m];
}

// minvrt (tet_hp/cns/update.cpp:146-347), Weiss-Smith branch.  Vertices: subtract the
// edge modes, scale by vprcn vdiag, switch to primitive variables at the vertex state,
// solve with the lower triangle of vpreconditioner, vertex Dirichlet (u, v, w, RT).
struct TetcMinvArgs {
    int npnt, nseg, em0, em;
    const int* ve_ptr;   // vertex -> (edge, end) codes 2*e + end, ascending edges
    const int* ve_code;
    const int* et_ptr;   // edge -> (element, local edge) codes 16*t + i, ascending elements
    const int* et_code;
    const int* tet_pnt;
    const int* seg_pnt;
    const unsigned char* vess;
    const unsigned char* eess;
    double* resv;
    double* rese;
    const double* vprcn;
    const double* eprcn;
    const double* iprcn;
    const double* vpc;   // npnt x 25
    const double* epc;   // nseg x 25
    const double* ugv;   // stage state (switch_variables)
    double vdiag, sfmv0, sfmv1, gm1, atm;
    const double* vfms;  // 4 x 10 vertex rows of the mass matrix
    // fused stage update ug = ug0 - cfa res: vertices only when there are no edge modes
    // (the edge sweep reads the vertex state), edges always
    double* ugv_out;
    double* uge_out;
    const double* ug0v;
    const double* ug0e;
    double cfa;
};

__global__ __launch_bounds__(256) void k_minv_vrtx_tetc(TetcMinvArgs a) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= a.npnt) return;
    double r[CNV], ug[CNV];
#pragma unroll
    for (int n = 0; n < CNV; ++n) { r[n] = a.resv[v * CNV + n]; ug[n] = a.ugv[v * CNV + n]; }
    if (a.em > 0)
        for (int k = a.ve_ptr[v]; k < a.ve_ptr[v + 1]; ++k) {
            const int c = a.ve_code[k], e = c >> 1;
            const double sf = (c & 1) ? a.sfmv1 : a.sfmv0;
#pragma unroll
            for (int n = 0; n < CNV; ++n) r[n] -= sf * a.rese[(e * a.em0) * CNV + n];
        }
#pragma unroll
    for (int n = 0; n < CNV; ++n) r[n] *= a.vprcn[v * CNV + n] * a.vdiag;
    tetc_switch_variables(ug, r, a.gm1, a.atm);
    tetc_lower_solve(a.vpc + 25L * v, r);
    const unsigned msk = a.vess[v];
    double u0[CNV];
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
        if ((msk >> n) & 1) r[n] = 0.0;
        u0[n] = a.ugv_out ? a.ug0v[v * CNV + n] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
        a.resv[v * CNV + n] = r[n];
        if (a.ugv_out) a.ugv_out[v * CNV + n] = u0[n] - a.cfa * r[n];
    }
}

// edges (em = 1): primitive switch at the mean vertex state and the epreconditioner
// solve on the raw edge residual, then the solved vertex values through vfms
// (iprcn = jcb), scale by eprcn ediag with ediag = 100, edge Dirichlet
__global__ __launch_bounds__(256) void k_minv_edge_tetc(TetcMinvArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.nseg) return;
    const int p0 = a.seg_pnt[2 * e], p1 = a.seg_pnt[2 * e + 1];
    double r[CNV], ua[CNV];
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
        r[n] = a.rese[(e * a.em0) * CNV + n];
        ua[n] = 0.5 * (a.ugv[p0 * CNV + n] + a.ugv[p1 * CNV + n]);
    }
    tetc_switch_variables(ua, r, a.gm1, a.atm);
    tetc_lower_solve(a.epc + 25L * e, r);
    for (int k = a.et_ptr[e]; k < a.et_ptr[e + 1]; ++k) {
        const int c = a.et_code[k], t = c >> 4, i = c & 15;
        double ip[CNV];
#pragma unroll
        for (int n = 0; n < CNV; ++n) ip[n] = a.iprcn[CNV * t + n];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = a.tet_pnt[4 * t + j];
            const double cf = a.vfms[j * 10 + 4 + i];
#pragma unroll
            for (int n = 0; n < CNV; ++n) r[n] -= cf * (a.resv[w * CNV + n] * ip[n]);
        }
    }
    const unsigned msk = a.eess[e];
    double u0[CNV];
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
        r[n] *= a.eprcn[e * CNV + n] * 100.0;
        if ((msk >> n) & 1) r[n] = 0.0;
        u0[n] = a.uge_out ? a.ug0e[(e * a.em0) * CNV + n] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < CNV; ++n) {
        a.rese[(e * a.em0) * CNV + n] = r[n];
        if (a.uge_out) a.uge_out[(e * a.em0) * CNV + n] = u0[n] - a.cfa * r[n];
    }
}

}  // namespace hpg
