// HIP kernels of the tri_hp_cns residual/relaxation path (gfx950): compressible
// Navier-Stokes in the primitive variables (p, u, v, RT), steady, straight-sided
// elements, diagonal preconditioner with the per-vertex 4x4 block preconditioner of
// the primitive-variable projection.
//
// The element residual is dominated by pointwise work: a 4x4 stabilisation matrix
// (eigen-decomposed flux Jacobians, an LU solve) at every Gauss point.  One element
// therefore maps to 32 lanes (half a wave), one Gauss point per lane (G <= 25):
// modal transforms are per-lane dot products against operator rows staged in LDS
// (stored point-contiguous, so the 32 lanes of a half-wave read consecutive words),
// and the derivative and integration contractions exchange point values through
// LDS.  The element loop is persistent: each workgroup stages the operators once
// and walks a contiguous range of 8-element octets of one colour, the ranges laid
// out so that each XCD works on a contiguous part of the colour.
#include <hip/hip_runtime.h>
#include "tri_kernels.h"

namespace hpg {

constexpr int NVC = 4;

struct CnsPhys {
    double mu, kcond, Rgas, gamma, adis;
};

// per-level operator pack for the cns kernels (doubles): proj operators transposed
// to [mode][32 points], derivative operators to [point'][32 points], integration
// operators to [point][16 modes]; padded entries are zero
template <int P>
struct CDims {
    static constexpr int TM = Dims<P>::TM, SM = Dims<P>::SM, IM = Dims<P>::IM, BM = Dims<P>::BM, G = Dims<P>::G;
    static constexpr int OP_P = 0, OP_PR = TM * 32, OP_PS = 2 * TM * 32, OP_DR = 3 * TM * 32;
    static constexpr int OP_DS = OP_DR + G * 32, OP_IR = OP_DS + G * 32, OP_IS = OP_IR + G * 16;
    static constexpr int OP_I = OP_IS + G * 16, SIZE = OP_I + G * 16;
};

// ---------------------------------------------------------------------------------
// 4x4 helpers
template <typename T>
__device__ __forceinline__ void mm4(const T A[4][4], const T B[4][4], T C[4][4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T t = T(0.0);
#pragma unroll
            for (int k = 0; k < 4; ++k) t += A[i][k] * B[k][j];
            C[i][j] = t;
        }
}

// Solves A X = B for the four columns of B (overwritten by X); Gaussian elimination
// with partial pivoting (first row of largest |pivot|), A is destroyed.  Row swaps
// are unrolled selects so that everything stays in registers.
template <typename T>
__device__ __forceinline__ void solve4(T a[4][4], T b[4][4]) {
    int piv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int p = k;
        T best = fabs(a[k][k]);
#pragma unroll
        for (int r = k + 1; r < 4; ++r) {
            const T v = fabs(a[r][k]);
            if (v > best) { best = v; p = r; }
        }
        piv[k] = p;
#pragma unroll
        for (int r = k + 1; r < 4; ++r)
            if (p == r) {
#pragma unroll
                for (int c = 0; c < 4; ++c) { const T tmp = a[k][c]; a[k][c] = a[r][c]; a[r][c] = tmp; }
            }
        const T inv = T(1.0) / a[k][k];
#pragma unroll
        for (int r = k + 1; r < 4; ++r) {
            a[r][k] *= inv;
            const T f = a[r][k];
            if (f != T(0.0)) {
#pragma unroll
                for (int c = k + 1; c < 4; ++c) a[r][c] -= f * a[k][c];
            }
        }
    }
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        T x[4] = {b[0][col], b[1][col], b[2][col], b[3][col]};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = k + 1; r < 4; ++r)
                if (piv[k] == r) { const T tmp = x[k]; x[k] = x[r]; x[r] = tmp; }
#pragma unroll
        for (int r = 1; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < r; ++c) x[r] -= a[r][c] * x[c];
#pragma unroll
        for (int r = 3; r >= 0; --r) {
#pragma unroll
            for (int c = r + 1; c < 4; ++c) x[r] -= a[r][c] * x[c];
            x[r] /= a[r][r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) b[r][col] = x[r];
    }
}

// utilities/matrix.cpp:16-37: 20 power iterations from x = 1 and the Rayleigh quotient
template <typename T>
__device__ __forceinline__ T spectral_radius4(const T A[4][4]) {
    T x[4] = {T(1.0), T(1.0), T(1.0), T(1.0)}, t[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
    for (int it = 0; it < 20; ++it) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            T s = T(0.0);
#pragma unroll
            for (int j = 0; j < 4; ++j) s += A[i][j] * x[j];
            t[i] = s;
        }
        T s = T(0.0);
#pragma unroll
        for (int i = 0; i < 4; ++i) s += t[i] * t[i];
        const T nrm = sqrt(s);
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = t[i] / nrm;
    }
    T num = T(0.0), den = T(0.0);
#pragma unroll
    for (int i = 0; i < 4; ++i) num += x[i] * t[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) den += x[i] * x[i];
    return fabs(num / den);
}

// tri_hp_cns::calculate_tau (tri_hp/cns/setup_preconditioner.cpp:241-515), steady
// (hdt = 0) and with alph = 0 as the reference forces it: Pinv (the inverse of the
// low-Mach preconditioner), Tau = Tinv^-1 P with Tinv = 2(|A| + |B| + h S)/h, and,
// when requested, timestep = 1/spectral_radius(Tinv).
template <typename T>
__device__ __forceinline__ void cns_tau(const CnsPhys& ph, T pr, T u, T v, T rt, T h,
                                        T Pinv[4][4], T Tau[4][4], T* timestep) {
    const T gam = T(ph.gamma), gm1 = gam - T(1.0), gogm1 = gam / gm1;
    const T rho = pr / rt, ke = T(0.5) * (u * u + v * v), c2 = gam * rt, c = sqrt(c2);
    const T nu = T(ph.mu) / rho, cp = gogm1 * T(ph.Rgas), alpha = T(ph.kcond) / (rho * cp);
    const T umag = sqrt(u * u + v * v);
    const T M = (umag / c > T(1.0e-5)) ? umag / c : T(1.0e-5);
    const T nuh = T(4.0) * nu / (h * c);
    const T alh = T(2.0) * alpha / (h * c);
    T b2 = T(1.0);
    if (M < T(0.8)) {
        const T s = M * M / (T(1.0) - M * M) + T(0.0) * T(0.0) + nuh * nuh + alh * alh;
        b2 = (s < T(1.0)) ? s : T(1.0);
    }
    // alph = 0: P and Pinv have b2 on (0,0) and the (3,0) coupling only
    T P[4][4] = {{b2, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {(b2 - T(1.0)) / (gogm1 * rho), 0, 0, 1}};
    const T Pi[4][4] = {{T(1.0) / b2, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {-(b2 - T(1.0)) / (gogm1 * rho * b2), 0, 0, 1}};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Pinv[i][j] = Pi[i][j];
    const T dpdc[4][4] = {{ke * gm1, -u * gm1, -v * gm1, gm1},
                               {-u / rho, T(1.0) / rho, 0, 0},
                               {-v / rho, 0, T(1.0) / rho, 0},
                               {(gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, gm1 / rho}};
    {
        T t[4][4];
        mm4(P, dpdc, t);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) P[i][j] = t[i][j];
    }
    T Tinv[4][4];
    {
        const T temp1 = sqrt(u * u * (T(1.0) - T(2.0) * b2 + b2 * b2) + T(4.0) * b2 * c2);
        const T V[4][4] = {{T(0.5) * (u * (b2 - T(1.0)) + temp1) * rho, T(0.5) * (u * (b2 - T(1.0)) - temp1) * rho, 0, 0},
                                {1, 1, 0, 0},
                                {0, 0, 1, 0},
                                {T(0.5) * (u * (b2 - T(1.0)) * gm1 + gm1 * temp1) / gam, T(0.5) * (u * (b2 - T(1.0)) * gm1 - gm1 * temp1) / gam, 0, 1}};
        const T e[4] = {fabs(T(0.5) * (u + u * b2 + temp1)), fabs(T(0.5) * (u + u * b2 - temp1)), fabs(u), fabs(u)};
        T W[4][4] = {{T(1.0) / (temp1 * rho), -T(0.5) * (u * (b2 - T(1.0)) - temp1) / temp1, 0, 0},
                          {-T(1.0) / (temp1 * rho), T(0.5) * (u * (b2 - T(1.0)) + temp1) / temp1, 0, 0},
                          {0, 0, 1, 0},
                          {-gm1 / (gam * rho), 0, 0, 1}};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) W[i][j] = e[i] * W[i][j];
        mm4(V, W, Tinv);  // |A|
    }
    {
        const T temp2 = sqrt(v * v * (T(1.0) - T(2.0) * b2 + b2 * b2) + T(4.0) * b2 * c2);
        const T V[4][4] = {{0, 0, T(0.5) * (v * (b2 - T(1.0)) + temp2) * rho, T(0.5) * (v * (b2 - T(1.0)) - temp2) * rho},
                                {0, 1, 0, 0},
                                {0, 0, 1, 1},
                                {1, 0, T(0.5) * (gm1 * v * (b2 - T(1.0)) + gm1 * temp2) / gam, T(0.5) * (gm1 * v * (b2 - T(1.0)) - gm1 * temp2) / gam}};
        const T e[4] = {fabs(v), fabs(v), fabs(T(0.5) * (v + v * b2 + temp2)), fabs(T(0.5) * (v + v * b2 - temp2))};
        T W[4][4] = {{-gm1 / (gam * rho), 0, 0, 1},
                          {0, 1, 0, 0},
                          {T(1.0) / (temp2 * rho), 0, T(0.5) * (-v * (b2 - T(1.0)) + temp2) / temp2, 0},
                          {-T(1.0) / (rho * temp2), 0, T(0.5) * (v * (b2 - T(1.0)) + temp2) / temp2, 0}};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) W[i][j] = e[i] * W[i][j];
        T B[4][4];
        mm4(V, W, B);
        const T Sd[4][4] = {{0, 0, 0, 0}, {0, nu / (h * h), 0, 0}, {0, 0, nu / (h * h), 0}, {0, 0, 0, alpha / (h * h)}};
        T S[4][4];
        mm4(P, Sd, S);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) Tinv[i][j] = T(2.0) * (Tinv[i][j] + B[i][j] + h * S[i][j]) / h;
    }
    if (timestep) *timestep = T(1.0) / spectral_radius4(Tinv);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Tau[i][j] = P[i][j];
    solve4(Tinv, Tau);
}

// conservative -> primitive change at state (p, u, v, RT) (tri_hp/cns/update.cpp:247-271)
__device__ __forceinline__ void switch_vars(const double w[4], double a[4], double gm1) {
    const double pr = w[0], u = w[1], v = w[2], rt = w[3];
    const double rho = pr / rt, ke = 0.5 * (u * u + v * v);
    const double d[4][4] = {{ke * gm1, -u * gm1, -v * gm1, gm1},
                            {-u / rho, 1.0 / rho, 0.0, 0.0},
                            {-v / rho, 0.0, 1.0 / rho, 0.0},
                            {(gm1 * ke - rt) / rho, -u * gm1 / rho, -v * gm1 / rho, gm1 / rho}};
    double t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) s += d[i][j] * a[j];
        t[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = t[i];
}

// forward substitution with the lower triangle (diagonal included) of a row-major 4x4
__device__ __forceinline__ void lower4(const double L[16], double r[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double l0 = r[j];
#pragma unroll
        for (int k = 0; k < j; ++k) l0 -= L[j * 4 + k] * r[k];
        r[j] = l0 / L[j * 4 + j];
    }
}

// ---------------------------------------------------------------------------------
// tri_hp_cns::element_rsdl, straight branch (tri_hp/cns/rsdl.cpp:312-530), steady,
// source "zero", CALC_TAU2 (tau at every Gauss point), for the elements
// order[begin:end) of one colour.
// T: arithmetic type of the kernel (double, or float in the FP32 inner solve of the
// defect correction); state and residual arrays are FP64 in both
template <typename T>
struct CnsElemArgsT {
    const double* pnts;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const int* order;
    int begin, end;
    const double* ugv;
    const double* ugs;
    const double* ugi;
    double* resv;
    double* ress;
    double* resi;
    double* rrv;
    double* rrs;
    double* rri;
    const T* ops;  // CDims<P> pack
    int sm0, im0;
    CnsPhys ph;
    double beta, keep;
    const unsigned char* fcv;
    const unsigned char* fcs;
    int colour;
    // FP32 kernel: constant source sfac[0] * src added on the first visit of every
    // coefficient (the scaled FP64 residual of the defect correction); null otherwise
    const double* srcv;
    const double* srcs;
    const double* srci;
    const double* sfac;
};
typedef CnsElemArgsT<double> CnsElemArgs;
typedef CnsElemArgsT<float> CnsElemArgs32;

// octet loop of the persistent element kernels: XCD x of the 8 walks the contiguous
// octet range [x*per, (x+1)*per) with the gridDim/8 blocks it is given
// LDS exchange among the lanes of one wave.  Every LDS buffer written inside the
// octet loops is private to a half-wave, so ordering the wave's own LDS accesses is
// enough: a wave's LDS instructions complete in issue order, and the fences and
// wave barrier keep the compiler from moving memory operations across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define HPG_OCTET_LOOP(noct, oc)                                                          \
    const int nbx_ = gridDim.x >> 3, xcd_ = blockIdx.x & 7, jb_ = blockIdx.x >> 3;       \
    const int per_ = ((noct) + 7) / 8;                                                     \
    const int oend_ = min((xcd_ + 1) * per_, (noct));                                      \
    for (int oc = xcd_ * per_ + jb_; oc < oend_; oc += nbx_)

template <typename T, int P, bool FULL>
__device__ __forceinline__ void elem_cns_body(const CnsElemArgsT<T>& a, int noct) {
    using D = CDims<P>;
    constexpr int TM = D::TM, G = D::G, NO = NVC * TM;
    constexpr int SMD = D::SM > 0 ? D::SM : 1;
    __shared__ T sops[D::SIZE];
    // rows padded by one word: the lanes of a half-wave read different variables' rows at
    // one Gauss point, and the two half-waves of a wave read their own buffers, which
    // unpadded sit a multiple of the LDS bank period apart
    __shared__ T suh[8][NVC][17];
    __shared__ T sf[8][8][33];
    for (int k = threadIdx.x; k < D::SIZE; k += blockDim.x) sops[k] = a.ops[k];
    const int hw = threadIdx.x >> 5, g = threadIdx.x & 31;
    const T gm1 = T(a.ph.gamma) - T(1.0), gogm1 = T(a.ph.gamma) / gm1;
    __syncthreads();  // operators staged

    HPG_OCTET_LOOP(noct, oc) {
        wave_lds_sync();  // previous octet done with this half-wave's LDS buffers
        const int idx = a.begin + oc * 8 + hw;
        const bool valid = idx < a.end;
        const int t = valid ? a.order[idx] : a.order[a.begin];
        int vtx[3], seg[3], sgn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vtx[k] = a.tri_pnt[3 * t + k];
            seg[k] = a.tri_seg[3 * t + k];
            sgn[k] = a.tri_sgn[3 * t + k];
        }
        const double x0 = a.pnts[2 * vtx[0]], y0 = a.pnts[2 * vtx[0] + 1];
        const double x1 = a.pnts[2 * vtx[1]], y1 = a.pnts[2 * vtx[1] + 1];
        const double x2 = a.pnts[2 * vtx[2]], y2 = a.pnts[2 * vtx[2] + 1];
        // ldcrd(n,0) = dx_n/dr, ldcrd(n,1) = dx_n/ds (rsdl.cpp:56-59)
        const T l00 = T(T(0.5) * (x2 - x1)), l01 = T(T(0.5) * (x0 - x1));
        const T l10 = T(T(0.5) * (y2 - y1)), l11 = T(T(0.5) * (y0 - y1));

        if (g < TM) {
            int kind;
            double sg;
            const int off = mode_offset<P, NVC>(g, t, vtx, seg, sgn, a.sm0, a.im0, kind, sg);
            const double* src = kind == 0 ? a.ugv : (kind == 1 ? a.ugs : a.ugi);
#pragma unroll
            for (int n = 0; n < NVC; ++n)
                suh[hw][n][g] = valid ? T(sg * src[off + n]) : ((g < 3 && (n == 0 || n == 3)) ? T(T(1.0)) : T(T(0.0)));
        }
        wave_lds_sync();

        // values (and derivatives of u, v, RT) at this lane's Gauss point
        T u[NVC], ur[NVC], us[NVC];
#pragma unroll
        for (int n = 0; n < NVC; ++n) { u[n] = T(0.0); ur[n] = T(0.0); us[n] = T(0.0); }
#pragma unroll 3
        for (int k = 0; k < TM; ++k) {
            const T pk = sops[D::OP_P + k * 32 + g];
#pragma unroll
            for (int n = 0; n < NVC; ++n) u[n] += pk * suh[hw][n][k];
            if constexpr (FULL) {
                const T prk = sops[D::OP_PR + k * 32 + g], psk = sops[D::OP_PS + k * 32 + g];
#pragma unroll
                for (int n = 1; n < NVC; ++n) { ur[n] += prk * suh[hw][n][k]; us[n] += psk * suh[hw][n][k]; }
            }
        }
        if (g >= G) { u[0] = T(1.0); u[3] = T(1.0); }  // padding lanes: keep the pointwise work finite

        // convective terms (rsdl.cpp:315-351)
        T cv[NVC][2];
        {
            const T rho = u[0] / u[3];
            const T fluxx = rho * (u[1] - T(0.0)), fluxy = rho * (u[2] - T(0.0));
            cv[0][0] = +l11 * fluxx - l01 * fluxy;
            cv[0][1] = -l10 * fluxx + l00 * fluxy;
            cv[1][0] = u[1] * cv[0][0] + l11 * u[0];
            cv[1][1] = u[1] * cv[0][1] - l10 * u[0];
            cv[2][0] = u[2] * cv[0][0] - l01 * u[0];
            cv[2][1] = u[2] * cv[0][1] + l00 * u[0];
            const T E = u[3] / gm1 + T(0.5) * (u[1] * u[1] + u[2] * u[2]);
            const T nfx = E * fluxx + u[0] * u[1], nfy = E * fluxy + u[0] * u[2];
            cv[3][0] = +l11 * nfx - l01 * nfy;
            cv[3][1] = -l10 * nfx + l00 * nfy;
        }
#pragma unroll
        for (int n = 0; n < NVC; ++n) { sf[hw][2 * n][g] = cv[n][0]; sf[hw][2 * n + 1][g] = cv[n][1]; }
        wave_lds_sync();
        // lf_im = intgrtrs of the convective fluxes; lane g owns outputs g and g + 32
        T li[2] = {T(0.0), T(0.0)};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int o = g + 32 * r;
            if (o < NO) {
                const int n = o / TM, k = o - n * TM;
                T s = T(0.0);
#pragma unroll 5
                for (int q = 0; q < G; ++q)
                    s += sops[D::OP_IR + q * 16 + k] * sf[hw][2 * n][q] + sops[D::OP_IS + q * 16 + k] * sf[hw][2 * n + 1][q];
                li[r] = s;
            }
        }
        T lr[2] = {T(0.0), T(0.0)};
        if constexpr (FULL) {
            // viscous stress and heat flux (rsdl.cpp:355-420), steady: rhorbd0 = 0
            const T cjcb = l00 * l11 - l10 * l01;
            const T mujcbi = T(a.ph.mu) / cjcb;
            const T kcjcbi = T(a.ph.kcond) / cjcb / T(a.ph.Rgas);
            const T v0000 = -mujcbi * (T(4.) / T(3.) * l11 * l11 + l01 * l01);
            const T v0011 = -mujcbi * (T(4.) / T(3.) * l10 * l10 + l00 * l00);
            const T v0001 = mujcbi * (T(4.) / T(3.) * l11 * l10 + l01 * l00);
            const T v1100 = -mujcbi * (l11 * l11 + T(4.) / T(3.) * l01 * l01);
            const T v1111 = -mujcbi * (l10 * l10 + T(4.) / T(3.) * l00 * l00);
            const T v1101 = mujcbi * (l11 * l10 + T(4.) / T(3.) * l01 * l00);
            const T v0100 = mujcbi * T(1.) / T(3.) * l01 * l11;
            const T v0111 = mujcbi * T(1.) / T(3.) * l00 * l10;
            const T v0101 = -mujcbi * (l01 * l10 - T(2.) / T(3.) * l00 * l11);
            const T v0110 = -mujcbi * (l00 * l11 - T(2.) / T(3.) * l01 * l10);
            const T k00 = -kcjcbi * (l11 * l11 + l01 * l01);
            const T k11 = -kcjcbi * (l10 * l10 + l00 * l00);
            const T k01 = kcjcbi * (l11 * l10 + l01 * l00);
            T df[NVC][2];
            df[0][0] = T(0.0); df[0][1] = T(0.0);
            df[1][0] = v0000 * ur[1] + v0100 * ur[2] + v0001 * us[1] + v0101 * us[2];
            df[1][1] = v0001 * ur[1] + v0110 * ur[2] + v0011 * us[1] + v0111 * us[2];
            df[2][0] = v0100 * ur[1] + v1100 * ur[2] + v0110 * us[1] + v1101 * us[2];
            df[2][1] = v0101 * ur[1] + v1101 * ur[2] + v0111 * us[1] + v1111 * us[2];
            df[3][0] = k00 * ur[3] + k01 * us[3];
            df[3][1] = k01 * ur[3] + k11 * us[3];
            df[3][0] += df[1][0] * u[1] + df[2][0] * u[2];
            df[3][1] += df[1][1] * u[1] + df[2][1] * u[2];
            wave_lds_sync();  // the half-wave is done reading the convective fluxes
#pragma unroll
            for (int n = 0; n < NVC; ++n) {
                sf[hw][2 * n][g] = cv[n][0] + df[n][0];
                sf[hw][2 * n + 1][g] = cv[n][1] + df[n][1];
            }
            wave_lds_sync();
            // strong residual: derivr/derivs of the total fluxes (rsdl.cpp:425-429)
            T res[NVC] = {T(0.0), T(0.0), T(0.0), T(0.0)};
#pragma unroll 5
            for (int q = 0; q < G; ++q) {
                const T dr = sops[D::OP_DR + q * 32 + g], ds = sops[D::OP_DS + q * 32 + g];
#pragma unroll
                for (int n = 0; n < NVC; ++n) res[n] += dr * sf[hw][2 * n][q] + ds * sf[hw][2 * n + 1][q];
            }
            // CALC_TAU2 (rsdl.cpp:431-452) and the stabilisation (:454-488)
            const T dx1 = T(x0 - x2), dy1 = T(y0 - y2), dx2 = T(x1 - x0), dy2 = T(y1 - y0);
            const T ar2 = dx1 * dy2 - dy1 * dx2;
            const T perim = sqrt(dx1 * dx1 + dy1 * dy1) + sqrt(dx2 * dx2 + dy2 * dy2) +
                                 sqrt((dx1 + dx2) * (dx1 + dx2) + (dy1 + dy2) * (dy1 + dy2));
            const T hmax = T(0.25) * ar2 / ((ar2 / perim) * (T(0.25) * (P + 1) * (P + 1)));
            {
                const T pr = u[0], uv = u[1], vv = u[2], rt = u[3];
                T Pi[4][4], tau[4][4];
                cns_tau<T>(a.ph, pr, uv, vv, rt, hmax, Pi, tau, nullptr);
                const T tf = T(a.ph.adis) / cjcb;
                T tres[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    T s = T(0.0);
#pragma unroll
                    for (int n = 0; n < 4; ++n) s += (tau[m][n] * tf) * res[n];
                    tres[m] = s;
                }
                const T rho = pr / rt, ke = T(0.5) * (uv * uv + vv * vv);
                const T A[4][4] = {
                    {uv / rt, rho, T(0.0), -rho * uv / rt},
                    {uv * uv / rt + T(1.0), T(2.0) * rho * uv, T(0.0), -rho * uv * uv / rt},
                    {uv * vv / rt, rho * vv, rho * uv, -rho * uv * vv / rt},
                    {uv * (gogm1 * rt + ke) / rt, rho * (gogm1 * rt + ke) + rho * uv * uv, rho * uv * vv,
                     -rho * uv * (gogm1 * rt + ke) / rt + rho * uv * gogm1}};
                const T B[4][4] = {
                    {vv / rt, T(0.0), rho, -rho * vv / rt},
                    {uv * vv / rt, rho * vv, rho * uv, -rho * uv * vv / rt},
                    {vv * vv / rt + T(1.0), T(0.0), T(2.0) * rho * vv, -rho * vv * vv / rt},
                    {vv * (gogm1 * rt + ke) / rt, rho * uv * vv, rho * (gogm1 * rt + ke) + rho * vv * vv,
                     -rho * vv * (gogm1 * rt + ke) / rt + rho * vv * gogm1}};
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        df[m][0] -= (+l11 * A[m][n] - l01 * B[m][n]) * tres[n];
                        df[m][1] -= (-l10 * A[m][n] + l00 * B[m][n]) * tres[n];
                    }
            }
            wave_lds_sync();  // the half-wave is done reading the total fluxes
#pragma unroll
            for (int n = 0; n < NVC; ++n) { sf[hw][2 * n][g] = df[n][0]; sf[hw][2 * n + 1][g] = df[n][1]; }
            wave_lds_sync();
            // lf_re = beta * intgrtrs(viscous - stabilisation fluxes) (rsdl.cpp:490-497)
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int o = g + 32 * r;
                if (o < NO) {
                    const int n = o / TM, k = o - n * TM;
                    T s = T(0.0);
#pragma unroll
                    for (int q = 0; q < G; ++q)
                        s += sops[D::OP_IR + q * 16 + k] * sf[hw][2 * n][q] + sops[D::OP_IS + q * 16 + k] * sf[hw][2 * n + 1][q];
                    lr[r] = s * T(a.beta);
                }
            }
        }

        // lftog with the first-colour initialisation of res / res_r (tri_hp/nstage.cpp:21-82)
        if (valid) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int o = g + 32 * r;
                if (o >= NO) continue;
                const int n = o / TM, k = o - n * TM;
                int kind;
                double sg;
                const int off = mode_offset<P, NVC>(k, t, vtx, seg, sgn, a.sm0, a.im0, kind, sg) + n;
                double* di = kind == 0 ? a.resv : (kind == 1 ? a.ress : a.resi);
                double* dr = kind == 0 ? a.rrv : (kind == 1 ? a.rrs : a.rri);
                const bool first = kind == 2 || (kind == 0 ? a.fcv[sel3(vtx, k)] : a.fcs[sel3(seg, (k - 3) / SMD)]) == a.colour;
                const T vli = T(sg) * li[r], vlr = T(sg) * lr[r];
                if (first) {
                    const double r0 = a.keep == 0.0 ? 0.0 : a.keep * dr[off];
                    if (FULL || a.keep != 1.0) dr[off] = r0 + vlr;
                    if constexpr (sizeof(T) == 4) {
                        const double* sr = kind == 0 ? a.srcv : (kind == 1 ? a.srcs : a.srci);
                        di[off] = r0 + vlr + vli + (sr ? sr[off] * a.sfac[0] : 0.0);
                    } else {
                        di[off] = r0 + vlr + vli;
                    }
                } else {
                    if constexpr (FULL) dr[off] += vlr;
                    di[off] += vlr + vli;
                }
            }
        }
    }
}

// two waves per SIMD: at p = 4 the full residual needs 295 registers unbounded (one wave
// per SIMD, its LDS round trips unhidden); bounded to 256 it spills 172 bytes per lane
template <int P, bool FULL>
__global__ __launch_bounds__(256, 2) void k_elem_cns(CnsElemArgs a, int noct) {
    elem_cns_body<double, P, FULL>(a, noct);
}

// FP32 arithmetic (the inner solve of the defect correction); FP64 state and residual
template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_cns32(CnsElemArgs32 a, int noct) {
    elem_cns_body<float, P, FULL>(a, noct);
}

// ---------------------------------------------------------------------------------
// edge-boundary flux residual (tri_hp/hp_boundary.cpp:973-1049) with the
// bdry_cns::generic flux (tri_hp/cns/bdry.cpp:192-245); p_bc = boundary pressure
struct CnsEdgeArgs {
    const int* eseg;
    int begin, end;
    const int* seg_pnt;
    const double* pnts;
    const double* ugv;
    const double* ugs;
    double* resv;
    double* ress;
    const double* g1;
    const double* dg1;
    const double* w1;
    int sm0;
    double gamma, pbc;
};

template <int P>
__global__ __launch_bounds__(64) void k_edge_cns(CnsEdgeArgs a) {
    constexpr int SM = Dims<P>::SM, N1 = SM + 2, GPX = P + 1;
    const int idx = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.end) return;
    const int sind = a.eseg[idx];
    const int v0 = a.seg_pnt[2 * sind], v1 = a.seg_pnt[2 * sind + 1];
    double uh[NVC][N1];
#pragma unroll
    for (int n = 0; n < NVC; ++n) {
        uh[n][0] = a.ugv[v0 * NVC + n];
        uh[n][1] = a.ugv[v1 * NVC + n];
#pragma unroll
        for (int m = 0; m < SM; ++m) uh[n][m + 2] = a.ugs[(sind * a.sm0 + m) * NVC + n];
    }
    const double X0 = a.pnts[2 * v0], Y0 = a.pnts[2 * v0 + 1], X1 = a.pnts[2 * v1], Y1 = a.pnts[2 * v1 + 1];
    const double gm1 = a.gamma - 1.0, pbc = a.pbc;
    double lf[NVC][N1];
#pragma unroll
    for (int n = 0; n < NVC; ++n)
#pragma unroll
        for (int k = 0; k < N1; ++k) lf[n][k] = 0.0;
#pragma unroll
    for (int ip = 0; ip < GPX; ++ip) {
        const double* g = a.g1 + ip * N1;
        const double* dg = a.dg1 + ip * N1;
        const double dxc = X0 * dg[0] + X1 * dg[1];
        const double dyc = Y0 * dg[0] + Y1 * dg[1];
        double u[NVC];
#pragma unroll
        for (int n = 0; n < NVC; ++n) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N1; ++k) s += uh[n][k] * g[k];
            u[n] = s;
        }
        double nx = dyc, ny = -dxc;
        const double jcb = sqrt(nx * nx + ny * ny);
        nx /= jcb; ny /= jcb;
        const double un = u[1] * nx + u[2] * ny;
        const double f0 = pbc / u[3] * un;
        const double rho = u[0] / u[3];
        const double E = pbc / (rho * gm1) + 0.5 * (u[1] * u[1] + u[2] * u[2]);
        const double fl[NVC] = {f0, f0 * u[1] + pbc * nx, f0 * u[2] + pbc * ny, rho * E * un + pbc * un};
#pragma unroll
        for (int n = 0; n < NVC; ++n) {
            const double r = fl[n] * jcb * a.w1[ip];
#pragma unroll
            for (int k = 0; k < N1; ++k) lf[n][k] += g[k] * r;
        }
    }
#pragma unroll
    for (int n = 0; n < NVC; ++n) {
        a.resv[v0 * NVC + n] += lf[n][0];
        a.resv[v1 * NVC + n] += lf[n][1];
#pragma unroll
        for (int m = 0; m < SM; ++m) a.ress[(sind * a.sm0 + m) * NVC + n] += lf[n][m + 2];
    }
}

// ---------------------------------------------------------------------------------
// setup_preconditioner (tri_hp/cns/setup_preconditioner.cpp:7-239, diagonal branch,
// straight elements): per element the low-Mach preconditioner inverse scaled by the
// inverse pseudo-time step, tprcn = jcb; then per-vertex / per-side sums
struct CnsSetupArgs {
    const double* pnts;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const double* ugv;
    const double* ugs;
    const double* ugi;
    const double* Pmat;  // G x tm
    double* tprcn;       // ntri x 4
    double* tpc;         // ntri x 16
    int* err;
    int ntri, sm0, im0;
    CnsPhys ph;
};

template <int P>
__global__ __launch_bounds__(128) void k_setup_elem_cns(CnsSetupArgs a) {
    constexpr int TM = Dims<P>::TM, SM = Dims<P>::SM, G = Dims<P>::G;
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    int vtx[3], seg[3], sgn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        vtx[k] = a.tri_pnt[3 * t + k];
        seg[k] = a.tri_seg[3 * t + k];
        sgn[k] = a.tri_sgn[3 * t + k];
    }
    double uh[NVC][TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        int kind;
        double sg;
        const int off = mode_offset<P, NVC>(k, t, vtx, seg, sgn, a.sm0, a.im0, kind, sg);
        const double* src = kind == 0 ? a.ugv : (kind == 1 ? a.ugs : a.ugi);
#pragma unroll
        for (int n = 0; n < NVC; ++n) uh[n][k] = sg * src[off + n];
    }
    double qmax = 0.0, umax[NVC] = {0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < G; ++g) {
        double w[NVC] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < TM; ++k) {
            const double pv = a.Pmat[g * TM + k];
#pragma unroll
            for (int n = 0; n < NVC; ++n) w[n] += pv * uh[n][k];
        }
        const double q = w[1] * w[1] + w[2] * w[2];
        qmax = (qmax > q) ? qmax : q;  // MAX of sys/param.h
#pragma unroll
        for (int n = 0; n < NVC; ++n) {
            const double f = fabs(w[n]);
            umax[n] = (umax[n] > f) ? umax[n] : f;
        }
    }
    const double X[3] = {a.pnts[2 * vtx[0]], a.pnts[2 * vtx[1]], a.pnts[2 * vtx[2]]};
    const double Y[3] = {a.pnts[2 * vtx[0] + 1], a.pnts[2 * vtx[1] + 1], a.pnts[2 * vtx[2] + 1]};
    const double dx1 = X[0] - X[2], dy1 = Y[0] - Y[2], dx2 = X[1] - X[0], dy2 = Y[1] - Y[0];
    const double jcb = 0.25 * (dx1 * dy2 - dy1 * dx2);
    double hmax = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int j1 = (j + 1) % 3;
        const double hh = (X[j] - X[j1]) * (X[j] - X[j1]) + (Y[j] - Y[j1]) * (Y[j] - Y[j1]);
        hmax = hh > hmax ? hh : hmax;
    }
    hmax = sqrt(hmax);
    const double pf = 0.25 * (P + 1) * (P + 1);
    const double h = 4. * jcb / (pf * hmax);
    hmax = hmax / pf;
    if (!(h > 0.0) || qmax != qmax) {
        atomicOr(a.err, 1);
#pragma unroll
        for (int n = 0; n < NVC; ++n) a.tprcn[NVC * t + n] = 1.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) a.tpc[16 * t + k] = (k % 5 == 0) ? 1.0 : 0.0;
        return;
    }
    double Pinv[4][4], tau[4][4], tstep;
    cns_tau(a.ph, umax[0], umax[1], umax[2], umax[3], hmax, Pinv, tau, &tstep);
    const double dtstari = 1.0 / tstep;
#pragma unroll
    for (int n = 0; n < NVC; ++n) a.tprcn[NVC * t + n] = jcb;
#pragma unroll
    for (int k = 0; k < 16; ++k) a.tpc[16 * t + k] = Pinv[k / 4][k % 4] * dtstari;
}

// vprcn = 1/(vdiag sum jcb), vpc = sum tpc*jcb over the incident elements (ascending)
__global__ __launch_bounds__(256) void k_setup_vrtx_cns(int npnt, const int* ptr, const int* tri, const double* tprcn, const double* tpc,
                                 double* vprcn, double* vpc, double vdiag) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= npnt) return;
    double s[NVC] = {0.0, 0.0, 0.0, 0.0}, m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = 0.0;
    for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
        const int t = tri[j];
        const double jcb = tprcn[NVC * t];
#pragma unroll
        for (int n = 0; n < NVC; ++n) s[n] += tprcn[NVC * t + n];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] += tpc[16 * t + k] * jcb;
    }
#pragma unroll
    for (int n = 0; n < NVC; ++n) vprcn[NVC * v + n] = 1.0 / (vdiag * s[n]);
#pragma unroll
    for (int k = 0; k < 16; ++k) vpc[16 * v + k] = m[k];
}

// sprcn = 1/sum jcb, spc = sum tpc over the (one or two) elements of each side
__global__ __launch_bounds__(256) void k_setup_side_cns(int nseg, const int* st_tri, const double* tprcn, const double* tpc, double* sprcn,
                                 double* spc) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double a[NVC] = {0.0, 0.0, 0.0, 0.0}, m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = 0.0;
    for (int j = 0; j < 2; ++j) {
        const int t = st_tri[2 * s + j];
        if (t < 0) continue;
#pragma unroll
        for (int n = 0; n < NVC; ++n) a[n] += tprcn[NVC * t + n];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] += tpc[16 * t + k];
    }
#pragma unroll
    for (int n = 0; n < NVC; ++n) sprcn[NVC * s + n] = 1.0 / a[n];
#pragma unroll
    for (int k = 0; k < 16; ++k) spc[16 * s + k] = m[k];
}

// ---------------------------------------------------------------------------------
// inflow boundaries: the conservative momentum and energy residuals follow from the
// mass residual and the prescribed u, v, RT (tri_hp/cns/bdry.cpp:247-340)
struct CnsBcArgs {
    double* resv;
    double* ress;
    const int* verts;  // inflow boundary vertices (each once)
    const int* sides;  // inflow boundary sides
    int nverts, nsides;
    const int* seg_pnt;
    const double* g1;
    const double* w1;
    const double* sdiag1d;  // upper band factor, kd = 2
    int sm0;
    double bu, bv, benergy;  // u, v, RT/(gamma-1) + |u|^2/2 of the boundary state
};

__global__ void k_mbr_vrtx(CnsBcArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nverts) return;
    const int v = a.verts[k];
    const double r0 = a.resv[v * NVC];
    a.resv[v * NVC + 1] = r0 * a.bu;
    a.resv[v * NVC + 2] = r0 * a.bv;
    a.resv[v * NVC + 3] = r0 * a.benergy;
}

// banded U^T U x = b, U upper band with bandwidth 2 stored per column (3 per row)
template <int N>
__device__ __forceinline__ void band2_solve(const double* ab, double b[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double s = b[j];
#pragma unroll
        for (int k = (j - 2 > 0 ? j - 2 : 0); k < j; ++k) s -= ab[j * 3 + 2 + k - j] * b[k];
        b[j] = s / ab[j * 3 + 2];
    }
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
        const double s = b[j] / ab[j * 3 + 2];
        b[j] = s;
#pragma unroll
        for (int k = (j - 2 > 0 ? j - 2 : 0); k < j; ++k) b[k] -= ab[j * 3 + 2 + k - j] * s;
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_mbr_side(CnsBcArgs a) {
    constexpr int SM = Dims<P>::SM, N1 = SM + 2, GPX = P + 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nsides) return;
    const int s = a.sides[k];
    const int v0 = a.seg_pnt[2 * s], v1 = a.seg_pnt[2 * s + 1];
    double rc[N1];
    rc[0] = a.resv[v0 * NVC];
    rc[1] = a.resv[v1 * NVC];
#pragma unroll
    for (int m = 0; m < SM; ++m) rc[m + 2] = a.ress[(s * a.sm0 + m) * NVC];
    const double bval[NVC] = {0.0, a.bu, a.bv, a.benergy};
#pragma unroll
    for (int n = 1; n < NVC; ++n) {
        double lf[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) lf[m] = 0.0;
        const double e0 = a.resv[v0 * NVC + n], e1 = a.resv[v1 * NVC + n];
#pragma unroll
        for (int i = 0; i < GPX; ++i) {
            const double* g = a.g1 + i * N1;
            double r1 = 0.0;
#pragma unroll
            for (int m = 0; m < N1; ++m) r1 += rc[m] * g[m];
            const double f = (e0 * g[0] + e1 * g[1]) - r1 * bval[n];
#pragma unroll
            for (int m = 0; m < N1; ++m) lf[m] += g[m] * a.w1[i] * f;
        }
        double x[SM];
#pragma unroll
        for (int m = 0; m < SM; ++m) x[m] = lf[m + 2];
        band2_solve<SM>(a.sdiag1d, x);
#pragma unroll
        for (int m = 0; m < SM; ++m) a.ress[(s * a.sm0 + m) * NVC + n] = -x[m];
    }
}

// ---------------------------------------------------------------------------------
// project_new_variables (tri_hp/cns/update.cpp:5-245): the conservative correction
// becomes a primitive one, r <- L^-1 dpdc(u) r with L the lower triangle of the
// vertex block preconditioner (interpolated linearly along sides and over elements),
// projected back onto the side and interior modes.
struct CnsProjArgs {
    const double* ugv;
    const double* ugs;
    const double* ugi;
    double* resv;
    double* ress;
    double* resi;
    double* rtv;  // copies of the conservative correction (res_temp)
    double* rts;
    const double* vpc;
    const unsigned char* vess;
    const unsigned char* sess;
    const int* seg_pnt;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const int* order;
    const double* g1;
    const double* w1;
    const double* sdiag1d;
    const double* idiag;  // interior band factor: row j = U(j-1,j), U(j,j)
    const double* ops;    // CDims<P> pack
    int npnt, nseg, ntri, sm0, im0;
    double gamma;
};

__global__ __launch_bounds__(256) void k_pnv_vrtx(CnsProjArgs a) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= a.npnt) return;
    double w[NVC], r[NVC];
#pragma unroll
    for (int n = 0; n < NVC; ++n) {
        w[n] = a.ugv[v * NVC + n];
        r[n] = a.resv[v * NVC + n];
        a.rtv[v * NVC + n] = r[n];
    }
    switch_vars(w, r, a.gamma - 1.0);
    double L[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) L[k] = a.vpc[16 * v + k];
    lower4(L, r);
    const unsigned msk = a.vess[v];
#pragma unroll
    for (int n = 0; n < NVC; ++n) a.resv[v * NVC + n] = ((msk >> n) & 1) ? 0.0 : r[n];
}

template <int P>
__global__ __launch_bounds__(256) void k_pnv_side(CnsProjArgs a) {
    constexpr int SM = Dims<P>::SM, N1 = SM + 2, GPX = P + 1;
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= a.nseg) return;
    const int p0 = a.seg_pnt[2 * s], p1 = a.seg_pnt[2 * s + 1];
    const double gm1 = a.gamma - 1.0;
    double uc[NVC][N1], rc[NVC][N1], tv0[NVC], tv1[NVC];
#pragma unroll
    for (int n = 0; n < NVC; ++n) {
        uc[n][0] = a.ugv[p0 * NVC + n]; uc[n][1] = a.ugv[p1 * NVC + n];
        rc[n][0] = a.rtv[p0 * NVC + n]; rc[n][1] = a.rtv[p1 * NVC + n];
        tv0[n] = a.resv[p0 * NVC + n]; tv1[n] = a.resv[p1 * NVC + n];
#pragma unroll
        for (int m = 0; m < SM; ++m) {
            uc[n][m + 2] = a.ugs[(s * a.sm0 + m) * NVC + n];
            rc[n][m + 2] = a.ress[(s * a.sm0 + m) * NVC + n];
        }
    }
    double lf[NVC][N1];
#pragma unroll
    for (int n = 0; n < NVC; ++n)
#pragma unroll
        for (int m = 0; m < N1; ++m) lf[n][m] = 0.0;
#pragma unroll
    for (int i = 0; i < GPX; ++i) {
        const double* g = a.g1 + i * N1;
        double w[NVC], r[NVC], L[16];
#pragma unroll
        for (int n = 0; n < NVC; ++n) {
            double su = 0.0, sr = 0.0;
#pragma unroll
            for (int m = 0; m < N1; ++m) { su += uc[n][m] * g[m]; sr += rc[n][m] * g[m]; }
            w[n] = su; r[n] = sr;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) L[k] = a.vpc[16 * p0 + k] * g[0] + a.vpc[16 * p1 + k] * g[1];
        switch_vars(w, r, gm1);
        lower4(L, r);
#pragma unroll
        for (int n = 0; n < NVC; ++n) {
            const double f = (tv0[n] * g[0] + tv1[n] * g[1]) - r[n];
#pragma unroll
            for (int m = 0; m < N1; ++m) lf[n][m] += g[m] * a.w1[i] * f;
        }
    }
    const unsigned msk = a.sess[s];
#pragma unroll
    for (int n = 0; n < NVC; ++n) {
        double x[SM];
#pragma unroll
        for (int m = 0; m < SM; ++m) x[m] = lf[n][m + 2];
        band2_solve<SM>(a.sdiag1d, x);
#pragma unroll
        for (int m = 0; m < SM; ++m) {
            const int ix = (s * a.sm0 + m) * NVC + n;
            a.rts[ix] = rc[n][m + 2];
            a.ress[ix] = ((msk >> n) & 1) ? 0.0 : -x[m];
        }
    }
}

// interiors: one element per half-wave, one Gauss point per lane
template <int P>
__global__ __launch_bounds__(256) void k_pnv_int(CnsProjArgs a, int noct) {
    using D = CDims<P>;
    constexpr int TM = D::TM, G = D::G, IM = D::IM, BM = D::BM, SM = D::SM;
    __shared__ double sP[TM * 32];
    __shared__ double sI[G * 16];
    // rows padded by one word (LDS banks: see elem_cns_body)
    __shared__ double sc[8][3][NVC][17];  // u, conservative correction, primitive boundary correction
    __shared__ double sf[8][NVC][33];
    __shared__ double so[8][NVC][IM];
    __shared__ double sL[8][3][16];   // the vertices' 4x4 blocks, loaded once per element
    __shared__ double sid[2 * (IM > 0 ? IM : 1)];
    for (int k = threadIdx.x; k < TM * 32; k += blockDim.x) sP[k] = a.ops[D::OP_P + k];
    for (int k = threadIdx.x; k < G * 16; k += blockDim.x) sI[k] = a.ops[D::OP_I + k];
    for (int k = threadIdx.x; k < 2 * IM; k += blockDim.x) sid[k] = a.idiag[k];
    const int hw = threadIdx.x >> 5, g = threadIdx.x & 31;
    const double gm1 = a.gamma - 1.0;
    __syncthreads();  // operators staged
    HPG_OCTET_LOOP(noct, oc) {
        wave_lds_sync();  // previous octet done with this half-wave's LDS buffers
        const int idx = oc * 8 + hw;
        const bool valid = idx < a.ntri;
        const int t = valid ? idx : 0;
        int vtx[3], seg[3], sgn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vtx[k] = a.tri_pnt[3 * t + k];
            seg[k] = a.tri_seg[3 * t + k];
            sgn[k] = a.tri_sgn[3 * t + k];
        }
        if (g < TM) {
            const int k = g;
#pragma unroll
            for (int n = 0; n < NVC; ++n) {
                double u, rr, tt;
                if (k < 3) {
                    const int ix = sel3(vtx, k) * NVC + n;
                    u = a.ugv[ix]; rr = a.rtv[ix]; tt = a.resv[ix];
                } else if (k < BM) {
                    const int i = (k - 3) / SM, m = (k - 3) - i * SM;
                    const double sg = ((m & 1) && sel3(sgn, i) < 0) ? -1.0 : 1.0;
                    const int ix = (sel3(seg, i) * a.sm0 + m) * NVC + n;
                    u = sg * a.ugs[ix]; rr = sg * a.rts[ix]; tt = sg * a.ress[ix];
                } else {
                    const int ix = (t * a.im0 + (k - BM)) * NVC + n;
                    u = a.ugi[ix]; rr = a.resi[ix]; tt = 0.0;
                }
                sc[hw][0][n][k] = u;
                sc[hw][1][n][k] = rr;
                sc[hw][2][n][k] = tt;
            }
        }
        for (int e = g; e < 48; e += 32) sL[hw][e >> 4][e & 15] = a.vpc[16 * sel3(vtx, e >> 4) + (e & 15)];
        wave_lds_sync();
        double w[NVC], r[NVC], tp[NVC];
#pragma unroll
        for (int n = 0; n < NVC; ++n) { w[n] = 0.0; r[n] = 0.0; tp[n] = 0.0; }
#pragma unroll 3
        for (int k = 0; k < TM; ++k) {
            const double pk = sP[k * 32 + g];
#pragma unroll
            for (int n = 0; n < NVC; ++n) {
                w[n] += pk * sc[hw][0][n][k];
                r[n] += pk * sc[hw][1][n][k];
                tp[n] += pk * sc[hw][2][n][k];
            }
        }
        if (g >= G) { w[0] = 1.0; w[3] = 1.0; }
        double L[16];
        {
            const double b0 = sP[0 * 32 + g], b1 = sP[1 * 32 + g], b2 = sP[2 * 32 + g];
#pragma unroll
            for (int k = 0; k < 16; ++k)
                L[k] = (k % 5 == 0 && g >= G) ? 1.0 : sL[hw][0][k] * b0 + sL[hw][1][k] * b1 + sL[hw][2][k] * b2;
        }
        switch_vars(w, r, gm1);
        lower4(L, r);
#pragma unroll
        for (int n = 0; n < NVC; ++n) sf[hw][n][g] = tp[n] - r[n];
        wave_lds_sync();
        // intgrt onto the interior modes, then the interior banded solve (dpbtrs)
        if (g < NVC * IM) {
            const int n = g / IM, m = g - n * IM, k = BM + m;
            double s = 0.0;
#pragma unroll 5
            for (int q = 0; q < G; ++q) s += sI[q * 16 + k] * sf[hw][n][q];
            so[hw][n][m] = s;
        }
        wave_lds_sync();
        if (valid && g < NVC) {
            const int n = g;
            double x[IM];
#pragma unroll
            for (int m = 0; m < IM; ++m) x[m] = so[hw][n][m];
#pragma unroll
            for (int j = 0; j < IM; ++j) x[j] = (x[j] - (j ? sid[j * 2] * x[j - 1] : 0.0)) / sid[j * 2 + 1];
#pragma unroll
            for (int j = IM - 1; j >= 0; --j) {
                x[j] /= sid[j * 2 + 1];
                if (j) x[j - 1] -= sid[j * 2] * x[j];
            }
#pragma unroll
            for (int m = 0; m < IM; ++m) a.resi[(t * a.im0 + m) * NVC + n] = -x[m];
        }
    }
}

}  // namespace hpg
