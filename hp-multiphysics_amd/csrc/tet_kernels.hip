// HIP kernels of the tet_hp / tet_hp_ins residual and relaxation path (gfx950),
// p in {1, 2}: vertex and edge modes, NV = 4 (u, v, w, p).
//
// Element residual at p = 2: k_elem_tet_mfma, 16 elements per wave, the transforms
// on the matrix cores (below).  At p = 1 (8 Gauss points, 4 modes): k_elem_tet, one
// element per half-wave, one Gauss point per lane; the transform operators sit in LDS,
// lane-contiguous, and the point fields a reduction needs are exchanged through a
// half-wave's private LDS rows.  Element residuals go to an element-major buffer
// lf[t][m][im/re][n] and the gather kernels sum them per vertex / edge over the
// incident elements in ascending element order, the order of the reference's element loop, so the
// assembly is deterministic without colouring.
#include <hip/hip_runtime.h>

namespace hpg {

constexpr int TNV = 4;

template <int P>
struct TDims {
    static constexpr int TM = (P + 1) * (P + 2) * (P + 3) / 6;
    static constexpr int GP = P + 1;
    static constexpr int G = GP * GP * GP;
    static constexpr int EM = P - 1;
    static constexpr int TM2 = 3 + 3 * EM;
    static constexpr int GF = GP * GP;
    // operator pack (doubles): P, Pr, Ps, Pt as [m][32] (lane = Gauss point);
    // Ir, Is, It as [g][16] (lane 3m + d reads direction d of mode m), each block padded
    // so that the three directions start 20 LDS banks apart; Dr, Ds, Dt as [g][DQ]: lane
    // g reads its own row D[g][g'] two points per 16-byte load, rows zero-padded to an
    // even length DQ > G whose 2 DQ dwords stride the banks with period 16 lanes
    static constexpr int IPAD = ((20 - (2 * G * 16) % 64) % 64 + 64) % 64 / 2;
    static constexpr int DQ = P == 2 ? 30 : 10;
    static constexpr int FROW = 34;   // flux row length: even (16-byte pairs), 4 banks apart
    static constexpr int OP_P = 0, OP_PR = TM * 32, OP_PS = 2 * TM * 32, OP_PT = 3 * TM * 32;
    static constexpr int OP_IR = 4 * TM * 32, OP_IS = OP_IR + G * 16 + IPAD, OP_IT = OP_IS + G * 16 + IPAD;
    static constexpr int OP_DR = OP_IT + G * 16, OP_DS = OP_DR + G * DQ, OP_DT = OP_DS + G * DQ;
    static constexpr int SIZE = OP_DT + G * DQ;
    static_assert(OP_DR % 2 == 0 && DQ % 2 == 0 && DQ > G, "16-byte pairs of the derivative rows");
};

// T: arithmetic of the element kernel (double, or float in the FP32 inner solve of the
// defect correction); state, tau and the element records are FP64 in both
template <typename T>
struct TetElemArgsT {
    const double* pnts;
    const int* tet_pnt;
    const int* tet_seg;
    const int* tet_sgn;
    int ntet, em0;      // ntet: number of elements to evaluate (of elist when set)
    const int* elist;   // optional list of elements; null: elements 0 .. ntet - 1
    const double* ugv;
    const double* uge;
    const double* tau;  // ntet x 4 (0: momentum, 3: continuity)
    double* lf;         // ntet x TM x 2 x 4: [t][m][0 = imaginary, 1 = real][n]
    const T* ops;
    double rho, mu, beta;
};
typedef TetElemArgsT<double> TetElemArgs;
typedef TetElemArgsT<float> TetElemArgs32;

// intgrtrst of a half-wave's point fluxes sf[3n + d][q] (d = r, s, t directions): lane
// 3m + d sums direction d of mode m over the Gauss points, and lane 3m adds the three
// partial sums (two shuffles inside the half-wave); lanes from 3 TM on carry zeros.
// Every lane of the half-wave takes part.  One operator read and NVn flux reads per
// point and lane, a third of the LDS traffic of a lane-per-mode sum.
template <class D, int NVn, int ROW, typename T>
__device__ __forceinline__ void tet_intgrtrst(const T* __restrict__ sops, const T (*sf)[ROW], int g,
                                              T out[NVn]) {
    const int m = g / 3, d = g - 3 * (g / 3);
    const bool act = g < 3 * D::TM;
    const T* op = sops + (d == 0 ? D::OP_IR : d == 1 ? D::OP_IS : D::OP_IT) + (act ? m : 0);
    T s[NVn];
#pragma unroll
    for (int n = 0; n < NVn; ++n) s[n] = T(0.0);
    if (act) {
#pragma unroll 3
        for (int q = 0; q < D::G; ++q) {
            const T iq = op[q * 16];
#pragma unroll
            for (int n = 0; n < NVn; ++n) s[n] += iq * sf[3 * n + d][q];
        }
    }
#pragma unroll
    for (int n = 0; n < NVn; ++n) {
        const T t1 = __shfl_down(s[n], 1, 64), t2 = __shfl_down(s[n], 2, 64);
        out[n] = s[n] + t1 + t2;
    }
}

// strong residual res[n] += sum_q (Dr[g][q] F[3n][q] + Ds[g][q] F[3n+1][q] + Dt[g][q] F[3n+2][q])
// of the Gauss point g, two points per 16-byte LDS load of the operator rows and of
// the (broadcast) flux rows; the zero pads at q = G close the odd count
template <class D, int NVn, typename T>
__device__ __forceinline__ void tet_strong_residual(const T* __restrict__ sops, const T (*sf)[D::FROW], int g,
                                                    T res[NVn]) {
    typedef T d2 __attribute__((ext_vector_type(2)));
    const d2* dr = reinterpret_cast<const d2*>(sops + D::OP_DR + g * D::DQ);
    const d2* ds = reinterpret_cast<const d2*>(sops + D::OP_DS + g * D::DQ);
    const d2* dt = reinterpret_cast<const d2*>(sops + D::OP_DT + g * D::DQ);
#pragma unroll 2
    for (int h = 0; h < (D::G + 1) / 2; ++h) {
        const d2 a = dr[h], b = ds[h], c = dt[h];
#pragma unroll
        for (int n = 0; n < NVn; ++n) {
            const d2 f0 = reinterpret_cast<const d2*>(sf[3 * n + 0])[h];
            const d2 f1 = reinterpret_cast<const d2*>(sf[3 * n + 1])[h];
            const d2 f2 = reinterpret_cast<const d2*>(sf[3 * n + 2])[h];
            res[n] += a.x * f0.x + b.x * f1.x + c.x * f2.x;
            res[n] += a.y * f0.y + b.y * f1.y + c.y * f2.y;
        }
    }
}

// tet_hp_ins::element_rsdl, straight branch (tet_hp/ins/rsdl.cpp:15-87, 442-785),
// steady, tau frozen by setup_preconditioner
template <typename T, int P, bool FULL>
__device__ __forceinline__ void elem_tet_body(const TetElemArgsT<T>& a, int noct) {
    using D = TDims<P>;
    constexpr int TM = D::TM, G = D::G;
    __shared__ __attribute__((aligned(16))) T sops[D::SIZE];
    __shared__ T suh[8][TNV][17];  // padded: the two half-waves' rows fall on different LDS banks
    __shared__ __attribute__((aligned(16))) T sf[8][12][D::FROW];  // point flux rows, padded
    for (int k = threadIdx.x; k < D::SIZE; k += blockDim.x) sops[k] = a.ops[k];
    const int hw = threadIdx.x >> 5, g = threadIdx.x & 31;
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
        // ugtouht (tet_hp/gtol.cpp:17-44); edge mode 0 carries sgn^0 = 1
        if (g < TM) {
#pragma unroll
            for (int n = 0; n < TNV; ++n) {
                double v;
                if (g < 4) {
                    const int ix = (g == 0 ? vtx[0] : g == 1 ? vtx[1] : g == 2 ? vtx[2] : vtx[3]) * TNV + n;
                    v = a.ugv[ix];
                } else {
                    const int e = g - 4;
                    const int s = e == 0 ? seg[0] : e == 1 ? seg[1] : e == 2 ? seg[2] : e == 3 ? seg[3] : e == 4 ? seg[4] : seg[5];
                    v = a.uge[(s * a.em0) * TNV + n];
                }
                suh[hw][n][g] = valid ? T(v) : (n == 0 && g < 4 ? T(1.0) : T(0.0));
            }
        }
        // coordinate derivatives (rsdl.cpp:58-62) and cofactors (:449-458); padding
        // half-waves use the reference tet
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
        // values (and velocity r, s, t derivatives) at this lane's Gauss point
        T u[TNV], du[3][3];
#pragma unroll
        for (int n = 0; n < TNV; ++n) u[n] = T(0.0);
#pragma unroll
        for (int n = 0; n < 3; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) du[n][c] = T(0.0);
        if (g < G) {
#pragma unroll
            for (int k = 0; k < TM; ++k) {
                const T pk = sops[D::OP_P + k * 32 + g];
#pragma unroll
                for (int n = 0; n < TNV; ++n) u[n] += pk * suh[hw][n][k];
                if constexpr (FULL) {
                    const T pr = sops[D::OP_PR + k * 32 + g], ps = sops[D::OP_PS + k * 32 + g],
                                 pt = sops[D::OP_PT + k * 32 + g];
#pragma unroll
                    for (int n = 0; n < 3; ++n) {
                        du[n][0] += pr * suh[hw][n][k];
                        du[n][1] += ps * suh[hw][n][k];
                        du[n][2] += pt * suh[hw][n][k];
                    }
                }
            }
        }
        // convective and pressure fluxes (rsdl.cpp:461-508)
        T cv[3][3], c3[3];
        {
            const T fx = T(a.rho) * u[0], fy = T(a.rho) * u[1], fz = T(a.rho) * u[2];
#pragma unroll
            for (int c = 0; c < 3; ++c) c3[c] = d[c][0] * fx + d[c][1] * fy + d[c][2] * fz;
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) cv[n][c] = u[n] * c3[c] + d[c][n] * u[3];
        }
#pragma unroll
        for (int n = 0; n < 3; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = cv[n][c];
#pragma unroll
        for (int c = 0; c < 3; ++c) sf[hw][9 + c][g] = c3[c];
        wave_lds_sync();
        // lf_im = intgrtrst of the fluxes (rsdl.cpp:509-511); mode g / 3 lands in lane g
        double* lfo = a.lf + (size_t)tt * (2 * TNV * TM);
        const bool owner = valid && g < 3 * TM && g % 3 == 0;
        {
            T li[TNV];
            tet_intgrtrst<D, TNV, D::FROW, T>(sops, sf[hw], g, li);
            if (owner)
#pragma unroll
                for (int n = 0; n < TNV; ++n) lfo[(g / 3) * (2 * TNV) + n] = li[n];
        }
        if constexpr (FULL) {
            // viscous fluxes (rsdl.cpp:518-674): visc(n,m)(a,b) = -mu/cjcb (delta_nm
            // sum_c d(a)(c) d(b)(c) + d(a)(m) d(b)(n)) applied to du(m,b)
            const T cjcbi = T(a.mu) / cjcb;
            T df[3][3];
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
                    T s = T(0.0);
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb) {
                        const T dd = d[aa][0] * d[bb][0] + d[aa][1] * d[bb][1] + d[aa][2] * d[bb][2];
#pragma unroll
                        for (int m = 0; m < 3; ++m) s += (d[aa][m] * d[bb][n] + (m == n ? dd : T(0.0))) * du[m][bb];
                    }
                    df[n][aa] = -cjcbi * s;
                }
            wave_lds_sync();  // the half-wave is done reading the convective fluxes
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = cv[n][c] + df[n][c];
            wave_lds_sync();
            // strong residual: derivr/derivs/derivt of the total fluxes (rsdl.cpp:683-692)
            T res[TNV] = {T(0.0), T(0.0), T(0.0), T(0.0)};
            if (g < G) tet_strong_residual<D, TNV, T>(sops, sf[hw], g, res);
            // stabilisation (rsdl.cpp:695-767)
            const T t0 = T(a.tau[4 * tt]) * res[0], t1 = T(a.tau[4 * tt]) * res[1], t2 = T(a.tau[4 * tt]) * res[2],
                         t3 = T(a.tau[4 * tt + 3]) * res[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                df[0][c] -= (d[c][0] * (T(2.) * u[0]) + d[c][1] * u[1] + d[c][2] * u[2]) * t0 + d[c][1] * u[0] * t1 +
                            d[c][2] * u[0] * t2 + d[c][0] * t3;
                df[1][c] -= d[c][0] * u[1] * t0 + (d[c][0] * u[0] + d[c][1] * (T(2.) * u[1]) + d[c][2] * u[2]) * t1 +
                            d[c][2] * u[1] * t2 + d[c][1] * t3;
                df[2][c] -= d[c][0] * u[2] * t0 + d[c][1] * u[2] * t1 +
                            (d[c][0] * u[0] + d[c][1] * u[1] + d[c][2] * (T(2.) * u[2])) * t2 + d[c][2] * t3;
                c3[c] = -d[c][0] * t0 - d[c][1] * t1 - d[c][2] * t2;
            }
            wave_lds_sync();  // the half-wave is done reading the total fluxes
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int c = 0; c < 3; ++c) sf[hw][3 * n + c][g] = g < G ? df[n][c] : T(0.0);
#pragma unroll
            for (int c = 0; c < 3; ++c) sf[hw][9 + c][g] = g < G ? c3[c] : T(0.0);
            wave_lds_sync();
            // lf_re = beta * intgrtrst (rsdl.cpp:768-776)
            T lr[TNV];
            tet_intgrtrst<D, TNV, D::FROW, T>(sops, sf[hw], g, lr);
            if (owner)
#pragma unroll
                for (int n = 0; n < TNV; ++n) lfo[(g / 3) * (2 * TNV) + TNV + n] = lr[n] * T(a.beta);
        } else {
            if (owner)
#pragma unroll
                for (int n = 0; n < TNV; ++n) lfo[(g / 3) * (2 * TNV) + TNV + n] = T(0.0);
        }
    }
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tet(TetElemArgs a, int noct) {
    elem_tet_body<double, P, FULL>(a, noct);
}

// FP32 arithmetic (the inner solve of the defect correction); FP64 state and records
template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tet32(TetElemArgs32 a, int noct) {
    elem_tet_body<float, P, FULL>(a, noct);
}

// ---------------------------------------------------------------------------------
// Matrix-core form of the FP64 element residual (p = 2): 16 tetrahedra per wave64,
// every transform a contraction on v_mfma_f64_16x16x4f64 with the layout of the tri
// kernels: columns are the wave's 16 elements, rows Gauss points (27 in two 16-row
// tiles) or modes (10 in one tile), A operands lane-ready from LDS, B operands the
// accumulator registers of the previous contraction (register i of lane l holds row
// 4 i + l / 16 of its tile, which is the B operand of K-step 4 tile + i when the
// operator's K index is laid out that way).  Point fluxes are formed in the lane that
// holds the point, right before the contraction that consumes them, so no field is
// exchanged between lanes and nothing but the operators sits in LDS.
template <int P>
struct TMDims {
    using D = TDims<P>;
    static constexpr int GT = (D::G + 15) / 16;  // Gauss-point tiles
    static constexpr int KT = (D::TM + 3) / 4;   // K-steps over the modes
    static constexpr int QK = 4 * GT;            // K-steps over the Gauss points
    // lane-ready A blocks of 64 doubles: P, Pr, Ps, Pt [op][tile][K-step];
    // Dr, Ds, Dt [op][tile][K-step]; Ir, Is, It [op][K-step]
    static constexpr int OFF_PROJ = 0;
    static constexpr int OFF_DER = 4 * GT * KT * 64;
    static constexpr int OFF_INT = OFF_DER + 3 * GT * QK * 64;
    static constexpr int SIZE = OFF_INT + 3 * QK * 64;
    static_assert(D::TM <= 16, "the modes fit one tile");
};

// tet_hp_ins::element_rsdl, straight branch (tet_hp/ins/rsdl.cpp:15-87, 442-785),
// steady, tau frozen by setup_preconditioner; writes lf like elem_tet_body.  ngrp
// groups of 16 elements; the grid is a multiple of 8 blocks and XCD x walks the
// contiguous group range [x * gpx, (x + 1) * gpx) (neighbouring elements share its L2)
template <typename T, int P, bool FULL>
__device__ __forceinline__ void elem_tet_mfma_body(const TetElemArgsT<T>& a, int ngrp) {
    using D = TDims<P>;
    using M = TMDims<P>;
    using V4 = typename Vec4<T>::type;
    constexpr int GT = M::GT, KT = M::KT, QK = M::QK, TM = D::TM;
    // output row of accumulator register i in lane l: 4 i + l / 16 (f64), 4 (l / 16) + i (f32)
    constexpr bool F64 = sizeof(T) == 8;
    // first Gauss point of K-step (tile kt, register i): its four points are
    // 16 kt + 4 i + l / 16 (f64) or 16 kt + 4 (l / 16) + i (f32)
    auto kfirst = [](int kt, int i) { return F64 ? 16 * kt + 4 * i : 16 * kt + i; };
    __shared__ T sops[M::SIZE];
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
    for (int grp = xcd * gpx + rank; grp < gend; grp += WX) {
        const int slot = grp * 16 + e;
        const bool valid = slot < a.ntet;
        const int tt = valid ? (a.elist ? a.elist[slot] : slot) : 0;
        // modal coefficients: K-step s of lane l is mode 4 s + l / 16 (vertices 0-3,
        // then edge mode 0 of edges 0-5, sign^0 = 1); padding elements are zero
        T b[TNV][KT];
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            const int k = 4 * s + q;
            const double* src = nullptr;
            if (valid && k < TM)
                src = k < 4 ? a.ugv + (size_t)a.tet_pnt[4 * tt + k] * TNV
                            : a.uge + (size_t)a.tet_seg[6 * tt + (k - 4)] * a.em0 * TNV;
#pragma unroll
            for (int n = 0; n < TNV; ++n) b[n][s] = src ? T(src[n]) : T(0);
        }
        // coordinate derivatives (rsdl.cpp:58-62) and cofactors (:449-458); padding
        // elements take the reference tetrahedron
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
        const T rho = T(a.rho);
        // values at the Gauss points (proj)
        V4 U[TNV][GT];
#pragma unroll
        for (int mt = 0; mt < GT; ++mt) {
            V4 c0 = z, c1 = z, c2 = z, c3 = z;
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const T av = sops[M::OFF_PROJ + (mt * KT + s) * 64 + lane];
                c0 = mfma(av, b[0][s], c0); c1 = mfma(av, b[1][s], c1);
                c2 = mfma(av, b[2][s], c2); c3 = mfma(av, b[3][s], c3);
            }
            U[0][mt] = c0; U[1][mt] = c1; U[2][mt] = c2; U[3][mt] = c3;
        }
        // convective and pressure fluxes at point (tile kt, register i) (rsdl.cpp:461-508)
        auto conv = [&](int kt, int i, T cv[3][3], T c3[3]) {
            const T u0 = U[0][kt][i], u1 = U[1][kt][i], u2 = U[2][kt][i], pr = U[3][kt][i];
            const T fx = rho * u0, fy = rho * u1, fz = rho * u2;
#pragma unroll
            for (int c = 0; c < 3; ++c) c3[c] = d[c][0] * fx + d[c][1] * fy + d[c][2] * fz;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cv[0][c] = u0 * c3[c] + d[c][0] * pr;
                cv[1][c] = u1 * c3[c] + d[c][1] * pr;
                cv[2][c] = u2 * c3[c] + d[c][2] * pr;
            }
        };
        // lf_im = intgrtrst of the fluxes (rsdl.cpp:509-511)
        {
            V4 L0 = z, L1 = z, L2 = z, L3 = z;
#pragma unroll
            for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T cv[3][3], c3[3];
                    conv(kt, i, cv, c3);
#pragma unroll
                    for (int dd = 0; dd < 3; ++dd) {
                        const T ad = sops[M::OFF_INT + (dd * QK + kk) * 64 + lane];
                        L0 = mfma(ad, cv[0][dd], L0); L1 = mfma(ad, cv[1][dd], L1);
                        L2 = mfma(ad, cv[2][dd], L2); L3 = mfma(ad, c3[dd], L3);
                    }
                }
            // register i holds mode 4 i + l / 16 (f64) or 4 (l / 16) + i (f32) of element e
            double* lfo = a.lf + (size_t)tt * (2 * TNV * TM);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = F64 ? 4 * i + q : 4 * q + i;
                if (valid && m < TM) {
                    double* o = lfo + m * (2 * TNV);
                    o[0] = double(L0[i]); o[1] = double(L1[i]); o[2] = double(L2[i]); o[3] = double(L3[i]);
                    if constexpr (!FULL) { o[4] = 0.0; o[5] = 0.0; o[6] = 0.0; o[7] = 0.0; }
                }
            }
        }
        if constexpr (FULL) {
            // viscous fluxes (rsdl.cpp:518-674): visc(n,m)(a,b) = -mu/cjcb (delta_nm
            // sum_c d(a)(c) d(b)(c) + d(a)(m) d(b)(n)) applied to du(m,b), summed as
            // -mu/cjcb (sum_b dd(a,b) du(n,b) + sum_m d(a)(m) sum_b d(b)(n) du(m,b))
            const T cjcbi = T(a.mu) / cjcb;
            T dd[3][3];
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) dd[x][y] = d[x][0] * d[y][0] + d[x][1] * d[y][1] + d[x][2] * d[y][2];
            // one Gauss-point tile at a time: the velocity derivatives du(m, b) of its
            // points (proj with r, s, t derivatives), the total fluxes of each point into
            // the strong residual (derivr/derivs/derivt, rsdl.cpp:683-692) and the
            // viscous fluxes into lf_re (intgrtrst, :768-776; the stabilisation part
            // follows once the strong residual is complete)
            V4 R[TNV][GT];
#pragma unroll
            for (int n = 0; n < TNV; ++n)
#pragma unroll
                for (int mt = 0; mt < GT; ++mt) R[n][mt] = z;
            V4 L0 = z, L1 = z, L2 = z, L3 = z;
#pragma unroll
            for (int kt = 0; kt < GT; ++kt) {
                V4 DU[3][3];
#pragma unroll
                for (int bb = 0; bb < 3; ++bb) {
                    V4 c0 = z, c1 = z, c2 = z;
#pragma unroll
                    for (int s = 0; s < KT; ++s) {
                        const T av = sops[M::OFF_PROJ + (((1 + bb) * GT + kt) * KT + s) * 64 + lane];
                        c0 = mfma(av, b[0][s], c0); c1 = mfma(av, b[1][s], c1); c2 = mfma(av, b[2][s], c2);
                    }
                    DU[0][bb] = c0; DU[1][bb] = c1; DU[2][bb] = c2;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T cv[3][3], c3[3], df[3][3], gm[3][3];
                    conv(kt, i, cv, c3);
#pragma unroll
                    for (int m = 0; m < 3; ++m)
#pragma unroll
                        for (int n = 0; n < 3; ++n)
                            gm[m][n] = d[0][n] * DU[m][0][i] + d[1][n] * DU[m][1][i] + d[2][n] * DU[m][2][i];
#pragma unroll
                    for (int n = 0; n < 3; ++n)
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const T h = dd[x][0] * DU[n][0][i] + dd[x][1] * DU[n][1][i] + dd[x][2] * DU[n][2][i];
                            df[n][x] = -cjcbi * (h + d[x][0] * gm[0][n] + d[x][1] * gm[1][n] + d[x][2] * gm[2][n]);
                        }
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const T ai = sops[M::OFF_INT + (x * QK + kk) * 64 + lane];
                        L0 = mfma(ai, df[0][x], L0); L1 = mfma(ai, df[1][x], L1); L2 = mfma(ai, df[2][x], L2);
                    }
#pragma unroll
                    for (int n = 0; n < 3; ++n)
#pragma unroll
                        for (int c = 0; c < 3; ++c) cv[n][c] += df[n][c];
#pragma unroll
                    for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const T ad = sops[M::OFF_DER + ((x * GT + mt) * QK + kk) * 64 + lane];
                            R[0][mt] = mfma(ad, cv[0][x], R[0][mt]); R[1][mt] = mfma(ad, cv[1][x], R[1][mt]);
                            R[2][mt] = mfma(ad, cv[2][x], R[2][mt]); R[3][mt] = mfma(ad, c3[x], R[3][mt]);
                        }
                }
            }
            // stabilisation (rsdl.cpp:695-767) into lf_re
            const T tm = T(a.tau[4 * tt]), tc = T(a.tau[4 * tt + 3]);
#pragma unroll
            for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = 4 * kt + i;
                    if (kfirst(kt, i) >= D::G) continue;  // K-step of padding points only
                    __builtin_amdgcn_sched_barrier(0);  // keep the K-steps' operand loads apart
                    T sf[3][3], c3[3];
                    const T u0 = U[0][kt][i], u1 = U[1][kt][i], u2 = U[2][kt][i];
                    const T t0 = tm * R[0][kt][i], t1 = tm * R[1][kt][i], t2 = tm * R[2][kt][i],
                                 t3 = tc * R[3][kt][i];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        sf[0][c] = -((d[c][0] * (T(2) * u0) + d[c][1] * u1 + d[c][2] * u2) * t0 + d[c][1] * u0 * t1 +
                                     d[c][2] * u0 * t2 + d[c][0] * t3);
                        sf[1][c] = -(d[c][0] * u1 * t0 + (d[c][0] * u0 + d[c][1] * (T(2) * u1) + d[c][2] * u2) * t1 +
                                     d[c][2] * u1 * t2 + d[c][1] * t3);
                        sf[2][c] = -(d[c][0] * u2 * t0 + d[c][1] * u2 * t1 +
                                     (d[c][0] * u0 + d[c][1] * u1 + d[c][2] * (T(2) * u2)) * t2 + d[c][2] * t3);
                        c3[c] = -d[c][0] * t0 - d[c][1] * t1 - d[c][2] * t2;
                    }
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const T ad = sops[M::OFF_INT + (x * QK + kk) * 64 + lane];
                        L0 = mfma(ad, sf[0][x], L0); L1 = mfma(ad, sf[1][x], L1);
                        L2 = mfma(ad, sf[2][x], L2); L3 = mfma(ad, c3[x], L3);
                    }
                }
            double* lfo = a.lf + (size_t)tt * (2 * TNV * TM);
            const T bt = T(a.beta);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = F64 ? 4 * i + q : 4 * q + i;
                if (valid && m < TM) {
                    double* o = lfo + m * (2 * TNV) + TNV;
                    o[0] = double(L0[i] * bt); o[1] = double(L1[i] * bt); o[2] = double(L2[i] * bt); o[3] = double(L3[i] * bt);
                }
            }
        }
    }
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_tet_mfma(TetElemArgs a, int ngrp) {
    elem_tet_mfma_body<double, P, FULL>(a, ngrp);
}

// FP32 arithmetic (the inner solve of the defect correction) on v_mfma_f32_16x16x4f32
template <int P, bool FULL>
__global__ __launch_bounds__(256, 2) void k_elem_tet32_mfma(TetElemArgs32 a, int ngrp) {
    elem_tet_mfma_body<float, P, FULL>(a, ngrp);
}

// inflow face flux, continuity only (tet_hp/ins/bdry.cpp:59-102, bdry_ins.h:122-134):
// flf[f][m] = integral over the face of phi_m rho u.n, n = dx/dr x dx/ds, which for a
// straight face is (p2 - p1) x (p0 - p1) / 4 at every Gauss point
template <int P>
__global__ __launch_bounds__(256) void k_face_tet(int nbf, const int* __restrict__ bf_pnt, const int* __restrict__ bf_seg,
                                                   const double* __restrict__ pnts, const double* __restrict__ ugv,
                                                   const double* __restrict__ uge, int em0, const double* __restrict__ P2,
                                                   const double* __restrict__ W2, double rho, double* __restrict__ flf) {
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
    double uh[3][6];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
#pragma unroll
        for (int j = 0; j < 3; ++j) uh[n][j] = ugv[pv[j] * TNV + n];
#pragma unroll
        for (int j = 0; j < 3 * D::EM; ++j) uh[n][3 + j] = uge[(se[j] * em0) * TNV + n];
    }
    double lf[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < D::GF; ++q) {
        double un[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < D::TM2; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) un[n] += P2[q * D::TM2 + m] * uh[n][m];
        const double flux = rho * (un[0] * nrm[0] + un[1] * nrm[1] + un[2] * nrm[2]) * W2[q];
#pragma unroll
        for (int m = 0; m < D::TM2; ++m) lf[m] += P2[q * D::TM2 + m] * flux;
    }
#pragma unroll
    for (int m = 0; m < D::TM2; ++m) flf[f * D::TM2 + m] = lf[m];
}

// assembly of one entity kind (vertices: mode = local vertex; edges: mode = 4 + local
// edge): res_r = omb res_r + sum lf_re; res = face fluxes + sum lf_im + res_r
// (tet_hp/update.cpp:13-31, 45-70, bdry.cpp:16-57); elements and faces ascending
struct TetGatherArgs {
    int nent, tm, tm2, mbase, fbase, stride;
    int fnv;          // face variables: 1 (continuity into the last variable) or NV, flf[f][fnv][tm2]
    const int* ptr;   // CSR entity -> (element, local index) codes 16*t + i
    const int* code;
    const int* fptr;  // CSR entity -> (face, local index) codes 8*f + j
    const int* fcode;
    const double* lf;
    const double* flf;
    double* res;
    double* rr;
    double omb;
};

template <int NV>
__global__ __launch_bounds__(256) void k_gather_tet(TetGatherArgs a) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= a.nent) return;
    double si[NV], sr[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) si[n] = 0.0;
    for (int k = a.fptr[v]; k < a.fptr[v + 1]; ++k) {
        const int c = a.fcode[k], f = c >> 3, j = c & 7;
        if (a.fnv == 1) {
            si[NV - 1] += a.flf[f * a.tm2 + a.fbase + j];
        } else {
#pragma unroll
            for (int n = 0; n < NV; ++n) si[n] += a.flf[(f * NV + n) * a.tm2 + a.fbase + j];
        }
    }
    double* R = a.res + (size_t)v * a.stride;
    double* RR = a.rr + (size_t)v * a.stride;
#pragma unroll
    for (int n = 0; n < NV; ++n) sr[n] = a.omb * RR[n];
    for (int k = a.ptr[v]; k < a.ptr[v + 1]; ++k) {
        const int c = a.code[k], t = c >> 4, i = c & 15;
        // one contiguous record of 2 NV doubles per (element, mode)
        const double* L = a.lf + (size_t)t * (2 * NV * a.tm) + (size_t)(a.mbase + i) * (2 * NV);
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            si[n] += L[n];
            sr[n] += L[NV + n];
        }
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        RR[n] = sr[n];
        R[n] = si[n] + sr[n];
    }
}

// setup_preconditioner element part (tet_hp/ins/setup_preconditioner.cpp:37-118)
struct TetSetupArgs {
    const double* pnts;
    const int* tet_pnt;
    const double* ugv;
    int ntet, p;
    double rho, mu, adis;
    double* iprcn;
    double* tau;
    int* err;
};

__global__ __launch_bounds__(256) void k_setup_tet(TetSetupArgs a) {
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
    // vol = |(v3-v2) x (v1-v2) . (v0-v2)| (tet_mesh/tet_prims.cpp:180-191), jcb = vol/8
    double d1[3], d2[3], d3[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) { d1[n] = X[3][n] - X[2][n]; d2[n] = X[1][n] - X[2][n]; d3[n] = X[0][n] - X[2][n]; }
    const double cx = d1[1] * d2[2] - d1[2] * d2[1], cy = d1[2] * d2[0] - d1[0] * d2[2], cz = d1[0] * d2[1] - d1[1] * d2[0];
    double jcb = 0.125 * fabs(cx * d3[0] + cy * d3[1] + cz * d3[2]);
    // largest / smallest face area, faces ordered as tet_mesh/setup.cpp:305-308
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
        const double ar = .5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        amax = ar > amax ? ar : amax;
        amin = ar < amin ? ar : amin;
    }
    const double pp = (a.p + 1) * (a.p + 1);
    const double h = 4.0 * jcb / (0.25 * pp * amax);
    const double hmax = 4.0 * jcb / (0.25 * pp * amin);
    double qmax = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double u = a.ugv[v[k] * TNV], w = a.ugv[v[k] * TNV + 1], z = a.ugv[v[k] * TNV + 2];
        const double q = u * u + w * w + z * z;
        qmax = (qmax > q) ? qmax : q;  // MAX of sys/param.h: a NaN q propagates
    }
    if (!(jcb > 0.0) || !(qmax >= 0.0)) atomicOr(a.err, 1);
    const double nu = a.mu / a.rho;
    double gam = 3.0 * qmax + (2. * nu / hmax) * (2. * nu / hmax);
    if (a.mu == 0.0) gam = fmax(gam, 0.1);
    const double q = sqrt(qmax);
    const double lam1 = q + sqrt(qmax + gam);
    a.tau[4 * t] = a.adis * h / (jcb * sqrt(gam));
    a.tau[4 * t + 3] = qmax * a.tau[4 * t];
    jcb *= 2. * nu * (1. / (h * h) + 1. / (h * h) + 1. / (h * h)) + 3 * lam1 / h;
    a.iprcn[4 * t + 0] = a.rho * jcb;
    a.iprcn[4 * t + 1] = a.rho * jcb;
    a.iprcn[4 * t + 2] = a.rho * jcb;
    a.iprcn[4 * t + 3] = jcb / gam;
}

// prcn(entity) = 1 / sum of the incident elements' iprcn, ascending elements
// (tet_hp/ins/setup_preconditioner.cpp:119-126, tet_hp/setup_preconditioner.cpp:455-476)
__global__ __launch_bounds__(256) void k_setup_gather_tet(int nent, const int* __restrict__ ptr,
                                                          const int* __restrict__ code, const double* __restrict__ iprcn,
                                                          double* __restrict__ prcn) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= nent) return;
    double s[TNV] = {0.0, 0.0, 0.0, 0.0};
    for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
        const int t = code[k] >> 4;
#pragma unroll
        for (int n = 0; n < TNV; ++n) s[n] += iprcn[4 * t + n];
    }
#pragma unroll
    for (int n = 0; n < TNV; ++n) prcn[v * TNV + n] = 1.0 / s[n];
}

// minvrt for p <= 2 (tet_hp/setup_preconditioner.cpp:11-210), diagonal preconditioner.
// pre_dirichlet: the update's vdirichlet/edirichlet before minvrt (update.cpp:218-223)
struct TetMinvArgs {
    int npnt, nseg, em0, em, pre_dirichlet;
    const int* ve_ptr;   // vertex -> (edge, end) codes 2*e + end, ascending edges
    const int* ve_code;
    const int* et_ptr;   // edge -> (element, local edge) codes 16*t + i, ascending elements
    const int* et_code;
    const int* tet_pnt;
    const unsigned char* vess;
    const unsigned char* eess;
    double* resv;
    double* rese;
    const double* vprcn;
    const double* eprcn;
    const double* iprcn;
    double vdiag, ediag, sfmv0, sfmv1;
    const double* vfms;  // 4 x 10 vertex rows of the mass matrix
    // fused stage update ug = ug0 - cfa res when ugv != null
    double* ugv;
    double* uge;
    const double* ug0v;
    const double* ug0e;
    double cfa;
};

__global__ __launch_bounds__(256) void k_minv_vrtx_tet(TetMinvArgs a) {
    const int v = xcd_block() * blockDim.x + threadIdx.x;
    if (v >= a.npnt) return;
    double r[TNV];
    const bool vd = a.vess[v] != 0;
#pragma unroll
    for (int n = 0; n < TNV; ++n) r[n] = (a.pre_dirichlet && vd && n < 3) ? 0.0 : a.resv[v * TNV + n];
    if (a.em > 0)
        for (int k = a.ve_ptr[v]; k < a.ve_ptr[v + 1]; ++k) {
            const int c = a.ve_code[k], e = c >> 1;
            const double sf = (c & 1) ? a.sfmv1 : a.sfmv0;
            const bool ed = a.pre_dirichlet && a.eess[e];
#pragma unroll
            for (int n = 0; n < TNV; ++n) r[n] -= sf * ((ed && n < 3) ? 0.0 : a.rese[(e * a.em0) * TNV + n]);
        }
    double x[TNV], u0[TNV];
#pragma unroll
    for (int n = 0; n < TNV; ++n) {
        x[n] = (vd && n < 3) ? 0.0 : r[n] * (a.vprcn[v * TNV + n] * a.vdiag);
        u0[n] = a.ugv ? a.ug0v[v * TNV + n] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < TNV; ++n) {
        a.resv[v * TNV + n] = x[n];
        if (a.ugv) a.ugv[v * TNV + n] = u0[n] - a.cfa * x[n];
    }
}

__global__ __launch_bounds__(256) void k_minv_edge_tet(TetMinvArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.nseg) return;
    const bool ed = a.eess[e] != 0;
    double r[TNV];
#pragma unroll
    for (int n = 0; n < TNV; ++n) r[n] = (a.pre_dirichlet && ed && n < 3) ? 0.0 : a.rese[(e * a.em0) * TNV + n];
    for (int k = a.et_ptr[e]; k < a.et_ptr[e + 1]; ++k) {
        const int c = a.et_code[k], t = c >> 4, i = c & 15;
        double ip[TNV];
#pragma unroll
        for (int n = 0; n < TNV; ++n) ip[n] = a.iprcn[4 * t + n];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = a.tet_pnt[4 * t + j];
            const double cf = a.vfms[j * 10 + 4 + i];
#pragma unroll
            for (int n = 0; n < TNV; ++n) r[n] -= cf * (a.resv[w * TNV + n] * ip[n]);
        }
    }
    double x[TNV], u0[TNV];
#pragma unroll
    for (int n = 0; n < TNV; ++n) {
        x[n] = (ed && n < 3) ? 0.0 : r[n] * (a.eprcn[e * TNV + n] * a.ediag);
        u0[n] = a.uge ? a.ug0e[(e * a.em0) * TNV + n] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < TNV; ++n) {
        a.rese[(e * a.em0) * TNV + n] = x[n];
        if (a.uge) a.uge[(e * a.em0) * TNV + n] = u0[n] - a.cfa * x[n];
    }
}

}  // namespace hpg
