// Device-side data structures shared by the tri_hp kernels and the host context.
#pragma once

namespace hpg {

constexpr int NV = 3;  // tri_hp_ins unknowns (u, v, p)

template <int P>
struct Dims {
    static constexpr int TM = (P + 1) * (P + 2) / 2;
    static constexpr int SM = P - 1;
    static constexpr int IM = (P - 1) * (P - 2) / 2;
    static constexpr int BM = 3 * P;
    static constexpr int G = (P + 1) * (P + 1);
    static constexpr int GT = (G + 15) / 16;  // 16-row tiles over Gauss points
    static constexpr int KT = (TM + 3) / 4;   // K-steps of 4 over modes
    // lane-ready operator pack (doubles): P, Pr, Ps | Dr, Ds | Ir, Is
    static constexpr int PROJ = GT * KT * 64;
    static constexpr int DER = GT * GT * 4 * 64;
    static constexpr int INT = GT * 4 * 64;
    static constexpr int OFF_P = 0, OFF_PR = PROJ, OFF_PS = 2 * PROJ;
    static constexpr int OFF_DR = 3 * PROJ, OFF_DS = 3 * PROJ + DER;
    static constexpr int OFF_IR = 3 * PROJ + 2 * DER, OFF_IS = OFF_IR + INT;
    static constexpr int PACK = OFF_IS + INT;
};

// Arguments of the element residual kernel: T = double (the FP64 residual) or float
// (the FP32 residual of the defect-correction inner solve)
template <typename T>
struct ElemArgsT {
    const double* pnts;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const int* order;
    int begin, end;
    const T* ugv;
    const T* ugs;
    const T* ugi;
    T* resv;  // imaginary part (lf_im) target
    T* ress;
    T* resi;
    T* rrv;   // real part (lf_re) target
    T* rrs;
    T* rri;
    const T* ops;
    int sm0, im0;
    T rho, mu, adis, beta;
    // first colour that touches each vertex / side: that colour initialises the
    // coefficient from the stage's carried real part instead of accumulating
    const unsigned char* fcv;
    const unsigned char* fcs;
    int colour;
    T keep;  // 1 - beta(stage) for stage < nstage, 0 for the full residual
    // optional (FP64): per-Gauss-point linearisation data (u, v, strong residual) in
    // FP32, stored per 16-element group in the accumulator layout; group = gbase + wave task
    float* lin;
    int gbase;
    // optional (FP32): constant source sfac[0] * src added on the first visit of every
    // coefficient (the scaled FP64 residual of the defect correction)
    const double* srcv;
    const double* srcs;
    const double* srci;
    const double* sfac;
    // curved elements (the curved launch only): per entry of order[begin:end), the
    // side-mode coefficients of the element's coordinates, (3 sides, sm0 modes, x/y)
    // with the side sign applied, zero on straight sides
    const double* crv;
};
typedef ElemArgsT<double> ElemArgs;
typedef ElemArgsT<float> ElemArgs32;

// FP32 linearised residual J(u0) * delta of one colour (mixed-precision inner solve)
struct LinArgs {
    const double* pnts;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const int* order;
    int begin, end;
    const float* dv;   // delta coefficients
    const float* ds;
    const float* di;
    float* resv;       // residual (imaginary + carried real part + source)
    float* ress;
    float* resi;
    float* rrv;        // carried real part
    float* rrs;
    float* rri;
    const double* srcv;  // constant source added once (the FP64 residual at u0), may be null
    const double* srcs;
    const double* srci;
    const float* ops;  // FP32 operator pack
    const float* lin;  // linearisation data written by the FP64 kernel
    int gbase;
    int sm0, im0;
    float rho, mu, adis, beta, keep;
    const unsigned char* fcv;
    const unsigned char* fcs;
    int colour;
    // probing of the element Jacobian diagonal: element-local unit vector on
    // (mode probe_m, variable probe_n); |J_e| diagonal entry accumulated into diag
    int probe_m, probe_n;
    double* diag;
};

}  // namespace hpg
