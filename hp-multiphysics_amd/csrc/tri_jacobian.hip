// Forward-difference element Jacobian of tri_hp::element_jacobian
// (tri_hp/nstage.cpp:131-165, fixed-mesh branch) for a batch of elements.
//
// The batch is split by element colour.  Elements of one colour share no vertex or
// side, so one column (mode, var) of every element of the colour is formed at once:
// k_jac_perturb adds dw(var) to the element's own coefficient of that mode in the
// global state (no other element of the colour reads it), the residual kernel of the
// physics runs over the colour's elements into zeroed scratch residual arrays (each
// scratch entry then holds exactly one element's lf_re + lf_im, lftog sign
// included), and k_jac_collect turns the scratch into the column
// (R - Rbar) / dw(var), clears the scratch it read and restores the coefficient
// from its saved value (bit-exact, unlike subtracting dw again).
#include <hip/hip_runtime.h>
#include "tri_kernels.h"

namespace hpg {

struct JacArgs {
    const int* jel;    // batch elements grouped by colour
    const int* jpos;   // position of jel[j] in the caller's batch (row of dw/base/K)
    int begin, end;    // this colour's range of jel
    const int *tri_pnt, *tri_seg, *tri_sgn;
    int sm0, im0;
    double *ugv, *ugs, *ugi;
    double *jrv, *jrs, *jri;   // scratch res of the colour
    double *jqv, *jqs, *jqi;   // scratch res_r of the colour
    double* dw;     // batch x NV
    double* save;   // batch: the perturbed coefficient's value
    double* base;   // batch x TM x NV: Rbar
    double* K;      // batch x N x N, N = TM * NV, K(row = i*NV + n, col = mode*NV + var)
    int mode, var;  // the column; mode < 0: the unperturbed evaluation Rbar
    double eps_r, eps_a;
};

struct JacElem {
    int t, jp;
    int vtx[3], seg[3], sgn[3];
};

__device__ __forceinline__ JacElem jac_elem(const JacArgs& a, int j) {
    JacElem e;
    e.t = a.jel[j];
    e.jp = a.jpos[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e.vtx[k] = a.tri_pnt[3 * e.t + k];
        e.seg[k] = a.tri_seg[3 * e.t + k];
        e.sgn[k] = a.tri_sgn[3 * e.t + k];
    }
    return e;
}

// dw(n) = eps_r * sum over the vertex modes of |uht(n)(i)| + eps_a (nstage.cpp:137-143)
template <int NVn>
__global__ void k_jac_dw(JacArgs a) {
    const int j = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.end) return;
    const JacElem e = jac_elem(a, j);
#pragma unroll
    for (int n = 0; n < NVn; ++n) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) s = s + fabs(a.ugv[e.vtx[i] * NVn + n]);
        a.dw[(size_t)e.jp * NVn + n] = s * a.eps_r + a.eps_a;
    }
}

// uht(var)(mode) += dw(var) (nstage.cpp:155): the global coefficient is sgn^m uht
template <int P, int NVn>
__global__ void k_jac_perturb(JacArgs a) {
    const int j = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.end) return;
    const JacElem e = jac_elem(a, j);
    int kind;
    double sg;
    const int off = mode_offset<P, NVn>(a.mode, e.t, e.vtx, e.seg, e.sgn, a.sm0, a.im0, kind, sg) + a.var;
    double* u = kind == 0 ? a.ugv : (kind == 1 ? a.ugs : a.ugi);
    const double x = u[off];
    a.save[e.jp] = x;
    u[off] = sg * (sg * x + a.dw[(size_t)e.jp * NVn + a.var]);
}

// one lane per (element, mode i, variable n) of the colour
template <int P, int NVn>
__global__ void k_jac_collect(JacArgs a) {
    constexpr int TM = Dims<P>::TM, N = TM * NVn;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nl = (long)(a.end - a.begin) * N;
    if (g >= nl) return;
    const int j = a.begin + (int)(g / N), r = (int)(g % N);
    const int i = r / NVn, n = r - i * NVn;
    const JacElem e = jac_elem(a, j);
    int kind;
    double sg;
    const int off = mode_offset<P, NVn>(i, e.t, e.vtx, e.seg, e.sgn, a.sm0, a.im0, kind, sg) + n;
    double* rv = kind == 0 ? a.jrv : (kind == 1 ? a.jrs : a.jri);
    double* qv = kind == 0 ? a.jqv : (kind == 1 ? a.jqs : a.jqi);
    const double R = sg * rv[off];
    rv[off] = 0.0;
    qv[off] = 0.0;
    double* rb = a.base + ((size_t)e.jp * TM + i) * NVn + n;
    if (a.mode < 0) {
        *rb = R;
        return;
    }
    const int col = a.mode * NVn + a.var;
    a.K[((size_t)e.jp * N + r) * N + col] = (R - *rb) / a.dw[(size_t)e.jp * NVn + a.var];
    if (i == a.mode && n == a.var) {
        double* u = kind == 0 ? a.ugv : (kind == 1 ? a.ugs : a.ugi);
        u[off] = a.save[e.jp];
    }
}

}  // namespace hpg
