// Forward-difference element Jacobian of tet_hp::element_jacobian
// (tet_hp/update.cpp:134-178) for a batch of tetrahedra, p <= 2.
//
// The element kernels write element-major records lf[t][m][im/re][n] for every
// tetrahedron, so one element launch yields lf_re + lf_im of each batch element
// directly.  The batch is split into sets that share no vertex (at p <= 2 they then
// share no edge either): for one column (mode, var) k_tjac_perturb adds dw to each
// set element's own coefficient, the element kernel runs, and k_tjac_collect forms
// (R - Rbar) / dw and restores the coefficient from its saved value.  As in the
// reference, dw is one value for all variables: eps_r times the sum of |uht(n)(i)|
// over the four vertices and every variable, plus eps_a.
#include <hip/hip_runtime.h>

namespace hpg {

struct TetJacArgs {
    const int* jel;    // batch elements grouped by set
    const int* jpos;   // position of jel[j] in the caller's batch
    int begin, end;    // this set's range of jel
    const int *tet_pnt, *tet_seg;
    int em0, tm, nv;   // edge-mode stride, modes and variables of the level
    double *ugv, *uge;
    const double* lf;  // element records [t][m][0 = im, 1 = re][n]
    double* dw;        // batch
    double* save;      // batch
    double* base;      // batch x tm x nv: Rbar
    double* K;         // batch x N x N, N = tm * nv
    int mode, var;     // the column; mode < 0: the unperturbed evaluation
    double eps_r, eps_a;
};

// address of coefficient (mode, var) of tetrahedron t: vertex modes 0-3, then edge
// mode 0 of the six edges (sign sgn^0 = 1)
__device__ __forceinline__ double* tjac_coef(const TetJacArgs& a, int t, int mode, int var) {
    if (mode < 4) return a.ugv + (size_t)a.tet_pnt[4 * t + mode] * a.nv + var;
    return a.uge + ((size_t)a.tet_seg[6 * t + (mode - 4)] * a.em0) * a.nv + var;
}

__global__ void k_tjac_dw(TetJacArgs a) {
    const int j = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.end) return;
    const int t = a.jel[j];
    double s = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int n = 0; n < a.nv; ++n) s = s + fabs(a.ugv[(size_t)a.tet_pnt[4 * t + i] * a.nv + n]);
    s = s * a.eps_r;
    a.dw[a.jpos[j]] = s + a.eps_a;
}

__global__ void k_tjac_perturb(TetJacArgs a) {
    const int j = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.end) return;
    double* u = tjac_coef(a, a.jel[j], a.mode, a.var);
    const int jp = a.jpos[j];
    a.save[jp] = *u;
    *u = *u + a.dw[jp];
}

// one lane per (set element, mode i, variable n)
__global__ void k_tjac_collect(TetJacArgs a) {
    const int N = a.tm * a.nv;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)(a.end - a.begin) * N) return;
    const int j = a.begin + (int)(g / N), r = (int)(g % N);
    const int i = r / a.nv, n = r - i * a.nv;
    const int t = a.jel[j], jp = a.jpos[j];
    const double* rec = a.lf + ((size_t)t * a.tm + i) * (2 * a.nv);
    const double R = rec[a.nv + n] + rec[n];
    double* rb = a.base + (size_t)jp * N + r;
    if (a.mode < 0) {
        *rb = R;
        return;
    }
    const int col = a.mode * a.nv + a.var;
    a.K[((size_t)jp * N + r) * N + col] = (R - *rb) / a.dw[jp];
    if (i == a.mode && n == a.var) *tjac_coef(a, t, a.mode, a.var) = a.save[jp];
}

}  // namespace hpg
