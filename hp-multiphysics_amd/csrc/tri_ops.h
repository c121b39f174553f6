// Host-side construction of the triangle basis operators used by the device kernels.
//
// The reference evaluates its modal basis through sum-factorised loops
// (tri_basis/proj.cpp:22-199, intgrt.cpp:22-158, proj.cpp:339-402).  Here every
// transform is built once as a dense matrix over the (x, eta) Gauss grid, derived
// directly from the basis definition (tri_basis/ptvalues.cpp:84-195 with the
// normalisation of tri_basis/initialize.cpp:362-425):
//   P  [g][m]  value of mode m at Gauss point g           (proj)
//   Pr [g][m]  d/dr, Ps[g][m] d/ds                         (proj with derivatives)
//   I  [m][g]  w_g phi_m(g)                                 (intgrt)
//   Ir [m][g] = -w_g dphi_m/dr, Is = -w_g dphi_m/ds        (intgrtrs, sign folded in)
//   Dr, Ds [g][g']  Lagrange-interpolant derivatives on the Gauss grid (derivr/derivs)
// Gauss points g = i*gpn + j (i along xi, j along eta), the reference's f(i,j) order.
// Gauss nodes come from Newton iteration on the monic Jacobi recurrence and the
// weights from the Christoffel sums, an independent route from the reference's
// Golub-Welsch eigen-solve (utilities/myblas.c:766-).  The lumped-mass factors
// follow the construction of tri_basis/initialize.cpp:590-958.
#pragma once
#include <vector>

namespace hpg {

struct TriOps {
    int p, sm, im, tm, bm, gpx, gpn, G;
    std::vector<double> xg, wx;      // Gauss-Legendre nodes/weights in xi (gpx)
    std::vector<double> eg, we;      // Gauss-Jacobi(1,0) nodes in eta, weights halved (gpn)
    std::vector<double> P, Pr, Ps;   // G x tm
    std::vector<double> I, Ir, Is;   // tm x G
    std::vector<double> Dr, Ds;      // G x G
    std::vector<double> g1, dg1, w1; // edge basis: gpx x (sm+2) values / d/dxi, weights
    // upper band (bandwidth 2, row j = column j of U: U(j-2,j), U(j-1,j), U(j,j)) Cholesky
    // factor of the side-mode block of the edge mass matrix (tri_basis/initialize.cpp:1097-1130)
    std::vector<double> sdiag1d;
    // lumped-mass inversion factors, layouts as in tri_basis.h:124-141
    double vdiag;
    std::vector<double> sfmv, ifmb, vfms, sdiag, sfms, bfmi, idiag;
    std::vector<double> mass;        // tm x tm
    explicit TriOps(int p);
};

// monic Jacobi recurrence p_{k+1} = (x - a_k) p_k - b_k p_{k-1}, b_0 = total mass
void monic_jacobi(int n, double al, double be, double* a, double* b);
void gauss_nodes(int n, double al, double be, double* x, double* w);

}  // namespace hpg
