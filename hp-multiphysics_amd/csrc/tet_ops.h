// Host-side construction of the tetrahedral basis operators used by the device
// kernels, p in {1, 2} (vertex and edge modes).
//
// The reference evaluates tet_basis through sum-factorised loops over the collapsed
// coordinates (tet_basis/proj.cpp:22-264, intgrt.cpp:23-609, proj.cpp:781-911).
// Here the transforms are dense matrices over the Gauss grid (i along x, j along y,
// k along z; g = (i*gp + j)*gp + k), built from closed forms of the modes: the
// vertex modes are the barycentric coordinates of the reference tet (vertex 2 at
// r=s=t=-1, 3 at r=1, 1 at s=1, 0 at t=1) and the p=2 edge mode of edge i is the
// product of the barycentrics of its end vertices (2,3), (3,1), (2,1), (1,0), (2,0),
// (3,0) (tet_basis/ptvalues.cpp:154-336 with em = 1):
//   P [g][m], Pr/Ps/Pt [g][m]   values and r, s, t derivatives          (proj)
//   Ir/Is/It [m][g] = -w_g dphi_m/d(r,s,t)                               (intgrtrst)
//   Dr/Ds/Dt [g][g']  derivatives of the collapsed-coordinate Lagrange interpolant
//                     (derivr/derivs/derivt)
// Face 0 (tet_basis/proj2d.cpp, intgrt2d.cpp): modes V1, V2, V3, E(V2,V3), E(V1,V3),
// E(V1,V2) on the (x, y) face grid, P2 [gf][m] and W2 [gf].
// Lumped-mass factors as tet_basis/initialize.cpp:521-800 and the p = 2 edge
// scaling of tet_hp/setup_preconditioner.cpp:199-200.
#pragma once
#include <vector>

namespace hpg {

struct TetOps {
    int p, em, tm, gp, G, tm2d, GF;
    std::vector<double> P, Pr, Ps, Pt;   // G x tm
    std::vector<double> Ir, Is, It;      // tm x G
    std::vector<double> Dr, Ds, Dt;      // G x G
    std::vector<double> P2, W2;          // GF x tm2d, GF
    std::vector<double> mass;            // tm x tm
    double vdiag;
    std::vector<double> sfmv;            // 2 x em
    std::vector<double> vfms;            // 4 x tm
    std::vector<double> ediag;           // em
    explicit TetOps(int p);
};

}  // namespace hpg
