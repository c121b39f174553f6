// Oracle (test infrastructure only): CPU restatement of tet_hp / tet_hp_ins for
// straight tetrahedra, steady, p in {1, 2} with p-multigrid 2 -> 1, inflow
// (Dirichlet velocity + continuity flux) on every boundary face.
// Follows tet_hp/update.cpp:9-131 (rsdl), :181-417 (update), tet_hp/gtol.cpp:17-85,
// 815-876 (ugtouht / lftog), tet_hp/setup_preconditioner.cpp:11-210 (minvrt, p <= 2),
// :428-525 (setup_preconditioner), tet_hp/mg_restrict.cpp:4-25,
// tet_hp/mg_prolongate.cpp:7-14, tet_hp/ins/rsdl.cpp:15-87, 442-785
// (element_rsdl, straight branch), tet_hp/ins/setup_preconditioner.cpp:7-135,
// tet_hp/ins/bdry.cpp:16-102 and tet_hp/ins/bdry_ins.h:118-183 (inflow face).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include <vector>

#include "otet_basis.h"

namespace oracle {

struct TetMesh {
    int npnt = 0, nseg = 0, ntet = 0, nbf = 0;
    std::vector<double> pnts;   // npnt x 3
    std::vector<int> seg_pnt;   // nseg x 2
    std::vector<int> tet_pnt;   // ntet x 4, positively oriented
    std::vector<int> tet_seg;   // ntet x 6, edge i joins local vertices vs(i)
    std::vector<int> tet_sgn;   // ntet x 6
    std::vector<int> bf_pnt;    // nbf x 3, outward normal ordering
    std::vector<int> bf_seg;    // nbf x 3, edge j opposite face vertex j
    std::vector<int> bf_sgn;    // nbf x 3
};

struct TetParams {
    double rho = 1.0, mu = 0.0, adis = 1.0, fadd = 1.0;
    int nstage = 5;
    // block multistage defaults (tri_mesh/blocks.cpp:1126-1141)
    double alpha[8] = {0.25, 1. / 6., .375, .5, 1.0, 1.0, 1.0, 1.0};
    double beta[8] = {1.0, 0.0, 5. / 9., 0.0, 4. / 9., 1.0, 1.0, 1.0};
    double cfl[4] = {2.2, 1.5, 1.0, 0.5};
    int log2pmax = 1;
};

class TetHpIns {
  public:
    static constexpr int NV = 4;
    TetMesh m;
    TetParams prm;
    std::vector<TetBasis> bases;   // log2p = 0, 1
    int log2p, em0;
    bool coarse_flag = false, isfrst = false;
    std::vector<double> ugv, uge, ug0v, ug0e, resv, rese, res_rv, res_re, res0v, res0e;
    std::vector<std::vector<double>> dresv, drese;
    std::vector<double> vprcn, eprcn, iprcn, tau, vol;
    std::vector<unsigned char> vess, eess;   // velocity Dirichlet on boundary-face vertices / edges

    TetHpIns(const TetMesh& mesh, const TetParams& p);
    const TetBasis& b() const { return bases[log2p]; }

    void ugtouht(int tind, double uht[NV][10]) const;
    void element_rsdl(int tind, int stage, const double uht[NV][10], double lf_re[NV][10], double lf_im[NV][10]) const;
    void face_rsdl(int f, double lf[NV][6]) const;
    void rsdl(int stage);
    int setup_preconditioner();
    void minvrt();
    void update();
    double maxres(double* per_var) const;
    void mg_restrict();
    void mg_prolongate();
    void vdirichlet();
    void edirichlet();
};

}  // namespace oracle
