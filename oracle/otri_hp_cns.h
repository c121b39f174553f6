// Oracle (test infrastructure only): CPU restatement of the reference's
// tri_hp_cns (compressible Navier-Stokes, primitive variables p, u, v, RT) for
// straight-sided steady problems with the diagonal preconditioner.  Generic parts
// (rsdl, minvrt, p-multigrid, cycle) are in otri_hp.h.
//
//   element_rsdl (cns)       tri_hp/cns/rsdl.cpp:16-35, 50-90, 312-530 (straight branch,
//                            CALC_TAU2, no BODYFORCE/SUTHERLAND, source "zero")
//   calculate_tau            tri_hp/cns/setup_preconditioner.cpp:241-515
//   spectral_radius          utilities/matrix.cpp:16-37
//   setup_preconditioner     tri_hp/cns/setup_preconditioner.cpp:7-239 (diagonal branch)
//                            + tri_hp/setup_preconditioner.cpp:9-96
//   edge boundary rsdl       tri_hp/hp_boundary.cpp:973-1049 with bdry_cns::generic::flux
//                            (tri_hp/cns/bdry.cpp:192-245)
//   inflow boundaries        essential u, v, RT (tri_hp/cns/bdry_cns.h:56-66);
//                            modify_boundary_residual tri_hp/cns/bdry.cpp:247-340
//   update                   tri_hp/cns/update.cpp:276-357
//   project_new_variables    tri_hp/cns/update.cpp:5-245, switch_variables :247-271
//   constants                tri_hp/cns/allocate.cpp:11-47 (kcond = mu/Pr R gamma/(gamma-1))
//
// Scope: bd(0) = 0 (steady: mvel = 0, dugdt = 0), every element straight, constant
// inflow boundary values (p, u, v, RT) = Params::bc, boundaries whose essential mask
// is 0b1110 are the inflow boundaries.  The eigenvalue magnitudes of calculate_tau
// are |.| of doubles.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include "otri_hp.h"

namespace oracle {

class TriHpCns : public TriHp<4> {
  public:
    static const int NV = 4;
    TriHpCns(const Mesh& mesh, const Params& prm);
    double kcond;
    // block preconditioners (npnt|nseg|ntri) x 4 x 4, row-major per entity
    std::vector<double> vpc, spc, tpc;
    std::vector<double> res_tv, res_ts, res_ti;   // res_temp of project_new_variables

    void element_rsdl(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                      double lf_im[NV][15]) override;
    void edge_rsdl() override;
    int setup_preconditioner() override;
    void update() override;
    void calculate_tau(const double pvu[4], double h, double Pinv[4][4], double Tau[4][4], double& timestep) const;
    void modify_boundary_residual();
    void project_new_variables();
    void switch_variables(const double pvu[4], double a[4]) const;
};

double spectral_radius4(const double A[4][4]);

}  // namespace oracle
