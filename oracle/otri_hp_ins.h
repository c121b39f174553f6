// Oracle (test infrastructure only): CPU restatement of the reference's
// tri_hp + tri_hp_ins residual/relaxation path for straight-sided steady problems
// with the diagonal preconditioner.
//
//   rsdl(stage)              tri_hp/nstage.cpp:12-110
//   mg_source                tri_hp/nstage.cpp:113-129
//   update                   tri_hp/nstage.cpp:256-427
//   ugtouht / lftog          tri_hp/gtol.cpp:14-55, 364-402; ugtouht1d gtol.cpp:156-171
//   element_rsdl (ins)       tri_hp/ins/rsdl.cpp:17-82, 259-433 (straight branch, CALC_TAU2)
//   edge boundary rsdl       tri_hp/hp_boundary.cpp:973-1049 with bdry_ins::generic::flux
//                            (tri_hp/ins/bdry_ins.h:34-49)
//   vdirichlet / sdirichlet  tri_hp/hp_boundary.cpp:680-704
//   setup_preconditioner     tri_hp/ins/setup_preconditioner.cpp:9-221 (diagonal, REFINED_WAY)
//                            + tri_hp/setup_preconditioner.cpp:9-96
//   minvrt                   tri_hp/setup_preconditioner.cpp:102-399 (diagonal branch)
//   maxres                   tri_hp/allocate.cpp:496-525
//   mg_restrict/prolongate   tri_hp/movco.cpp:4-40, tri_hp/movfin.cpp:7-14 (p-levels)
//   element_jacobian         tri_hp/nstage.cpp:131-165 (fixed mesh)
//   cycle                    tri_mesh/blocks.cpp:1267-1348 (one grid, extra_finest_levels)
//   area, inscribedradius    tri_mesh/tprims.cpp:66-80, 224-242
//
// Scope: bd(0) = 0 (steady), so mvel = 0 and dugdt = 0 (tri_hp/ins/allocate.cpp:65-115
// gives cjcb = -bd(0)*... = 0); no AXISYMMETRIC/MESH_REF_VEL; every element straight.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include "otri_basis.h"
#include <memory>
#include <vector>

namespace oracle {

struct Mesh {
    int npnt = 0, nseg = 0, ntri = 0;
    std::vector<double> pnts;      // npnt x 2
    std::vector<int> seg_pnt;      // nseg x 2
    std::vector<int> tri_pnt;      // ntri x 3
    std::vector<int> tri_seg;      // ntri x 3
    std::vector<int> tri_sgn;      // ntri x 3
    // edge boundaries (ordered segment lists) and their essential-variable bit masks
    std::vector<int> ebdry_start;  // nebd+1
    std::vector<int> ebdry_seg;
    std::vector<int> ebdry_essential;
};

struct Params {
    double rho = 1.0, mu = 0.0, adis = 1.0;
    double bc_pressure = 0.0;      // freestream ibc f(NV-1) on the edge flux (getnewibc.cpp:21-32)
    int nstage = 5;
    double alpha[8] = {0.25, 1. / 6., .375, .5, 1.0, 1.0, 1.0, 1.0};
    double beta[8] = {1.0, 0.0, 5. / 9., 0.0, 4. / 9., 1.0, 1.0, 1.0};
    double cfl[4] = {2.5, 1.5, 1.0, 0.5};
    double fadd = 1.0;
    int log2pmax = 2;
};

class TriHpIns {
  public:
    static const int NV = 3;
    TriHpIns(const Mesh& mesh, const Params& prm);

    Mesh m;
    Params prm;
    std::vector<std::unique_ptr<TriBasis>> bases;  // log2p = 0,1,2 -> p = 1,2,4
    int log2p, sm0, im0, smf, imf;
    bool coarse_flag = false, isfrst = false;

    // solution and work arrays, reference layouts v(npnt,NV), s(nseg,sm,NV), i(ntri,im,NV)
    std::vector<double> ugv, ugs, ugi;
    std::vector<double> ug0v, ug0s, ug0i;
    std::vector<double> resv, ress, resi, res_rv, res_rs, res_ri, res0v, res0s, res0i;
    std::vector<std::vector<double>> dresv, dress, dresi;
    std::vector<double> vprcn, sprcn, tprcn;  // (npnt|nseg|ntri) x NV
    std::vector<double> tau;                  // ntri x NV
    double uht[NV][15], lf[NV][15];

    const TriBasis& b() const { return *bases[log2p]; }
    double area(int t) const;
    double inscribedradius(int t) const;
    void ugtouht(int tind);
    void lftog(int tind, double* gv, double* gs, double* gi, const double lfl[NV][15]);
    void element_rsdl(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15], double lf_im[NV][15]);
    void edge_rsdl();
    void rsdl(int stage);
    void mg_source();
    int setup_preconditioner();
    void assemble_prcn(int ntri_done);
    void vdirichlet();
    void sdirichlet(int mode);
    void minvrt();
    void update();
    double maxres(double* per_var = nullptr) const;
    void mg_restrict();
    void mg_prolongate();
    void element_jacobian(int tind, double eps_r, double eps_a, double* K);
    double cycle(int vw, int lvl, int mglvls, int itercrsn, int iterrfne);
};

}  // namespace oracle
