// Oracle (test infrastructure only): CPU restatement of the reference's generic
// tri_hp residual/relaxation machinery for steady problems with the diagonal
// preconditioner, shared by the ins and cns restatements.
//
//   rsdl(stage)              tri_hp/nstage.cpp:12-110
//   mg_source                tri_hp/nstage.cpp:113-129
//   update                   tri_hp/nstage.cpp:256-427
//   ugtouht / lftog          tri_hp/gtol.cpp:14-55, 364-402
//   vdirichlet / sdirichlet  tri_hp/hp_boundary.cpp:680-704
//   preconditioner assembly  tri_hp/setup_preconditioner.cpp:9-96
//   minvrt                   tri_hp/setup_preconditioner.cpp:102-399 (diagonal branch)
//   maxres                   tri_hp/allocate.cpp:496-525
//   mg_restrict/prolongate   tri_hp/movco.cpp:4-106, tri_hp/movfin.cpp:7-82 (p-levels, and
//                            h-levels through the fine/coarse vertex connections)
//   element_jacobian         tri_hp/nstage.cpp:131-165 (fixed mesh)
//   cycle                    tri_mesh/blocks.cpp:1267-1348 (one grid, extra_finest_levels)
//   area, inscribedradius    tri_mesh/tprims.cpp:66-80, 224-242
//   crdtocht / crdtocht1d    tri_hp/gtol.cpp:204-245, 296-320 (curved boundary sides)
//   periodic exchanges       tri_hp/setup_preconditioner.cpp:24-36, 171-176, 255-258, 331-334,
//                            tri_mesh/tri_boundary.cpp:344-470, tri_mesh/boundary.h:594-660
//
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
    // ins
    double rho = 1.0, mu = 0.0, adis = 1.0;
    double bc_pressure = 0.0;      // freestream ibc f(NV-1) on the edge flux (getnewibc.cpp:21-32)
    // cns (tri_hp/cns/allocate.cpp:32-40): gamma, prandtl, gas constant; mu above;
    // free-stream boundary values (p, u, v, RT) of the inflow boundaries
    double gamma = 1.4, prandtl = 0.713, Rgas = 287.058;
    double bc[4] = {1.0, 0.0, 0.0, 1.0};
    int nstage = 5;
    double alpha[8] = {0.25, 1. / 6., .375, .5, 1.0, 1.0, 1.0, 1.0};
    double beta[8] = {1.0, 0.0, 5. / 9., 0.0, 4. / 9., 1.0, 1.0, 1.0};
    double cfl[4] = {2.5, 1.5, 1.0, 0.5};
    double fadd = 1.0;
    int log2pmax = 2;
};

template <int NV>
class TriHp {
  public:
    TriHp(const Mesh& mesh, const Params& prm);
    virtual ~TriHp() = default;

    Mesh m;
    Params prm;
    std::vector<std::unique_ptr<TriBasis>> bases;  // log2p = 0,1,2 -> p = 1,2,4
    int log2p, sm0, im0, smf, imf;
    bool coarse_flag = false, isfrst = false;
    // h-multigrid: a coarse mesh level linked to its fine neighbour.  ccnnct of the
    // fine mesh (tri_mesh.h:197-201): for every fine vertex the coarse triangle that
    // holds it and its barycentric weights; fcnnct of this mesh: for every coarse
    // vertex the fine triangle and weights.  vug_frst keeps the restricted state.
    bool coarse_level = false;
    TriHp* fine = nullptr;
    std::vector<int> cc_tri, fc_tri;
    std::vector<double> cc_wt, fc_wt, vug_frst;
    void link_fine(TriHp* f, const int* cctri, const double* ccwt, const int* fctri, const double* fcwt);
    // curved boundary sides: hp_edge_bdry::crds (tri_hp/hp_boundary.cpp:621-675), the
    // row of each side in scrd (-1: straight), rows of smf modes x 2 in the side's
    // direction; tri_crv marks the elements that own one when log2pmax > 0
    // (tri_hp::metric::setinfo, tri_hp/metric.cpp:45-70)
    std::vector<int> seg_crv;
    std::vector<double> scrd;
    std::vector<char> tri_crv;
    void set_curved(int nside, const int* sides, const double* crds);
    bool curved(int t) const { return !tri_crv.empty() && tri_crv[t]; }
    void crdtocht(int tind, double cht[2][15]) const;
    void crdtocht1d(int sind, double cht[2][8]) const;
    // periodic boundaries (prdc, tri_mesh/tri_boundary.h:113-148): side j of the first
    // edge boundary of a pair matched with side n-1-j of the second (the comm
    // boundaries' reverse order, tri_mesh/tri_boundary.cpp:344-470); vertex copies and
    // side pairs (a, b, relative direction).  Every copy keeps its partial residual; the
    // preconditioner sums and the solved vertex and side values are averaged over the
    // copies where tri_hp exchanges them (tri_hp/setup_preconditioner.cpp:24-36, 171-176,
    // 255-258, 331-334); the periodic boundaries carry no flux and no essential mask.
    std::vector<std::vector<int>> prdc_groups;
    std::vector<int> prdc_a, prdc_b, prdc_o;
    std::vector<char> prdc_bdry;
    void set_periodic(int npair, const int* pairs);
    void prdc_sync();                         // state copies from the first copy / side a
    void vexchange(double* data);             // average over each vertex group, NV per vertex
    void sexchange(double* data, int stride, int mode, bool sign);  // side pairs, one mode

    // solution and work arrays, reference layouts v(npnt,NV), s(nseg,sm,NV), i(ntri,im,NV)
    std::vector<double> ugv, ugs, ugi;
    std::vector<double> ug0v, ug0s, ug0i;
    std::vector<double> resv, ress, resi, res_rv, res_rs, res_ri, res0v, res0s, res0i;
    std::vector<std::vector<double>> dresv, dress, dresi;
    std::vector<double> vprcn, sprcn, tprcn;  // (npnt|nseg|ntri) x NV
    double uht[NV][15], lf[NV][15];
    std::vector<double> bre, bim;  // element residuals of one chunk of rsdl's parallel loop

    const TriBasis& b() const { return *bases[log2p]; }
    double area(int t) const;
    double inscribedradius(int t) const;
    void ugtouht(int tind);
    void ugtouht_to(int tind, double uht[NV][15]) const;
    void lftog(int tind, double* gv, double* gs, double* gi, const double lfl[NV][15]);
    virtual void element_rsdl(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                              double lf_im[NV][15]) = 0;
    virtual void edge_rsdl() = 0;
    virtual int setup_preconditioner() = 0;
    void rsdl(int stage);
    void mg_source();
    void assemble_prcn(int ntri_done);
    void vdirichlet();
    void sdirichlet(int mode);
    void minvrt();
    virtual void update();
    double maxres(double* per_var = nullptr) const;
    void mg_restrict();
    void mg_prolongate();
    void element_jacobian(int tind, double eps_r, double eps_a, double* K);
    double cycle(int vw, int lvl, int mglvls, int itercrsn, int iterrfne);
};

extern template class TriHp<3>;
extern template class TriHp<4>;

}  // namespace oracle
