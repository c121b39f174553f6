// Oracle (test infrastructure only): CPU restatement of the reference's
// tri_hp_ins residual/relaxation path for steady problems with the diagonal
// preconditioner, straight and curved elements (generic parts in otri_hp.h).
//
//   element_rsdl (ins)       tri_hp/ins/rsdl.cpp:17-82, 259-433 (straight branch, CALC_TAU2)
//                            and the curved branch :83-257 (metrics tri_hp/metric.cpp:12-21)
//   edge boundary rsdl       tri_hp/hp_boundary.cpp:973-1049 with the generic, euler and
//                            characteristic fluxes (tri_hp/ins/bdry_ins.h:34-113,
//                            tri_hp/ins/bdry.cpp:245-320)
//   setup_preconditioner     tri_hp/ins/setup_preconditioner.cpp:9-221 (diagonal, REFINED_WAY,
//                            curved branch :44-90)
//                            + tri_hp/setup_preconditioner.cpp:9-96
//
// Scope: bd(0) = 0 (steady), so mvel = 0 and dugdt = 0 (tri_hp/ins/allocate.cpp:65-115
// gives cjcb = -bd(0)*... = 0); no AXISYMMETRIC/MESH_REF_VEL; curved elements are those
// owning a curved boundary side (the default metric, not the mapped ones).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include "otri_hp.h"

namespace oracle {

// boundary flux types of tri_hp_ins::getnewedgeobject (tri_hp/ins/getnewbdry.cpp:70-110)
enum { EDGE_GENERIC = 0, EDGE_EULER = 1, EDGE_CHARACTERISTIC = 2 };

// flux of one boundary point (steady, mesh velocity zero): u the interior state, nrm the
// unit normal, side_length the point's length factor, p the basis degree, ub the
// freestream state (u, v, p) of the ibc
//   generic          tri_hp/ins/bdry_ins.h:34-49
//   euler            tri_hp/ins/bdry_ins.h:96-113
//   characteristic   tri_hp/ins/bdry.cpp:245-320 (natural fluxes zero)
void ins_edge_flux(int type, const double u[3], const double nrm[2], double side_length, int p, double rho,
                   double mu, const double ub[3], double flx[3]);

class TriHpIns : public TriHp<3> {
  public:
    static const int NV = 3;
    TriHpIns(const Mesh& mesh, const Params& prm);
    std::vector<double> tau;  // ntri x NV
    // flux type of each edge boundary (empty: all generic) and the freestream velocity;
    // the freestream pressure is prm.bc_pressure
    std::vector<int> ebdry_flux;
    double ibc_u[2] = {0.0, 0.0};
    void element_rsdl(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                      double lf_im[NV][15]) override;
    void element_rsdl_curved(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                             double lf_im[NV][15]);
    void edge_rsdl() override;
    int setup_preconditioner() override;
};

}  // namespace oracle
