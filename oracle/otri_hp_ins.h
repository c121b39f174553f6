// Oracle (test infrastructure only): CPU restatement of the reference's
// tri_hp_ins residual/relaxation path for steady problems with the diagonal
// preconditioner, straight and curved elements (generic parts in otri_hp.h).
//
//   element_rsdl (ins)       tri_hp/ins/rsdl.cpp:17-82, 259-433 (straight branch, CALC_TAU2)
//                            and the curved branch :83-257 (metrics tri_hp/metric.cpp:12-21)
//   edge boundary rsdl       tri_hp/hp_boundary.cpp:973-1049 with bdry_ins::generic::flux
//                            (tri_hp/ins/bdry_ins.h:34-49)
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

class TriHpIns : public TriHp<3> {
  public:
    static const int NV = 3;
    TriHpIns(const Mesh& mesh, const Params& prm);
    std::vector<double> tau;  // ntri x NV
    void element_rsdl(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                      double lf_im[NV][15]) override;
    void element_rsdl_curved(int tind, int stage, const double u_ht[NV][15], double lf_re[NV][15],
                             double lf_im[NV][15]);
    void edge_rsdl() override;
    int setup_preconditioner() override;
};

}  // namespace oracle
