// Oracle (test infrastructure only): CPU restatement of tet_hp with the tet_hp_ins and
// tet_hp_cns physics for straight tetrahedra, steady, p in {1, 2} with p-multigrid
// 2 -> 1, inflow on every boundary face.
// Generic part (TetHp<NV>): tet_hp/update.cpp:9-131 (rsdl), :134-178 (element_jacobian), tet_hp/gtol.cpp:17-85,
// 815-876 (ugtouht / lftog), tet_hp/allocate.cpp:530-559 (maxres),
// tet_hp/mg_restrict.cpp:4-25, tet_hp/mg_prolongate.cpp:7-14.
// tet_hp_ins (TetHpIns): tet_hp/update.cpp:181-417 (update),
// tet_hp/setup_preconditioner.cpp:11-210 (minvrt, p <= 2), :428-525,
// tet_hp/ins/rsdl.cpp:15-87, 442-785 (element_rsdl, straight branch),
// tet_hp/ins/setup_preconditioner.cpp:7-135, tet_hp/ins/bdry.cpp:16-102 and
// tet_hp/ins/bdry_ins.h:118-183 (inflow face: Dirichlet velocity + continuity flux).
// tet_hp_cns (TetHpCns): tet_hp/cns/rsdl.cpp:16-87, 472-843 (element_rsdl, straight
// branch), tet_hp/cns/setup_preconditioner.cpp:9-489 (Weiss-Smith preconditioner and
// the element tau), tet_hp/cns/update.cpp:18-375 (update, minvrt, switch_variables),
// tet_hp/cns/bdry.cpp:15-233 and bdry_cns.h:70-98 (inflow face: Dirichlet u, v, w, RT
// and the face flux with the boundary values).
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
    // tet_hp_cns (tet_hp/cns/allocate.cpp:29-50): gamma, prandtl, R, atm_pressure and
    // the inflow boundary values (p unused, u, v, w, RT); Weiss-Smith preconditioner
    double gamma = 1.4, prandtl = 0.713, R = 287.058, atm_pressure = 0.0;
    double bc[5] = {0.0, 0.0, 0.0, 0.0, 1.0};
    int preconditioner = 1;
};

template <int NV_>
class TetHp {
  public:
    static constexpr int NV = NV_;
    TetMesh m;
    TetParams prm;
    std::vector<TetBasis> bases;   // log2p = 0, 1
    int log2p, em0;
    bool coarse_flag = false, isfrst = false;
    std::vector<double> ugv, uge, ug0v, ug0e, resv, rese, res_rv, res_re, res0v, res0e;
    std::vector<std::vector<double>> dresv, drese;
    std::vector<double> vprcn, eprcn, iprcn, vol;
    std::vector<unsigned char> vess, eess;   // boundary-face vertices / edges

    TetHp(const TetMesh& mesh, const TetParams& p);
    virtual ~TetHp() = default;
    const TetBasis& b() const { return bases[log2p]; }

    void ugtouht(int tind, double uht[NV][10]) const;
    virtual void element_rsdl(int tind, int stage, const double uht[NV][10], double lf_re[NV][10],
                              double lf_im[NV][10]) const = 0;
    virtual void face_rsdl(int f, double lf[NV][6]) const = 0;
    void rsdl(int stage);
    virtual int setup_preconditioner() = 0;
    virtual void minvrt() = 0;
    virtual void update() = 0;
    double maxres(double* per_var) const;
    // tet_hp/update.cpp:134-178: forward-difference element Jacobian, K[(i*NV+n)*N + mode*NV+var]
    void element_jacobian(int tind, double eps_r, double eps_a, double* K);
    void mg_restrict();
    void mg_prolongate();
};

class TetHpIns : public TetHp<4> {
  public:
    std::vector<double> tau;   // ntet x 4 (0: momentum, 3: continuity)
    TetHpIns(const TetMesh& mesh, const TetParams& p);
    void element_rsdl(int tind, int stage, const double uht[NV][10], double lf_re[NV][10],
                      double lf_im[NV][10]) const override;
    void face_rsdl(int f, double lf[NV][6]) const override;
    int setup_preconditioner() override;
    void minvrt() override;
    void update() override;
    void vdirichlet();
    void edirichlet();
};

class TetHpCns : public TetHp<5> {
  public:
    std::vector<double> tau;    // ntet x 5 x 5: adis * Tau / jcb
    std::vector<double> vpc;    // npnt x 5 x 5: vpreconditioner
    std::vector<double> epc;    // nseg x 5 x 5: epreconditioner
    std::vector<double> tstep;  // ntet: pseudo-time step factor of the element
    TetHpCns(const TetMesh& mesh, const TetParams& p);
    void element_rsdl(int tind, int stage, const double uht[NV][10], double lf_re[NV][10],
                      double lf_im[NV][10]) const override;
    void face_rsdl(int f, double lf[NV][6]) const override;
    int setup_preconditioner() override;
    void minvrt() override;
    void update() override;
    void vdirichlet();
    void edirichlet();
    // tet_hp/cns/setup_preconditioner.cpp:206-489 for one element
    void preconditioner_tau_timestep(const double pvu[5], double h, double hmax, double Pinv[25], double Tau[25],
                                     double& timestep) const;
    // dF/dw, dG/dw, dH/dw at one point, u[n * stride] (tet_hp/cns/rsdl.cpp:799-826)
    void flux_jacobians(const double* u, int stride, double A[25], double B[25], double C[25]) const;
    // |P dpdc A|, |P dpdc B|, |P dpdc C| (setup_preconditioner.cpp:289-400)
    void abs_flux_jacobians(const double pvu[5], double b2, double A[25], double B[25], double C[25]) const;
    // tet_hp/cns/update.cpp:349-375: a <- dpdc(pvu) a
    void switch_variables(const double pvu[5], double a[5]) const;
};

}  // namespace oracle
