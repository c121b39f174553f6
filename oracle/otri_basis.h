// Oracle (test infrastructure only): CPU restatement of the reference's modal
// triangle basis, tri_basis<p,0> (tri_basis/tri_basis.h:213-492), with p in {1,2,4}.
// Follows tri_basis/initialize.cpp:161-455 (tables), :590-958 (lumped-mass
// factors), tri_basis/ptvalues.cpp:84-195 (point values), tri_basis/proj.cpp:22-402
// and tri_basis/intgrt.cpp:22-158 (transforms), tri_basis/proj1d.cpp:19-59 and
// tri_basis/intgrt1d.cpp:17-43 (edge transforms), tri_basis/initialize.cpp:1085-1130
// (1-D side mass factor sdiag1d).  The recurrence coefficients
// restate ORTHPOL's recur (utilities/myblas.c:257-459) from the closed forms of the
// monic Jacobi recurrence; Gauss rules come from oracle::gauss_rule.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include <vector>

namespace oracle {

struct TriBasis {
    int p, sm, im, tm, bm, nmodx, gpx, gpn;
    static const int ibwth = 1;
    static const int sbwth = 2;
    int arow;                              // row length of the recurrence tables
    std::vector<double> a0, b0;            // (sm+2) x (gpx+1)
    std::vector<double> gx, dgx;           // gpx x nmodx
    std::vector<double> wtx, xp, x0;       // gpx
    std::vector<double> gxwtx, dgxwtx;     // nmodx x gpx
    std::vector<double> dltx, dltx1, dltn1;// gpx, gpx x gpx, gpx x gpx
    std::vector<double> gn, dgn;           // gpn x tm
    std::vector<double> wtn, np, n0;       // gpn
    std::vector<double> gnwtn, gnwtnn0, dgnwtn;  // tm x gpn
    std::vector<double> dltn, dltn2;       // gpn, gpn x gpn
    std::vector<double> norm;              // tm
    // lumped-mass inversion factors (tri_basis.h:124-141)
    std::vector<double> sfmv;              // 2 x sm
    std::vector<double> ifmb;              // bm x im
    double vdiag;
    std::vector<double> vfms;              // 3 x sm
    std::vector<double> sdiag;             // sm
    std::vector<double> sfms;              // (sm-1) x sm x 3
    std::vector<double> bfmi;              // bm x im
    std::vector<double> idiag;             // im x (ibwth+1), factored
    std::vector<double> sdiag1d;           // sm x (sbwth+1), factored 1-D side mass
    mutable std::vector<double> pgx, dpgx, pgn, dpgn;

    explicit TriBasis(int p);

    void ptvalues_deriv(double x, double eta) const;
    // stride-addressed transforms, argument conventions of tri_basis_interface
    void proj(const double* lin, double* f, double* dx, double* dy, int stride) const;
    void proj(const double* lin, double* f, int stride) const;
    void proj(double u1, double u2, double u3, double* f, int stride) const;
    void derivr(const double* f, double* dr, int stride) const;
    void derivs(const double* f, double* ds, int stride) const;
    void intgrt(double* rslt, const double* f, int stride) const;
    void intgrtrs(double* rslt, const double* dx, const double* dy, int stride) const;
    void proj1d(const double* lin, double* f, double* dx) const;
    void proj1d(const double* lin, double* f) const;
    void proj1d(double u1, double u2, double* f) const;
    void intgrt1d(double* rslt, const double* f) const;

  private:
    void initialize_values();
    void lumpinv();
    void lumpinv1d();
};

// utilities/myblas.c:257-459 restated (ipoly 1 = Legendre, 6 = Jacobi(al,be)).
void recur_coeffs(int n, int ipoly, double al, double be, double* a, double* b);

}  // namespace oracle
