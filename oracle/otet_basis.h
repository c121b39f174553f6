// Oracle (test infrastructure only): CPU restatement of the reference's modal
// tetrahedral basis, tet_basis (tet_basis/tet_basis.h:34-262), for p in {1, 2}
// (vertex and edge modes; face and interior modes first appear at p = 3).
// Follows tet_basis/initialize.cpp:72-404 (Gauss rules and 1-D tables),
// :521-800 (mass matrix, vdiag, sfmv, vfms, ediag), tet_basis/ptvalues.cpp:154-336
// (point values), tet_basis/proj.cpp:22-264, 266-416, 781-911 (transforms and
// Gauss-point derivatives), tet_basis/intgrt.cpp:23-290, 292-609 (integrations),
// tet_basis/proj2d.cpp:20-120 and tet_basis/intgrt2d.cpp:22-113 (face 0
// transforms used by the boundary flux).
//
// Gauss-point arrays are (gpx, gpy, gpz) with the z index fastest, the layout of
// the reference's f(i,j,k) at stridex = gpy*gpz, stridey = gpz.  Face arrays are
// (gpx, gpy) with y fastest.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
#pragma once
#include <vector>

namespace oracle {

struct TetBasis {
    int p, em, fm, im, tm, bm, nmodx, nmody, nmodz, gpx, gpy, gpz, G;
    int tm2d;  // face modes: 3 vertices + 3 edges (p <= 2)
    std::vector<double> a0, b0;               // Jacobi(2,2) recurrence of the edge functions
    std::vector<double> xp, x0, wtx;          // Gauss-Legendre, x0 = (1+x)/2
    std::vector<double> yp, y0, y1, wty;      // Gauss-Jacobi(1,0), y0 = 2/(1-y), y1 = (1+y)/2
    std::vector<double> zp, z0, wtz;          // Gauss-Jacobi(2,0), z0 = 2/(1-z)
    std::vector<double> dltx, dltx2, dltx3;   // barycentric Lagrange factors
    std::vector<double> dlty, dlty2, dlty3;
    std::vector<double> dltz, dltz2;
    std::vector<double> gx, dgx;              // gpx x nmodx
    std::vector<double> gy, dgy;              // gpy x nmody
    std::vector<double> gz, dgz;              // gpz x nmodz
    std::vector<double> gxwtx, dgxwtx;        // nmodx x gpx
    std::vector<double> gywty, dgywty;        // nmody x gpy
    std::vector<double> gzwtz, dgzwtz;        // nmodz x gpz
    // lumped-mass inversion factors used by minvrt (tet_basis.h:95-137)
    std::vector<double> mm;                   // tm x tm mass matrix
    double vdiag;
    std::vector<double> sfmv;                 // 2 x em
    std::vector<double> vfms;                 // 4 x tm (vertex rows)
    std::vector<double> ediag;                // em

    explicit TetBasis(int p);

    void ptvalues_deriv(double x, double y, double z, double* pgx, double* dpgx, double* pgy, double* dpgy,
                        double* pgz, double* dpgz) const;
    // value and r, s, t derivatives at the Gauss points (tet_basis/proj.cpp:22-264)
    void proj(const double* lin, double* f, double* dx, double* dy, double* dz) const;
    void proj(const double* lin, double* f) const;
    void proj(double u0, double u1, double u2, double u3, double* f) const;
    // rslt = integral of f against every mode (tet_basis/intgrt.cpp:23-290)
    void intgrt(double* rslt, const double* f) const;
    // rslt -= integral of (dphi/dr dx + dphi/ds dy + dphi/dt dz) (intgrt.cpp:292-609)
    void intgrtrst(double* rslt, const double* dx, const double* dy, const double* dz) const;
    // d += r, s, t derivative of the Gauss-point field f (proj.cpp:781-911)
    void derivr(const double* f, double* d) const;
    void derivs(const double* f, double* d) const;
    void derivt(const double* f, double* d) const;
    // face 0 transforms (modes V1, V2, V3, E1, E2, E3), face Gauss points gpx x gpy
    void proj2d(const double* lin, double* f, double* dx, double* dy) const;
    void proj2d(const double* lin, double* f) const;
    void intgrt2d(double* rslt, const double* f) const;

  private:
    void lumpinv();
};

}  // namespace oracle
