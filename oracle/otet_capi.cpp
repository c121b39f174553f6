// Oracle (test infrastructure only): C entry points of the tet_hp_ins and tet_hp_cns
// restatements, bound by oracle/oracle_tet.py.
#include <cstring>
#include <memory>
#include <string>

#include "otet_basis.h"
#include "otet_hp.h"

using namespace oracle;

static const TetBasis& tb(int p) {
    static std::unique_ptr<TetBasis> b1, b2;
    if (p == 1) { if (!b1) b1.reset(new TetBasis(1)); return *b1; }
    if (!b2) b2.reset(new TetBasis(2));
    return *b2;
}

extern "C" {

// p, em, tm, gpx, G, tm2d
int otetb_info(int p, int* out) {
    const TetBasis& B = tb(p);
    out[0] = B.p; out[1] = B.em; out[2] = B.tm; out[3] = B.gpx; out[4] = B.G; out[5] = B.tm2d;
    return 0;
}
int otetb_get(int p, const char* name, double* out) {
    const TetBasis& B = tb(p);
    const std::string n(name);
    const std::vector<double>* v = nullptr;
    if (n == "xp") v = &B.xp; else if (n == "yp") v = &B.yp; else if (n == "zp") v = &B.zp;
    else if (n == "wtx") v = &B.wtx; else if (n == "wty") v = &B.wty; else if (n == "wtz") v = &B.wtz;
    else if (n == "mm") v = &B.mm; else if (n == "sfmv") v = &B.sfmv; else if (n == "vfms") v = &B.vfms;
    else if (n == "ediag") v = &B.ediag;
    else if (n == "vdiag") { out[0] = B.vdiag; return 1; }
    if (!v) return -1;
    if (!v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (int)v->size();
}
void otetb_proj(int p, const double* lin, double* f, double* dx, double* dy, double* dz) { tb(p).proj(lin, f, dx, dy, dz); }
void otetb_proj_val(int p, const double* lin, double* f) { tb(p).proj(lin, f); }
void otetb_proj_lin(int p, double u0, double u1, double u2, double u3, double* f) { tb(p).proj(u0, u1, u2, u3, f); }
void otetb_intgrt(int p, double* r, const double* f) { tb(p).intgrt(r, f); }
void otetb_intgrtrst(int p, double* r, const double* dx, const double* dy, const double* dz) { tb(p).intgrtrst(r, dx, dy, dz); }
void otetb_derivr(int p, const double* f, double* d) { tb(p).derivr(f, d); }
void otetb_derivs(int p, const double* f, double* d) { tb(p).derivs(f, d); }
void otetb_derivt(int p, const double* f, double* d) { tb(p).derivt(f, d); }
void otetb_proj2d(int p, const double* lin, double* f, double* dx, double* dy) { tb(p).proj2d(lin, f, dx, dy); }
void otetb_intgrt2d(int p, double* r, const double* f) { tb(p).intgrt2d(r, f); }

// prm4: rho, mu, adis, fadd
void* otet_create(int npnt, const double* pnts, int nseg, const int* seg_pnt, int ntet, const int* tet_pnt,
                  const int* tet_seg, const int* tet_sgn, int nbf, const int* bf_pnt, const int* bf_seg,
                  const int* bf_sgn, const double* prm4, int log2p, int nstage, const double* alpha,
                  const double* beta, const double* cfl) {
    try {
        TetMesh m;
        m.npnt = npnt; m.nseg = nseg; m.ntet = ntet; m.nbf = nbf;
        m.pnts.assign(pnts, pnts + 3 * npnt);
        m.seg_pnt.assign(seg_pnt, seg_pnt + 2 * nseg);
        m.tet_pnt.assign(tet_pnt, tet_pnt + 4 * ntet);
        m.tet_seg.assign(tet_seg, tet_seg + 6 * ntet);
        m.tet_sgn.assign(tet_sgn, tet_sgn + 6 * ntet);
        m.bf_pnt.assign(bf_pnt, bf_pnt + 3 * nbf);
        m.bf_seg.assign(bf_seg, bf_seg + 3 * nbf);
        m.bf_sgn.assign(bf_sgn, bf_sgn + 3 * nbf);
        TetParams p;
        p.rho = prm4[0]; p.mu = prm4[1]; p.adis = prm4[2]; p.fadd = prm4[3];
        p.log2pmax = log2p;
        p.nstage = nstage;
        for (int s = 0; s < nstage && s < 8; ++s) { p.alpha[s] = alpha[s]; p.beta[s] = beta[s]; }
        for (int l = 0; l <= log2p; ++l) p.cfl[l] = cfl[l];
        return new TetHpIns(m, p);
    } catch (...) {
        return nullptr;
    }
}
void otet_destroy(void* h) { delete static_cast<TetHpIns*>(h); }
static TetHpIns& X(void* h) { return *static_cast<TetHpIns*>(h); }
static void cp_in(std::vector<double>& d, const double* s) { if (!d.empty()) std::memcpy(d.data(), s, d.size() * 8); }
static void cp_out(const std::vector<double>& s, double* d) { if (!s.empty()) std::memcpy(d, s.data(), s.size() * 8); }
void otet_set_state(void* h, const double* v, const double* e) { cp_in(X(h).ugv, v); cp_in(X(h).uge, e); }
void otet_get_state(void* h, double* v, double* e) { cp_out(X(h).ugv, v); cp_out(X(h).uge, e); }
void otet_get_res(void* h, double* v, double* e) { cp_out(X(h).resv, v); cp_out(X(h).rese, e); }
void otet_set_res(void* h, const double* v, const double* e) { cp_in(X(h).resv, v); cp_in(X(h).rese, e); }
void otet_get_prcn(void* h, double* vp, double* ep, double* ip, double* tau) {
    cp_out(X(h).vprcn, vp); cp_out(X(h).eprcn, ep); cp_out(X(h).iprcn, ip); cp_out(X(h).tau, tau);
}
void otet_set_tau(void* h, const double* tau) { cp_in(X(h).tau, tau); }
void otet_rsdl(void* h, int stage) { X(h).rsdl(stage); }
int otet_setup_preconditioner(void* h) { return X(h).setup_preconditioner(); }
void otet_minvrt(void* h) { X(h).minvrt(); }
void otet_update(void* h) { X(h).update(); }
double otet_maxres(void* h, double* per_var) { return X(h).maxres(per_var); }
void otet_element_jacobian(void* h, int t, double er, double ea, double* K) { X(h).element_jacobian(t, er, ea, K); }
void otet_mg_restrict(void* h) { X(h).mg_restrict(); }
void otet_mg_prolongate(void* h) { X(h).mg_prolongate(); }
int otet_log2p(void* h) { return X(h).log2p; }
void otet_element_rsdl(void* h, int tind, int stage, const double* uh, double* lre, double* lim) {
    const int T = X(h).b().tm;
    double u[4][10], a[4][10], b[4][10];
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < T; ++i) u[n][i] = uh[n * T + i];
    X(h).element_rsdl(tind, stage, u, a, b);
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < T; ++i) { lre[n * T + i] = a[n][i]; lim[n * T + i] = b[n][i]; }
}
void otet_face_rsdl(void* h, int f, double* lf) {
    double l[4][6];
    X(h).face_rsdl(f, l);
    const int TF = X(h).b().tm2d;
    for (int n = 0; n < 4; ++n)
        for (int i = 0; i < TF; ++i) lf[n * TF + i] = l[n][i];
}

// ---- tet_hp_cns ----------------------------------------------------------------------
// prm: mu, adis, fadd, gamma, prandtl, R, bc[5] (p unused, u, v, w, RT), preconditioner
void* otetc_create(int npnt, const double* pnts, int nseg, const int* seg_pnt, int ntet, const int* tet_pnt,
                   const int* tet_seg, const int* tet_sgn, int nbf, const int* bf_pnt, const int* bf_seg,
                   const int* bf_sgn, const double* prm, int log2p, int nstage, const double* alpha,
                   const double* beta, const double* cfl) {
    try {
        TetMesh m;
        m.npnt = npnt; m.nseg = nseg; m.ntet = ntet; m.nbf = nbf;
        m.pnts.assign(pnts, pnts + 3 * npnt);
        m.seg_pnt.assign(seg_pnt, seg_pnt + 2 * nseg);
        m.tet_pnt.assign(tet_pnt, tet_pnt + 4 * ntet);
        m.tet_seg.assign(tet_seg, tet_seg + 6 * ntet);
        m.tet_sgn.assign(tet_sgn, tet_sgn + 6 * ntet);
        m.bf_pnt.assign(bf_pnt, bf_pnt + 3 * nbf);
        m.bf_seg.assign(bf_seg, bf_seg + 3 * nbf);
        m.bf_sgn.assign(bf_sgn, bf_sgn + 3 * nbf);
        TetParams p;
        p.mu = prm[0]; p.adis = prm[1]; p.fadd = prm[2]; p.gamma = prm[3]; p.prandtl = prm[4]; p.R = prm[5];
        for (int k = 0; k < 5; ++k) p.bc[k] = prm[6 + k];
        p.preconditioner = (int)prm[11];
        p.log2pmax = log2p;
        p.nstage = nstage;
        for (int s = 0; s < nstage && s < 8; ++s) { p.alpha[s] = alpha[s]; p.beta[s] = beta[s]; }
        for (int l = 0; l <= log2p; ++l) p.cfl[l] = cfl[l];
        return new TetHpCns(m, p);
    } catch (...) {
        return nullptr;
    }
}
void otetc_destroy(void* h) { delete static_cast<TetHpCns*>(h); }
static TetHpCns& XC(void* h) { return *static_cast<TetHpCns*>(h); }
void otetc_set_state(void* h, const double* v, const double* e) { cp_in(XC(h).ugv, v); cp_in(XC(h).uge, e); }
void otetc_get_state(void* h, double* v, double* e) { cp_out(XC(h).ugv, v); cp_out(XC(h).uge, e); }
void otetc_get_res(void* h, double* v, double* e) { cp_out(XC(h).resv, v); cp_out(XC(h).rese, e); }
void otetc_set_res(void* h, const double* v, const double* e) { cp_in(XC(h).resv, v); cp_in(XC(h).rese, e); }
// vprcn, eprcn (npnt/nseg x 5), tau (ntet x 25), vpc (npnt x 25), epc (nseg x 25), tstep (ntet)
void otetc_get_prcn(void* h, double* vp, double* ep, double* tau, double* vpc, double* epc, double* tstep) {
    cp_out(XC(h).vprcn, vp); cp_out(XC(h).eprcn, ep); cp_out(XC(h).tau, tau);
    cp_out(XC(h).vpc, vpc); cp_out(XC(h).epc, epc); cp_out(XC(h).tstep, tstep);
}
void otetc_set_tau(void* h, const double* tau) { cp_in(XC(h).tau, tau); }
void otetc_rsdl(void* h, int stage) { XC(h).rsdl(stage); }
int otetc_setup_preconditioner(void* h) { return XC(h).setup_preconditioner(); }
void otetc_minvrt(void* h) { XC(h).minvrt(); }
void otetc_update(void* h) { XC(h).update(); }
double otetc_maxres(void* h, double* per_var) { return XC(h).maxres(per_var); }
void otetc_element_jacobian(void* h, int t, double er, double ea, double* K) { XC(h).element_jacobian(t, er, ea, K); }
void otetc_mg_restrict(void* h) { XC(h).mg_restrict(); }
void otetc_mg_prolongate(void* h) { XC(h).mg_prolongate(); }
int otetc_log2p(void* h) { return XC(h).log2p; }
void otetc_element_rsdl(void* h, int tind, int stage, const double* uh, double* lre, double* lim) {
    const int T = XC(h).b().tm;
    double u[5][10], a[5][10], b[5][10];
    for (int n = 0; n < 5; ++n)
        for (int i = 0; i < T; ++i) u[n][i] = uh[n * T + i];
    XC(h).element_rsdl(tind, stage, u, a, b);
    for (int n = 0; n < 5; ++n)
        for (int i = 0; i < T; ++i) { lre[n * T + i] = a[n][i]; lim[n * T + i] = b[n][i]; }
}
void otetc_face_rsdl(void* h, int f, double* lf) {
    double l[5][6];
    XC(h).face_rsdl(f, l);
    const int TF = XC(h).b().tm2d;
    for (int n = 0; n < 5; ++n)
        for (int i = 0; i < TF; ++i) lf[n * TF + i] = l[n][i];
}
// Tau and Pinv of one element state (test support)
void otetc_tau_timestep(void* h, const double* pvu, double hmin, double hmax, double* pinv, double* tau, double* ts) {
    XC(h).preconditioner_tau_timestep(pvu, hmin, hmax, pinv, tau, *ts);
}
void otetc_switch_variables(void* h, const double* pvu, double* a) { XC(h).switch_variables(pvu, a); }
void otetc_flux_jacobians(void* h, const double* w, double* A, double* B, double* C) {
    XC(h).flux_jacobians(w, 1, A, B, C);
}
void otetc_abs_jacobians(void* h, const double* pvu, double b2, double* A, double* B, double* C) {
    XC(h).abs_flux_jacobians(pvu, b2, A, B, C);
}

}  // extern "C"
