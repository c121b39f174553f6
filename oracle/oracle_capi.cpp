// Oracle (test infrastructure only): C entry points over the CPU restatement so
// that tests/ (ctypes) and bench.py's cpu_baseline leg can drive it.  The product
// library never links this file.
#include <omp.h>
#include "otri_basis.h"
#include "otri_hp_ins.h"
#include "otri_hp_cns.h"
#include "olinalg.h"
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

using namespace oracle;

static const TriBasis& basis_for(int p) {
    static std::map<int, std::unique_ptr<TriBasis>> cache;
    auto it = cache.find(p);
    if (it == cache.end()) it = cache.emplace(p, std::unique_ptr<TriBasis>(new TriBasis(p))).first;
    return *it->second;
}

extern "C" {

// OpenMP threads of the element loops (rsdl, minvrt's interior sweep); returns the
// number in effect.  Results do not depend on it.
int oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void oracle_recur(int n, int ipoly, double al, double be, double* a, double* b) { recur_coeffs(n, ipoly, al, be, a, b); }
void oracle_gauss(int n, const double* a, const double* b, double* x, double* w) { gauss_rule(n, a, b, x, w); }

int obasis_info(int p, int* out) {
    const TriBasis& B = basis_for(p);
    out[0] = B.sm; out[1] = B.im; out[2] = B.tm; out[3] = B.bm; out[4] = B.nmodx; out[5] = B.gpx; out[6] = B.gpn;
    return 0;
}

// Copies a named table; returns its length (or -1 for an unknown name).
int obasis_get(int p, const char* name, double* out) {
    const TriBasis& B = basis_for(p);
    const std::map<std::string, const std::vector<double>*> tabs = {
        {"gx", &B.gx}, {"dgx", &B.dgx}, {"wtx", &B.wtx}, {"xp", &B.xp}, {"x0", &B.x0},
        {"gxwtx", &B.gxwtx}, {"dgxwtx", &B.dgxwtx}, {"dltx", &B.dltx}, {"dltx1", &B.dltx1},
        {"dltn1", &B.dltn1}, {"gn", &B.gn}, {"dgn", &B.dgn}, {"wtn", &B.wtn}, {"np", &B.np},
        {"n0", &B.n0}, {"gnwtn", &B.gnwtn}, {"gnwtnn0", &B.gnwtnn0}, {"dgnwtn", &B.dgnwtn},
        {"dltn", &B.dltn}, {"dltn2", &B.dltn2}, {"norm", &B.norm}, {"sfmv", &B.sfmv},
        {"ifmb", &B.ifmb}, {"vfms", &B.vfms}, {"sdiag", &B.sdiag}, {"sfms", &B.sfms},
        {"bfmi", &B.bfmi}, {"idiag", &B.idiag}, {"sdiag1d", &B.sdiag1d}, {"a0", &B.a0}, {"b0", &B.b0}};
    if (std::string(name) == "vdiag") { if (out) out[0] = B.vdiag; return 1; }
    auto it = tabs.find(name);
    if (it == tabs.end()) return -1;
    if (out) std::memcpy(out, it->second->data(), it->second->size() * sizeof(double));
    return int(it->second->size());
}

void obasis_ptvalues_deriv(int p, double x, double eta, double* pgx, double* dpgx, double* pgn, double* dpgn) {
    const TriBasis& B = basis_for(p);
    B.ptvalues_deriv(x, eta);
    std::memcpy(pgx, B.pgx.data(), B.nmodx * sizeof(double));
    std::memcpy(dpgx, B.dpgx.data(), B.nmodx * sizeof(double));
    std::memcpy(pgn, B.pgn.data(), B.tm * sizeof(double));
    std::memcpy(dpgn, B.dpgn.data(), B.tm * sizeof(double));
}

// Transforms on gpx x gpn point arrays (stride gpn), argument conventions of tri_basis_interface.
void obasis_proj(int p, const double* lin, double* f, double* dx, double* dy) { const TriBasis& B = basis_for(p); B.proj(lin, f, dx, dy, B.gpn); }
void obasis_proj_val(int p, const double* lin, double* f) { const TriBasis& B = basis_for(p); B.proj(lin, f, B.gpn); }
void obasis_proj_lin(int p, double u1, double u2, double u3, double* f) { const TriBasis& B = basis_for(p); B.proj(u1, u2, u3, f, B.gpn); }
void obasis_derivr(int p, const double* f, double* dr) { const TriBasis& B = basis_for(p); B.derivr(f, dr, B.gpn); }
void obasis_derivs(int p, const double* f, double* ds) { const TriBasis& B = basis_for(p); B.derivs(f, ds, B.gpn); }
void obasis_intgrt(int p, double* rslt, const double* f) { const TriBasis& B = basis_for(p); B.intgrt(rslt, f, B.gpn); }
void obasis_intgrtrs(int p, double* rslt, const double* dx, const double* dy) { const TriBasis& B = basis_for(p); B.intgrtrs(rslt, dx, dy, B.gpn); }
void obasis_proj1d(int p, const double* lin, double* f, double* dx) { basis_for(p).proj1d(lin, f, dx); }
void obasis_intgrt1d(int p, double* rslt, const double* f) { basis_for(p).intgrt1d(rslt, f); }
void oracle_dpbtrsnu2(const double* abd, int stride, int ordr, int ofdg, double* b, int rhs) { dpbtrsnu2_literal(abd, stride, ordr, ofdg, b, rhs); }

// ---- tri_hp_ins --------------------------------------------------------------
// prm: rho, mu, adis, bc_pressure, fadd
void* otri_create(int npnt, const double* pnts, int nseg, const int* seg_pnt, int ntri, const int* tri_pnt,
                  const int* tri_seg, const int* tri_sgn, int nebd, const int* ebdry_start, const int* ebdry_seg,
                  const int* ebdry_essential, const double* prm5, int log2p, int nstage, const double* alpha,
                  const double* beta, const double* cfl) {
    try {
        Mesh m;
        m.npnt = npnt; m.nseg = nseg; m.ntri = ntri;
        m.pnts.assign(pnts, pnts + 2 * npnt);
        m.seg_pnt.assign(seg_pnt, seg_pnt + 2 * nseg);
        m.tri_pnt.assign(tri_pnt, tri_pnt + 3 * ntri);
        m.tri_seg.assign(tri_seg, tri_seg + 3 * ntri);
        m.tri_sgn.assign(tri_sgn, tri_sgn + 3 * ntri);
        m.ebdry_start.assign(ebdry_start, ebdry_start + nebd + 1);
        m.ebdry_seg.assign(ebdry_seg, ebdry_seg + ebdry_start[nebd]);
        m.ebdry_essential.assign(ebdry_essential, ebdry_essential + nebd);
        Params p;
        p.rho = prm5[0]; p.mu = prm5[1]; p.adis = prm5[2]; p.bc_pressure = prm5[3]; p.fadd = prm5[4];
        p.log2pmax = log2p;
        p.nstage = nstage;
        for (int s = 0; s < nstage && s < 7; ++s) { p.alpha[s] = alpha[s]; p.beta[s] = beta[s]; }
        p.alpha[nstage] = 1.0; p.beta[nstage] = 1.0;
        for (int l = 0; l <= log2p; ++l) p.cfl[l] = cfl[l];
        return new TriHpIns(m, p);
    } catch (...) {
        return nullptr;
    }
}
void otri_destroy(void* h) { delete static_cast<TriHpIns*>(h); }

static TriHpIns& T(void* h) { return *static_cast<TriHpIns*>(h); }

void otri_set_state(void* h, const double* v, const double* s, const double* i) {
    TriHpIns& x = T(h);
    std::memcpy(x.ugv.data(), v, x.ugv.size() * 8);
    if (!x.ugs.empty()) std::memcpy(x.ugs.data(), s, x.ugs.size() * 8);
    if (!x.ugi.empty()) std::memcpy(x.ugi.data(), i, x.ugi.size() * 8);
    x.prdc_sync();
}
void otri_get_state(void* h, double* v, double* s, double* i) {
    TriHpIns& x = T(h);
    std::memcpy(v, x.ugv.data(), x.ugv.size() * 8);
    if (!x.ugs.empty()) std::memcpy(s, x.ugs.data(), x.ugs.size() * 8);
    if (!x.ugi.empty()) std::memcpy(i, x.ugi.data(), x.ugi.size() * 8);
}
void otri_get_res(void* h, double* v, double* s, double* i) {
    TriHpIns& x = T(h);
    std::memcpy(v, x.resv.data(), x.resv.size() * 8);
    if (!x.ress.empty()) std::memcpy(s, x.ress.data(), x.ress.size() * 8);
    if (!x.resi.empty()) std::memcpy(i, x.resi.data(), x.resi.size() * 8);
}
void otri_set_res(void* h, const double* v, const double* s, const double* i) {
    TriHpIns& x = T(h);
    std::memcpy(x.resv.data(), v, x.resv.size() * 8);
    if (!x.ress.empty()) std::memcpy(x.ress.data(), s, x.ress.size() * 8);
    if (!x.resi.empty()) std::memcpy(x.resi.data(), i, x.resi.size() * 8);
}
void otri_get_prcn(void* h, double* vp, double* sp, double* tp) {
    TriHpIns& x = T(h);
    std::memcpy(vp, x.vprcn.data(), x.vprcn.size() * 8);
    std::memcpy(sp, x.sprcn.data(), x.sprcn.size() * 8);
    std::memcpy(tp, x.tprcn.data(), x.tprcn.size() * 8);
}
void otri_set_tprcn(void* h, const double* tp) {
    TriHpIns& x = T(h);
    std::memcpy(x.tprcn.data(), tp, x.tprcn.size() * 8);
}
void otri_assemble_prcn(void* h) { T(h).assemble_prcn(T(h).m.ntri); }
void otri_rsdl(void* h, int stage) { T(h).rsdl(stage); }
int otri_setup_preconditioner(void* h) { return T(h).setup_preconditioner(); }
void otri_minvrt(void* h) { T(h).minvrt(); }
void otri_update(void* h) { T(h).update(); }
double otri_maxres(void* h, double* per_var) { return T(h).maxres(per_var); }
void otri_mg_restrict(void* h) { T(h).mg_restrict(); }
void otri_mg_prolongate(void* h) { T(h).mg_prolongate(); }
int otri_log2p(void* h) { return T(h).log2p; }
double otri_cycle(void* h, int vw, int mglvls, int itercrsn, int iterrfne) {
    try { return T(h).cycle(vw, 0, mglvls, itercrsn, iterrfne); } catch (...) { return -1.0; }
}
void otri_element_rsdl(void* h, int tind, int stage, const double* uh, double* lf_re, double* lf_im) {
    TriHpIns& x = T(h);
    double u[3][15], lr[3][15], li[3][15];
    const int tm = x.b().tm;
    for (int n = 0; n < 3; ++n) for (int k = 0; k < tm; ++k) u[n][k] = uh[n * tm + k];
    x.element_rsdl(tind, stage, u, lr, li);
    for (int n = 0; n < 3; ++n) for (int k = 0; k < tm; ++k) { lf_re[n * tm + k] = lr[n][k]; lf_im[n * tm + k] = li[n][k]; }
}
void otri_element_jacobian(void* h, int tind, double eps_r, double eps_a, double* K) { T(h).element_jacobian(tind, eps_r, eps_a, K); }

// curved boundary sides (ins): crds (nside, smf, 2) as hp_edge_bdry::crds
void otri_set_curved(void* h, int nside, const int* sides, const double* crds) { T(h).set_curved(nside, sides, crds); }

// periodic pairs of edge boundaries (ins)
int otri_set_periodic(void* h, int npair, const int* pairs) {
    try {
        T(h).set_periodic(npair, pairs);
    } catch (...) {
        return -1;
    }
    return 0;
}

// boundary flux types per edge boundary and the freestream state (u, v, p)
void otri_set_edge_flux(void* h, int nebd, const int* types, const double* ibc) {
    TriHpIns& x = T(h);
    x.ebdry_flux.assign(types, types + nebd);
    x.ibc_u[0] = ibc[0];
    x.ibc_u[1] = ibc[1];
    x.prm.bc_pressure = ibc[2];
}
// one boundary point's flux (test support)
void otri_edge_flux(int type, const double* u, const double* nrm, double side_length, int p, double rho, double mu,
                    const double* ub, double* flx) {
    ins_edge_flux(type, u, nrm, side_length, p, rho, mu, ub, flx);
}

// h-multigrid link of a coarse level to its fine neighbour (ins and cns)
void otri_get_mg(void* h, double* res0v, double* vug) { auto& x = T(h); std::memcpy(res0v, x.res0v.data(), x.m.npnt * 3 * 8); if (!x.vug_frst.empty()) std::memcpy(vug, x.vug_frst.data(), x.vug_frst.size() * 8); }
void otri_link_fine(void* coarse, void* fine, const int* cc_tri, const double* cc_wt, const int* fc_tri, const double* fc_wt) {
    static_cast<TriHpIns*>(coarse)->link_fine(static_cast<TriHpIns*>(fine), cc_tri, cc_wt, fc_tri, fc_wt);
}

// ---- tri_hp_cns --------------------------------------------------------------
// prm: mu, adis, gamma, prandtl, Rgas, fadd, bc p, bc u, bc v, bc RT
void* ocns_create(int npnt, const double* pnts, int nseg, const int* seg_pnt, int ntri, const int* tri_pnt,
                  const int* tri_seg, const int* tri_sgn, int nebd, const int* ebdry_start, const int* ebdry_seg,
                  const int* ebdry_essential, const double* prm10, int log2p, int nstage, const double* alpha,
                  const double* beta, const double* cfl) {
    try {
        Mesh m;
        m.npnt = npnt; m.nseg = nseg; m.ntri = ntri;
        m.pnts.assign(pnts, pnts + 2 * npnt);
        m.seg_pnt.assign(seg_pnt, seg_pnt + 2 * nseg);
        m.tri_pnt.assign(tri_pnt, tri_pnt + 3 * ntri);
        m.tri_seg.assign(tri_seg, tri_seg + 3 * ntri);
        m.tri_sgn.assign(tri_sgn, tri_sgn + 3 * ntri);
        m.ebdry_start.assign(ebdry_start, ebdry_start + nebd + 1);
        m.ebdry_seg.assign(ebdry_seg, ebdry_seg + ebdry_start[nebd]);
        m.ebdry_essential.assign(ebdry_essential, ebdry_essential + nebd);
        Params p;
        p.mu = prm10[0]; p.adis = prm10[1]; p.gamma = prm10[2]; p.prandtl = prm10[3]; p.Rgas = prm10[4];
        p.fadd = prm10[5];
        for (int n = 0; n < 4; ++n) p.bc[n] = prm10[6 + n];
        p.log2pmax = log2p;
        p.nstage = nstage;
        for (int s = 0; s < nstage && s < 7; ++s) { p.alpha[s] = alpha[s]; p.beta[s] = beta[s]; }
        p.alpha[nstage] = 1.0; p.beta[nstage] = 1.0;
        for (int l = 0; l <= log2p; ++l) p.cfl[l] = cfl[l];
        return new TriHpCns(m, p);
    } catch (...) {
        return nullptr;
    }
}
void ocns_destroy(void* h) { delete static_cast<TriHpCns*>(h); }
static TriHpCns& Cn(void* h) { return *static_cast<TriHpCns*>(h); }
static void put(std::vector<double>& dst, const double* src) { if (!dst.empty()) std::memcpy(dst.data(), src, dst.size() * 8); }
static void take(const std::vector<double>& src, double* dst) { if (!src.empty()) std::memcpy(dst, src.data(), src.size() * 8); }
void ocns_set_state(void* h, const double* v, const double* s, const double* i) { auto& x = Cn(h); put(x.ugv, v); put(x.ugs, s); put(x.ugi, i); }
void ocns_get_state(void* h, double* v, double* s, double* i) { auto& x = Cn(h); take(x.ugv, v); take(x.ugs, s); take(x.ugi, i); }
void ocns_get_res(void* h, double* v, double* s, double* i) { auto& x = Cn(h); take(x.resv, v); take(x.ress, s); take(x.resi, i); }
void ocns_set_res(void* h, const double* v, const double* s, const double* i) { auto& x = Cn(h); put(x.resv, v); put(x.ress, s); put(x.resi, i); }
void ocns_get_prcn(void* h, double* vp, double* sp, double* tp, double* vpc, double* spc, double* tpc) {
    auto& x = Cn(h);
    take(x.vprcn, vp); take(x.sprcn, sp); take(x.tprcn, tp); take(x.vpc, vpc); take(x.spc, spc); take(x.tpc, tpc);
}
// vertex 4x4 blocks (row-major), as left by setup_preconditioner; lets the projection of
// project_new_variables be checked with chosen blocks
void ocns_set_vpc(void* h, const double* vpc) { put(Cn(h).vpc, vpc); }
void ocns_rsdl(void* h, int stage) { Cn(h).rsdl(stage); }
int ocns_setup_preconditioner(void* h) { return Cn(h).setup_preconditioner(); }
void ocns_minvrt(void* h) { Cn(h).minvrt(); }
void ocns_modify_boundary_residual(void* h) { Cn(h).modify_boundary_residual(); }
void ocns_project_new_variables(void* h) { Cn(h).project_new_variables(); }
void ocns_update(void* h) { Cn(h).update(); }
void ocns_get_mg(void* h, double* res0v, double* vug) { auto& x = Cn(h); std::memcpy(res0v, x.res0v.data(), x.m.npnt * 4 * 8); if (!x.vug_frst.empty()) std::memcpy(vug, x.vug_frst.data(), x.vug_frst.size() * 8); }
double ocns_maxres(void* h, double* per_var) { return Cn(h).maxres(per_var); }
void ocns_mg_restrict(void* h) { Cn(h).mg_restrict(); }
void ocns_mg_prolongate(void* h) { Cn(h).mg_prolongate(); }
int ocns_log2p(void* h) { return Cn(h).log2p; }
double ocns_cycle(void* h, int vw, int mglvls, int itercrsn, int iterrfne) {
    try { return Cn(h).cycle(vw, 0, mglvls, itercrsn, iterrfne); } catch (...) { return -1.0; }
}
void ocns_element_rsdl(void* h, int tind, int stage, const double* uh, double* lf_re, double* lf_im) {
    auto& x = Cn(h);
    double u[4][15], lr[4][15], li[4][15];
    const int tm = x.b().tm;
    for (int n = 0; n < 4; ++n) for (int k = 0; k < tm; ++k) u[n][k] = uh[n * tm + k];
    x.element_rsdl(tind, stage, u, lr, li);
    for (int n = 0; n < 4; ++n) for (int k = 0; k < tm; ++k) { lf_re[n * tm + k] = lr[n][k]; lf_im[n * tm + k] = li[n][k]; }
}
void ocns_element_jacobian(void* h, int tind, double eps_r, double eps_a, double* K) { Cn(h).element_jacobian(tind, eps_r, eps_a, K); }
// calculate_tau at one state: Pinv (16), Tau (16) row-major, timestep
void ocns_calculate_tau(void* h, const double* pvu, double hh, double* Pinv, double* Tau, double* timestep) {
    double P[4][4], T4[4][4];
    Cn(h).calculate_tau(pvu, hh, P, T4, *timestep);
    for (int a = 0; a < 16; ++a) { Pinv[a] = P[a / 4][a % 4]; Tau[a] = T4[a / 4][a % 4]; }
}
void ocns_link_fine(void* coarse, void* fine, const int* cc_tri, const double* cc_wt, const int* fc_tri, const double* fc_wt) {
    static_cast<TriHpCns*>(coarse)->link_fine(static_cast<TriHpCns*>(fine), cc_tri, cc_wt, fc_tri, fc_wt);
}
double ocns_spectral_radius(const double* A16) {
    double A[4][4];
    for (int a = 0; a < 16; ++a) A[a / 4][a % 4] = A16[a];
    return spectral_radius4(A);
}

}  // extern "C"
