// Oracle (test infrastructure only).  See otri_hp.h for the reference map.
#include "otri_hp.h"
#include "olinalg.h"
#include <cmath>
#include <algorithm>
#include <stdexcept>

namespace oracle {

static const int S = 6;  // MXGP stride of the reference work arrays (tri_hp/tri_hp.h:31-33)

template <int NV>
TriHp<NV>::TriHp(const Mesh& mesh, const Params& p) : m(mesh), prm(p) {
    for (int l = 0; l < 3; ++l) bases.emplace_back(new TriBasis(1 << l));
    log2p = prm.log2pmax;
    sm0 = smf = bases[log2p]->sm;
    im0 = imf = bases[log2p]->im;
    ugv.assign(m.npnt * NV, 0.0); ugs.assign(m.nseg * sm0 * NV, 0.0); ugi.assign(m.ntri * im0 * NV, 0.0);
    ug0v = ugv; ug0s = ugs; ug0i = ugi;
    resv = ugv; ress = ugs; resi = ugi;
    res_rv = resv; res_rs = ress; res_ri = resi;
    res0v = resv; res0s = ress; res0i = resi;
    dresv.assign(3, resv); dress.assign(3, ress); dresi.assign(3, resi);
    vprcn.assign(m.npnt * NV, 0.0); sprcn.assign(m.nseg * NV, 0.0); tprcn.assign(m.ntri * NV, 0.0);
}

// tri_mesh/tprims.cpp:66-80 (twice the signed area)
template <int NV>
double TriHp<NV>::area(int t) const {
    const int p0 = m.tri_pnt[3 * t], p1 = m.tri_pnt[3 * t + 1], p2 = m.tri_pnt[3 * t + 2];
    const double dx1 = m.pnts[2 * p0] - m.pnts[2 * p2], dy1 = m.pnts[2 * p0 + 1] - m.pnts[2 * p2 + 1];
    const double dx2 = m.pnts[2 * p1] - m.pnts[2 * p0], dy2 = m.pnts[2 * p1 + 1] - m.pnts[2 * p0 + 1];
    return dx1 * dy2 - dy1 * dx2;
}

// tri_mesh/tprims.cpp:224-242
template <int NV>
double TriHp<NV>::inscribedradius(int t) const {
    const int p0 = m.tri_pnt[3 * t], p1 = m.tri_pnt[3 * t + 1], p2 = m.tri_pnt[3 * t + 2];
    const double dx1 = m.pnts[2 * p0] - m.pnts[2 * p2], dy1 = m.pnts[2 * p0 + 1] - m.pnts[2 * p2 + 1];
    const double dx2 = m.pnts[2 * p1] - m.pnts[2 * p0], dy2 = m.pnts[2 * p1 + 1] - m.pnts[2 * p0 + 1];
    const double ar = dx1 * dy2 - dy1 * dx2;
    const double perim = std::sqrt(dx1 * dx1 + dy1 * dy1) + std::sqrt(dx2 * dx2 + dy2 * dy2) +
                         std::sqrt((dx1 + dx2) * (dx1 + dx2) + (dy1 + dy2) * (dy1 + dy2));
    return ar / perim;
}

// curved boundary sides; an element is curved when it owns one and the finest p > 1
template <int NV>
void TriHp<NV>::set_curved(int nside, const int* sides, const double* crds) {
    seg_crv.assign(m.nseg, -1);
    scrd.assign(crds, crds + (size_t)nside * smf * 2);
    for (int r = 0; r < nside; ++r) seg_crv[sides[r]] = r;
    tri_crv.assign(m.ntri, 0);
    if (prm.log2pmax > 0)
        for (int t = 0; t < m.ntri; ++t)
            for (int i = 0; i < 3; ++i)
                if (seg_crv[m.tri_seg[3 * t + i]] >= 0) tri_crv[t] = 1;
}

// tri_hp/gtol.cpp:204-245: vertex coordinates, the side modes of curved boundary
// sides (as stored, no side sign), zero elsewhere
template <int NV>
void TriHp<NV>::crdtocht(int tind, double cht[2][15]) const {
    const TriBasis& B = b();
    for (int n = 0; n < 2; ++n)
        for (int k = 0; k < 15; ++k) cht[n][k] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int n = 0; n < 2; ++n) cht[n][i] = m.pnts[2 * m.tri_pnt[3 * tind + i] + n];
    if (B.sm == 0) return;
    int cnt = 3;
    for (int i = 0; i < 3; ++i) {
        const int sind = m.tri_seg[3 * tind + i];
        const int r = seg_crv.empty() ? -1 : seg_crv[sind];
        for (int mm = 0; mm < B.sm; ++mm) {
            for (int n = 0; n < 2; ++n) cht[n][cnt] = r >= 0 ? scrd[((size_t)r * smf + mm) * 2 + n] : 0.0;
            ++cnt;
        }
    }
}

// tri_hp/gtol.cpp:296-320
template <int NV>
void TriHp<NV>::crdtocht1d(int sind, double cht[2][8]) const {
    const TriBasis& B = b();
    const int v0 = m.seg_pnt[2 * sind], v1 = m.seg_pnt[2 * sind + 1];
    const int r = seg_crv.empty() ? -1 : seg_crv[sind];
    for (int n = 0; n < 2; ++n) {
        cht[n][0] = m.pnts[2 * v0 + n];
        cht[n][1] = m.pnts[2 * v1 + n];
        for (int mm = 0; mm < B.sm; ++mm) cht[n][mm + 2] = r >= 0 ? scrd[((size_t)r * smf + mm) * 2 + n] : 0.0;
    }
}

template <int NV>
void TriHp<NV>::set_periodic(int npair, const int* pairs) {
    const int nebd = int(m.ebdry_start.size()) - 1;
    std::vector<int> parent(m.npnt);
    for (int v = 0; v < m.npnt; ++v) parent[v] = v;
    auto find = [&](int v) { while (parent[v] != v) v = parent[v] = parent[parent[v]]; return v; };
    prdc_bdry.assign(nebd, 0);
    for (int k = 0; k < npair; ++k) {
        const int b1 = pairs[2 * k], b2 = pairs[2 * k + 1];
        prdc_bdry[b1] = prdc_bdry[b2] = 1;
        m.ebdry_essential[b1] = m.ebdry_essential[b2] = 0;
        const int o1 = m.ebdry_start[b1], o2 = m.ebdry_start[b2], n = m.ebdry_start[b1 + 1] - o1;
        if (n != m.ebdry_start[b2 + 1] - o2) throw std::runtime_error("periodic pair: side counts differ");
        for (int j = 0; j < n; ++j) {
            const int sa = m.ebdry_seg[o1 + j], sb = m.ebdry_seg[o2 + n - 1 - j];
            const int a0 = m.seg_pnt[2 * sa], a1 = m.seg_pnt[2 * sa + 1];
            const int q0 = m.seg_pnt[2 * sb], q1 = m.seg_pnt[2 * sb + 1];
            auto dx = [&](int p, int q, int r, int t, int c) {
                return (m.pnts[2 * p + c] - m.pnts[2 * q + c]) - (m.pnts[2 * r + c] - m.pnts[2 * t + c]);
            };
            const double L = std::hypot(m.pnts[2 * a1] - m.pnts[2 * a0], m.pnts[2 * a1 + 1] - m.pnts[2 * a0 + 1]);
            int m0, m1, o;
            if (std::hypot(dx(q1, a0, q0, a1, 0), dx(q1, a0, q0, a1, 1)) <= 1e-8 * L) { m0 = q1; m1 = q0; o = -1; }
            else if (std::hypot(dx(q0, a0, q1, a1, 0), dx(q0, a0, q1, a1, 1)) <= 1e-8 * L) { m0 = q0; m1 = q1; o = 1; }
            else throw std::runtime_error("periodic pair: sides are not translates of each other");
            prdc_a.push_back(sa); prdc_b.push_back(sb); prdc_o.push_back(o);
            parent[find(a0)] = find(m0);
            parent[find(a1)] = find(m1);
        }
    }
    std::vector<std::vector<int>> mem(m.npnt);
    for (int v = 0; v < m.npnt; ++v) mem[find(v)].push_back(v);
    prdc_groups.clear();
    for (int v = 0; v < m.npnt; ++v)
        if (mem[v].size() > 1) prdc_groups.push_back(mem[v]);
    prdc_sync();
}

template <int NV>
void TriHp<NV>::prdc_sync() {
    for (const auto& g : prdc_groups)
        for (size_t k = 1; k < g.size(); ++k)
            for (int n = 0; n < NV; ++n) ugv[g[k] * NV + n] = ugv[g[0] * NV + n];
    for (size_t k = 0; k < prdc_a.size(); ++k)
        for (int mm = 0; mm < smf; ++mm)
            for (int n = 0; n < NV; ++n)
                ugs[(prdc_b[k] * smf + mm) * NV + n] =
                    ((mm & 1) && prdc_o[k] < 0 ? -1.0 : 1.0) * ugs[(prdc_a[k] * smf + mm) * NV + n];
}

// comm_finish with boundary::average (tri_mesh/boundary.h:636-656): own value plus the
// partners', times 1/matches, the same on every copy
template <int NV>
void TriHp<NV>::vexchange(double* data) {
    for (const auto& g : prdc_groups)
        for (int n = 0; n < NV; ++n) {
            double sum = 0.0;
            for (int v : g) sum += data[v * NV + n];
            sum *= 1. / g.size();
            for (int v : g) data[v * NV + n] = sum;
        }
}

// side pairs: the partner's odd modes change sign when the two copies run opposite
// ways (hp_edge_bdry::smatchsolution_rcv, tri_hp/hp_boundary.cpp:748-790)
template <int NV>
void TriHp<NV>::sexchange(double* data, int stride, int mode, bool sign) {
    for (size_t k = 0; k < prdc_a.size(); ++k) {
        const double sg = (sign && (mode & 1) && prdc_o[k] < 0) ? -1.0 : 1.0;
        for (int n = 0; n < NV; ++n) {
            double& a = data[(prdc_a[k] * stride + mode) * NV + n];
            double& bb = data[(prdc_b[k] * stride + mode) * NV + n];
            const double avg = (a + sg * bb) * 0.5;
            a = avg;
            bb = sg * avg;
        }
    }
}

// tri_hp/gtol.cpp:14-55
template <int NV>
void TriHp<NV>::ugtouht(int tind) { ugtouht_to(tind, uht); }

template <int NV>
void TriHp<NV>::ugtouht_to(int tind, double uht[NV][15]) const {
    const TriBasis& B = b();
    for (int i = 0; i < 3; ++i) {
        const int indx = m.tri_pnt[3 * tind + i];
        for (int n = 0; n < NV; ++n) uht[n][i] = ugv[indx * NV + n];
    }
    int cnt = 3;
    for (int i = 0; i < 3; ++i) {
        const int sind = m.tri_seg[3 * tind + i], sign = m.tri_sgn[3 * tind + i];
        int msgn = 1;
        for (int mm = 0; mm < B.sm; ++mm) {
            for (int n = 0; n < NV; ++n) uht[n][cnt] = msgn * ugs[(sind * sm0 + mm) * NV + n];
            msgn *= sign;
            ++cnt;
        }
    }
    if (B.im > 0) {
        int indx = 0;
        for (int mm = 1; mm < B.sm; ++mm) {
            for (int k = 0; k < B.sm - mm; ++k) {
                for (int n = 0; n < NV; ++n) uht[n][cnt] = ugi[(tind * im0 + indx) * NV + n];
                ++cnt; ++indx;
            }
            indx += sm0 - B.sm;
        }
    }
}

// tri_hp/gtol.cpp:364-402 (global residual arrays use the finest-level strides)
template <int NV>
void TriHp<NV>::lftog(int tind, double* gv, double* gs, double* gi, const double lfl[NV][15]) {
    const TriBasis& B = b();
    for (int k = 0; k < 3; ++k) {
        const int g = m.tri_pnt[3 * tind + k];
        for (int n = 0; n < NV; ++n) gv[g * NV + n] += lfl[n][k];
    }
    if (B.p > 1) {
        int indx = 3;
        for (int i = 0; i < 3; ++i) {
            const int sind = m.tri_seg[3 * tind + i], sgn = m.tri_sgn[3 * tind + i];
            int msgn = 1;
            for (int mm = 0; mm < B.sm; ++mm) {
                for (int n = 0; n < NV; ++n) gs[(sind * smf + mm) * NV + n] += msgn * lfl[n][indx];
                msgn *= sgn;
                ++indx;
            }
        }
        indx = B.bm;
        for (int mm = 0; mm < B.im; ++mm) {
            for (int n = 0; n < NV; ++n) gi[(tind * imf + mm) * NV + n] += lfl[n][indx];
            ++indx;
        }
    }
}

// tri_hp/nstage.cpp:12-110
template <int NV>
void TriHp<NV>::rsdl(int stage) {
    const TriBasis& B = b();
    const int nst = prm.nstage;
    const double f = (stage < nst) ? 1.0 - prm.beta[stage] : 0.0;
    for (int i = 0; i < m.npnt * NV; ++i) { resv[i] = 0.0; res_rv[i] = (stage < nst) ? res_rv[i] * f : 0.0; }
    if (B.sm)
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) {
                    const int ix = (s * smf + k) * NV + n;
                    ress[ix] = 0.0;
                    res_rs[ix] = (stage < nst) ? res_rs[ix] * f : 0.0;
                }
    if (B.sm && B.im)
        for (int t = 0; t < m.ntri; ++t)
            for (int k = 0; k < B.im; ++k)
                for (int n = 0; n < NV; ++n) {
                    const int ix = (t * imf + k) * NV + n;
                    resi[ix] = 0.0;
                    res_ri[ix] = (stage < nst) ? res_ri[ix] * f : 0.0;
                }
    // Element residuals of a chunk are computed in parallel (OpenMP; each is a pure
    // function of the element's coefficients), then scattered by one thread in element
    // order, the order of the reference's loop: the sums are bitwise those of a serial run.
    const int CH = std::min(m.ntri, 8192);
    if (bre.size() < (size_t)CH * NV * 15) {
        bre.resize((size_t)CH * NV * 15);
        bim.resize((size_t)CH * NV * 15);
    }
    for (int t0 = 0; t0 < m.ntri; t0 += CH) {
        const int t1 = std::min(m.ntri, t0 + CH);
#pragma omp parallel for schedule(static)
        for (int tind = t0; tind < t1; ++tind) {
            double uh[NV][15];
            ugtouht_to(tind, uh);
            double(*lre)[15] = reinterpret_cast<double(*)[15]>(&bre[(size_t)(tind - t0) * NV * 15]);
            double(*lim)[15] = reinterpret_cast<double(*)[15]>(&bim[(size_t)(tind - t0) * NV * 15]);
            element_rsdl(tind, stage, uh, lre, lim);
        }
        for (int tind = t0; tind < t1; ++tind) {
            const double(*lre)[15] = reinterpret_cast<const double(*)[15]>(&bre[(size_t)(tind - t0) * NV * 15]);
            const double(*lim)[15] = reinterpret_cast<const double(*)[15]>(&bim[(size_t)(tind - t0) * NV * 15]);
            lftog(tind, resv.data(), ress.data(), resi.data(), lim);
            lftog(tind, res_rv.data(), res_rs.data(), res_ri.data(), lre);
        }
    }
    for (int i = 0; i < m.npnt * NV; ++i) resv[i] += res_rv[i];
    if (B.sm > 0) {
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) ress[(s * smf + k) * NV + n] += res_rs[(s * smf + k) * NV + n];
        if (B.im > 0)
            for (int t = 0; t < m.ntri; ++t)
                for (int k = 0; k < B.im; ++k)
                    for (int n = 0; n < NV; ++n) resi[(t * imf + k) * NV + n] += res_ri[(t * imf + k) * NV + n];
    }
    edge_rsdl();
    if (coarse_flag) mg_source();
}

// tri_hp/nstage.cpp:113-129
template <int NV>
void TriHp<NV>::mg_source() {
    const TriBasis& B = b();
    std::vector<double>& dv = dresv[log2p];
    std::vector<double>& ds = dress[log2p];
    std::vector<double>& di = dresi[log2p];
    if (isfrst) {
        for (int i = 0; i < m.npnt * NV; ++i) dv[i] = prm.fadd * res0v[i] - resv[i];
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) {
                    const int ix = (s * smf + k) * NV + n;
                    ds[ix] = prm.fadd * res0s[ix] - ress[ix];
                }
        for (int t = 0; t < m.ntri; ++t)
            for (int k = 0; k < B.im; ++k)
                for (int n = 0; n < NV; ++n) {
                    const int ix = (t * imf + k) * NV + n;
                    di[ix] = prm.fadd * res0i[ix] - resi[ix];
                }
        isfrst = false;
    }
    for (int i = 0; i < m.npnt * NV; ++i) resv[i] += dv[i];
    for (int s = 0; s < m.nseg; ++s)
        for (int k = 0; k < B.sm; ++k)
            for (int n = 0; n < NV; ++n) ress[(s * smf + k) * NV + n] += ds[(s * smf + k) * NV + n];
    for (int t = 0; t < m.ntri; ++t)
        for (int k = 0; k < B.im; ++k)
            for (int n = 0; n < NV; ++n) resi[(t * imf + k) * NV + n] += di[(t * imf + k) * NV + n];
}

// tri_hp/ins/setup_preconditioner.cpp:213-219 then tri_hp/setup_preconditioner.cpp:68-76
// (the reference's loop stops at the failing element, so only the first ntri_done
// elements contribute)
template <int NV>
void TriHp<NV>::assemble_prcn(int ntri_done) {
    const TriBasis& B = b();
    std::fill(vprcn.begin(), vprcn.end(), 0.0);
    if (B.sm > 0) std::fill(sprcn.begin(), sprcn.end(), 0.0);
    for (int tind = 0; tind < ntri_done; ++tind)
        for (int i = 0; i < 3; ++i) {
            const int vv = m.tri_pnt[3 * tind + i];
            for (int n = 0; n < NV; ++n) vprcn[vv * NV + n] += tprcn[tind * NV + n];
            if (B.sm > 0) {
                const int side = m.tri_seg[3 * tind + i];
                for (int n = 0; n < NV; ++n) sprcn[side * NV + n] += tprcn[tind * NV + n];
            }
        }
    // tri_hp/setup_preconditioner.cpp:24-36: the sums averaged over periodic copies
    vexchange(vprcn.data());
    if (log2p) sexchange(sprcn.data(), 1, 0, false);
    for (int i = 0; i < m.npnt * NV; ++i) vprcn[i] = 1.0 / (B.vdiag * vprcn[i]);
    if (log2p)
        for (int i = 0; i < m.nseg * NV; ++i) sprcn[i] = 1.0 / sprcn[i];
}

// tri_hp/hp_boundary.cpp:680-693
template <int NV>
void TriHp<NV>::vdirichlet() {
    const int nebd = int(m.ebdry_start.size()) - 1;
    for (int e = 0; e < nebd; ++e) {
        const int mask = m.ebdry_essential[e];
        if (!mask) continue;
        int sind = -1;
        for (int j = m.ebdry_start[e]; j < m.ebdry_start[e + 1]; ++j) {
            sind = m.ebdry_seg[j];
            const int v0 = m.seg_pnt[2 * sind];
            for (int n = 0; n < NV; ++n) if (mask & (1 << n)) resv[v0 * NV + n] = 0.0;
        }
        const int v1 = m.seg_pnt[2 * sind + 1];
        for (int n = 0; n < NV; ++n) if (mask & (1 << n)) resv[v1 * NV + n] = 0.0;
    }
}

// tri_hp/hp_boundary.cpp:695-704
template <int NV>
void TriHp<NV>::sdirichlet(int mode) {
    const int nebd = int(m.ebdry_start.size()) - 1;
    for (int e = 0; e < nebd; ++e) {
        const int mask = m.ebdry_essential[e];
        for (int j = m.ebdry_start[e]; j < m.ebdry_start[e + 1]; ++j) {
            const int sind = m.ebdry_seg[j];
            for (int n = 0; n < NV; ++n) if (mask & (1 << n)) ress[(sind * smf + mode) * NV + n] = 0.0;
        }
    }
}

// tri_hp/setup_preconditioner.cpp:102-373 (diagonal preconditioner branch)
template <int NV>
void TriHp<NV>::minvrt() {
    const TriBasis& B = b();
    const int sm = B.sm, im = B.im;
    if (sm > 0) {
        for (int sind = 0; sind < m.nseg; ++sind)
            for (int k = 0; k < sm; ++k)
                for (int i = 0; i < 2; ++i) {
                    const int v0 = m.seg_pnt[2 * sind + i];
                    for (int n = 0; n < NV; ++n)
                        resv[v0 * NV + n] -= B.sfmv[i * sm + k] * ress[(sind * smf + k) * NV + n];
                }
        if (im > 0)
            for (int tind = 0; tind < m.ntri; ++tind) {
                int indx2 = 3;
                for (int i = 0; i < 3; ++i) {
                    const int v0 = m.tri_pnt[3 * tind + i];
                    for (int k = 0; k < im; ++k)
                        for (int n = 0; n < NV; ++n)
                            resv[v0 * NV + n] -= B.ifmb[i * im + k] * resi[(tind * imf + k) * NV + n];
                    const int sind = m.tri_seg[3 * tind + i], sgn = m.tri_sgn[3 * tind + i];
                    int msgn = 1;
                    for (int j = 0; j < sm; ++j) {
                        for (int k = 0; k < im; ++k)
                            for (int n = 0; n < NV; ++n)
                                ress[(sind * smf + j) * NV + n] -= msgn * B.ifmb[indx2 * im + k] * resi[(tind * imf + k) * NV + n];
                        msgn *= sgn;
                        ++indx2;
                    }
                }
            }
    }
    for (int i = 0; i < m.npnt * NV; ++i) resv[i] *= vprcn[i];
    vexchange(resv.data());  // setup_preconditioner.cpp:171-176
    vdirichlet();
    if (sm == 0) return;

    for (int tind = 0; tind < m.ntri; ++tind) {
        double uh[NV][3];
        for (int i = 0; i < 3; ++i) {
            const int v0 = m.tri_pnt[3 * tind + i];
            for (int n = 0; n < NV; ++n) uh[n][i] = resv[v0 * NV + n] * tprcn[tind * NV + n];
        }
        for (int i = 0; i < 3; ++i) {
            const int sind = m.tri_seg[3 * tind + i], sgn = m.tri_sgn[3 * tind + i];
            for (int j = 0; j < 3; ++j) {
                const int indx1 = (i + j) % 3;
                int msgn = 1;
                for (int k = 0; k < sm; ++k) {
                    for (int n = 0; n < NV; ++n)
                        ress[(sind * smf + k) * NV + n] -= msgn * B.vfms[j * sm + k] * uh[n][indx1];
                    msgn *= sgn;
                }
            }
        }
    }

    for (int mode = 0; mode < sm - 1; ++mode) {
        for (int s = 0; s < m.nseg; ++s)
            for (int n = 0; n < NV; ++n) ress[(s * smf + mode) * NV + n] *= sprcn[s * NV + n] * B.sdiag[mode];
        sexchange(ress.data(), smf, mode, true);  // :255-258
        sdirichlet(mode);
        for (int tind = 0; tind < m.ntri; ++tind) {
            int side[3], sign[3];
            double uh[NV][3];
            for (int i = 0; i < 3; ++i) {
                side[i] = m.tri_seg[3 * tind + i];
                sign[i] = m.tri_sgn[3 * tind + i];
                const int sgn = (mode % 2 ? sign[i] : 1);
                for (int n = 0; n < NV; ++n) uh[n][i] = sgn * ress[(side[i] * smf + mode) * NV + n] * tprcn[tind * NV + n];
            }
            for (int i = 0; i < 3; ++i) {
                int msgn = ((mode + 1) % 2 ? sign[i] : 1);
                for (int mm = mode + 1; mm < sm; ++mm) {
                    for (int j = 0; j < 3; ++j) {
                        const int indx = (i + j) % 3;
                        for (int n = 0; n < NV; ++n)
                            ress[(side[i] * smf + mm) * NV + n] -= msgn * B.sfms[(mode * sm + mm) * 3 + j] * uh[n][indx];
                    }
                    msgn *= sign[i];
                }
            }
        }
    }
    const int mode = sm - 1;
    for (int s = 0; s < m.nseg; ++s)
        for (int n = 0; n < NV; ++n) ress[(s * smf + mode) * NV + n] *= sprcn[s * NV + n] * B.sdiag[mode];
    sexchange(ress.data(), smf, mode, true);  // :331-334
    sdirichlet(mode);
    if (im == 0) return;

    // each element reads the solved vertex and side values and writes only its own
    // interior modes: independent, so OpenMP threads share the loop
#pragma omp parallel for schedule(static)
    for (int tind = 0; tind < m.ntri; ++tind) {
        dpbtrsnu2_literal(B.idiag.data(), B.ibwth + 1, im, B.ibwth, &resi[tind * imf * NV], NV);
        // restouht_bdry (tri_hp/setup_preconditioner.cpp:375-399)
        double uh[NV][15];
        for (int i = 0; i < 3; ++i) {
            const int indx = m.tri_pnt[3 * tind + i];
            for (int n = 0; n < NV; ++n) uh[n][i] = resv[indx * NV + n];
        }
        int cnt = 3;
        for (int i = 0; i < 3; ++i) {
            const int indx = m.tri_seg[3 * tind + i], sign = m.tri_sgn[3 * tind + i];
            int msgn = 1;
            for (int mm = 0; mm < sm; ++mm) {
                for (int n = 0; n < NV; ++n) uh[n][cnt] = msgn * ress[(indx * smf + mm) * NV + n];
                msgn *= sign;
                ++cnt;
            }
        }
        for (int k = 0; k < im; ++k) {
            for (int n = 0; n < NV; ++n) resi[(tind * imf + k) * NV + n] /= tprcn[tind * NV + n];
            for (int i = 0; i < B.bm; ++i)
                for (int n = 0; n < NV; ++n) resi[(tind * imf + k) * NV + n] -= B.bfmi[i * im + k] * uh[n][i];
        }
    }
}

// tri_hp/nstage.cpp:256-427
template <int NV>
void TriHp<NV>::update() {
    ug0v = ugv; ug0s = ugs; ug0i = ugi;
    const TriBasis& B = b();
    for (int stage = 0; stage < prm.nstage; ++stage) {
        rsdl(stage);
        minvrt();
        const double cflalpha = prm.alpha[stage] * prm.cfl[log2p];
        for (int i = 0; i < m.npnt * NV; ++i) ugv[i] = ug0v[i] - cflalpha * resv[i];
        if (B.sm > 0) {
            for (int s = 0; s < m.nseg; ++s)
                for (int k = 0; k < B.sm; ++k)
                    for (int n = 0; n < NV; ++n)
                        ugs[(s * sm0 + k) * NV + n] = ug0s[(s * sm0 + k) * NV + n] - cflalpha * ress[(s * smf + k) * NV + n];
            if (B.im > 0)
                for (int t = 0; t < m.ntri; ++t) {
                    int indx = 0, indx1 = 0;
                    for (int mm = 1; mm < B.sm; ++mm) {
                        for (int k = 0; k < B.sm - mm; ++k) {
                            for (int n = 0; n < NV; ++n)
                                ugi[(t * im0 + indx1) * NV + n] = ug0i[(t * im0 + indx1) * NV + n] - cflalpha * resi[(t * imf + indx) * NV + n];
                            ++indx; ++indx1;
                        }
                        indx1 += sm0 - B.sm;
                    }
                }
        }
    }
}

// tri_hp/allocate.cpp:496-525
template <int NV>
double TriHp<NV>::maxres(double* per_var) const {
    double mxr[NV] = {0, 0, 0};
    for (int i = 0; i < m.npnt; ++i)
        for (int n = 0; n < NV; ++n) mxr[n] = std::max(mxr[n], std::fabs(resv[i * NV + n]));
    double e = 0.0;
    for (int n = 0; n < NV; ++n) { e = std::max(e, mxr[n]); if (per_var) per_var[n] = mxr[n]; }
    return e;
}

// tri_hp/movco.cpp:4-40 (p-level branch)
template <int NV>
void TriHp<NV>::link_fine(TriHp* f, const int* cctri, const double* ccwt, const int* fctri, const double* fcwt) {
    fine = f;
    coarse_level = true;
    cc_tri.assign(cctri, cctri + f->m.npnt);
    cc_wt.assign(ccwt, ccwt + 3 * f->m.npnt);
    fc_tri.assign(fctri, fctri + m.npnt);
    fc_wt.assign(fcwt, fcwt + 3 * m.npnt);
    vug_frst.assign(m.npnt * NV, 0.0);
}

// tri_hp/movco.cpp:4-106.  p-levels (!coarse_level): truncate the residual into res0.
// h-levels: res0 = sum over fine vertices of weight x fine residual on the coarse
// triangle that holds the fine vertex; ug = fine ug interpolated at the coarse
// vertices, kept in vug_frst.  (The reference shares the residual arrays between the
// levels of a block; here the fine residual is read from the fine level.)
template <int NV>
void TriHp<NV>::mg_restrict() {
    isfrst = true;
    coarse_flag = true;
    if (coarse_level) {
        std::fill(res0v.begin(), res0v.end(), 0.0);
        for (int i = 0; i < fine->m.npnt; ++i) {
            const int tind = cc_tri[i];
            for (int j = 0; j < 3; ++j) {
                const int v0 = m.tri_pnt[3 * tind + j];
                for (int n = 0; n < NV; ++n) res0v[v0 * NV + n] += cc_wt[3 * i + j] * fine->resv[i * NV + n];
            }
        }
        for (int i = 0; i < m.npnt; ++i) {
            const int tind = fc_tri[i];
            for (int n = 0; n < NV; ++n) ugv[i * NV + n] = 0.0;
            for (int j = 0; j < 3; ++j) {
                const int fv = fine->m.tri_pnt[3 * tind + j];
                for (int n = 0; n < NV; ++n) ugv[i * NV + n] += fc_wt[3 * i + j] * fine->ugv[fv * NV + n];
            }
            for (int n = 0; n < NV; ++n) vug_frst[i * NV + n] = ugv[i * NV + n];
        }
        return;
    }
    --log2p;
    const TriBasis& B = b();
    res0v = resv;
    if (B.p > 1)
        for (int s = 0; s < m.nseg; ++s)
            for (int k = 0; k < B.sm; ++k)
                for (int n = 0; n < NV; ++n) res0s[(s * smf + k) * NV + n] = ress[(s * smf + k) * NV + n];
    if (B.p > 2)
        for (int t = 0; t < m.ntri; ++t) {
            int indx = 0, indx1 = 0;
            for (int mm = 1; mm < B.sm; ++mm) {
                for (int k = 0; k < B.sm - mm; ++k) {
                    for (int n = 0; n < NV; ++n) res0i[(t * imf + indx) * NV + n] = resi[(t * imf + indx1) * NV + n];
                    ++indx; ++indx1;
                }
                indx1 += B.p;
            }
        }
}

// tri_hp/movfin.cpp:7-82.  p-levels: back to the finer p.  h-levels: the coarse
// correction (vug_frst - ug, negated) is interpolated to the fine vertices through
// the fine vertices' coarse triangles, written into the fine residual and added to
// the fine ug.
template <int NV>
void TriHp<NV>::mg_prolongate() {
    if (coarse_level) {
        for (int i = 0; i < m.npnt * NV; ++i) vug_frst[i] -= ugv[i];
        for (int i = 0; i < fine->m.npnt; ++i) {
            const int tind = cc_tri[i];
            for (int n = 0; n < NV; ++n) fine->resv[i * NV + n] = 0.0;
            for (int j = 0; j < 3; ++j) {
                const int ind = m.tri_pnt[3 * tind + j];
                for (int n = 0; n < NV; ++n) fine->resv[i * NV + n] -= cc_wt[3 * i + j] * vug_frst[ind * NV + n];
            }
        }
        for (int i = 0; i < fine->m.npnt * NV; ++i) fine->ugv[i] += fine->resv[i];
        return;
    }
    ++log2p;
    if (log2p == prm.log2pmax) coarse_flag = false;
}

template <int NV>
void TriHp<NV>::element_jacobian(int tind, double eps_r, double eps_a, double* K) {
    const TriBasis& B = b();
    const int T = B.tm, N = T * NV;
    ugtouht(tind);
    double dw[NV] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
        for (int n = 0; n < NV; ++n) dw[n] = dw[n] + std::fabs(uht[n][i]);
    for (int n = 0; n < NV; ++n) dw[n] = dw[n] * eps_r + eps_a;
    double lre[NV][15], lim[NV][15], Rbar[NV][15];
    element_rsdl(tind, 0, uht, lre, lim);
    for (int i = 0; i < T; ++i)
        for (int n = 0; n < NV; ++n) Rbar[n][i] = lre[n][i] + lim[n][i];
    int kcol = 0;
    for (int mode = 0; mode < T; ++mode)
        for (int var = 0; var < NV; ++var) {
            uht[var][mode] += dw[var];
            element_rsdl(tind, 0, uht, lre, lim);
            int krow = 0;
            for (int i = 0; i < T; ++i)
                for (int n = 0; n < NV; ++n) K[(krow++) * N + kcol] = (lre[n][i] + lim[n][i] - Rbar[n][i]) / dw[var];
            ++kcol;
            uht[var][mode] -= dw[var];
        }
}

// tri_mesh/blocks.cpp:1267-1348 with ngrid = 1 (p-levels only)
template <int NV>
double TriHp<NV>::cycle(int vw, int lvl, int mglvls, int itercrsn, int iterrfne) {
    double error = 0.0;
    for (int vcount = 0; vcount < vw; ++vcount) {
        if (setup_preconditioner()) throw std::runtime_error("setup_preconditioner error");
        for (int it = 0; it < itercrsn; ++it) update();
        if (lvl + vcount == 0) {
            error = maxres();
            if (!(error > 0.0)) throw std::runtime_error("maxres not positive");
        }
        if (lvl == mglvls - 1) return error;
        rsdl(prm.nstage);
        mg_restrict();
        cycle(vw, lvl + 1, mglvls, itercrsn, iterrfne);
        mg_prolongate();
        if (iterrfne) {
            if (setup_preconditioner()) throw std::runtime_error("setup_preconditioner error");
        }
        for (int it = 0; it < iterrfne; ++it) update();
    }
    return error;
}

template class TriHp<3>;
template class TriHp<4>;

}  // namespace oracle
