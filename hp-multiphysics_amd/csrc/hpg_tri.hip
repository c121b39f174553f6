// Context and C-ABI of the tri_hp residual/relaxation library (include/hpgpu.h).
//
// One context = one solution block.  The context owns device mirrors of the
// block's mesh and solution arrays, the per-level operator packs, the element
// colouring used for the deterministic scatter, and one HIP stream.  Everything
// between hpg_set_state and hpg_get_state stays resident in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hpgpu.h"
#include "tri_kernels.hip"
#include "tri_ops.h"

using namespace hpg;

namespace {

struct LevelDims {
    int p, tm, sm, im, bm, G, GT, KT, PROJ, DER, INT, PACK;
    explicit LevelDims(int p_) : p(p_) {
        tm = (p + 1) * (p + 2) / 2; sm = p - 1; im = (p - 1) * (p - 2) / 2; bm = 3 * p;
        G = (p + 1) * (p + 1); GT = (G + 15) / 16; KT = (tm + 3) / 4;
        PROJ = GT * KT * 64; DER = GT * GT * 4 * 64; INT = GT * 4 * 64;
        PACK = 3 * PROJ + 2 * DER + 2 * INT;
    }
};

// Lane-ready operator pack: A operand of every MFMA K-step, 64 doubles per step.
std::vector<double> build_pack(const TriOps& o) {
    const LevelDims d(o.p);
    std::vector<double> pk(d.PACK, 0.0);
    const std::vector<double>* proj[3] = {&o.P, &o.Pr, &o.Ps};
    for (int op = 0; op < 3; ++op)
        for (int mt = 0; mt < d.GT; ++mt)
            for (int s = 0; s < d.KT; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int g = 16 * mt + (l & 15), k = 4 * s + (l >> 4);
                    pk[op * d.PROJ + (mt * d.KT + s) * 64 + l] = (g < d.G && k < d.tm) ? (*proj[op])[g * d.tm + k] : 0.0;
                }
    const std::vector<double>* der[2] = {&o.Dr, &o.Ds};
    for (int op = 0; op < 2; ++op)
        for (int mt = 0; mt < d.GT; ++mt)
            for (int kt = 0; kt < d.GT; ++kt)
                for (int i = 0; i < 4; ++i)
                    for (int l = 0; l < 64; ++l) {
                        const int g = 16 * mt + (l & 15), k = 16 * kt + 4 * i + (l >> 4);
                        pk[3 * d.PROJ + op * d.DER + ((mt * d.GT + kt) * 4 + i) * 64 + l] =
                            (g < d.G && k < d.G) ? (*der[op])[g * d.G + k] : 0.0;
                    }
    const std::vector<double>* in[2] = {&o.Ir, &o.Is};
    for (int op = 0; op < 2; ++op)
        for (int kt = 0; kt < d.GT; ++kt)
            for (int i = 0; i < 4; ++i)
                for (int l = 0; l < 64; ++l) {
                    const int m = l & 15, k = 16 * kt + 4 * i + (l >> 4);
                    pk[3 * d.PROJ + 2 * d.DER + op * d.INT + (kt * 4 + i) * 64 + l] =
                        (m < d.tm && k < d.G) ? (*in[op])[m * d.G + k] : 0.0;
                }
    return pk;
}

// Greedy colouring: items whose vertex sets intersect get different colours.
// Returns the item order grouped by colour (ascending inside a colour) and offsets.
void colour(int nitem, int per, const int* verts, int npnt, std::vector<int>& order, std::vector<int>& off) {
    std::vector<uint64_t> used(npnt, 0);
    std::vector<int> col(nitem);
    int ncol = 0;
    for (int t = 0; t < nitem; ++t) {
        uint64_t m = 0;
        for (int k = 0; k < per; ++k) m |= used[verts[per * t + k]];
        int c = 0;
        while (c < 64 && ((m >> c) & 1)) ++c;
        if (c == 64) throw std::runtime_error("colouring needs more than 64 colours");
        col[t] = c;
        ncol = std::max(ncol, c + 1);
        for (int k = 0; k < per; ++k) used[verts[per * t + k]] |= (uint64_t(1) << c);
    }
    off.assign(ncol + 1, 0);
    for (int t = 0; t < nitem; ++t) off[col[t] + 1]++;
    for (int c = 0; c < ncol; ++c) off[c + 1] += off[c];
    order.resize(nitem);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int t = 0; t < nitem; ++t) order[fill[col[t]]++] = t;
}

struct Level {
    int p = 0;
    std::unique_ptr<TriOps> ops;
    double* pack = nullptr;
    double* Pmat = nullptr;
    double* lump = nullptr;
    double* edge = nullptr;  // g1 | dg1 | w1
    int o_sfmv, o_ifmb, o_vfms, o_sdiag, o_sfms, o_bfmi, o_idiag;
};

}  // namespace

struct hpg_ctx {
    hipStream_t stream = nullptr;
    hpg_params prm;
    int log2pmax = 0, log2p = 0;
    bool coarse_flag = false, isfrst = false;
    int precision = HPG_FP64, prcn_mode = HPG_PRCN_ANALYTIC;
    int npnt = 0, nseg = 0, ntri = 0, sm0 = 0, im0 = 0;
    std::vector<void*> allocs;
    double* pnts = nullptr;
    int *tri_pnt = nullptr, *tri_seg = nullptr, *tri_sgn = nullptr, *seg_pnt = nullptr;
    int* order = nullptr;
    std::vector<int> coff;
    int* eseg = nullptr;
    std::vector<int> eoff;
    int *vs_ptr, *vs_side, *vs_end, *vt_ptr, *vt_tri, *vt_loc, *st_tri, *st_loc;
    unsigned char *vess, *sess;
    double *ugv, *ugs, *ugi, *ug0v, *ug0s, *ug0i;
    double *resv, *ress, *resi, *rrv, *rrs, *rri, *r0v, *r0s, *r0i;
    double *dresv[3], *dress[3], *dresi[3];
    double *vprcn, *sprcn, *tprcn;
    int* d_err;
    unsigned long long* d_max;
    Level lv[3];
    std::string err;
    // optional per-launch timing of the element-residual kernel (HIP events on `stream`)
    bool timing = false;
    std::vector<hipEvent_t> events;
    struct Rec { int full, nelem; };
    std::vector<Rec> recs;

    hipEvent_t event(size_t k) {
        while (events.size() <= k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
            events.push_back(e);
        }
        return events[k];
    }

    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        allocs.push_back(p);
        hipMemsetAsync(p, 0, n * sizeof(T), stream);
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const T* h, size_t n) {
        T* d = alloc<T>(n);
        if (n && hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream) != hipSuccess)
            throw std::runtime_error("hipMemcpy H2D failed");
        return d;
    }
    ~hpg_ctx() {
        if (stream) hipStreamSynchronize(stream);
        for (void* p : allocs) hipFree(p);
        for (hipEvent_t e : events) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
};

#define HPG_TRY(ctx, ...)                                                      \
    try {                                                                      \
        __VA_ARGS__;                                                           \
    } catch (const std::exception& ex) {                                       \
        if (ctx) (ctx)->err = ex.what();                                       \
        return -1;                                                             \
    }

static void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static inline unsigned nblk(long n, int bs = 256) { return unsigned((n + bs - 1) / bs); }

namespace {

void launch_modes(hpg_ctx* c, int op, long nent, int nm, int mstride, double* x, double* y, const double* z, double k) {
    if (nent <= 0 || nm <= 0) return;
    const long n = nent * nm * NV;
    hipLaunchKernelGGL(k_modes, dim3(nblk(n)), dim3(256), 0, c->stream, op, nent, nm, mstride, x, y, z, k);
}

// apply a k_modes op over the active v/s/i modes of the current level
void modes_vsi(hpg_ctx* c, int op, double* xv, double* xs, double* xi, double* yv, double* ys, double* yi,
               const double* zv, const double* zs, const double* zi, double k) {
    const LevelDims d(1 << c->log2p);
    launch_modes(c, op, c->npnt, 1, 1, xv, yv, zv, k);
    if (d.sm > 0) launch_modes(c, op, c->nseg, d.sm, c->sm0, xs, ys, zs, k);
    if (d.im > 0) launch_modes(c, op, c->ntri, d.im, c->im0, xi, yi, zi, k);
}

template <int P>
void launch_elem(hpg_ctx* c, const ElemArgs& a0, bool full) {
    for (size_t col = 0; col + 1 < c->coff.size(); ++col) {
        ElemArgs a = a0;
        a.begin = c->coff[col];
        a.end = c->coff[col + 1];
        const int nlogical = (a.end - a.begin + 63) / 64;
        if (nlogical == 0) continue;
        const int nb = ((nlogical + 7) / 8) * 8;
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full)
            hipLaunchKernelGGL((k_elem_ins<P, true>), dim3(nb), dim3(256), 0, c->stream, a, nlogical);
        else
            hipLaunchKernelGGL((k_elem_ins<P, false>), dim3(nb), dim3(256), 0, c->stream, a, nlogical);
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 1 : 0, a.end - a.begin});
        }
    }
}

MinvArgs minv_args(hpg_ctx* c) {
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    MinvArgs a;
    a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.vs_ptr = c->vs_ptr; a.vs_side = c->vs_side; a.vs_end = c->vs_end;
    a.vt_ptr = c->vt_ptr; a.vt_tri = c->vt_tri; a.vt_loc = c->vt_loc;
    a.st_tri = c->st_tri; a.st_loc = c->st_loc;
    a.vess = c->vess; a.sess = c->sess;
    a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.vprcn = c->vprcn; a.sprcn = c->sprcn; a.tprcn = c->tprcn;
    a.lump = L.lump;
    a.npnt = c->npnt; a.nseg = c->nseg; a.ntri = c->ntri;
    a.sm = d.sm; a.im = d.im; a.bm = d.bm; a.sm0 = c->sm0; a.im0 = c->im0;
    a.o_sfmv = L.o_sfmv; a.o_ifmb = L.o_ifmb; a.o_vfms = L.o_vfms; a.o_sdiag = L.o_sdiag;
    a.o_sfms = L.o_sfms; a.o_bfmi = L.o_bfmi; a.o_idiag = L.o_idiag;
    return a;
}

void do_rsdl(hpg_ctx* c, int stage) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    modes_vsi(c, 0, c->resv, c->ress, c->resi, c->rrv, c->rrs, c->rri, nullptr, nullptr, nullptr, f);
    const Level& L = c->lv[c->log2p];
    ElemArgs a;
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi;
    a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.rrv = c->rrv; a.rrs = c->rrs; a.rri = c->rri;
    a.ops = L.pack; a.sm0 = c->sm0; a.im0 = c->im0;
    a.rho = c->prm.rho; a.mu = c->prm.mu; a.adis = c->prm.adis; a.beta = beta;
    const bool full = beta > 0.0;
    switch (L.p) {
        case 1: launch_elem<1>(c, a, full); break;
        case 2: launch_elem<2>(c, a, full); break;
        case 4: launch_elem<4>(c, a, full); break;
        default: throw std::runtime_error("unsupported p");
    }
    modes_vsi(c, 1, c->resv, c->ress, c->resi, c->rrv, c->rrs, c->rri, nullptr, nullptr, nullptr, 0.0);
    const LevelDims d(L.p);
    EdgeArgs e;
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.resv = c->resv; e.ress = c->ress;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.gpx = d.p + 1; e.sm = d.sm; e.sm0 = c->sm0; e.rho = c->prm.rho; e.bc_pressure = c->prm.bc_pressure;
    for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
        e.begin = c->eoff[col];
        e.end = c->eoff[col + 1];
        if (e.end > e.begin) hipLaunchKernelGGL(k_edge_ins, dim3(nblk(e.end - e.begin, 64)), dim3(64), 0, c->stream, e);
    }
    if (c->coarse_flag) {
        const int l = c->log2p;
        if (c->isfrst) {
            modes_vsi(c, 2, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], c->r0v, c->r0s, c->r0i,
                      c->prm.fadd);
            c->isfrst = false;
        } else {
            modes_vsi(c, 1, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], nullptr, nullptr,
                      nullptr, 0.0);
        }
    }
    check(hipGetLastError(), "rsdl launch");
}

void do_minvrt(hpg_ctx* c) {
    MinvArgs a = minv_args(c);
    hipLaunchKernelGGL(k_minv_vrtx, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, a);
    if (a.sm > 0) {
        hipLaunchKernelGGL(k_minv_side, dim3(nblk(c->nseg)), dim3(256), 0, c->stream, a);
        for (int mode = 0; mode < a.sm - 1; ++mode)
            hipLaunchKernelGGL(k_minv_sub, dim3(nblk(c->nseg)), dim3(256), 0, c->stream, a, mode);
        hipLaunchKernelGGL(k_minv_sscale, dim3(nblk(c->nseg)), dim3(256), 0, c->stream, a);
        if (a.im > 0) hipLaunchKernelGGL(k_minv_int, dim3(nblk(c->ntri)), dim3(256), 0, c->stream, a);
    }
    check(hipGetLastError(), "minvrt launch");
}

int do_setup(hpg_ctx* c) {
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    check(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream), "memset");
    SetupArgs s;
    s.pnts = c->pnts; s.tri_pnt = c->tri_pnt; s.tri_seg = c->tri_seg; s.tri_sgn = c->tri_sgn;
    s.ugv = c->ugv; s.ugs = c->ugs; s.ugi = c->ugi; s.Pmat = L.Pmat; s.tprcn = c->tprcn; s.err = c->d_err;
    s.ntri = c->ntri; s.p = d.p; s.sm = d.sm; s.im = d.im; s.tm = d.tm; s.G = d.G; s.sm0 = c->sm0; s.im0 = c->im0;
    s.rho = c->prm.rho; s.mu = c->prm.mu; s.adis = c->prm.adis;
    hipLaunchKernelGGL(k_setup_elem, dim3(nblk(c->ntri, 128)), dim3(128), 0, c->stream, s);
    hipLaunchKernelGGL(k_setup_vrtx, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->vt_ptr, c->vt_tri,
                       c->tprcn, c->vprcn, L.ops->vdiag);
    if (d.sm > 0)
        hipLaunchKernelGGL(k_setup_side, dim3(nblk(c->nseg)), dim3(256), 0, c->stream, c->nseg, c->st_tri, c->tprcn,
                           c->sprcn);
    int herr = 0;
    check(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "memcpy err");
    check(hipStreamSynchronize(c->stream), "sync");
    return herr;
}

// max over (nent, nm of mstride, NV) of |x|, per variable
void absmax(hpg_ctx* c, long nent, int nm, int mstride, const double* x, const unsigned char* mask, double out[NV]) {
    unsigned long long h[NV] = {0, 0, 0};
    if (nent > 0 && nm > 0) {
        check(hipMemsetAsync(c->d_max, 0, NV * sizeof(unsigned long long), c->stream), "memset");
        const long n = nent * nm;
        const unsigned nb = std::min<unsigned>(nblk(n), 2048u);
        hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(256), 0, c->stream, nent, nm, mstride, x, mask, c->d_max);
        check(hipMemcpyAsync(h, c->d_max, sizeof(h), hipMemcpyDeviceToHost, c->stream), "memcpy");
        check(hipStreamSynchronize(c->stream), "sync");
    }
    for (int k = 0; k < NV; ++k) {
        double v;
        std::memcpy(&v, &h[k], sizeof(double));
        out[k] = v;
    }
}

}  // namespace

extern "C" {

int hpg_create(const hpg_tri_mesh* m, const hpg_params* prm, hpg_ctx** out) {
    if (!m || !prm || !out) return -2;
    *out = nullptr;
    if (prm->physics != HPG_PHYSICS_INS || prm->log2p < 0 || prm->log2p > 2 || prm->nstage < 1 || prm->nstage > 7)
        return -2;
    std::unique_ptr<hpg_ctx> c(new hpg_ctx());
    try {
        check(hipStreamCreate(&c->stream), "hipStreamCreate");
        c->prm = *prm;
        c->prm.alpha[prm->nstage] = 1.0;
        c->prm.beta[prm->nstage] = 1.0;
        c->log2pmax = c->log2p = prm->log2p;
        const int p0 = 1 << prm->log2p;
        c->sm0 = p0 - 1;
        c->im0 = (p0 - 1) * (p0 - 2) / 2;
        c->npnt = m->npnt; c->nseg = m->nseg; c->ntri = m->ntri;
        const int np = m->npnt, ns = m->nseg, nt = m->ntri;

        c->pnts = c->upload(m->pnts, 2 * (size_t)np);
        c->tri_pnt = c->upload(m->tri_pnt, 3 * (size_t)nt);
        c->tri_seg = c->upload(m->tri_seg, 3 * (size_t)nt);
        c->tri_sgn = c->upload(m->tri_sgn, 3 * (size_t)nt);
        c->seg_pnt = c->upload(m->seg_pnt, 2 * (size_t)ns);

        std::vector<int> order, off;
        colour(nt, 3, m->tri_pnt, np, order, off);
        c->order = c->upload(order.data(), order.size());
        c->coff = off;
        const int neseg = m->ebdry_start[m->nebd];
        std::vector<int> ev(2 * (size_t)neseg);
        for (int j = 0; j < neseg; ++j) {
            ev[2 * j] = m->seg_pnt[2 * m->ebdry_seg[j]];
            ev[2 * j + 1] = m->seg_pnt[2 * m->ebdry_seg[j] + 1];
        }
        std::vector<int> eorder, eoffv, eseg(neseg);
        colour(neseg, 2, ev.data(), np, eorder, eoffv);
        for (int j = 0; j < neseg; ++j) eseg[j] = m->ebdry_seg[eorder[j]];
        c->eseg = c->upload(eseg.data(), eseg.size());
        c->eoff = eoffv;

        // adjacency in ascending order (the reference's accumulation order)
        std::vector<int> vs_ptr(np + 1, 0), vt_ptr(np + 1, 0);
        for (int s = 0; s < ns; ++s) { vs_ptr[m->seg_pnt[2 * s] + 1]++; vs_ptr[m->seg_pnt[2 * s + 1] + 1]++; }
        for (int t = 0; t < nt; ++t) for (int k = 0; k < 3; ++k) vt_ptr[m->tri_pnt[3 * t + k] + 1]++;
        for (int v = 0; v < np; ++v) { vs_ptr[v + 1] += vs_ptr[v]; vt_ptr[v + 1] += vt_ptr[v]; }
        std::vector<int> vs_side(vs_ptr[np]), vs_end(vs_ptr[np]), vt_tri(vt_ptr[np]), vt_loc(vt_ptr[np]);
        {
            std::vector<int> f(vs_ptr.begin(), vs_ptr.end() - 1);
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < 2; ++i) {
                    const int v = m->seg_pnt[2 * s + i];
                    vs_side[f[v]] = s; vs_end[f[v]] = i; f[v]++;
                }
            std::vector<int> g(vt_ptr.begin(), vt_ptr.end() - 1);
            for (int t = 0; t < nt; ++t)
                for (int i = 0; i < 3; ++i) {
                    const int v = m->tri_pnt[3 * t + i];
                    vt_tri[g[v]] = t; vt_loc[g[v]] = i; g[v]++;
                }
        }
        std::vector<int> st_tri(2 * (size_t)ns, -1), st_loc(2 * (size_t)ns, -1);
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < 3; ++i) {
                const int s = m->tri_seg[3 * t + i];
                const int slot = st_tri[2 * s] < 0 ? 0 : 1;
                if (slot == 1 && st_tri[2 * s + 1] >= 0) throw std::runtime_error("segment shared by 3 triangles");
                st_tri[2 * s + slot] = t; st_loc[2 * s + slot] = i;
            }
        std::vector<unsigned char> vess(np, 0), sess(ns, 0);
        for (int b = 0; b < m->nebd; ++b) {
            const unsigned char mk = (unsigned char)(m->ebdry_essential[b] & 7);
            for (int j = m->ebdry_start[b]; j < m->ebdry_start[b + 1]; ++j) {
                const int s = m->ebdry_seg[j];
                sess[s] |= mk;
                vess[m->seg_pnt[2 * s]] |= mk;
                vess[m->seg_pnt[2 * s + 1]] |= mk;
            }
        }
        c->vs_ptr = c->upload(vs_ptr.data(), vs_ptr.size());
        c->vs_side = c->upload(vs_side.data(), vs_side.size());
        c->vs_end = c->upload(vs_end.data(), vs_end.size());
        c->vt_ptr = c->upload(vt_ptr.data(), vt_ptr.size());
        c->vt_tri = c->upload(vt_tri.data(), vt_tri.size());
        c->vt_loc = c->upload(vt_loc.data(), vt_loc.size());
        c->st_tri = c->upload(st_tri.data(), st_tri.size());
        c->st_loc = c->upload(st_loc.data(), st_loc.size());
        c->vess = c->upload(vess.data(), vess.size());
        c->sess = c->upload(sess.data(), sess.size());

        const size_t nv = (size_t)np * NV, nsz = (size_t)ns * c->sm0 * NV, ni = (size_t)nt * c->im0 * NV;
        double** vsi[] = {&c->ugv, &c->ug0v, &c->resv, &c->rrv, &c->r0v};
        double** ssi[] = {&c->ugs, &c->ug0s, &c->ress, &c->rrs, &c->r0s};
        double** isi[] = {&c->ugi, &c->ug0i, &c->resi, &c->rri, &c->r0i};
        for (auto p : vsi) *p = c->alloc<double>(nv);
        for (auto p : ssi) *p = c->alloc<double>(nsz);
        for (auto p : isi) *p = c->alloc<double>(ni);
        for (int l = 0; l <= c->log2pmax; ++l) {
            c->dresv[l] = c->alloc<double>(nv);
            c->dress[l] = c->alloc<double>(nsz);
            c->dresi[l] = c->alloc<double>(ni);
        }
        c->vprcn = c->alloc<double>(nv);
        c->sprcn = c->alloc<double>((size_t)ns * NV);
        c->tprcn = c->alloc<double>((size_t)nt * NV);
        c->d_err = c->alloc<int>(1);
        c->d_max = c->alloc<unsigned long long>(NV);

        for (int l = 0; l <= c->log2pmax; ++l) {
            Level& L = c->lv[l];
            L.p = 1 << l;
            L.ops.reset(new TriOps(L.p));
            const TriOps& o = *L.ops;
            const std::vector<double> pk = build_pack(o);
            L.pack = c->upload(pk.data(), pk.size());
            L.Pmat = c->upload(o.P.data(), o.P.size());
            std::vector<double> lump;
            auto app = [&](const std::vector<double>& v) { const int o2 = (int)lump.size(); lump.insert(lump.end(), v.begin(), v.end()); return o2; };
            L.o_sfmv = app(o.sfmv); L.o_ifmb = app(o.ifmb); L.o_vfms = app(o.vfms); L.o_sdiag = app(o.sdiag);
            L.o_sfms = app(o.sfms); L.o_bfmi = app(o.bfmi); L.o_idiag = app(o.idiag);
            L.lump = c->upload(lump.data(), lump.size());
            std::vector<double> ed;
            ed.insert(ed.end(), o.g1.begin(), o.g1.end());
            ed.insert(ed.end(), o.dg1.begin(), o.dg1.end());
            ed.insert(ed.end(), o.w1.begin(), o.w1.end());
            L.edge = c->upload(ed.data(), ed.size());
        }
        check(hipStreamSynchronize(c->stream), "create sync");
    } catch (const std::exception& ex) {
        return -1;
    }
    *out = c.release();
    return 0;
}

int hpg_destroy(hpg_ctx* c) {
    delete c;
    return 0;
}

static int copy3(hpg_ctx* c, bool h2d, double* dv, double* ds, double* di, const double* hv, const double* hs,
                 const double* hi, double* ov, double* os, double* oi) {
    HPG_TRY(c, {
        const size_t nv = (size_t)c->npnt * NV * 8, ns = (size_t)c->nseg * c->sm0 * NV * 8,
                     ni = (size_t)c->ntri * c->im0 * NV * 8;
        if (h2d) {
            check(hipMemcpyAsync(dv, hv, nv, hipMemcpyHostToDevice, c->stream), "H2D");
            if (ns) check(hipMemcpyAsync(ds, hs, ns, hipMemcpyHostToDevice, c->stream), "H2D");
            if (ni) check(hipMemcpyAsync(di, hi, ni, hipMemcpyHostToDevice, c->stream), "H2D");
        } else {
            check(hipMemcpyAsync(ov, dv, nv, hipMemcpyDeviceToHost, c->stream), "D2H");
            if (ns) check(hipMemcpyAsync(os, ds, ns, hipMemcpyDeviceToHost, c->stream), "D2H");
            if (ni) check(hipMemcpyAsync(oi, di, ni, hipMemcpyDeviceToHost, c->stream), "D2H");
        }
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_set_state(hpg_ctx* c, const double* v, const double* s, const double* i) {
    if (!c) return -2;
    return copy3(c, true, c->ugv, c->ugs, c->ugi, v, s, i, nullptr, nullptr, nullptr);
}
int hpg_get_state(hpg_ctx* c, double* v, double* s, double* i) {
    if (!c) return -2;
    return copy3(c, false, c->ugv, c->ugs, c->ugi, nullptr, nullptr, nullptr, v, s, i);
}
int hpg_get_residual(hpg_ctx* c, double* v, double* s, double* i) {
    if (!c) return -2;
    return copy3(c, false, c->resv, c->ress, c->resi, nullptr, nullptr, nullptr, v, s, i);
}
int hpg_set_residual(hpg_ctx* c, const double* v, const double* s, const double* i) {
    if (!c) return -2;
    return copy3(c, true, c->resv, c->ress, c->resi, v, s, i, nullptr, nullptr, nullptr);
}
int hpg_get_preconditioner(hpg_ctx* c, double* vp, double* sp, double* tp) {
    if (!c) return -2;
    HPG_TRY(c, {
        check(hipMemcpyAsync(vp, c->vprcn, (size_t)c->npnt * NV * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(sp, c->sprcn, (size_t)c->nseg * NV * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(tp, c->tprcn, (size_t)c->ntri * NV * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}
int hpg_set_tprcn(hpg_ctx* c, const double* tp) {
    if (!c) return -2;
    HPG_TRY(c, {
        check(hipMemcpyAsync(c->tprcn, tp, (size_t)c->ntri * NV * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_rsdl(hpg_ctx* c, int stage) {
    if (!c || stage < 0 || stage > c->prm.nstage) return -2;
    HPG_TRY(c, do_rsdl(c, stage));
    return 0;
}

int hpg_setup_preconditioner(hpg_ctx* c) {
    if (!c) return -2;
    int e = 0;
    HPG_TRY(c, e = do_setup(c));
    return e;
}

int hpg_minvrt(hpg_ctx* c) {
    if (!c) return -2;
    HPG_TRY(c, do_minvrt(c));
    return 0;
}

int hpg_update(hpg_ctx* c) {
    if (!c) return -2;
    if (c->precision != HPG_FP64) { c->err = "precision mode not available"; return -3; }
    HPG_TRY(c, {
        const size_t nv = (size_t)c->npnt * NV * 8, ns = (size_t)c->nseg * c->sm0 * NV * 8,
                     ni = (size_t)c->ntri * c->im0 * NV * 8;
        check(hipMemcpyAsync(c->ug0v, c->ugv, nv, hipMemcpyDeviceToDevice, c->stream), "D2D");
        if (ns) check(hipMemcpyAsync(c->ug0s, c->ugs, ns, hipMemcpyDeviceToDevice, c->stream), "D2D");
        if (ni) check(hipMemcpyAsync(c->ug0i, c->ugi, ni, hipMemcpyDeviceToDevice, c->stream), "D2D");
        for (int stage = 0; stage < c->prm.nstage; ++stage) {
            do_rsdl(c, stage);
            do_minvrt(c);
            const double ca = c->prm.alpha[stage] * c->prm.cfl[c->log2p];
            modes_vsi(c, 4, c->ugv, c->ugs, c->ugi, c->ug0v, c->ug0s, c->ug0i, c->resv, c->ress, c->resi, ca);
        }
        check(hipGetLastError(), "update launch");
    });
    return 0;
}

int hpg_maxres(hpg_ctx* c, double* per_var, double* out) {
    if (!c) return -2;
    HPG_TRY(c, {
        double mx[NV];
        absmax(c, c->npnt, 1, 1, c->resv, nullptr, mx);
        double e = 0.0;
        for (int n = 0; n < NV; ++n) {
            e = std::max(e, mx[n]);
            if (per_var) per_var[n] = mx[n];
        }
        if (out) *out = e;
    });
    return 0;
}

int hpg_resnorm(hpg_ctx* c, double* out) {
    if (!c || !out) return -2;
    HPG_TRY(c, {
        const LevelDims d(1 << c->log2p);
        double a[NV], b[NV] = {0, 0, 0}, e[NV] = {0, 0, 0};
        absmax(c, c->npnt, 1, 1, c->resv, c->vess, a);
        if (d.sm) absmax(c, c->nseg, d.sm, c->sm0, c->ress, c->sess, b);
        if (d.im) absmax(c, c->ntri, d.im, c->im0, c->resi, nullptr, e);
        double r = 0.0;
        for (int n = 0; n < NV; ++n) r = std::max(r, std::max(a[n], std::max(b[n], e[n])));
        *out = r;
    });
    return 0;
}

int hpg_mg_restrict(hpg_ctx* c) {
    if (!c || c->log2p == 0) return -2;
    HPG_TRY(c, {
        c->isfrst = true;
        c->coarse_flag = true;
        --c->log2p;
        modes_vsi(c, 3, c->resv, c->ress, c->resi, c->r0v, c->r0s, c->r0i, nullptr, nullptr, nullptr, 0.0);
        check(hipGetLastError(), "restrict");
    });
    return 0;
}

int hpg_mg_prolongate(hpg_ctx* c) {
    if (!c || c->log2p >= c->log2pmax) return -2;
    ++c->log2p;
    if (c->log2p == c->log2pmax) c->coarse_flag = false;
    return 0;
}

int hpg_get_log2p(hpg_ctx* c, int* l) {
    if (!c || !l) return -2;
    *l = c->log2p;
    return 0;
}

int hpg_set_precision(hpg_ctx* c, int mode) {
    if (!c || (mode != HPG_FP64 && mode != HPG_FP32_DEFECT)) return -2;
    c->precision = mode;
    return 0;
}

int hpg_set_preconditioner(hpg_ctx* c, int mode) {
    if (!c || (mode != HPG_PRCN_ANALYTIC && mode != HPG_PRCN_LEARNED)) return -2;
    c->prcn_mode = mode;
    return 0;
}

int hpg_kernel_timing(hpg_ctx* c, int enable) {
    if (!c) return -2;
    HPG_TRY(c, {
        check(hipStreamSynchronize(c->stream), "sync");
        c->timing = enable != 0;
        c->recs.clear();
    });
    return 0;
}

int hpg_kernel_timing_read(hpg_ctx* c, int full, double* total_ms, long long* nlaunch, long long* nelem) {
    if (!c || !total_ms || !nlaunch || !nelem) return -2;
    HPG_TRY(c, {
        check(hipStreamSynchronize(c->stream), "sync");
        double tot = 0.0;
        long long nl = 0, ne = 0;
        for (size_t r = 0; r < c->recs.size(); ++r) {
            if (c->recs[r].full != (full ? 1 : 0)) continue;
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, c->events[2 * r], c->events[2 * r + 1]), "hipEventElapsedTime");
            tot += ms;
            ++nl;
            ne += c->recs[r].nelem;
        }
        *total_ms = tot; *nlaunch = nl; *nelem = ne;
    });
    return 0;
}

int hpg_synchronize(hpg_ctx* c) {
    if (!c) return -2;
    return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : -1;
}

const char* hpg_last_error(hpg_ctx* c) {
    static thread_local std::string s;
    s = c ? c->err : std::string("null context");
    return s.c_str();
}

// Host-built basis operators (test support; no device work): copies the named
// operator of the p basis into out (if non-NULL) and returns its length, -1 if unknown.
int hpg_basis_operator(int p, const char* name, double* out) {
    if (p != 1 && p != 2 && p != 4) return -1;
    try {
        const TriOps o(p);
        const std::string n(name ? name : "");
        std::vector<double> v;
        if (n == "P") v = o.P; else if (n == "Pr") v = o.Pr; else if (n == "Ps") v = o.Ps;
        else if (n == "I") v = o.I; else if (n == "Ir") v = o.Ir; else if (n == "Is") v = o.Is;
        else if (n == "Dr") v = o.Dr; else if (n == "Ds") v = o.Ds;
        else if (n == "xg") v = o.xg; else if (n == "wx") v = o.wx; else if (n == "eg") v = o.eg; else if (n == "we") v = o.we;
        else if (n == "g1") v = o.g1; else if (n == "dg1") v = o.dg1;
        else if (n == "vdiag") v = {o.vdiag};
        else if (n == "sfmv") v = o.sfmv; else if (n == "ifmb") v = o.ifmb; else if (n == "vfms") v = o.vfms;
        else if (n == "sdiag") v = o.sdiag; else if (n == "sfms") v = o.sfms; else if (n == "bfmi") v = o.bfmi;
        else if (n == "idiag") v = o.idiag; else if (n == "mass") v = o.mass;
        else return -1;
        if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(double));
        return (int)v.size();
    } catch (...) {
        return -1;
    }
}

// MFMA f64 16x16x4 lane-layout probe (test support; A 16x4, B 4x16 row-major, out 64x4)
int hpg_selftest_mfma(const double* A, const double* B, double* out) {
    double *dA = nullptr, *dB = nullptr, *dO = nullptr;
    if (hipMalloc(&dA, 64 * 8) != hipSuccess || hipMalloc(&dB, 64 * 8) != hipSuccess || hipMalloc(&dO, 256 * 8) != hipSuccess)
        return -1;
    hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice);
    hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dA, dB, dO);
    const hipError_t e = hipMemcpy(out, dO, 256 * 8, hipMemcpyDeviceToHost);
    hipFree(dA); hipFree(dB); hipFree(dO);
    return e == hipSuccess ? 0 : -1;
}

}  // extern "C"
