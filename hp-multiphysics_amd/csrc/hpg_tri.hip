// Context and C-ABI of the tri_hp residual/relaxation library (include/hpgpu.h).
//
// One context = one solution block.  The context owns device mirrors of the
// block's mesh and solution arrays, the per-level operator packs, the element
// colouring used for the deterministic scatter, and one HIP stream.  Everything
// between hpg_set_state and hpg_get_state stays resident in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hpgpu.h"
#include "tri_kernels.hip"
#include "tri_kernels_lin.hip"
#include "tri_kernels_cns.hip"
#include "tri_jacobian.hip"
#include "tri_ops.h"
#include "mlp.h"

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

// Lane-ready operator pack: A operand of every MFMA K-step, 64 values per step.
// `row(i, q)` is the accumulator row held by register i of lane group q (FP64:
// 4i + q; FP32: f32_row), which fixes the K permutation of the chained operators.
// `moderow(r)` is the mode that A row r of Ir/Is produces (identity unless the
// accumulator layout of the output tile differs from the FP64 one: the FP32 element
// kernel maps row 4q + i to mode 4i + q, so that its register s of lane group q holds
// mode 4s + q as in the FP64 kernel).
template <class RowFn>
std::vector<double> build_pack(const TriOps& o, RowFn row, bool transpose_modes = false) {
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
                        const int g = 16 * mt + (l & 15), k = 16 * kt + row(i, l >> 4);
                        pk[3 * d.PROJ + op * d.DER + ((mt * d.GT + kt) * 4 + i) * 64 + l] =
                            (g < d.G && k < d.G) ? (*der[op])[g * d.G + k] : 0.0;
                    }
    const std::vector<double>* in[2] = {&o.Ir, &o.Is};
    for (int op = 0; op < 2; ++op)
        for (int kt = 0; kt < d.GT; ++kt)
            for (int i = 0; i < 4; ++i)
                for (int l = 0; l < 64; ++l) {
                    const int r = l & 15, m = transpose_modes ? 4 * (r & 3) + (r >> 2) : r;
                    const int k = 16 * kt + row(i, l >> 4);
                    pk[3 * d.PROJ + 2 * d.DER + op * d.INT + (kt * 4 + i) * 64 + l] =
                        (m < d.tm && k < d.G) ? (*in[op])[m * d.G + k] : 0.0;
                }
    return pk;
}

// Greedy colouring: items whose vertex sets intersect get different colours (at most
// 128, a bit mask of two words per vertex).  Returns the item order grouped by colour
// (ascending inside a colour) and offsets.  O(nitem * per).
void colour(int nitem, int per, const int* verts, int npnt, std::vector<int>& order, std::vector<int>& off,
            std::vector<int>& col) {
    constexpr int WORDS = 2;
    std::vector<uint64_t> used((size_t)WORDS * npnt, 0);
    col.assign(nitem, 0);
    int ncol = 0;
    for (int t = 0; t < nitem; ++t) {
        uint64_t m[WORDS] = {0, 0};
        for (int k = 0; k < per; ++k)
            for (int w = 0; w < WORDS; ++w) m[w] |= used[(size_t)WORDS * verts[per * t + k] + w];
        int c = 0;
        while (c < 64 * WORDS && ((m[c >> 6] >> (c & 63)) & 1)) ++c;
        if (c == 64 * WORDS) throw std::runtime_error("colouring needs more than 128 colours");
        col[t] = c;
        ncol = std::max(ncol, c + 1);
        for (int k = 0; k < per; ++k) used[(size_t)WORDS * verts[per * t + k] + (c >> 6)] |= (uint64_t(1) << (c & 63));
    }
    off.assign(ncol + 1, 0);
    for (int t = 0; t < nitem; ++t) off[col[t] + 1]++;
    for (int c = 0; c < ncol; ++c) off[c + 1] += off[c];
    order.resize(nitem);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int t = 0; t < nitem; ++t) order[fill[col[t]]++] = t;
}

// relative size |dhat| / |u0| that the scale of the FP32 defect correction aims at:
// FP32 residual differences then resolve J(u0) dhat to about 1e-3 (rounding ~ eps32 / h,
// nonlinearity ~ h)
constexpr double FD_H = 1e-3;

struct Level {
    int p = 0;
    std::unique_ptr<TriOps> ops;
    double* pack = nullptr;
    double* op1 = nullptr;  // p = 1: dense P, Pr, Ps, Ir, Is, Dr, Ds (k_elem_ins1_valu layout)
    double* Pmat = nullptr;
    double* lump = nullptr;
    double* edge = nullptr;  // g1 | dg1 | w1
    float* pack32 = nullptr;   // FP32 linearised kernel (output rows in the FP32 layout)
    float* packn32 = nullptr;  // FP32 element residual (output rows permuted to the FP64 layout)
    float* lump32 = nullptr;
    int o_sfmv, o_ifmb, o_vfms, o_sdiag, o_sfms, o_bfmi, o_idiag;
    // cns: point-per-lane operator pack (CDims layout), edge side-mass and interior band factors
    double* cpack = nullptr;
    float* cpack32 = nullptr;  // the same operators in FP32 (defect-correction inner solve)
    double* sdiag1d = nullptr;
    double* idiag = nullptr;
};

// cns operator pack of one level: P, Pr, Ps as [mode][32], Dr, Ds as [point'][32],
// Ir, Is, I as [point][16], zero padded
std::vector<double> build_cns_pack(const TriOps& o) {
    const int tm = o.tm, G = o.G;
    const int OP_PR = tm * 32, OP_PS = 2 * tm * 32, OP_DR = 3 * tm * 32, OP_DS = OP_DR + G * 32;
    const int OP_IR = OP_DS + G * 32, OP_IS = OP_IR + G * 16, OP_I = OP_IS + G * 16, SIZE = OP_I + G * 16;
    std::vector<double> pk(SIZE, 0.0);
    for (int k = 0; k < tm; ++k)
        for (int g = 0; g < G; ++g) {
            pk[k * 32 + g] = o.P[g * tm + k];
            pk[OP_PR + k * 32 + g] = o.Pr[g * tm + k];
            pk[OP_PS + k * 32 + g] = o.Ps[g * tm + k];
        }
    for (int q = 0; q < G; ++q)
        for (int g = 0; g < G; ++g) {
            pk[OP_DR + q * 32 + g] = o.Dr[g * G + q];
            pk[OP_DS + q * 32 + g] = o.Ds[g * G + q];
        }
    for (int q = 0; q < G; ++q)
        for (int k = 0; k < tm; ++k) {
            pk[OP_IR + q * 16 + k] = o.Ir[k * G + q];
            pk[OP_IS + q * 16 + k] = o.Is[k * G + q];
            pk[OP_I + q * 16 + k] = o.I[k * G + q];
        }
    return pk;
}

}  // namespace

struct hpg_ctx {
    hipStream_t stream = nullptr;
    hpg_params prm;
    int log2pmax = 0, log2p = 0;
    bool coarse_flag = false, isfrst = false;
    int precision = HPG_FP64, prcn_mode = HPG_PRCN_ANALYTIC;
    int npnt = 0, nseg = 0, ntri = 0, sm0 = 0, im0 = 0;
    int physics = HPG_PHYSICS_INS, nv = NV;  // unknowns per vertex/mode: 3 (ins) or 4 (cns)
    // cns: physical constants, inflow boundary state, block preconditioners, res_temp
    CnsPhys cph{};
    double cbc[4] = {1.0, 0.0, 0.0, 1.0};
    double *vpc = nullptr, *spc = nullptr, *tpc = nullptr, *rtv = nullptr, *rts = nullptr;
    int *inflow_v = nullptr, *inflow_s = nullptr;
    int n_inflow_v = 0, n_inflow_s = 0;
    // h-multigrid: this context is a coarse mesh level linked to `fine`
    hpg_ctx* fine = nullptr;
    bool coarse_level = false;
    hipEvent_t xev = nullptr;  // orders this context's stream against a neighbour level's
    int *hr_ptr = nullptr, *hr_fidx = nullptr, *hi_fv = nullptr, *hp_cv = nullptr;
    double *hr_wt = nullptr, *hi_wt = nullptr, *hp_wt = nullptr, *vug_frst = nullptr;
    std::vector<void*> allocs;
    double* pnts = nullptr;
    int *tri_pnt = nullptr, *tri_seg = nullptr, *tri_sgn = nullptr, *seg_pnt = nullptr;
    int* order = nullptr;
    std::vector<int> coff;
    std::vector<int> tcol;  // colour of each element (host)
    int* eseg = nullptr;
    std::vector<int> eoff;
    int neseg = 0;
    std::vector<int> h_eseg;  // eseg on the host (boundary side of each edge index)
    std::vector<int> h_ebid;  // edge boundary of each edge index
    std::vector<int> h_ebdry_start, h_ebdry_seg, h_ebdry_ess;  // the mesh's edge boundaries
    // periodic boundaries (ins): vertex copies in groups (pg_ptr/pg_v, ascending), side
    // pairs (ps_a, ps_b and the relative direction ps_o = +-1); the vertex and side sweeps
    // of setup and minvrt read the union adjacency of every copy (vtu_*, vsu_*; the own
    // lists when nothing is periodic)
    bool periodic = false;
    int npg = 0, nps = 0;
    int *pg_ptr = nullptr, *pg_v = nullptr, *ps_a = nullptr, *ps_b = nullptr, *ps_o = nullptr;
    int *vtu_ptr = nullptr, *vtu_tri = nullptr, *vtu_loc = nullptr;
    int *vsu_ptr = nullptr, *vsu_side = nullptr, *vsu_end = nullptr;
    // boundary flux type of each edge index (null: all generic) and the freestream
    // velocity of the euler / characteristic fluxes (ins)
    unsigned char* etype = nullptr;
    double ibc_u[2] = {0.0, 0.0};
    // curved boundary sides (ins, p >= 2): the elements owning one (clist) and the rest
    // in mesh order (slist), the elements' coordinate side modes (crv, ncurv x 3 x sm0
    // x 2), each edge's row in scrd (ecrv) and the rows (the sides' coordinate modes)
    int ncurv = 0;
    int *clist = nullptr, *slist = nullptr, *ecrv = nullptr;
    double *crv = nullptr, *scrd = nullptr;
    double *PRmat[3] = {nullptr, nullptr, nullptr}, *PSmat[3] = {nullptr, nullptr, nullptr};
    // p = 1 vertex path (ins): element-major element and boundary-edge results and the
    // boundary-edge endpoints at each vertex (entries 2*idx + k into eout)
    double *eres = nullptr, *eout = nullptr;
    float* eres32 = nullptr;  // the FP32 inner solve's element-major residual
    int *vb_ptr = nullptr, *vb_ent = nullptr;
    int *vs_ptr, *vs_side, *vs_end, *vt_ptr, *vt_tri, *vt_loc, *st_tri, *st_loc;
    unsigned char *vess, *sess;
    unsigned char *fcv, *fcs;  // first colour touching each vertex / side
    double *ugv, *ugs, *ugi, *ug0v, *ug0s, *ug0i;
    double *resv, *ress, *resi, *rrv, *rrs, *rri, *r0v, *r0s, *r0i;
    double *dresv[3], *dress[3], *dresi[3];
    double *vprcn, *sprcn, *tprcn;
    int* d_err;
    // mixed precision: per-colour group offsets, linearisation data, FP32 inner arrays
    std::vector<int> gbase;
    int ngroups = 0;
    int ninner = 2;
    float* lin = nullptr;
    float *dlv = nullptr, *dls = nullptr, *dli = nullptr, *d0v = nullptr, *d0s = nullptr, *d0i = nullptr;
    float *r32v = nullptr, *r32s = nullptr, *r32i = nullptr, *rr32v = nullptr, *rr32s = nullptr, *rr32i = nullptr;
    float *vp32 = nullptr, *sp32 = nullptr, *tp32 = nullptr;
    // FP32 defect correction: base state fl32(u0) of the outer iteration; device scalars
    // (scale c of the inner correction, 1/c - 1) and maxima (|u0|, |v - b|) as float bits
    float *b32v = nullptr, *b32s = nullptr, *b32i = nullptr;
    double* fd_scal = nullptr;
    unsigned* fd_max = nullptr;
    bool fd_reset = true;
    // cns: the FP64 residual R64(u0) and the state u0 of the outer iteration
    double *srcv = nullptr, *srcs = nullptr, *srci = nullptr, *u0cv = nullptr, *u0cs = nullptr, *u0ci = nullptr;
    // learned preconditioner (the MLP is owned by the caller)
    hpg_mlp* mlp = nullptr;
    float* featX = nullptr;
    float* mlpY = nullptr;
    double* jdiag = nullptr;
    // element scalings of the vertex, side and interior rows when the learned
    // preconditioner is in effect on the finest level (typed_prcn), FP64 and FP32
    double *tprcn_vx = nullptr, *tprcn_sd = nullptr, *tprcn_in = nullptr;
    float *tp32s = nullptr, *tp32i = nullptr;
    bool typed_prcn = false;
    bool user_y = false;  // mlpY holds log scalings set by the caller (hpg_set_prcn_log_scalings)
    unsigned long long* d_max;
    // hpg_mg_cycle as a captured HIP graph (on the fine context of a stack): replayed
    // while the signature of the settings that fix the enqueued work is unchanged
    bool graphs = true;
    std::vector<unsigned long long> cg_sig;
    bool cg_seen = false;  // the signature was run eagerly once; capture at the next call
    hipGraphExec_t cg_exec = nullptr;
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
        if (cg_exec) hipGraphExecDestroy(cg_exec);
        for (void* p : allocs) hipFree(p);
        for (hipEvent_t e : events) hipEventDestroy(e);
        if (xev) hipEventDestroy(xev);
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
// block count rounded up to a multiple of 8 for kernels that use the XCD-aware block map
static inline unsigned nblk8(long n, int bs = 256) { return ((nblk(n, bs) + 7u) / 8u) * 8u; }

namespace {

// device-to-device copy of `bytes` (a multiple of 4) as a kernel launch, so that a
// captured cycle holds kernel nodes only
void dcopy(hpg_ctx* c, void* dst, const void* src, size_t bytes) {
    const long n = (long)(bytes / 4);
    if (n <= 0) return;
    const unsigned nb = std::min<unsigned>(nblk(n), 4096u);
    hipLaunchKernelGGL(k_copy_words, dim3(nb), dim3(256), 0, c->stream, n, (const unsigned*)src, (unsigned*)dst);
}

void launch_modes(hpg_ctx* c, int op, long nent, int nm, int mstride, double* x, double* y, const double* z, double k) {
    if (nent <= 0 || nm <= 0) return;
    const long n = nent * nm * c->nv;
    if (c->nv == NVC)
        hipLaunchKernelGGL(k_modes<NVC>, dim3(nblk(n)), dim3(256), 0, c->stream, op, nent, nm, mstride, x, y, z, k);
    else
        hipLaunchKernelGGL(k_modes<NV>, dim3(nblk(n)), dim3(256), 0, c->stream, op, nent, nm, mstride, x, y, z, k);
}

// apply a k_modes op over the active v/s/i modes of the current level
void modes_vsi(hpg_ctx* c, int op, double* xv, double* xs, double* xi, double* yv, double* ys, double* yi,
               const double* zv, const double* zs, const double* zi, double k) {
    const LevelDims d(1 << c->log2p);
    launch_modes(c, op, c->npnt, 1, 1, xv, yv, zv, k);
    if (d.sm > 0) launch_modes(c, op, c->nseg, d.sm, c->sm0, xs, ys, zs, k);
    if (d.im > 0) launch_modes(c, op, c->ntri, d.im, c->im0, xi, yi, zi, k);
}

// blocks of 256 threads of `kernel` that the whole device holds at once (a multiple
// of 8: one share per XCD); persistent kernels launch at most this many
static int resident_blocks(const void* kernel) {
    int dev = 0, cus = 0, per = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0), "hipOccupancy");
    return std::max(8, (cus * std::max(per, 1)) / 8 * 8);
}

// grid of a persistent launch over `want` blocks' worth of work
static int persistent_grid(int want, int cap) { return std::min((want + 7) / 8 * 8, cap); }

template <int P>
void launch_elem(hpg_ctx* c, const ElemArgs& a0, bool full, const std::vector<int>* offs = nullptr) {
    // offs: colour offsets into a0.order (default: every element, c->order)
    const std::vector<int>& off = offs ? *offs : c->coff;
    for (size_t col = 0; col + 1 < off.size(); ++col) {
        ElemArgs a = a0;
        a.colour = (int)col;
        a.begin = off[col];
        a.end = off[col + 1];
        a.gbase = offs ? 0 : c->gbase[col];
        const int nlogical = (a.end - a.begin + 63) / 64;
        if (nlogical == 0) continue;
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full) {
            static const int cap = resident_blocks((const void*)k_elem_ins<P, true>);
            hipLaunchKernelGGL((k_elem_ins<P, true>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream, a);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_ins<P, false>);
            hipLaunchKernelGGL((k_elem_ins<P, false>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream,
                               a);
        }
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 1 : 0, a.end - a.begin});
        }
    }
}

// element-major residual over a0.order[begin:end) (null order: mesh order) in one
// persistent launch (timing records 1 / 0 like the colour launches)
template <int P>
void launch_elem_em(hpg_ctx* c, const ElemArgs& a, bool full) {
    const int nlogical = (a.end - a.begin + 63) / 64;
    if (nlogical == 0) return;
    const size_t r = c->recs.size();
    if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
    if (full) {
        static const int cap = resident_blocks((const void*)k_elem_ins_em<P, true>);
        hipLaunchKernelGGL((k_elem_ins_em<P, true>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream,
                           a, c->eres);
    } else {
        static const int cap = resident_blocks((const void*)k_elem_ins_em<P, false>);
        hipLaunchKernelGGL((k_elem_ins_em<P, false>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream,
                           a, c->eres);
    }
    if (c->timing) {
        check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
        c->recs.push_back({full ? 1 : 0, a.end - a.begin});
    }
}

template <int P>
void launch_elem32_em(hpg_ctx* c, const ElemArgs32& a, bool full) {
    const int nlogical = (a.end - a.begin + 63) / 64;
    if (nlogical == 0) return;
    const size_t r = c->recs.size();
    if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
    if (full) {
        static const int cap = resident_blocks((const void*)k_elem_ins32_em<P, true>);
        hipLaunchKernelGGL((k_elem_ins32_em<P, true>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0,
                           c->stream, a, c->eres32);
    } else {
        static const int cap = resident_blocks((const void*)k_elem_ins32_em<P, false>);
        hipLaunchKernelGGL((k_elem_ins32_em<P, false>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0,
                           c->stream, a, c->eres32);
    }
    if (c->timing) {
        check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
        c->recs.push_back({full ? 3 : 2, a.end - a.begin});
    }
}

// FP32 element residual of the defect-correction inner solve, colour by colour (timing
// records 3 / 2 like the linearised kernel)
template <int P>
void launch_elem32(hpg_ctx* c, const ElemArgs32& a0, bool full) {
    for (size_t col = 0; col + 1 < c->coff.size(); ++col) {
        ElemArgs32 a = a0;
        a.colour = (int)col;
        a.begin = c->coff[col];
        a.end = c->coff[col + 1];
        a.gbase = 0;
        const int nlogical = (a.end - a.begin + 63) / 64;
        if (nlogical == 0) continue;
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full) {
            static const int cap = resident_blocks((const void*)k_elem_ins32<P, true>);
            hipLaunchKernelGGL((k_elem_ins32<P, true>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream,
                               a);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_ins32<P, false>);
            hipLaunchKernelGGL((k_elem_ins32<P, false>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0,
                               c->stream, a);
        }
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 3 : 2, a.end - a.begin});
        }
    }
}

template <int P>
void launch_lin(hpg_ctx* c, const LinArgs& a0, bool full) {
    for (size_t col = 0; col + 1 < c->coff.size(); ++col) {
        LinArgs a = a0;
        a.colour = (int)col;
        a.begin = c->coff[col];
        a.end = c->coff[col + 1];
        a.gbase = c->gbase[col];
        const int nlogical = (a.end - a.begin + 63) / 64;
        if (nlogical == 0) continue;
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full) {
            static const int cap = resident_blocks((const void*)k_elem_lin<P, true>);
            hipLaunchKernelGGL((k_elem_lin<P, true>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream, a);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_lin<P, false>);
            hipLaunchKernelGGL((k_elem_lin<P, false>), dim3(persistent_grid(nlogical, cap)), dim3(256), 0, c->stream,
                               a);
        }
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 3 : 2, a.end - a.begin});
        }
    }
}

// minvrt arguments over the FP64 residual (f32 = false) or the FP32 inner residual
template <typename T>
MinvArgsT<T> minv_args(hpg_ctx* c) {
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    MinvArgsT<T> a;
    a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    // the vertex sweep couples every side and element of a vertex's periodic copies
    a.vs_ptr = c->vsu_ptr; a.vs_side = c->vsu_side; a.vs_end = c->vsu_end;
    a.vt_ptr = c->vtu_ptr; a.vt_tri = c->vtu_tri; a.vt_loc = c->vtu_loc;
    a.st_tri = c->st_tri; a.st_loc = c->st_loc;
    a.vess = c->vess; a.sess = c->sess;
    if constexpr (sizeof(T) == 8) {
        a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
        a.vprcn = c->vprcn; a.sprcn = c->sprcn;
        a.tprcn_s = c->typed_prcn ? c->tprcn_sd : c->tprcn;
        a.tprcn_i = c->typed_prcn ? c->tprcn_in : c->tprcn;
        a.lump = L.lump;
    } else {
        a.resv = c->r32v; a.ress = c->r32s; a.resi = c->r32i;
        a.vprcn = c->vp32; a.sprcn = c->sp32;
        a.tprcn_s = c->typed_prcn ? c->tp32s : c->tp32;
        a.tprcn_i = c->typed_prcn ? c->tp32i : c->tp32;
        a.lump = L.lump32;
    }
    a.npnt = c->npnt; a.nseg = c->nseg; a.ntri = c->ntri;
    a.sm = d.sm; a.im = d.im; a.bm = d.bm; a.sm0 = c->sm0; a.im0 = c->im0;
    a.o_sfmv = L.o_sfmv; a.o_ifmb = L.o_ifmb; a.o_vfms = L.o_vfms; a.o_sdiag = L.o_sdiag;
    a.o_sfms = L.o_sfms; a.o_bfmi = L.o_bfmi; a.o_idiag = L.o_idiag;
    a.ugv = a.ugs = a.ugi = nullptr;
    a.ug0v = a.ug0s = a.ug0i = nullptr;
    a.cfa = T(0);
    return a;
}

void do_rsdl_cns(hpg_ctx* c, int stage);

// ins stage at p = 1 in two launches: element residuals and boundary-edge fluxes in
// element-major form (no colouring), then one lane per (vertex, variable) sums them in
// a fixed order, adds the multigrid source and, with cfa >= 0, applies the vertex
// minvrt and the stage update (save_u0: first stage, ug0 = ug)
// every copy of a periodic vertex takes the sum over its group (after the residual)
template <typename T>
void periodic_combine(hpg_ctx* c, T* resv) {
    if (!c->periodic || c->npg == 0) return;
    hipLaunchKernelGGL((k_periodic_combine<T, NV>), dim3(nblk((long)c->npg * NV)), dim3(256), 0, c->stream, c->npg,
                       c->pg_ptr, c->pg_v, resv);
}

// the state's periodic copies from the first copy of each group and side a of each pair
void periodic_sync(hpg_ctx* c) {
    if (!c->periodic) return;
    const long n = (long)c->npg * NV + (long)c->nps * std::max(c->sm0, 0) * NV;
    if (n == 0) return;
    hipLaunchKernelGGL((k_periodic_sync<double, NV>), dim3(nblk(n)), dim3(256), 0, c->stream, c->npg, c->pg_ptr,
                       c->pg_v, c->nps, c->ps_a, c->ps_b, c->ps_o, c->sm0, c->sm0, c->ugv, c->ugs);
    check(hipGetLastError(), "periodic sync");
}

// multigrid source of a coarse p-level (tri_hp/nstage.cpp:113-129): the first rsdl
// after mg_restrict sets dres = fadd r0 - res, later ones add dres
void mg_source_vsi(hpg_ctx* c) {
    if (!c->coarse_flag) return;
    const int l = c->log2p;
    if (c->isfrst) {
        modes_vsi(c, 2, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], c->r0v, c->r0s, c->r0i,
                  c->prm.fadd);
        c->isfrst = false;
    } else {
        modes_vsi(c, 1, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], nullptr, nullptr, nullptr,
                  0.0);
    }
}

void vstage_p1(hpg_ctx* c, int stage, double cfa, bool save_u0) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const bool full = beta > 0.0;
    const Level& L = c->lv[c->log2p];
    ElemArgs a{};
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = nullptr; a.begin = 0; a.end = c->ntri;   // mesh order: no colouring needed
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi;
    a.ops = L.pack; a.sm0 = c->sm0; a.im0 = c->im0;
    a.rho = c->prm.rho; a.mu = c->prm.mu; a.adis = c->prm.adis; a.beta = beta;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0; a.keep = f;
    const LevelDims d(1);
    EdgeArgs e{};
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.resv = c->resv; e.ress = c->ress;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.gpx = d.p + 1; e.sm = d.sm; e.sm0 = c->sm0; e.rho = c->prm.rho; e.bc_pressure = c->prm.bc_pressure;
    e.etype = c->etype; e.ub0 = c->ibc_u[0]; e.ub1 = c->ibc_u[1]; e.mu = c->prm.mu;
    e.begin = 0; e.end = c->neseg;
    // element blocks (one thread per element, mesh order) followed by the boundary-edge
    // blocks
    const int eg = (int)nblk(c->ntri, 256);
    const int eblk = (int)nblk(c->neseg, 256);
    const size_t r = c->recs.size();
    if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
    if (full)
        hipLaunchKernelGGL(k_elem_ins1_valu<true>, dim3(eg + eblk), dim3(256), 0, c->stream, a, L.op1, c->eres, e,
                           c->eout, eg);
    else
        hipLaunchKernelGGL(k_elem_ins1_valu<false>, dim3(eg + eblk), dim3(256), 0, c->stream, a, L.op1, c->eres, e,
                           c->eout, eg);
    if (c->timing) {
        check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
        c->recs.push_back({full ? 1 : 0, c->ntri});
    }
    VStageArgs v{};
    v.npnt = c->npnt; v.vt_ptr = c->vt_ptr; v.vt_tri = c->vt_tri; v.vt_loc = c->vt_loc;
    v.vb_ptr = c->vb_ptr; v.vb_ent = c->vb_ent; v.eres = c->eres; v.eout = c->eout;
    v.resv = c->resv; v.rrv = c->rrv; v.keep = f; v.full = full ? 1 : 0;
    if (c->coarse_flag && !c->periodic) {
        const int l = c->log2p;
        v.mg = c->isfrst ? 2 : 1;
        v.dres = c->dresv[l]; v.r0 = c->r0v; v.fadd = c->prm.fadd;
        c->isfrst = false;
    }
    if (cfa >= 0.0) {
        v.vess = c->vess; v.vprcn = c->vprcn; v.ug = c->ugv; v.ug0 = c->ug0v; v.cfa = cfa;
        v.save_u0 = save_u0 ? 1 : 0;
    }
    hipLaunchKernelGGL(k_vstage1<NV>, dim3(nblk8((long)c->npnt * NV)), dim3(256), 0, c->stream, v);
    check(hipGetLastError(), "p1 stage launch");
}

void do_rsdl(hpg_ctx* c, int stage, bool store_lin = false) {
    if (c->physics == HPG_PHYSICS_CNS) { do_rsdl_cns(c, stage); return; }
    if (c->lv[c->log2p].p == 1 && !store_lin) {
        vstage_p1(c, stage, -1.0, false);
        if (c->periodic) {
            periodic_combine<double>(c, c->resv);
            mg_source_vsi(c);
        }
        return;
    }
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const Level& L = c->lv[c->log2p];
    ElemArgs a{};
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi;
    a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.rrv = c->rrv; a.rrs = c->rrs; a.rri = c->rri;
    a.ops = L.pack; a.sm0 = c->sm0; a.im0 = c->im0;
    a.rho = c->prm.rho; a.mu = c->prm.mu; a.adis = c->prm.adis; a.beta = beta;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0; a.keep = f;
    a.lin = (store_lin && beta > 0.0) ? c->lin : nullptr;
    a.gbase = 0;
    const bool full = beta > 0.0;
    if (!store_lin) {
        // element-major over all elements in mesh order (the straight ones, then the
        // curved ones, when there are curved boundary sides), then the fixed-order gather
        a.order = c->ncurv ? c->slist : nullptr; a.begin = 0; a.end = c->ntri - c->ncurv;
        switch (L.p) {
            case 2: launch_elem_em<2>(c, a, full); break;
            case 4: launch_elem_em<4>(c, a, full); break;
            default: throw std::runtime_error("unsupported p");
        }
        if (c->ncurv) {
            ElemArgs b = a;
            b.order = c->clist; b.end = c->ncurv; b.crv = c->crv;
            const dim3 grid(persistent_grid((c->ncurv + 63) / 64, 1 << 20));
            if (L.p == 2) {
                if (full) hipLaunchKernelGGL((k_elem_ins_em_curved<2, true>), grid, dim3(256), 0, c->stream, b, c->eres);
                else hipLaunchKernelGGL((k_elem_ins_em_curved<2, false>), grid, dim3(256), 0, c->stream, b, c->eres);
            } else {
                if (full) hipLaunchKernelGGL((k_elem_ins_em_curved<4, true>), grid, dim3(256), 0, c->stream, b, c->eres);
                else hipLaunchKernelGGL((k_elem_ins_em_curved<4, false>), grid, dim3(256), 0, c->stream, b, c->eres);
            }
        }
        const LevelDims dd(L.p);
        GatherArgsT<double> ga{};
        ga.npnt = c->npnt; ga.nseg = c->nseg; ga.sm = dd.sm; ga.sm0 = c->sm0; ga.bm = dd.bm;
        ga.vt_ptr = c->vt_ptr; ga.vt_tri = c->vt_tri; ga.vt_loc = c->vt_loc;
        ga.st_tri = c->st_tri; ga.st_loc = c->st_loc; ga.eres = c->eres;
        ga.resv = c->resv; ga.rrv = c->rrv; ga.ress = c->ress; ga.rrs = c->rrs;
        ga.keep = f; ga.full = full ? 1 : 0;
        const long lanes = ((long)c->npnt + (long)c->nseg * dd.sm) * NV;
        hipLaunchKernelGGL((k_gather_vs<double, NV>), dim3(nblk(lanes)), dim3(256), 0, c->stream, ga);
    } else {
        switch (L.p) {
            case 1: launch_elem<1>(c, a, full); break;
            case 2: launch_elem<2>(c, a, full); break;
            case 4: launch_elem<4>(c, a, full); break;
            default: throw std::runtime_error("unsupported p");
        }
    }
    const LevelDims d(L.p);
    EdgeArgs e{};
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.resv = c->resv; e.ress = c->ress;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.gpx = d.p + 1; e.sm = d.sm; e.sm0 = c->sm0; e.rho = c->prm.rho; e.bc_pressure = c->prm.bc_pressure;
    e.ecrv = c->ecrv; e.scrd = c->scrd;
    e.etype = c->etype; e.ub0 = c->ibc_u[0]; e.ub1 = c->ibc_u[1]; e.mu = c->prm.mu;
    for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
        e.begin = c->eoff[col];
        e.end = c->eoff[col + 1];
        if (e.end > e.begin) {
            const dim3 be(nblk(e.end - e.begin, 64)), te(64);
            if (L.p == 1) hipLaunchKernelGGL(k_edge_ins<1>, be, te, 0, c->stream, e);
            else if (L.p == 2) hipLaunchKernelGGL(k_edge_ins<2>, be, te, 0, c->stream, e);
            else hipLaunchKernelGGL(k_edge_ins<4>, be, te, 0, c->stream, e);
        }
    }
    periodic_combine<double>(c, c->resv);
    mg_source_vsi(c);
    check(hipGetLastError(), "rsdl launch");
}

// FP32 residual of the defect-correction inner solve at the FP32 state v (dlv/dls/dli):
// r32 = R32(v) + (1/c - 1) R64(u0), multistage blended as rsdl(stage) blends R
void do_rsdl32(hpg_ctx* c, int stage) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const Level& L = c->lv[c->log2p];
    ElemArgs32 a{};
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.ugv = c->dlv; a.ugs = c->dls; a.ugi = c->dli;
    a.resv = c->r32v; a.ress = c->r32s; a.resi = c->r32i;
    a.rrv = c->rr32v; a.rrs = c->rr32s; a.rri = c->rr32i;
    a.ops = L.packn32; a.sm0 = c->sm0; a.im0 = c->im0;
    a.rho = (float)c->prm.rho; a.mu = (float)c->prm.mu; a.adis = (float)c->prm.adis; a.beta = (float)beta;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0; a.keep = (float)f;
    a.lin = nullptr; a.gbase = 0;
    // periodic: the vertex source enters after the copies are summed (once per group)
    a.srcv = c->periodic ? nullptr : c->resv; a.srcs = c->ress; a.srci = c->resi; a.sfac = c->fd_scal + 1;
    const bool full = beta > 0.0;
    if (L.p >= 2) {
        // element-major over all elements in mesh order, then the fixed-order gather
        // (which adds the source of the vertex and side modes)
        a.order = c->ncurv ? c->slist : nullptr; a.begin = 0; a.end = c->ntri - c->ncurv;
        if (L.p == 2) launch_elem32_em<2>(c, a, full);
        else launch_elem32_em<4>(c, a, full);
        if (c->ncurv) {
            ElemArgs32 b = a;
            b.order = c->clist; b.end = c->ncurv; b.crv = c->crv;
            const dim3 grid(persistent_grid((c->ncurv + 63) / 64, 1 << 20));
            if (L.p == 2) {
                if (full) hipLaunchKernelGGL((k_elem_ins32_em_curved<2, true>), grid, dim3(256), 0, c->stream, b, c->eres32);
                else hipLaunchKernelGGL((k_elem_ins32_em_curved<2, false>), grid, dim3(256), 0, c->stream, b, c->eres32);
            } else {
                if (full) hipLaunchKernelGGL((k_elem_ins32_em_curved<4, true>), grid, dim3(256), 0, c->stream, b, c->eres32);
                else hipLaunchKernelGGL((k_elem_ins32_em_curved<4, false>), grid, dim3(256), 0, c->stream, b, c->eres32);
            }
        }
        const LevelDims dd(L.p);
        GatherArgsT<float> ga{};
        ga.npnt = c->npnt; ga.nseg = c->nseg; ga.sm = dd.sm; ga.sm0 = c->sm0; ga.bm = dd.bm;
        ga.vt_ptr = c->vt_ptr; ga.vt_tri = c->vt_tri; ga.vt_loc = c->vt_loc;
        ga.st_tri = c->st_tri; ga.st_loc = c->st_loc; ga.eres = c->eres32;
        ga.resv = c->r32v; ga.rrv = c->rr32v; ga.ress = c->r32s; ga.rrs = c->rr32s;
        ga.keep = (float)f; ga.full = full ? 1 : 0;
        // periodic: the vertex source enters after the copies are summed (once per group)
        ga.srcv = c->periodic ? nullptr : c->resv; ga.srcs = c->ress; ga.sfac = c->fd_scal + 1;
        const long lanes = ((long)c->npnt + (long)c->nseg * dd.sm) * NV;
        hipLaunchKernelGGL((k_gather_vs<float, NV>), dim3(nblk(lanes)), dim3(256), 0, c->stream, ga);
    } else {
        launch_elem32<1>(c, a, full);
    }
    const LevelDims d(L.p);
    EdgeArgsT<float> e{};
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->dlv; e.ugs = c->dls;
    e.resv = c->r32v; e.ress = c->r32s;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.gpx = d.p + 1; e.sm = d.sm; e.sm0 = c->sm0; e.rho = c->prm.rho; e.bc_pressure = c->prm.bc_pressure;
    e.ecrv = c->ecrv; e.scrd = c->scrd;
    e.etype = c->etype; e.ub0 = c->ibc_u[0]; e.ub1 = c->ibc_u[1]; e.mu = c->prm.mu;
    for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
        e.begin = c->eoff[col];
        e.end = c->eoff[col + 1];
        if (e.end > e.begin) {
            const dim3 be(nblk(e.end - e.begin, 64)), te(64);
            if (L.p == 1) hipLaunchKernelGGL((k_edge_ins<1, float>), be, te, 0, c->stream, e);
            else if (L.p == 2) hipLaunchKernelGGL((k_edge_ins<2, float>), be, te, 0, c->stream, e);
            else hipLaunchKernelGGL((k_edge_ins<4, float>), be, te, 0, c->stream, e);
        }
    }
    if (c->periodic) {
        periodic_combine<float>(c, c->r32v);
        hipLaunchKernelGGL(k_add_src32, dim3(nblk((long)c->npnt * NV)), dim3(256), 0, c->stream, (long)c->npnt * NV,
                           c->resv, c->fd_scal + 1, c->r32v);
    }
    check(hipGetLastError(), "rsdl32 launch");
}

// FP32 linearised residual of the inner solve: r32 = src + J(u0) delta (multistage blended)
void do_lin_rsdl(hpg_ctx* c, int stage, bool with_source) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const Level& L = c->lv[c->log2p];
    LinArgs a;
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.dv = c->dlv; a.ds = c->dls; a.di = c->dli;
    a.resv = c->r32v; a.ress = c->r32s; a.resi = c->r32i;
    a.rrv = c->rr32v; a.rrs = c->rr32s; a.rri = c->rr32i;
    a.srcv = with_source ? c->resv : nullptr;
    a.srcs = with_source ? c->ress : nullptr;
    a.srci = with_source ? c->resi : nullptr;
    a.ops = L.pack32; a.lin = c->lin; a.gbase = 0; a.sm0 = c->sm0; a.im0 = c->im0;
    a.rho = (float)c->prm.rho; a.mu = (float)c->prm.mu; a.adis = (float)c->prm.adis;
    a.beta = (float)beta; a.keep = (float)f;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0;
    const bool full = beta > 0.0;
    switch (L.p) {
        case 1: launch_lin<1>(c, a, full); break;
        case 2: launch_lin<2>(c, a, full); break;
        case 4: launch_lin<4>(c, a, full); break;
        default: throw std::runtime_error("unsupported p");
    }
    const LevelDims d(L.p);
    EdgeLinArgs e;
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.dv = c->dlv; e.ds = c->dls; e.resv = c->r32v; e.ress = c->r32s;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.gpx = d.p + 1; e.sm = d.sm; e.sm0 = c->sm0; e.rho = c->prm.rho;
    for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
        e.begin = c->eoff[col];
        e.end = c->eoff[col + 1];
        if (e.end > e.begin) {
            const dim3 be(nblk(e.end - e.begin, 64)), te(64);
            if (L.p == 1) hipLaunchKernelGGL(k_edge_lin<1>, be, te, 0, c->stream, e);
            else if (L.p == 2) hipLaunchKernelGGL(k_edge_lin<2>, be, te, 0, c->stream, e);
            else hipLaunchKernelGGL(k_edge_lin<4>, be, te, 0, c->stream, e);
        }
    }
    check(hipGetLastError(), "lin rsdl launch");
}

template <typename T, int NVn>
void launch_minv_nv(hpg_ctx* c, const MinvArgsT<T>& a) {
    // one lane per (entity, variable); k_minv_sscale one lane per (side, mode, variable)
    const dim3 bv(nblk8((long)c->npnt * NVn)), bs(nblk8((long)c->nseg * NVn)), bt(nblk8((long)c->ntri * NVn));
    const dim3 bq(nblk8((long)c->nseg * a.sm * NVn)), th(256);
    switch (a.sm + 1) {
        case 1:
            hipLaunchKernelGGL((k_minv_vrtx<T, 1, NVn>), bv, th, 0, c->stream, a);
            break;
        case 2:
            hipLaunchKernelGGL((k_minv_vrtx<T, 2, NVn>), bv, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_side<T, 2, NVn>), bs, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_sscale<T, 2, NVn>), bq, th, 0, c->stream, a);
            break;
        case 4:
            hipLaunchKernelGGL((k_minv_vrtx<T, 4, NVn>), bv, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_side<T, 4, NVn>), bs, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_sub<T, 4, 0, NVn>), bs, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_sub<T, 4, 1, NVn>), bs, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_sscale<T, 4, NVn>), bq, th, 0, c->stream, a);
            hipLaunchKernelGGL((k_minv_int<T, 4, NVn>), bt, th, 0, c->stream, a);
            break;
        default:
            throw std::runtime_error("minvrt: unsupported p");
    }
    check(hipGetLastError(), "minvrt launch");
}

template <typename T>
void launch_minv(hpg_ctx* c, const MinvArgsT<T>& a) {
    if (c->nv == NVC) launch_minv_nv<T, NVC>(c, a);
    else launch_minv_nv<T, NV>(c, a);
}

// minvrt; with update_cfa >= 0 the stage update ug = ug0 - update_cfa*res is fused
// into the sweeps that produce the final vertex, side and interior values
void do_minvrt(hpg_ctx* c, double update_cfa = -1.0) {
    MinvArgsT<double> a = minv_args<double>(c);
    if (update_cfa >= 0.0) {
        a.ugv = c->ugv; a.ug0v = c->ug0v; a.cfa = update_cfa;
        if (a.sm > 0) { a.ugs = c->ugs; a.ug0s = c->ug0s; }
        if (a.im > 0) { a.ugi = c->ugi; a.ug0i = c->ug0i; }
    }
    launch_minv<double>(c, a);
}

// FP32 minvrt of the inner residual with the fused delta update
void do_minvrt32(hpg_ctx* c, float cfa) {
    MinvArgsT<float> a = minv_args<float>(c);
    a.ugv = c->dlv; a.ug0v = c->d0v; a.cfa = cfa;
    if (a.sm > 0) { a.ugs = c->dls; a.ug0s = c->d0s; }
    if (a.im > 0) { a.ugi = c->dli; a.ug0i = c->d0i; }
    launch_minv<float>(c, a);
}

// ---- tri_hp_cns drivers ------------------------------------------------------------
template <int P>
void launch_elem_cns(hpg_ctx* c, const CnsElemArgs& a0, bool full, const std::vector<int>* offs = nullptr) {
    const std::vector<int>& off = offs ? *offs : c->coff;
    for (size_t col = 0; col + 1 < off.size(); ++col) {
        CnsElemArgs a = a0;
        a.colour = (int)col;
        a.begin = off[col];
        a.end = off[col + 1];
        const int noct = (a.end - a.begin + 7) / 8;
        if (noct == 0) continue;
        const int nb = std::min(((noct + 7) / 8) * 8, 1024);
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full)
            hipLaunchKernelGGL((k_elem_cns<P, true>), dim3(nb), dim3(256), 0, c->stream, a, noct);
        else
            hipLaunchKernelGGL((k_elem_cns<P, false>), dim3(nb), dim3(256), 0, c->stream, a, noct);
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 1 : 0, a.end - a.begin});
        }
    }
}

template <int P>
void launch_elem_cns32(hpg_ctx* c, const CnsElemArgs32& a0, bool full) {
    for (size_t col = 0; col + 1 < c->coff.size(); ++col) {
        CnsElemArgs32 a = a0;
        a.colour = (int)col;
        a.begin = c->coff[col];
        a.end = c->coff[col + 1];
        const int noct = (a.end - a.begin + 7) / 8;
        if (noct == 0) continue;
        const int nb = std::min(((noct + 7) / 8) * 8, 1024);
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        if (full)
            hipLaunchKernelGGL((k_elem_cns32<P, true>), dim3(nb), dim3(256), 0, c->stream, a, noct);
        else
            hipLaunchKernelGGL((k_elem_cns32<P, false>), dim3(nb), dim3(256), 0, c->stream, a, noct);
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({full ? 3 : 2, a.end - a.begin});
        }
    }
}

// coarse p-level source (tri_hp/nstage.cpp:113-129)
void mg_source(hpg_ctx* c) {
    const int l = c->log2p;
    if (c->isfrst) {
        modes_vsi(c, 2, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], c->r0v, c->r0s, c->r0i,
                  c->prm.fadd);
        c->isfrst = false;
    } else {
        modes_vsi(c, 1, c->resv, c->ress, c->resi, c->dresv[l], c->dress[l], c->dresi[l], nullptr, nullptr, nullptr,
                  0.0);
    }
}

void do_rsdl_cns(hpg_ctx* c, int stage) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const Level& L = c->lv[c->log2p];
    CnsElemArgs a;
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi;
    a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.rrv = c->rrv; a.rrs = c->rrs; a.rri = c->rri;
    a.ops = L.cpack; a.sm0 = c->sm0; a.im0 = c->im0;
    a.ph = c->cph; a.beta = beta; a.keep = f;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0;
    const bool full = beta > 0.0;
    const LevelDims d(L.p);
    CnsEdgeArgs e;
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.resv = c->resv; e.ress = c->ress;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.sm0 = c->sm0; e.gamma = c->cph.gamma; e.pbc = c->cbc[0];
    auto edges = [&](auto kern) {
        for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
            e.begin = c->eoff[col];
            e.end = c->eoff[col + 1];
            if (e.end > e.begin) hipLaunchKernelGGL(kern, dim3(nblk(e.end - e.begin, 64)), dim3(64), 0, c->stream, e);
        }
    };
    switch (L.p) {
        case 1: launch_elem_cns<1>(c, a, full); edges(k_edge_cns<1>); break;
        case 2: launch_elem_cns<2>(c, a, full); edges(k_edge_cns<2>); break;
        case 4: launch_elem_cns<4>(c, a, full); edges(k_edge_cns<4>); break;
        default: throw std::runtime_error("unsupported p");
    }
    if (c->coarse_flag) mg_source(c);
    check(hipGetLastError(), "cns rsdl launch");
}

// cns residual of the defect correction's inner solve at the FP64 state v (ug): the
// element residual in FP32 plus the scaled defect source (1/c - 1) R64(u0) (srcv...),
// the boundary fluxes in FP64
void do_rsdl_cns32(hpg_ctx* c, int stage) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const double f = stage < nst ? 1.0 - beta : 0.0;
    const Level& L = c->lv[c->log2p];
    CnsElemArgs32 a{};
    a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.begin = 0; a.end = 0;
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi;
    a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.rrv = c->rrv; a.rrs = c->rrs; a.rri = c->rri;
    a.ops = L.cpack32; a.sm0 = c->sm0; a.im0 = c->im0;
    a.ph = c->cph; a.beta = beta; a.keep = f;
    a.fcv = c->fcv; a.fcs = c->fcs; a.colour = 0;
    a.srcv = c->srcv; a.srcs = c->srcs; a.srci = c->srci; a.sfac = c->fd_scal + 1;
    const bool full = beta > 0.0;
    const LevelDims d(L.p);
    CnsEdgeArgs e;
    e.eseg = c->eseg; e.seg_pnt = c->seg_pnt; e.pnts = c->pnts; e.ugv = c->ugv; e.ugs = c->ugs;
    e.resv = c->resv; e.ress = c->ress;
    e.g1 = L.edge; e.dg1 = L.edge + (d.p + 1) * (d.sm + 2); e.w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    e.sm0 = c->sm0; e.gamma = c->cph.gamma; e.pbc = c->cbc[0];
    auto edges = [&](auto kern) {
        for (size_t col = 0; col + 1 < c->eoff.size(); ++col) {
            e.begin = c->eoff[col];
            e.end = c->eoff[col + 1];
            if (e.end > e.begin) hipLaunchKernelGGL(kern, dim3(nblk(e.end - e.begin, 64)), dim3(64), 0, c->stream, e);
        }
    };
    switch (L.p) {
        case 1: launch_elem_cns32<1>(c, a, full); edges(k_edge_cns<1>); break;
        case 2: launch_elem_cns32<2>(c, a, full); edges(k_edge_cns<2>); break;
        case 4: launch_elem_cns32<4>(c, a, full); edges(k_edge_cns<4>); break;
        default: throw std::runtime_error("unsupported p");
    }
    check(hipGetLastError(), "cns rsdl32 launch");
}

// sync = false (the fused multigrid cycle): the error flag accumulates on the device and
// is read once after the cycle; nothing waits for the device here
int do_setup_cns(hpg_ctx* c, bool sync = true) {
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    if (sync) check(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream), "memset");
    CnsSetupArgs s;
    s.pnts = c->pnts; s.tri_pnt = c->tri_pnt; s.tri_seg = c->tri_seg; s.tri_sgn = c->tri_sgn;
    s.ugv = c->ugv; s.ugs = c->ugs; s.ugi = c->ugi; s.Pmat = L.Pmat; s.tprcn = c->tprcn; s.tpc = c->tpc;
    s.err = c->d_err; s.ntri = c->ntri; s.sm0 = c->sm0; s.im0 = c->im0; s.ph = c->cph;
    const dim3 be(nblk8(c->ntri, 128)), te(128);
    switch (d.p) {
        case 1: hipLaunchKernelGGL(k_setup_elem_cns<1>, be, te, 0, c->stream, s); break;
        case 2: hipLaunchKernelGGL(k_setup_elem_cns<2>, be, te, 0, c->stream, s); break;
        default: hipLaunchKernelGGL(k_setup_elem_cns<4>, be, te, 0, c->stream, s); break;
    }
    hipLaunchKernelGGL(k_setup_vrtx_cns, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->vt_ptr, c->vt_tri,
                       c->tprcn, c->tpc, c->vprcn, c->vpc, L.ops->vdiag);
    if (d.sm > 0)
        hipLaunchKernelGGL(k_setup_side_cns, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, c->nseg, c->st_tri,
                           c->tprcn, c->tpc, c->sprcn, c->spc);
    if (!sync) return 0;
    int herr = 0;
    check(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "memcpy err");
    check(hipStreamSynchronize(c->stream), "sync");
    return herr;
}

// inflow modify_boundary_residual, then project_new_variables (tri_hp/cns/update.cpp:276-357)
void cns_boundary_and_projection(hpg_ctx* c) {
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    const double* g1 = L.edge;
    const double* w1 = L.edge + 2 * (d.p + 1) * (d.sm + 2);
    CnsBcArgs b;
    b.resv = c->resv; b.ress = c->ress; b.verts = c->inflow_v; b.sides = c->inflow_s;
    b.nverts = c->n_inflow_v; b.nsides = c->n_inflow_s; b.seg_pnt = c->seg_pnt;
    b.g1 = g1; b.w1 = w1; b.sdiag1d = L.sdiag1d; b.sm0 = c->sm0;
    b.bu = c->cbc[1]; b.bv = c->cbc[2];
    b.benergy = c->cbc[3] / (c->cph.gamma - 1.0) + 0.5 * (c->cbc[1] * c->cbc[1] + c->cbc[2] * c->cbc[2]);
    if (b.nverts > 0) hipLaunchKernelGGL(k_mbr_vrtx, dim3(nblk(b.nverts)), dim3(256), 0, c->stream, b);
    if (b.nsides > 0 && d.sm > 0) {
        if (d.p == 2) hipLaunchKernelGGL(k_mbr_side<2>, dim3(nblk(b.nsides)), dim3(256), 0, c->stream, b);
        else hipLaunchKernelGGL(k_mbr_side<4>, dim3(nblk(b.nsides)), dim3(256), 0, c->stream, b);
    }
    CnsProjArgs a;
    a.ugv = c->ugv; a.ugs = c->ugs; a.ugi = c->ugi; a.resv = c->resv; a.ress = c->ress; a.resi = c->resi;
    a.rtv = c->rtv; a.rts = c->rts; a.vpc = c->vpc; a.vess = c->vess; a.sess = c->sess;
    a.seg_pnt = c->seg_pnt; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
    a.order = c->order; a.g1 = g1; a.w1 = w1; a.sdiag1d = L.sdiag1d; a.idiag = L.idiag; a.ops = L.cpack;
    a.npnt = c->npnt; a.nseg = c->nseg; a.ntri = c->ntri; a.sm0 = c->sm0; a.im0 = c->im0; a.gamma = c->cph.gamma;
    hipLaunchKernelGGL(k_pnv_vrtx, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, a);
    if (d.sm > 0) {
        if (d.p == 2) hipLaunchKernelGGL(k_pnv_side<2>, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, a);
        else hipLaunchKernelGGL(k_pnv_side<4>, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, a);
    }
    if (d.im > 0) {
        const int noct = (c->ntri + 7) / 8;
        static const int cap = resident_blocks((const void*)k_pnv_int<4>);
        hipLaunchKernelGGL(k_pnv_int<4>, dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
    }
    check(hipGetLastError(), "cns projection launch");
}

void update_cns(hpg_ctx* c) {
    const size_t nv = (size_t)c->npnt * c->nv * 8, ns = (size_t)c->nseg * c->sm0 * c->nv * 8,
                 ni = (size_t)c->ntri * c->im0 * c->nv * 8;
    dcopy(c, c->ug0v, c->ugv, nv);
    if (ns) dcopy(c, c->ug0s, c->ugs, ns);
    if (ni) dcopy(c, c->ug0i, c->ugi, ni);
    for (int stage = 0; stage < c->prm.nstage; ++stage) {
        do_rsdl_cns(c, stage);
        do_minvrt(c);
        cns_boundary_and_projection(c);
        modes_vsi(c, 4, c->ugv, c->ugs, c->ugi, c->ug0v, c->ug0s, c->ug0i, c->resv, c->ress, c->resi,
                  c->prm.alpha[stage] * c->prm.cfl[c->log2p]);
    }
    check(hipGetLastError(), "cns update launch");
}

// FP32 defect correction of tri_hp_cns::update: the outer iteration of mixed_update with
// FP64 storage.  R64(u0) (judge and defect) is kept in src; the inner state v = fl32(u0)
// + dhat lives in ug, so minvrt, the inflow boundary residual, the primitive projection
// and the stage update run as in update_cns; stage 0 is F(b) = R64(u0)/c and the later
// stages evaluate the element residual in FP32; finally u = u0 + c (v - fl32(u0)).
void mixed_update_cns(hpg_ctx* c) {
    const long nv = (long)c->npnt * c->nv, ns = (long)c->nseg * c->sm0 * c->nv, ni = (long)c->ntri * c->im0 * c->nv;
    do_rsdl_cns(c, c->prm.nstage);
    auto copy = [&](double* dst, const double* src, long n) {
        if (n) dcopy(c, dst, src, n * 8);
    };
    copy(c->srcv, c->resv, nv); copy(c->srcs, c->ress, ns); copy(c->srci, c->resi, ni);
    copy(c->u0cv, c->ugv, nv); copy(c->u0cs, c->ugs, ns); copy(c->u0ci, c->ugi, ni);
    hipLaunchKernelGGL(k_round32, dim3(std::min<unsigned>(nblk(nv), RED_BLOCKS)), dim3(256), 0, c->stream, nv, c->ugv, c->fd_max);
    if (ns) hipLaunchKernelGGL(k_round32, dim3(std::min<unsigned>(nblk(ns), RED_BLOCKS)), dim3(256), 0, c->stream, ns, c->ugs, c->fd_max);
    if (ni) hipLaunchKernelGGL(k_round32, dim3(std::min<unsigned>(nblk(ni), RED_BLOCKS)), dim3(256), 0, c->stream, ni, c->ugi, c->fd_max);
    hipLaunchKernelGGL(k_fd_scale, dim3(1), dim3(1), 0, c->stream, c->fd_scal, c->fd_max, c->fd_reset ? 1 : 0, FD_H);
    c->fd_reset = false;
    for (int it = 0; it < c->ninner; ++it) {
        copy(c->ug0v, c->ugv, nv); copy(c->ug0s, c->ugs, ns); copy(c->ug0i, c->ugi, ni);
        for (int stage = 0; stage < c->prm.nstage; ++stage) {
            if (it == 0 && stage == 0) {
                // res and res_r hold R64(u0) and its viscous part: the residual at v = b
                hipLaunchKernelGGL(k_scale64, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->resv, c->fd_scal);
                if (ns) hipLaunchKernelGGL(k_scale64, dim3(nblk(ns)), dim3(256), 0, c->stream, ns, c->ress, c->fd_scal);
                if (ni) hipLaunchKernelGGL(k_scale64, dim3(nblk(ni)), dim3(256), 0, c->stream, ni, c->resi, c->fd_scal);
            } else {
                do_rsdl_cns32(c, stage);
            }
            do_minvrt(c);
            cns_boundary_and_projection(c);
            modes_vsi(c, 4, c->ugv, c->ugs, c->ugi, c->ug0v, c->ug0s, c->ug0i, c->resv, c->ress, c->resi,
                      c->prm.alpha[stage] * c->prm.cfl[c->log2p]);
        }
    }
    hipLaunchKernelGGL(k_fd_final64, dim3(std::min<unsigned>(nblk(nv), RED_BLOCKS)), dim3(256), 0, c->stream, nv, c->ugv, c->u0cv, c->fd_scal, c->fd_max + 1);
    if (ns) hipLaunchKernelGGL(k_fd_final64, dim3(std::min<unsigned>(nblk(ns), RED_BLOCKS)), dim3(256), 0, c->stream, ns, c->ugs, c->u0cs, c->fd_scal, c->fd_max + 1);
    if (ni) hipLaunchKernelGGL(k_fd_final64, dim3(std::min<unsigned>(nblk(ni), RED_BLOCKS)), dim3(256), 0, c->stream, ni, c->ugi, c->u0ci, c->fd_scal, c->fd_max + 1);
    check(hipGetLastError(), "cns mixed update launch");
}

int do_setup(hpg_ctx* c, bool sync = true) {
    if (c->physics == HPG_PHYSICS_CNS) return do_setup_cns(c, sync);
    const Level& L = c->lv[c->log2p];
    const LevelDims d(L.p);
    if (sync) check(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream), "memset");
    SetupArgs s{};
    s.pnts = c->pnts; s.tri_pnt = c->tri_pnt; s.tri_seg = c->tri_seg; s.tri_sgn = c->tri_sgn;
    s.ugv = c->ugv; s.ugs = c->ugs; s.ugi = c->ugi; s.Pmat = L.Pmat; s.tprcn = c->tprcn; s.err = c->d_err;
    s.ntri = c->ntri; s.p = d.p; s.sm = d.sm; s.im = d.im; s.tm = d.tm; s.G = d.G; s.sm0 = c->sm0; s.im0 = c->im0;
    s.rho = c->prm.rho; s.mu = c->prm.mu; s.adis = c->prm.adis;
    switch (d.p) {
        case 1: hipLaunchKernelGGL(k_setup_elem<1>, dim3(nblk8(c->ntri, 128)), dim3(128), 0, c->stream, s); break;
        case 2: hipLaunchKernelGGL(k_setup_elem<2>, dim3(nblk8(c->ntri, 128)), dim3(128), 0, c->stream, s); break;
        default: hipLaunchKernelGGL(k_setup_elem<4>, dim3(nblk8(c->ntri, 128)), dim3(128), 0, c->stream, s); break;
    }
    if (c->ncurv && d.p >= 2) {
        // the curved branch overwrites its elements' tprcn (the reference's setinfo marks
        // curved elements only for p >= 2, tri_hp/metric.cpp:45-70)
        s.clist = c->clist; s.ncurv = c->ncurv; s.crv = c->crv;
        s.PRmat = c->PRmat[c->log2p]; s.PSmat = c->PSmat[c->log2p];
        if (d.p == 2) hipLaunchKernelGGL(k_setup_elem_curved<2>, dim3(nblk(c->ncurv, 64)), dim3(64), 0, c->stream, s);
        else hipLaunchKernelGGL(k_setup_elem_curved<4>, dim3(nblk(c->ncurv, 64)), dim3(64), 0, c->stream, s);
    }
    // learned preconditioner on this level: element features -> MLP (or the caller's
    // log scalings) -> separate scalings of the vertex, side and interior rows
    c->typed_prcn = c->prcn_mode == HPG_PRCN_LEARNED && (c->mlp || c->user_y);
    if (c->typed_prcn) {
        if (!c->user_y) {
            hipLaunchKernelGGL(k_mlp_features, dim3(nblk(c->ntri)), dim3(256), 0, c->stream, c->ntri, c->pnts,
                               c->tri_pnt, c->ugv, c->tprcn, c->prm.mu, c->prm.rho, c->featX);
            mlp_forward_device(c->mlp, c->featX, c->ntri, c->mlpY, c->stream);
        }
        hipLaunchKernelGGL(k_mlp_apply, dim3(nblk((long)c->ntri * NV)), dim3(256), 0, c->stream, c->ntri, c->mlpY,
                           c->tprcn, c->tprcn_vx, c->tprcn_sd, c->tprcn_in);
    }
    hipLaunchKernelGGL(k_setup_vrtx, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->vtu_ptr, c->vtu_tri,
                       c->typed_prcn ? c->tprcn_vx : c->tprcn, c->vprcn, L.ops->vdiag);
    if (d.sm > 0)
        hipLaunchKernelGGL(k_setup_side, dim3(nblk(c->nseg)), dim3(256), 0, c->stream, c->nseg, c->st_tri,
                           c->typed_prcn ? c->tprcn_sd : c->tprcn, c->sprcn);
    if (!sync) return 0;
    int herr = 0;
    check(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "memcpy err");
    check(hipStreamSynchronize(c->stream), "sync");
    return herr;
}

// max over (nent, nm of mstride, NV) of |x|, per variable
void absmax(hpg_ctx* c, long nent, int nm, int mstride, const double* x, const unsigned char* mask, double out[NVC]) {
    unsigned long long h[NVC] = {0, 0, 0, 0};
    if (nent > 0 && nm > 0) {
        check(hipMemsetAsync(c->d_max, 0, NVC * sizeof(unsigned long long), c->stream), "memset");
        const long n = nent * nm;
        const unsigned nb = std::min<unsigned>(nblk(n), 2048u);
        if (c->nv == NVC)
            hipLaunchKernelGGL(k_absmax<NVC>, dim3(nb), dim3(256), 0, c->stream, nent, nm, mstride, x, mask, c->d_max);
        else
            hipLaunchKernelGGL(k_absmax<NV>, dim3(nb), dim3(256), 0, c->stream, nent, nm, mstride, x, mask, c->d_max);
        check(hipMemcpyAsync(h, c->d_max, sizeof(h), hipMemcpyDeviceToHost, c->stream), "memcpy");
        check(hipStreamSynchronize(c->stream), "sync");
    }
    for (int k = 0; k < NVC; ++k) {
        double v;
        std::memcpy(&v, &h[k], sizeof(double));
        out[k] = v;
    }
}

}  // namespace

// device buffers of one call, released when the call ends
struct CallBuffers {
    hipStream_t stream;
    std::vector<void*> p;
    explicit CallBuffers(hipStream_t s) : stream(s) {}
    template <class T>
    T* get(size_t n, bool zero = false) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        p.push_back(q);
        if (zero) check(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), stream), "memset");
        return static_cast<T*>(q);
    }
    ~CallBuffers() {
        hipStreamSynchronize(stream);
        for (void* q : p) hipFree(q);
    }
};

template <int P, int NVn>
void jacobian_colour(hpg_ctx* c, const JacArgs& ja, int N) {
    const int nel = ja.end - ja.begin;
    const dim3 b1(nblk(nel, 64)), t1(64), bN(nblk((long)nel * N)), tN(256);
    const Level& L = c->lv[c->log2p];
    const double beta = c->prm.beta[0];
    const bool full = beta > 0.0;
    // one launch over the colour's batch elements.  The kernel's first-visit flags are
    // keyed to colour 0 here, not to the elements' own colour; that does not matter:
    // every scratch entry starts at zero, so both lftog paths store lf_re + lf_im
    const std::vector<int> offs = {ja.begin, ja.end};
    hipLaunchKernelGGL(k_jac_dw<NVn>, b1, t1, 0, c->stream, ja);
    for (int col = -1; col < N; ++col) {
        JacArgs a = ja;
        a.mode = col < 0 ? -1 : col / NVn;
        a.var = col < 0 ? 0 : col % NVn;
        if (col >= 0) hipLaunchKernelGGL((k_jac_perturb<P, NVn>), b1, t1, 0, c->stream, a);
        if (c->physics == HPG_PHYSICS_CNS) {
            CnsElemArgs e;
            e.pnts = c->pnts; e.tri_pnt = c->tri_pnt; e.tri_seg = c->tri_seg; e.tri_sgn = c->tri_sgn;
            e.order = ja.jel; e.begin = 0; e.end = 0;
            e.ugv = c->ugv; e.ugs = c->ugs; e.ugi = c->ugi;
            e.resv = ja.jrv; e.ress = ja.jrs; e.resi = ja.jri;
            e.rrv = ja.jqv; e.rrs = ja.jqs; e.rri = ja.jqi;
            e.ops = L.cpack; e.sm0 = c->sm0; e.im0 = c->im0;
            e.ph = c->cph; e.beta = beta; e.keep = 0.0;
            e.fcv = c->fcv; e.fcs = c->fcs; e.colour = 0;
            launch_elem_cns<P>(c, e, full, &offs);
        } else {
            ElemArgs e{};
            e.pnts = c->pnts; e.tri_pnt = c->tri_pnt; e.tri_seg = c->tri_seg; e.tri_sgn = c->tri_sgn;
            e.order = ja.jel; e.begin = 0; e.end = 0;
            e.ugv = c->ugv; e.ugs = c->ugs; e.ugi = c->ugi;
            e.resv = ja.jrv; e.ress = ja.jrs; e.resi = ja.jri;
            e.rrv = ja.jqv; e.rrs = ja.jqs; e.rri = ja.jqi;
            e.ops = L.pack; e.sm0 = c->sm0; e.im0 = c->im0;
            e.rho = c->prm.rho; e.mu = c->prm.mu; e.adis = c->prm.adis; e.beta = beta;
            e.fcv = c->fcv; e.fcs = c->fcs; e.colour = 0; e.keep = 0.0;
            e.lin = nullptr; e.gbase = 0;
            launch_elem<P>(c, e, full, &offs);
        }
        hipLaunchKernelGGL((k_jac_collect<P, NVn>), bN, tN, 0, c->stream, a);
    }
    check(hipGetLastError(), "jacobian launch");
}

extern "C" {

namespace {
// the boundary-edge lists of the flux kernels: `sides` (boundary sides, edge boundary
// bid of each) coloured so that no two edges of a colour share a vertex, each edge's
// edge boundary, and the edge endpoints at each vertex (entries 2*idx + k of eout)
void build_edges(hpg_ctx* c, const int* seg_pnt, int np, const std::vector<int>& sides, const std::vector<int>& bid) {
    const int neseg = (int)sides.size();
    std::vector<int> ev(2 * (size_t)neseg);
    for (int j = 0; j < neseg; ++j) {
        ev[2 * j] = seg_pnt[2 * sides[j]];
        ev[2 * j + 1] = seg_pnt[2 * sides[j] + 1];
    }
    std::vector<int> eorder, eoffv, ecol, eseg(neseg);
    colour(neseg, 2, ev.data(), np, eorder, eoffv, ecol);
    c->h_ebid.assign(neseg, 0);
    for (int j = 0; j < neseg; ++j) {
        eseg[j] = sides[eorder[j]];
        c->h_ebid[j] = bid[eorder[j]];
    }
    c->eseg = c->upload(eseg.data(), eseg.size());
    c->h_eseg = eseg;
    c->eoff = eoffv;
    c->neseg = neseg;
    std::vector<int> vb_ptr(np + 1, 0), vb_ent(2 * (size_t)neseg);
    for (int j = 0; j < neseg; ++j)
        for (int k = 0; k < 2; ++k) vb_ptr[seg_pnt[2 * eseg[j] + k] + 1]++;
    for (int v = 0; v < np; ++v) vb_ptr[v + 1] += vb_ptr[v];
    std::vector<int> f(vb_ptr.begin(), vb_ptr.end() - 1);
    for (int j = 0; j < neseg; ++j)
        for (int k = 0; k < 2; ++k) vb_ent[f[seg_pnt[2 * eseg[j] + k]]++] = 2 * j + k;
    c->vb_ptr = c->upload(vb_ptr.data(), vb_ptr.size());
    c->vb_ent = c->upload(vb_ent.data(), vb_ent.size());
}
}  // namespace

int hpg_create(const hpg_tri_mesh* m, const hpg_params* prm, hpg_ctx** out) {
    if (!m || !prm || !out) return -2;
    *out = nullptr;
    if ((prm->physics != HPG_PHYSICS_INS && prm->physics != HPG_PHYSICS_CNS) || prm->log2p < 0 || prm->log2p > 2 ||
        prm->nstage < 1 || prm->nstage > 7)
        return -2;
    std::unique_ptr<hpg_ctx> c(new hpg_ctx());
    try {
        check(hipStreamCreate(&c->stream), "hipStreamCreate");
        c->prm = *prm;
        c->prm.alpha[prm->nstage] = 1.0;
        c->prm.beta[prm->nstage] = 1.0;
        c->physics = prm->physics;
        c->nv = prm->physics == HPG_PHYSICS_CNS ? NVC : NV;
        if (c->physics == HPG_PHYSICS_CNS) {
            // tri_hp/cns/allocate.cpp:45: kcond = mu/Pr * R gamma/(gamma-1)
            c->cph.mu = prm->mu;
            c->cph.gamma = prm->gamma;
            c->cph.Rgas = prm->Rgas;
            c->cph.adis = prm->adis;
            c->cph.kcond = prm->mu / prm->prandtl * prm->Rgas * prm->gamma / (prm->gamma - 1.0);
            for (int n = 0; n < 4; ++n) c->cbc[n] = prm->bc[n];
            if (!(prm->gamma > 1.0) || !(prm->prandtl > 0.0) || !(prm->Rgas > 0.0)) throw std::runtime_error("cns constants");
        }
        c->log2pmax = c->log2p = prm->log2p;
        const int p0 = 1 << prm->log2p;
        c->sm0 = p0 - 1;
        c->im0 = (p0 - 1) * (p0 - 2) / 2;
        c->npnt = m->npnt; c->nseg = m->nseg; c->ntri = m->ntri;
        const int np = m->npnt, ns = m->nseg, nt = m->ntri;
        // host checks of every index array the kernels gather through
        if (np <= 0 || ns <= 0 || nt <= 0 || m->nebd < 0) throw std::runtime_error("empty mesh");
        if (!m->pnts || !m->seg_pnt || !m->tri_pnt || !m->tri_seg || !m->tri_sgn) throw std::runtime_error("null mesh array");
        for (long k = 0; k < 3L * nt; ++k) {
            if (m->tri_pnt[k] < 0 || m->tri_pnt[k] >= np) throw std::runtime_error("tri_pnt out of range");
            if (m->tri_seg[k] < 0 || m->tri_seg[k] >= ns) throw std::runtime_error("tri_seg out of range");
            if (m->tri_sgn[k] != 1 && m->tri_sgn[k] != -1) throw std::runtime_error("tri_sgn must be +-1");
        }
        for (long k = 0; k < 2L * ns; ++k)
            if (m->seg_pnt[k] < 0 || m->seg_pnt[k] >= np) throw std::runtime_error("seg_pnt out of range");
        // a mesh without edge boundaries may pass nebd = 0 and null boundary arrays
        int neseg = 0;
        if (m->nebd > 0) {
            if (!m->ebdry_start || !m->ebdry_essential || m->ebdry_start[0] != 0) throw std::runtime_error("ebdry_start");
            for (int b = 0; b < m->nebd; ++b)
                if (m->ebdry_start[b + 1] < m->ebdry_start[b]) throw std::runtime_error("ebdry_start not ascending");
            neseg = m->ebdry_start[m->nebd];
            if (neseg > 0 && !m->ebdry_seg) throw std::runtime_error("null ebdry_seg");
            for (int k = 0; k < neseg; ++k)
                if (m->ebdry_seg[k] < 0 || m->ebdry_seg[k] >= ns) throw std::runtime_error("ebdry_seg out of range");
        }

        c->pnts = c->upload(m->pnts, 2 * (size_t)np);
        c->tri_pnt = c->upload(m->tri_pnt, 3 * (size_t)nt);
        c->tri_seg = c->upload(m->tri_seg, 3 * (size_t)nt);
        c->tri_sgn = c->upload(m->tri_sgn, 3 * (size_t)nt);
        c->seg_pnt = c->upload(m->seg_pnt, 2 * (size_t)ns);

        std::vector<int> order, off, tcol;
        colour(nt, 3, m->tri_pnt, np, order, off, tcol);
        std::vector<unsigned char> fcv(np, 255), fcs(ns, 255);
        for (int t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) {
                unsigned char& a = fcv[m->tri_pnt[3 * t + k]];
                unsigned char& b = fcs[m->tri_seg[3 * t + k]];
                a = std::min<unsigned char>(a, (unsigned char)tcol[t]);
                b = std::min<unsigned char>(b, (unsigned char)tcol[t]);
            }
        c->fcv = c->upload(fcv.data(), fcv.size());
        c->fcs = c->upload(fcs.data(), fcs.size());
        c->order = c->upload(order.data(), order.size());
        c->coff = off;
        c->tcol = tcol;
        {
            std::vector<int> sides, bid(neseg, 0);
            if (neseg > 0) sides.assign(m->ebdry_seg, m->ebdry_seg + neseg);
            for (int b = 0; b < m->nebd; ++b)
                for (int j = m->ebdry_start[b]; j < m->ebdry_start[b + 1]; ++j) bid[j] = b;
            build_edges(c.get(), m->seg_pnt, np, sides, bid);
        }
        // a mesh without edge boundaries may pass null arrays (checked above)
        if (m->nebd > 0) {
            c->h_ebdry_start.assign(m->ebdry_start, m->ebdry_start + m->nebd + 1);
            c->h_ebdry_ess.assign(m->ebdry_essential, m->ebdry_essential + m->nebd);
            if (neseg > 0) c->h_ebdry_seg.assign(m->ebdry_seg, m->ebdry_seg + neseg);
        } else {
            c->h_ebdry_start.assign(1, 0);
        }

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
            const unsigned char mk = (unsigned char)(m->ebdry_essential[b] & (c->nv == NVC ? 15 : 7));
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
        c->vtu_ptr = c->vt_ptr; c->vtu_tri = c->vt_tri; c->vtu_loc = c->vt_loc;
        c->vsu_ptr = c->vs_ptr; c->vsu_side = c->vs_side; c->vsu_end = c->vs_end;
        c->st_tri = c->upload(st_tri.data(), st_tri.size());
        c->st_loc = c->upload(st_loc.data(), st_loc.size());
        c->vess = c->upload(vess.data(), vess.size());
        c->sess = c->upload(sess.data(), sess.size());

        const int NVX = c->nv;
        const size_t nv = (size_t)np * NVX, nsz = (size_t)ns * c->sm0 * NVX, ni = (size_t)nt * c->im0 * NVX;
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
        c->sprcn = c->alloc<double>((size_t)ns * NVX);
        c->tprcn = c->alloc<double>((size_t)nt * NVX);
        c->d_err = c->alloc<int>(1);
        c->d_max = c->alloc<unsigned long long>(NVC);
        if (c->physics == HPG_PHYSICS_CNS) {
            c->vpc = c->alloc<double>((size_t)np * 16);
            c->spc = c->alloc<double>((size_t)ns * 16);
            c->tpc = c->alloc<double>((size_t)nt * 16);
            c->rtv = c->alloc<double>(nv);
            c->rts = c->alloc<double>(nsz);
            // inflow boundaries (essential u, v, RT): their vertices (once each) and sides
            std::vector<int> iv, is;
            std::vector<unsigned char> seen(np, 0);
            for (int b = 0; b < m->nebd; ++b) {
                if (m->ebdry_essential[b] != 0b1110) continue;
                for (int j = m->ebdry_start[b]; j < m->ebdry_start[b + 1]; ++j) {
                    const int sg = m->ebdry_seg[j];
                    is.push_back(sg);
                    for (int e = 0; e < 2; ++e) {
                        const int v = m->seg_pnt[2 * sg + e];
                        if (!seen[v]) { seen[v] = 1; iv.push_back(v); }
                    }
                }
            }
            c->n_inflow_v = (int)iv.size();
            c->n_inflow_s = (int)is.size();
            c->inflow_v = c->upload(iv.data(), iv.size());
            c->inflow_s = c->upload(is.data(), is.size());
        }

        // mixed precision: 16-element groups per colour, FP32 inner arrays
        c->gbase.assign(c->coff.size(), 0);
        for (size_t col = 0; col + 1 < c->coff.size(); ++col)
            c->gbase[col + 1] = c->gbase[col] + ((c->coff[col + 1] - c->coff[col] + 63) / 64) * 4;
        c->ngroups = c->gbase.back();
        const LevelDims dmax(p0);
        c->lin = c->alloc<float>((size_t)c->ngroups * 5 * dmax.GT * 4 * 64);
        float** f32v[] = {&c->dlv, &c->d0v, &c->r32v, &c->rr32v};
        float** f32s[] = {&c->dls, &c->d0s, &c->r32s, &c->rr32s};
        float** f32i[] = {&c->dli, &c->d0i, &c->r32i, &c->rr32i};
        for (auto p : f32v) *p = c->alloc<float>(nv);
        for (auto p : f32s) *p = c->alloc<float>(nsz);
        for (auto p : f32i) *p = c->alloc<float>(ni);
        c->b32v = c->alloc<float>(nv);
        c->b32s = c->alloc<float>(nsz);
        c->b32i = c->alloc<float>(ni);
        c->fd_scal = c->alloc<double>(2);
        c->fd_max = c->alloc<unsigned>(2);
        c->vp32 = c->alloc<float>(nv);
        c->sp32 = c->alloc<float>((size_t)ns * NV);
        c->tp32 = c->alloc<float>((size_t)nt * NV);
        c->featX = c->alloc<float>((size_t)nt * MLP_IN);
        c->mlpY = c->alloc<float>((size_t)nt * MLP_OUT);
        c->jdiag = c->alloc<double>((size_t)nt * 3 * NV);
        if (c->physics == HPG_PHYSICS_INS) {
            // element-major vertex and side modes of the finest level (3p per element)
            c->eres = c->alloc<double>((size_t)nt * 3 * (1 << c->log2pmax) * 2 * NV);
            c->eres32 = c->alloc<float>((size_t)nt * 3 * (1 << c->log2pmax) * 2 * NV);
            c->eout = c->alloc<double>((size_t)neseg * 2 * NV);
            c->tprcn_vx = c->alloc<double>((size_t)nt * NV);
            c->tprcn_sd = c->alloc<double>((size_t)nt * NV);
            c->tprcn_in = c->alloc<double>((size_t)nt * NV);
            c->tp32s = c->alloc<float>((size_t)nt * NV);
            c->tp32i = c->alloc<float>((size_t)nt * NV);
        }

        for (int l = 0; l <= c->log2pmax; ++l) {
            Level& L = c->lv[l];
            L.p = 1 << l;
            L.ops.reset(new TriOps(L.p));
            const TriOps& o = *L.ops;
            const std::vector<double> pk = build_pack(o, [](int i, int q) { return 4 * i + q; });
            L.pack = c->upload(pk.data(), pk.size());
            if (L.p == 1) {
                std::vector<double> d1(OP1_SIZE);
                std::copy(o.P.begin(), o.P.end(), d1.begin() + OP1_P);
                std::copy(o.Pr.begin(), o.Pr.end(), d1.begin() + OP1_PR);
                std::copy(o.Ps.begin(), o.Ps.end(), d1.begin() + OP1_PS);
                std::copy(o.Ir.begin(), o.Ir.end(), d1.begin() + OP1_IR);
                std::copy(o.Is.begin(), o.Is.end(), d1.begin() + OP1_IS);
                std::copy(o.Dr.begin(), o.Dr.end(), d1.begin() + OP1_DR);
                std::copy(o.Ds.begin(), o.Ds.end(), d1.begin() + OP1_DS);
                if (o.P.size() != 12 || o.Ir.size() != 12 || o.Dr.size() != 16)
                    throw std::runtime_error("p = 1 operator sizes");
                L.op1 = c->upload(d1.data(), d1.size());
            }
            const std::vector<double> pk32d = build_pack(o, [](int i, int q) { return f32_row(i, q); });
            const std::vector<float> pk32(pk32d.begin(), pk32d.end());
            L.pack32 = c->upload(pk32.data(), pk32.size());
            const std::vector<double> pkn32d = build_pack(o, [](int i, int q) { return f32_row(i, q); }, true);
            const std::vector<float> pkn32(pkn32d.begin(), pkn32d.end());
            L.packn32 = c->upload(pkn32.data(), pkn32.size());
            L.Pmat = c->upload(o.P.data(), o.P.size());
            std::vector<double> lump;
            auto app = [&](const std::vector<double>& v) { const int o2 = (int)lump.size(); lump.insert(lump.end(), v.begin(), v.end()); return o2; };
            L.o_sfmv = app(o.sfmv); L.o_ifmb = app(o.ifmb); L.o_vfms = app(o.vfms); L.o_sdiag = app(o.sdiag);
            L.o_sfms = app(o.sfms); L.o_bfmi = app(o.bfmi); L.o_idiag = app(o.idiag);
            L.lump = c->upload(lump.data(), lump.size());
            const std::vector<float> lump32(lump.begin(), lump.end());
            L.lump32 = c->upload(lump32.data(), lump32.size());
            std::vector<double> ed;
            ed.insert(ed.end(), o.g1.begin(), o.g1.end());
            ed.insert(ed.end(), o.dg1.begin(), o.dg1.end());
            ed.insert(ed.end(), o.w1.begin(), o.w1.end());
            L.edge = c->upload(ed.data(), ed.size());
            if (c->physics == HPG_PHYSICS_CNS) {
                const std::vector<double> cp = build_cns_pack(o);
                L.cpack = c->upload(cp.data(), cp.size());
                const std::vector<float> cp32(cp.begin(), cp.end());
                L.cpack32 = c->upload(cp32.data(), cp32.size());
                L.sdiag1d = c->upload(o.sdiag1d.data(), o.sdiag1d.size());
                const std::vector<double> idg = o.idiag.empty() ? std::vector<double>(2, 1.0) : o.idiag;
                L.idiag = c->upload(idg.data(), idg.size());
            }
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
        const size_t nv = (size_t)c->npnt * c->nv * 8, ns = (size_t)c->nseg * c->sm0 * c->nv * 8,
                     ni = (size_t)c->ntri * c->im0 * c->nv * 8;
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
    c->fd_reset = true;
    const int rc = copy3(c, true, c->ugv, c->ugs, c->ugi, v, s, i, nullptr, nullptr, nullptr);
    if (rc == 0 && c->periodic) HPG_TRY(c, periodic_sync(c));
    return rc;
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
        check(hipMemcpyAsync(vp, c->vprcn, (size_t)c->npnt * c->nv * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(sp, c->sprcn, (size_t)c->nseg * c->nv * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(tp, c->tprcn, (size_t)c->ntri * c->nv * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}
int hpg_get_block_preconditioner(hpg_ctx* c, double* vpc, double* spc, double* tpc) {
    if (!c || c->physics != HPG_PHYSICS_CNS) return -2;
    HPG_TRY(c, {
        if (vpc) check(hipMemcpyAsync(vpc, c->vpc, (size_t)c->npnt * 16 * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        if (spc) check(hipMemcpyAsync(spc, c->spc, (size_t)c->nseg * 16 * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        if (tpc) check(hipMemcpyAsync(tpc, c->tpc, (size_t)c->ntri * 16 * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_set_tprcn(hpg_ctx* c, const double* tp) {
    if (!c) return -2;
    HPG_TRY(c, {
        check(hipMemcpyAsync(c->tprcn, tp, (size_t)c->ntri * c->nv * 8, hipMemcpyHostToDevice, c->stream), "H2D");
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

namespace {

void launch_f32(hpg_ctx* c, long n, const double* x, float* y) {
    if (n > 0) hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, c->stream, n, x, y);
}

// One outer defect-correction iteration: the FP64 residual R64(u0) (the judge and the
// defect), then `ninner` multistage iterations of the reference's scheme in FP32 on
// F(v) = R32(v) + (1/c - 1) R64(u0) from v = b = fl32(u0) with the FP32 minvrt, then
// u0 += c (v - b) in FP64.  F(b + dhat) ~ R64(u0)/c + J(u0) dhat, so the inner solve
// is the scaled defect equation; c keeps |dhat| near FD_H |u0| (k_fd_scale).  Stage 0
// of the first inner iteration is the full residual at v = b, i.e. R64(u0)/c with the
// FP64 viscous part carried into the stage blend, so it costs no FP32 evaluation: an
// update evaluates one FP64 and nstage - 1 FP32 residuals, as many as the FP64 scheme.
void mixed_update(hpg_ctx* c) {
    const long nv = (long)c->npnt * NV, nsz = (long)c->nseg * c->sm0 * NV, ni = (long)c->ntri * c->im0 * NV;
    do_rsdl(c, c->prm.nstage);
    launch_f32(c, nv, c->vprcn, c->vp32);
    launch_f32(c, (long)c->nseg * NV, c->sprcn, c->sp32);
    launch_f32(c, (long)c->ntri * NV, c->tprcn, c->tp32);
    if (c->typed_prcn) {
        launch_f32(c, (long)c->ntri * NV, c->tprcn_sd, c->tp32s);
        launch_f32(c, (long)c->ntri * NV, c->tprcn_in, c->tp32i);
    }
    hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(nv), RED_BLOCKS)), dim3(256), 0, c->stream, nv, c->ugv, c->b32v, c->dlv, c->fd_max);
    if (nsz) hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(nsz), RED_BLOCKS)), dim3(256), 0, c->stream, nsz, c->ugs, c->b32s, c->dls, c->fd_max);
    if (ni) hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(ni), RED_BLOCKS)), dim3(256), 0, c->stream, ni, c->ugi, c->b32i, c->dli, c->fd_max);
    hipLaunchKernelGGL(k_fd_scale, dim3(1), dim3(1), 0, c->stream, c->fd_scal, c->fd_max, c->fd_reset ? 1 : 0, FD_H);
    c->fd_reset = false;
    for (int it = 0; it < c->ninner; ++it) {
        dcopy(c, c->d0v, c->dlv, nv * 4);
        if (nsz) dcopy(c, c->d0s, c->dls, nsz * 4);
        if (ni) dcopy(c, c->d0i, c->dli, ni * 4);
        for (int stage = 0; stage < c->prm.nstage; ++stage) {
            if (it == 0 && stage == 0) {
                hipLaunchKernelGGL(k_scale_f32, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->resv, c->fd_scal, c->r32v);
                if (nsz) hipLaunchKernelGGL(k_scale_f32, dim3(nblk(nsz)), dim3(256), 0, c->stream, nsz, c->ress, c->fd_scal, c->r32s);
                if (ni) hipLaunchKernelGGL(k_scale_f32, dim3(nblk(ni)), dim3(256), 0, c->stream, ni, c->resi, c->fd_scal, c->r32i);
                launch_f32(c, nv, c->rrv, c->rr32v);
                launch_f32(c, nsz, c->rrs, c->rr32s);
                launch_f32(c, ni, c->rri, c->rr32i);
            } else {
                do_rsdl32(c, stage);
            }
            do_minvrt32(c, (float)(c->prm.alpha[stage] * c->prm.cfl[c->log2p]));
        }
    }
    // |dhat| of this iteration (read by the next k_fd_scale), then the FP64 update
    hipLaunchKernelGGL(k_diffmax32, dim3(std::min<unsigned>(nblk(nv), RED_BLOCKS)), dim3(256), 0, c->stream, nv, c->dlv, c->b32v, c->fd_max + 1);
    if (nsz) hipLaunchKernelGGL(k_diffmax32, dim3(std::min<unsigned>(nblk(nsz), RED_BLOCKS)), dim3(256), 0, c->stream, nsz, c->dls, c->b32s, c->fd_max + 1);
    if (ni) hipLaunchKernelGGL(k_diffmax32, dim3(std::min<unsigned>(nblk(ni), RED_BLOCKS)), dim3(256), 0, c->stream, ni, c->dli, c->b32i, c->fd_max + 1);
    hipLaunchKernelGGL(k_add_scaled, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->ugv, c->dlv, c->b32v, c->fd_scal);
    if (nsz) hipLaunchKernelGGL(k_add_scaled, dim3(nblk(nsz)), dim3(256), 0, c->stream, nsz, c->ugs, c->dls, c->b32s, c->fd_scal);
    if (ni) hipLaunchKernelGGL(k_add_scaled, dim3(nblk(ni)), dim3(256), 0, c->stream, ni, c->ugi, c->dli, c->b32i, c->fd_scal);
    check(hipGetLastError(), "mixed update launch");
}

}  // namespace

// FP32 residual R32(fl32(u)) of the current state (finest level, no defect source):
// the kernels of the defect correction's inner solve, for tests
int hpg_residual_fp32(hpg_ctx* c, double* ov, double* os, double* oi) {
    if (!c || !ov) return -2;
    if (c->log2p != c->log2pmax) { c->err = "finest level only"; return -2; }
    if (c->physics == HPG_PHYSICS_CNS) {
        // FP64 state and residual arrays, FP32 element arithmetic; the residual arrays
        // are overwritten
        HPG_TRY(c, {
            hipLaunchKernelGGL(k_fd_scale, dim3(1), dim3(1), 0, c->stream, c->fd_scal, c->fd_max, 1, 1.0);
            c->fd_reset = true;
            do_rsdl_cns32(c, c->prm.nstage);
        });
        return hpg_get_residual(c, ov, os, oi);
    }
    HPG_TRY(c, {
        const long nv = (long)c->npnt * NV, nsz = (long)c->nseg * c->sm0 * NV, ni = (long)c->ntri * c->im0 * NV;
        hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(nv), RED_BLOCKS)), dim3(256), 0, c->stream, nv, c->ugv, c->b32v, c->dlv, c->fd_max);
        if (nsz) hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(nsz), RED_BLOCKS)), dim3(256), 0, c->stream, nsz, c->ugs, c->b32s, c->dls, c->fd_max);
        if (ni) hipLaunchKernelGGL(k_base32, dim3(std::min<unsigned>(nblk(ni), RED_BLOCKS)), dim3(256), 0, c->stream, ni, c->ugi, c->b32i, c->dli, c->fd_max);
        // scale 1: the source factor 1/c - 1 is zero
        hipLaunchKernelGGL(k_fd_scale, dim3(1), dim3(1), 0, c->stream, c->fd_scal, c->fd_max, 1, 1.0);
        c->fd_reset = true;
        do_rsdl32(c, c->prm.nstage);
        std::vector<float> h((size_t)std::max(nv, std::max(nsz, ni)));
        auto get = [&](const float* d, long n, double* out) {
            if (n == 0) return;
            check(hipMemcpyAsync(h.data(), d, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
            check(hipStreamSynchronize(c->stream), "sync");
            for (long k = 0; k < n; ++k) out[k] = h[k];
        };
        get(c->r32v, nv, ov);
        get(c->r32s, nsz, os);
        get(c->r32i, ni, oi);
    });
    return 0;
}

int hpg_set_defect_inner(hpg_ctx* c, int ninner) {
    if (!c || ninner < 1 || ninner > 1000) return -2;
    c->ninner = ninner;
    return 0;
}

int hpg_apply_jacobian(hpg_ctx* c, const double* dv, const double* ds, const double* di, double* ov, double* os,
                       double* oi) {
    if (!c || !dv || !ov) return -2;
    if (c->ncurv && c->lv[c->log2p].p >= 2) { c->err = "the linearised residual has no curved branch"; return -2; }
    if (c->periodic) { c->err = "the linearised residual does not join periodic copies"; return -2; }
    if (c->physics != HPG_PHYSICS_INS) { c->err = "ins only"; return -2; }
    if (c->log2p != c->log2pmax) { c->err = "apply_jacobian runs on the finest level"; return -2; }
    HPG_TRY(c, {
        const size_t nv = (size_t)c->npnt * NV, nsz = (size_t)c->nseg * c->sm0 * NV, ni = (size_t)c->ntri * c->im0 * NV;
        do_rsdl(c, c->prm.nstage, true);
        std::vector<float> hv(dv, dv + nv), hs(nsz), hi(ni);
        for (size_t k = 0; k < nsz; ++k) hs[k] = (float)ds[k];
        for (size_t k = 0; k < ni; ++k) hi[k] = (float)di[k];
        check(hipMemcpyAsync(c->dlv, hv.data(), nv * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        if (nsz) check(hipMemcpyAsync(c->dls, hs.data(), nsz * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        if (ni) check(hipMemcpyAsync(c->dli, hi.data(), ni * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        do_lin_rsdl(c, c->prm.nstage, false);
        check(hipMemcpyAsync(hv.data(), c->r32v, nv * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        if (nsz) check(hipMemcpyAsync(hs.data(), c->r32s, nsz * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        if (ni) check(hipMemcpyAsync(hi.data(), c->r32i, ni * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
        for (size_t k = 0; k < nv; ++k) ov[k] = hv[k];
        for (size_t k = 0; k < nsz; ++k) os[k] = hs[k];
        for (size_t k = 0; k < ni; ++k) oi[k] = hi[k];
    });
    return 0;
}

int hpg_set_mlp(hpg_ctx* c, hpg_mlp* m) {
    if (!c) return -2;
    c->mlp = m;
    return 0;
}

int hpg_set_prcn_log_scalings(hpg_ctx* c, const float* y) {
    if (!c || c->physics != HPG_PHYSICS_INS) return -2;
    if (!y) { c->user_y = false; return 0; }
    HPG_TRY(c, {
        check(hipMemcpyAsync(c->mlpY, y, (size_t)c->ntri * MLP_OUT * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        check(hipStreamSynchronize(c->stream), "sync");
        c->user_y = true;
    });
    return 0;
}

int hpg_mlp_features(hpg_ctx* c, float* out) {
    if (!c || !out) return -2;
    if (c->physics != HPG_PHYSICS_INS) { c->err = "ins only"; return -2; }
    HPG_TRY(c, {
        hipLaunchKernelGGL(k_mlp_features, dim3(nblk(c->ntri)), dim3(256), 0, c->stream, c->ntri, c->pnts, c->tri_pnt,
                           c->ugv, c->tprcn, c->prm.mu, c->prm.rho, c->featX);
        check(hipMemcpyAsync(out, c->featX, (size_t)c->ntri * MLP_IN * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_measure_jacobian_diag(hpg_ctx* c, double* out) {
    if (!c || !out) return -2;
    if (c->ncurv && c->lv[c->log2p].p >= 2) { c->err = "the linearised residual has no curved branch"; return -2; }
    if (c->periodic) { c->err = "the linearised residual does not join periodic copies"; return -2; }
    if (c->physics != HPG_PHYSICS_INS) { c->err = "ins only"; return -2; }
    if (c->log2p != c->log2pmax) { c->err = "measure_jacobian_diag runs on the finest level"; return -2; }
    HPG_TRY(c, {
        do_rsdl(c, c->prm.nstage, true);
        check(hipMemsetAsync(c->jdiag, 0, (size_t)c->ntri * 3 * NV * 8, c->stream), "memset");
        const Level& L = c->lv[c->log2p];
        const LevelDims d(L.p);
        LinArgs a{};
        a.pnts = c->pnts; a.tri_pnt = c->tri_pnt; a.tri_seg = c->tri_seg; a.tri_sgn = c->tri_sgn;
        a.order = c->order; a.ops = L.pack32; a.lin = c->lin; a.sm0 = c->sm0; a.im0 = c->im0;
        a.rho = (float)c->prm.rho; a.mu = (float)c->prm.mu; a.adis = (float)c->prm.adis;
        a.beta = 1.f; a.keep = 0.f; a.fcv = c->fcv; a.fcs = c->fcs; a.diag = c->jdiag;
        for (int mode = 0; mode < d.tm; ++mode)
            for (int n = 0; n < NV; ++n) {
                a.probe_m = mode; a.probe_n = n;
                for (size_t col = 0; col + 1 < c->coff.size(); ++col) {
                    LinArgs b = a;
                    b.colour = (int)col; b.begin = c->coff[col]; b.end = c->coff[col + 1]; b.gbase = c->gbase[col];
                    const int nlogical = (b.end - b.begin + 63) / 64;
                    if (nlogical == 0) continue;
                    const int nb = ((nlogical + 7) / 8) * 8;
                    switch (L.p) {
                        case 1: hipLaunchKernelGGL((k_elem_lin<1, true, true>), dim3(nb), dim3(256), 0, c->stream, b); break;
                        case 2: hipLaunchKernelGGL((k_elem_lin<2, true, true>), dim3(nb), dim3(256), 0, c->stream, b); break;
                        default: hipLaunchKernelGGL((k_elem_lin<4, true, true>), dim3(nb), dim3(256), 0, c->stream, b); break;
                    }
                }
            }
        check(hipGetLastError(), "probe launch");
        check(hipMemcpyAsync(out, c->jdiag, (size_t)c->ntri * 3 * NV * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_element_jacobian(hpg_ctx* c, const int* tind, int nel, double eps_r, double eps_a, double* K) {
    if (!c || nel < 0) return -2;
    if (c->ncurv && c->lv[c->log2p].p >= 2) { c->err = "element_jacobian: no curved branch"; return -2; }
    if (nel == 0) return 0;
    if (!tind || !K) return -2;
    HPG_TRY(c, {
        const int nvv = c->nv;
        const Level& L = c->lv[c->log2p];
        const LevelDims d(L.p);
        const int N = d.tm * nvv;
        std::vector<char> seen(c->ntri, 0);
        for (int j = 0; j < nel; ++j) {
            if (tind[j] < 0 || tind[j] >= c->ntri) throw std::runtime_error("element index out of range");
            if (seen[tind[j]]) throw std::runtime_error("element listed twice");
            seen[tind[j]] = 1;
        }
        // group the batch by colour (stable in the caller's order)
        const int ncol = (int)c->coff.size() - 1;
        std::vector<int> cnt(ncol + 1, 0), jel(nel), jpos(nel);
        for (int j = 0; j < nel; ++j) ++cnt[c->tcol[tind[j]] + 1];
        for (int k = 0; k < ncol; ++k) cnt[k + 1] += cnt[k];
        const std::vector<int> joff = cnt;
        for (int j = 0; j < nel; ++j) {
            const int q = cnt[c->tcol[tind[j]]]++;
            jel[q] = tind[j];
            jpos[q] = j;
        }
        CallBuffers B(c->stream);
        JacArgs ja{};
        ja.jel = B.get<int>(nel);
        ja.jpos = B.get<int>(nel);
        check(hipMemcpyAsync((void*)ja.jel, jel.data(), (size_t)nel * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        check(hipMemcpyAsync((void*)ja.jpos, jpos.data(), (size_t)nel * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        ja.tri_pnt = c->tri_pnt; ja.tri_seg = c->tri_seg; ja.tri_sgn = c->tri_sgn;
        ja.sm0 = c->sm0; ja.im0 = c->im0;
        ja.ugv = c->ugv; ja.ugs = c->ugs; ja.ugi = c->ugi;
        const size_t nv = (size_t)c->npnt * nvv, ns = (size_t)c->nseg * c->sm0 * nvv, ni = (size_t)c->ntri * c->im0 * nvv;
        ja.jrv = B.get<double>(nv, true); ja.jrs = B.get<double>(ns, true); ja.jri = B.get<double>(ni, true);
        ja.jqv = B.get<double>(nv, true); ja.jqs = B.get<double>(ns, true); ja.jqi = B.get<double>(ni, true);
        ja.dw = B.get<double>((size_t)nel * nvv);
        ja.save = B.get<double>(nel);
        ja.base = B.get<double>((size_t)nel * N);
        ja.K = B.get<double>((size_t)nel * N * N);
        ja.eps_r = eps_r; ja.eps_a = eps_a;
        for (int col = 0; col < ncol; ++col) {
            if (joff[col + 1] == joff[col]) continue;
            JacArgs a = ja;
            a.begin = joff[col];
            a.end = joff[col + 1];
            const bool cns = c->physics == HPG_PHYSICS_CNS;
            switch (L.p) {
                case 1: cns ? jacobian_colour<1, 4>(c, a, N) : jacobian_colour<1, 3>(c, a, N); break;
                case 2: cns ? jacobian_colour<2, 4>(c, a, N) : jacobian_colour<2, 3>(c, a, N); break;
                case 4: cns ? jacobian_colour<4, 4>(c, a, N) : jacobian_colour<4, 3>(c, a, N); break;
                default: throw std::runtime_error("unsupported p");
            }
        }
        check(hipMemcpyAsync(K, ja.K, (size_t)nel * N * N * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

namespace {
// tri_hp::update / tri_hp_cns::update, or their FP32 defect-correction forms on the
// finest level; enqueued on the context's stream, no host synchronisation
void do_update(hpg_ctx* c) {
    if (c->physics == HPG_PHYSICS_CNS) {
        if (c->precision == HPG_FP32_DEFECT && !c->coarse_flag) mixed_update_cns(c);
        else update_cns(c);
        return;
    }
    if (c->precision == HPG_FP32_DEFECT && !c->coarse_flag) {
        mixed_update(c);
        return;
    }
    if (c->lv[c->log2p].p == 1 && !c->periodic) {
        // vertex modes only: every stage is one vstage_p1 (ug0 saved by the first)
        for (int stage = 0; stage < c->prm.nstage; ++stage)
            vstage_p1(c, stage, c->prm.alpha[stage] * c->prm.cfl[c->log2p], stage == 0);
        return;
    }
    const size_t nv = (size_t)c->npnt * NV * 8, ns = (size_t)c->nseg * c->sm0 * NV * 8,
                 ni = (size_t)c->ntri * c->im0 * NV * 8;
    dcopy(c, c->ug0v, c->ugv, nv);
    if (ns) dcopy(c, c->ug0s, c->ugs, ns);
    if (ni) dcopy(c, c->ug0i, c->ugi, ni);
    for (int stage = 0; stage < c->prm.nstage; ++stage) {
        do_rsdl(c, stage);
        do_minvrt(c, c->prm.alpha[stage] * c->prm.cfl[c->log2p]);
    }
    check(hipGetLastError(), "update launch");
}
}  // namespace

int hpg_update(hpg_ctx* c) {
    if (!c) return -2;
    HPG_TRY(c, do_update(c));
    return 0;
}

int hpg_maxres(hpg_ctx* c, double* per_var, double* out) {
    if (!c) return -2;
    HPG_TRY(c, {
        double mx[NVC];
        absmax(c, c->npnt, 1, 1, c->resv, nullptr, mx);
        double e = 0.0;
        for (int n = 0; n < c->nv; ++n) {
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
        double a[NVC], b[NVC] = {0, 0, 0, 0}, e[NVC] = {0, 0, 0, 0};
        absmax(c, c->npnt, 1, 1, c->resv, c->vess, a);
        if (d.sm) absmax(c, c->nseg, d.sm, c->sm0, c->ress, c->sess, b);
        if (d.im) absmax(c, c->ntri, d.im, c->im0, c->resi, nullptr, e);
        double r = 0.0;
        for (int n = 0; n < c->nv; ++n) r = std::max(r, std::max(a[n], std::max(b[n], e[n])));
        *out = r;
    });
    return 0;
}

int hpg_link_fine(hpg_ctx* c, hpg_ctx* f, const int* cc_tri, const double* cc_wt, const int* fc_tri,
                  const double* fc_wt) {
    if (!c || !f || c == f || !cc_tri || !cc_wt || !fc_tri || !fc_wt) return -2;
    if (c->periodic || f->periodic) { c->err = "h-multigrid levels of periodic meshes are not supported"; return -2; }
    if (c->nv != f->nv || c->physics != f->physics || c->log2pmax != 0) {
        c->err = "a coarse level has the fine level's physics and runs at p = 1";
        return -2;
    }
    HPG_TRY(c, {
        const int nf = f->npnt, nc = c->npnt;
        std::vector<int> th(3 * (size_t)c->ntri);
        check(hipMemcpy(th.data(), c->tri_pnt, th.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        std::vector<int> fh(3 * (size_t)f->ntri);
        check(hipMemcpy(fh.data(), f->tri_pnt, fh.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        // restriction gather: coarse vertex <- (fine vertex, weight), ascending fine vertex
        std::vector<int> ptr(nc + 1, 0), cv(3 * (size_t)nf);
        for (int i = 0; i < nf; ++i) {
            const int t = cc_tri[i];
            if (t < 0 || t >= c->ntri) throw std::runtime_error("ccnnct triangle out of range");
            for (int j = 0; j < 3; ++j) { cv[3 * i + j] = th[3 * t + j]; ptr[th[3 * t + j] + 1]++; }
        }
        for (int v = 0; v < nc; ++v) ptr[v + 1] += ptr[v];
        std::vector<int> fidx(ptr[nc]);
        std::vector<double> wt(ptr[nc]);
        {
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (int i = 0; i < nf; ++i)
                for (int j = 0; j < 3; ++j) {
                    const int v = cv[3 * i + j];
                    fidx[fill[v]] = i;
                    wt[fill[v]] = cc_wt[3 * i + j];
                    fill[v]++;
                }
        }
        std::vector<int> fv(3 * (size_t)nc);
        for (int v = 0; v < nc; ++v) {
            const int t = fc_tri[v];
            if (t < 0 || t >= f->ntri) throw std::runtime_error("fcnnct triangle out of range");
            for (int j = 0; j < 3; ++j) fv[3 * v + j] = fh[3 * t + j];
        }
        c->hr_ptr = c->upload(ptr.data(), ptr.size());
        c->hr_fidx = c->upload(fidx.data(), fidx.size());
        c->hr_wt = c->upload(wt.data(), wt.size());
        c->hi_fv = c->upload(fv.data(), fv.size());
        c->hi_wt = c->upload(fc_wt, 3 * (size_t)nc);
        c->hp_cv = c->upload(cv.data(), cv.size());
        c->hp_wt = c->upload(cc_wt, 3 * (size_t)nf);
        c->vug_frst = c->alloc<double>((size_t)nc * c->nv);
        if (!c->xev) check(hipEventCreateWithFlags(&c->xev, hipEventDisableTiming), "hipEventCreate");
        if (!f->xev) check(hipEventCreateWithFlags(&f->xev, hipEventDisableTiming), "hipEventCreate");
        check(hipStreamSynchronize(c->stream), "sync");
        c->fine = f;
        c->coarse_level = true;
    });
    return 0;
}

int hpg_mg_restrict(hpg_ctx* c) {
    if (!c) return -2;
    if (c->coarse_level) {
        // tri_hp/movco.cpp:41-102: fine residual and state onto this coarse mesh
        HPG_TRY(c, {
            hpg_ctx* f = c->fine;
            check(hipEventRecord(f->xev, f->stream), "hipEventRecord");
            check(hipStreamWaitEvent(c->stream, f->xev, 0), "hipStreamWaitEvent");
            c->isfrst = true;
            c->coarse_flag = true;
            if (c->nv == NVC) {
                hipLaunchKernelGGL(k_hrestrict<NVC>, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->hr_ptr,
                                   c->hr_fidx, c->hr_wt, f->resv, c->r0v);
                hipLaunchKernelGGL(k_hinterp<NVC>, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->hi_fv,
                                   c->hi_wt, f->ugv, c->ugv, c->vug_frst);
            } else {
                hipLaunchKernelGGL(k_hrestrict<NV>, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->hr_ptr,
                                   c->hr_fidx, c->hr_wt, f->resv, c->r0v);
                hipLaunchKernelGGL(k_hinterp<NV>, dim3(nblk(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->hi_fv,
                                   c->hi_wt, f->ugv, c->ugv, c->vug_frst);
            }
            check(hipGetLastError(), "h restrict");
        });
        return 0;
    }
    if (c->log2p == 0) return -2;
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
    if (!c) return -2;
    if (c->coarse_level) {
        // tri_hp/movfin.cpp:16-82: coarse correction interpolated onto the fine vertices
        HPG_TRY(c, {
            hpg_ctx* f = c->fine;
            const long n = (long)c->npnt * c->nv;
            hipLaunchKernelGGL(k_hcorr, dim3(nblk(n)), dim3(256), 0, c->stream, n, c->vug_frst, c->ugv);
            if (c->nv == NVC)
                hipLaunchKernelGGL(k_hprolong<NVC>, dim3(nblk(f->npnt)), dim3(256), 0, c->stream, f->npnt, c->hp_cv,
                                   c->hp_wt, c->vug_frst, f->resv, f->ugv);
            else
                hipLaunchKernelGGL(k_hprolong<NV>, dim3(nblk(f->npnt)), dim3(256), 0, c->stream, f->npnt, c->hp_cv,
                                   c->hp_wt, c->vug_frst, f->resv, f->ugv);
            check(hipGetLastError(), "h prolongate");
            check(hipEventRecord(c->xev, c->stream), "hipEventRecord");
            check(hipStreamWaitEvent(f->stream, c->xev, 0), "hipStreamWaitEvent");
        });
        return 0;
    }
    if (c->log2p >= c->log2pmax) return -2;
    ++c->log2p;
    if (c->log2p == c->log2pmax) c->coarse_flag = false;
    return 0;
}

namespace {
struct CycleRun {
    hpg_ctx* const* grids;
    int ngrid, extra, mglvls, vw, itercrsn, iterrfne;
};

// block::cycle (tri_mesh/blocks.cpp:1267-1348): level lvl uses grid
// clamp(lvl - extra, 0, ngrid - 1) (:1274-1275); p-levels of the fine grid first
void cycle_level(const CycleRun& r, int lvl) {
    const int gl = std::min(std::max(lvl - r.extra, 0), r.ngrid - 1);
    const int glp = std::min(std::max(lvl - r.extra + 1, 0), r.ngrid - 1);
    hpg_ctx* g = r.grids[gl];
    hpg_ctx* gp = r.grids[glp];
    auto ok = [](hpg_ctx* x, int rc, const char* what) {
        if (rc < 0) throw std::runtime_error(std::string(what) + ": " + x->err);
    };
    for (int vcount = 0; vcount < r.vw; ++vcount) {
        do_setup(g, false);
        for (int k = 0; k < r.itercrsn; ++k) do_update(g);
        if (lvl + vcount == 0) {
            // maxres of the entry level after its first smoothing (tri_hp/allocate.cpp:496-525),
            // read after the cycle
            hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, g->stream, 2 * NVC, (unsigned*)g->d_max);
            const dim3 nb(std::min<unsigned>(nblk((long)g->npnt), 2048u));
            if (g->nv == NVC)
                hipLaunchKernelGGL(k_absmax<NVC>, nb, dim3(256), 0, g->stream, (long)g->npnt, 1, 1, g->resv, nullptr, g->d_max);
            else
                hipLaunchKernelGGL(k_absmax<NV>, nb, dim3(256), 0, g->stream, (long)g->npnt, 1, 1, g->resv, nullptr, g->d_max);
        }
        if (lvl == r.mglvls - 1) return;
        do_rsdl(g, g->prm.nstage);
        ok(gp, hpg_mg_restrict(gp), "mg_restrict");
        cycle_level(r, lvl + 1);
        ok(gp, hpg_mg_prolongate(gp), "mg_prolongate");
        if (r.iterrfne) do_setup(g, false);
        for (int k = 0; k < r.iterrfne; ++k) do_update(g);
    }
}

// the host-side settings that fix what a cycle enqueues (kernels and their arguments)
std::vector<unsigned long long> cycle_signature(hpg_ctx* const* grids, int ngrid, int extra, int vw, int ic, int ir) {
    std::vector<unsigned long long> sg = {(unsigned long long)ngrid, (unsigned long long)extra,
                                          (unsigned long long)vw, (unsigned long long)ic, (unsigned long long)ir};
    for (int k = 0; k < ngrid; ++k) {
        const hpg_ctx* g = grids[k];
        sg.push_back((unsigned long long)(uintptr_t)g);
        sg.push_back((unsigned long long)(uintptr_t)g->mlp);
        sg.push_back((unsigned long long)(uintptr_t)g->crv);
        sg.push_back((unsigned long long)(uintptr_t)g->etype);
        sg.push_back((unsigned long long)(uintptr_t)g->st_tri);
        sg.push_back((unsigned long long)(uintptr_t)g->eseg);
        for (double x : {g->ibc_u[0], g->ibc_u[1], g->prm.bc_pressure}) {
            unsigned long long b;
            std::memcpy(&b, &x, sizeof(b));
            sg.push_back(b);
        }
        sg.push_back(((unsigned long long)g->precision << 40) | ((unsigned long long)g->ninner << 24) |
                     ((unsigned long long)g->prcn_mode << 16) | ((unsigned long long)g->user_y << 12) |
                     ((unsigned long long)g->fd_reset << 8) | ((unsigned long long)g->timing << 4) |
                     (unsigned long long)g->log2p);
        for (int l = 0; l < 4; ++l) {
            unsigned long long b;
            std::memcpy(&b, &g->prm.cfl[l], sizeof(b));
            sg.push_back(b);
        }
    }
    return sg;
}

// clears the error flags of up to 16 contexts, from the fine context's stream
struct FlagList { unsigned* p[16]; int n; };
__global__ void k_clear_flags(FlagList f) {
    const int k = threadIdx.x;
    if (k < f.n) f.p[k][0] = 0u;
}
}  // namespace

// block::cycle over a stack of grids, enqueued without host round trips inside the cycle.
// The second call with the same settings captures the cycle into a HIP graph (the
// coarse contexts' streams join through the restriction and prolongation events);
// later calls with the same settings replay it.
int hpg_mg_cycle(hpg_ctx* const* grids, int ngrid, int extra_finest_levels, int vw, int itercrsn, int iterrfne,
                 double* maxres) {
    if (!grids || ngrid < 1 || extra_finest_levels < 0 || vw < 1 || itercrsn < 0 || iterrfne < 0) return -2;
    for (int k = 0; k < ngrid; ++k)
        if (!grids[k] || (k > 0 && (!grids[k]->coarse_level || grids[k]->fine != grids[k - 1]))) return -2;
    hpg_ctx* f = grids[0];
    if (f->log2p != f->log2pmax || extra_finest_levels > f->log2pmax) { f->err = "cycle starts on the finest p-level"; return -2; }
    if (ngrid > 16) { f->err = "at most 16 meshes in a cycle"; return -2; }
    int nerr = 0;
    HPG_TRY(f, {
        FlagList fl{};
        fl.n = ngrid;
        for (int k = 0; k < ngrid; ++k) fl.p[k] = (unsigned*)grids[k]->d_err;
        CycleRun r{grids, ngrid, extra_finest_levels, ngrid + extra_finest_levels, vw, itercrsn, iterrfne};
        bool use_graph = f->graphs;
        for (int k = 0; k < ngrid; ++k) use_graph = use_graph && !grids[k]->timing;
        const std::vector<unsigned long long> sg =
            use_graph ? cycle_signature(grids, ngrid, extra_finest_levels, vw, itercrsn, iterrfne)
                      : std::vector<unsigned long long>();
        if (use_graph && f->cg_exec && sg == f->cg_sig) {
            check(hipGraphLaunch(f->cg_exec, f->stream), "hipGraphLaunch");
        } else if (use_graph && f->cg_seen && sg == f->cg_sig) {
            // second call with these settings: capture the cycle, then run the graph
            if (f->cg_exec) { check(hipGraphExecDestroy(f->cg_exec), "hipGraphExecDestroy"); f->cg_exec = nullptr; }
            hipGraph_t graph = nullptr;
            check(hipStreamBeginCapture(f->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            try {
                hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, f->stream, fl);
                cycle_level(r, 0);
            } catch (...) {
                hipStreamEndCapture(f->stream, &graph);
                if (graph) hipGraphDestroy(graph);
                throw;
            }
            check(hipStreamEndCapture(f->stream, &graph), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&f->cg_exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            check(e, "hipGraphInstantiate");
            check(hipGraphLaunch(f->cg_exec, f->stream), "hipGraphLaunch");
        } else {
            hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, f->stream, fl);
            cycle_level(r, 0);
            if (use_graph) {
                if (f->cg_exec && sg != f->cg_sig) { check(hipGraphExecDestroy(f->cg_exec), "hipGraphExecDestroy"); f->cg_exec = nullptr; }
                f->cg_sig = sg;
                f->cg_seen = true;
            }
        }
        check(hipGetLastError(), "cycle launch");
        unsigned long long mx[NVC] = {0, 0, 0, 0};
        check(hipMemcpyAsync(mx, f->d_max, sizeof(mx), hipMemcpyDeviceToHost, f->stream), "D2H");
        for (int k = 0; k < ngrid; ++k) {
            int e = 0;
            check(hipMemcpyAsync(&e, grids[k]->d_err, sizeof(int), hipMemcpyDeviceToHost, grids[k]->stream), "D2H");
            check(hipStreamSynchronize(grids[k]->stream), "sync");
            nerr += e != 0;
        }
        double m = 0.0;
        for (int n = 0; n < f->nv; ++n) {
            double v;
            std::memcpy(&v, &mx[n], sizeof(double));
            m = std::max(m, v);
        }
        if (maxres) *maxres = m;
    });
    return nerr;
}

int hpg_set_cycle_graph(hpg_ctx* c, int enable) {
    if (!c) return -2;
    c->graphs = enable != 0;
    if (!c->graphs) c->cg_seen = false;
    return 0;
}

int hpg_get_log2p(hpg_ctx* c, int* l) {
    if (!c || !l) return -2;
    *l = c->log2p;
    return 0;
}

int hpg_set_precision(hpg_ctx* c, int mode) {
    if (!c || (mode != HPG_FP64 && mode != HPG_FP32_DEFECT)) return -2;
    if (c->physics == HPG_PHYSICS_CNS && mode == HPG_FP32_DEFECT && !c->srcv) {
        // the FP64 defect and u0 copies of the cns outer iteration, allocated on first use
        HPG_TRY(c, {
            const size_t nv = (size_t)c->npnt * c->nv, ns = (size_t)c->nseg * c->sm0 * c->nv,
                         ni = (size_t)c->ntri * c->im0 * c->nv;
            c->srcv = c->alloc<double>(nv); c->srcs = c->alloc<double>(ns); c->srci = c->alloc<double>(ni);
            c->u0cv = c->alloc<double>(nv); c->u0cs = c->alloc<double>(ns); c->u0ci = c->alloc<double>(ni);
            check(hipStreamSynchronize(c->stream), "sync");
        });
    }
    c->precision = mode;
    c->fd_reset = true;
    return 0;
}

int hpg_set_preconditioner(hpg_ctx* c, int mode) {
    if (!c || (mode != HPG_PRCN_ANALYTIC && mode != HPG_PRCN_LEARNED)) return -2;
    if (c->physics != HPG_PHYSICS_INS && mode != HPG_PRCN_ANALYTIC) { c->err = "the learned preconditioner is built for ins"; return -2; }
    c->prcn_mode = mode;
    return 0;
}

int hpg_set_cfl(hpg_ctx* c, const double* cfl, int n) {
    if (!c || !cfl || n < 1 || n > 4) return -2;
    for (int k = 0; k < n; ++k)
        if (!(cfl[k] >= 0.0) || !std::isfinite(cfl[k])) return -2;
    for (int k = 0; k < n; ++k) c->prm.cfl[k] = cfl[k];
    return 0;
}

int hpg_set_periodic(hpg_ctx* c, int npair, const int* pairs) {
    if (!c || npair < 0 || (npair > 0 && !pairs)) return -2;
    if (c->physics != HPG_PHYSICS_INS) { c->err = "periodic boundaries: ins only"; return -2; }
    if (c->periodic) { c->err = "periodic boundaries are already set"; return -2; }
    if (c->coarse_level || c->fine) { c->err = "periodic boundaries on an h-multigrid level"; return -2; }
    if (c->etype || c->ecrv) { c->err = "set periodic boundaries before edge flux types and curved sides"; return -2; }
    const int nebd = (int)c->h_ebdry_start.size() - 1;
    for (int k = 0; k < 2 * npair; ++k)
        if (pairs[k] < 0 || pairs[k] >= nebd) { c->err = "periodic pair: edge boundary out of range"; return -2; }
    if (npair == 0) return 0;
    HPG_TRY(c, {
        const int np = c->npnt, ns = c->nseg, nt = c->ntri;
        check(hipStreamSynchronize(c->stream), "sync");
        std::vector<double> X(2 * (size_t)np);
        std::vector<int> sp(2 * (size_t)ns), tp(3 * (size_t)nt), tsg(3 * (size_t)nt), st(2 * (size_t)ns),
            sl(2 * (size_t)ns);
        check(hipMemcpy(X.data(), c->pnts, X.size() * 8, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(sp.data(), c->seg_pnt, sp.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(tp.data(), c->tri_pnt, tp.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(tsg.data(), c->tri_seg, tsg.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(st.data(), c->st_tri, st.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(sl.data(), c->st_loc, sl.size() * 4, hipMemcpyDeviceToHost), "D2H");
        // pair the sides as the reference's comm boundaries do (tri_mesh/tri_boundary.cpp:
        // 344-470): side j of the first boundary with side n-1-j of the second; the
        // second must be the first moved by one translation
        std::vector<int> parent(np);
        for (int v = 0; v < np; ++v) parent[v] = v;
        auto find = [&](int v) { while (parent[v] != v) v = parent[v] = parent[parent[v]]; return v; };
        std::vector<char> pbdry(nebd, 0), pside(ns, 0);
        std::vector<int> psa, psb, pso;
        for (int k = 0; k < npair; ++k) {
            const int b1 = pairs[2 * k], b2 = pairs[2 * k + 1];
            if (b1 == b2 || pbdry[b1] || pbdry[b2]) throw std::runtime_error("periodic pair: an edge boundary twice");
            pbdry[b1] = pbdry[b2] = 1;
            const int o1 = c->h_ebdry_start[b1], o2 = c->h_ebdry_start[b2];
            const int n = c->h_ebdry_start[b1 + 1] - o1;
            if (n != c->h_ebdry_start[b2 + 1] - o2 || n == 0) throw std::runtime_error("periodic pair: side counts differ");
            double T0[2] = {0.0, 0.0};
            for (int j = 0; j < n; ++j) {
                const int sa = c->h_ebdry_seg[o1 + j], sb = c->h_ebdry_seg[o2 + n - 1 - j];
                const int a0 = sp[2 * sa], a1 = sp[2 * sa + 1], q0 = sp[2 * sb], q1 = sp[2 * sb + 1];
                const double L = std::hypot(X[2 * a1] - X[2 * a0], X[2 * a1 + 1] - X[2 * a0 + 1]);
                const double tol = 1e-8 * std::max(L, 1e-300);
                auto d = [&](int p, int q, int r, int t) {
                    return std::hypot((X[2 * p] - X[2 * q]) - (X[2 * r] - X[2 * t]),
                                      (X[2 * p + 1] - X[2 * q + 1]) - (X[2 * r + 1] - X[2 * t + 1]));
                };
                int m0, m1, o;
                if (d(q1, a0, q0, a1) <= tol) { m0 = q1; m1 = q0; o = -1; }
                else if (d(q0, a0, q1, a1) <= tol) { m0 = q0; m1 = q1; o = 1; }
                else throw std::runtime_error("periodic pair: sides are not translates of each other");
                const double T[2] = {X[2 * m0] - X[2 * a0], X[2 * m0 + 1] - X[2 * a0 + 1]};
                if (j == 0) { T0[0] = T[0]; T0[1] = T[1]; }
                else if (std::hypot(T[0] - T0[0], T[1] - T0[1]) > tol)
                    throw std::runtime_error("periodic pair: not one translation");
                if (st[2 * sa + 1] >= 0 || st[2 * sb + 1] >= 0 || st[2 * sa] < 0 || st[2 * sb] < 0)
                    throw std::runtime_error("periodic pair: not boundary sides");
                psa.push_back(sa); psb.push_back(sb); pso.push_back(o);
                pside[sa] = pside[sb] = 1;
                parent[find(a0)] = find(m0);
                parent[find(a1)] = find(m1);
            }
        }
        // vertex groups, copies ascending
        std::vector<std::vector<int>> members(np);
        for (int v = 0; v < np; ++v) members[find(v)].push_back(v);
        std::vector<int> grp_of(np, -1), pg_ptr(1, 0), pg_v;
        for (int v = 0; v < np; ++v)
            if (members[v].size() > 1) {
                const int g = (int)pg_ptr.size() - 1;
                for (int u : members[v]) { pg_v.push_back(u); grp_of[u] = g; }
                pg_ptr.push_back((int)pg_v.size());
            }
        // own adjacency (ascending), as hpg_create builds it
        std::vector<std::vector<std::pair<int, int>>> vt(np), vs(np);
        for (int t = 0; t < nt; ++t)
            for (int i = 0; i < 3; ++i) vt[tp[3 * t + i]].push_back({t, i});
        for (int q = 0; q < ns; ++q)
            for (int i = 0; i < 2; ++i) vs[sp[2 * q + i]].push_back({q, i});
        // a slave side of a pair stands for its master: (sa, end of the same vertex in sa)
        std::vector<int> master(ns, -1), morient(ns, 1);
        for (size_t k = 0; k < psb.size(); ++k) { master[psb[k]] = psa[k]; morient[psb[k]] = pso[k]; }
        std::vector<int> vtu_ptr(np + 1, 0), vtu_tri, vtu_loc, vsu_ptr(np + 1, 0), vsu_side, vsu_end;
        for (int v = 0; v < np; ++v) {
            std::vector<std::pair<int, int>> a = vt[v], b;
            const int g = grp_of[v];
            const std::vector<int> one(1, v);
            const std::vector<int>& cp = g >= 0 ? members[find(v)] : one;
            if (g >= 0) {
                a.clear();
                for (int u : cp) a.insert(a.end(), vt[u].begin(), vt[u].end());
                std::sort(a.begin(), a.end());
            }
            for (int u : cp)
                for (auto e : vs[u]) {
                    int q = e.first, end = e.second;
                    if (master[q] >= 0) { end = morient[q] > 0 ? end : 1 - end; q = master[q]; }
                    b.push_back({q, end});
                }
            std::sort(b.begin(), b.end());
            b.erase(std::unique(b.begin(), b.end()), b.end());
            for (auto e : a) { vtu_tri.push_back(e.first); vtu_loc.push_back(e.second); }
            for (auto e : b) { vsu_side.push_back(e.first); vsu_end.push_back(e.second); }
            vtu_ptr[v + 1] = (int)vtu_tri.size();
            vsu_ptr[v + 1] = (int)vsu_side.size();
        }
        // each side of a pair: both elements ascending; the partner's flagged when the
        // two copies run opposite ways
        for (size_t k = 0; k < psa.size(); ++k) {
            const int sa = psa[k], sb = psb[k];
            const int ta = st[2 * sa], la = sl[2 * sa], tb = st[2 * sb], lb = sl[2 * sb];
            const int fl = pso[k] < 0 ? 4 : 0;
            const bool a_first = ta < tb;
            st[2 * sa] = a_first ? ta : tb; sl[2 * sa] = a_first ? la : (lb | fl);
            st[2 * sa + 1] = a_first ? tb : ta; sl[2 * sa + 1] = a_first ? (lb | fl) : la;
            st[2 * sb] = a_first ? ta : tb; sl[2 * sb] = a_first ? (la | fl) : lb;
            st[2 * sb + 1] = a_first ? tb : ta; sl[2 * sb + 1] = a_first ? lb : (la | fl);
        }
        // Dirichlet masks of the other boundaries, shared by every copy; no flux on the
        // periodic boundaries (their sides are inside the domain)
        std::vector<unsigned char> vess(np, 0), sess(ns, 0);
        std::vector<int> esides, ebid;
        for (int b = 0; b < nebd; ++b) {
            if (pbdry[b]) continue;
            const unsigned char mk = (unsigned char)(c->h_ebdry_ess[b] & 7);
            for (int j = c->h_ebdry_start[b]; j < c->h_ebdry_start[b + 1]; ++j) {
                const int q = c->h_ebdry_seg[j];
                sess[q] |= mk;
                vess[sp[2 * q]] |= mk;
                vess[sp[2 * q + 1]] |= mk;
                esides.push_back(q);
                ebid.push_back(b);
            }
        }
        for (size_t g = 0; g + 1 < pg_ptr.size(); ++g) {
            unsigned char mk = 0;
            for (int k = pg_ptr[g]; k < pg_ptr[g + 1]; ++k) mk |= vess[pg_v[k]];
            for (int k = pg_ptr[g]; k < pg_ptr[g + 1]; ++k) vess[pg_v[k]] = mk;
        }
        build_edges(c, sp.data(), np, esides, ebid);
        c->st_tri = c->upload(st.data(), st.size());
        c->st_loc = c->upload(sl.data(), sl.size());
        c->vess = c->upload(vess.data(), vess.size());
        c->sess = c->upload(sess.data(), sess.size());
        c->vtu_ptr = c->upload(vtu_ptr.data(), vtu_ptr.size());
        c->vtu_tri = c->upload(vtu_tri.data(), vtu_tri.size());
        c->vtu_loc = c->upload(vtu_loc.data(), vtu_loc.size());
        c->vsu_ptr = c->upload(vsu_ptr.data(), vsu_ptr.size());
        c->vsu_side = c->upload(vsu_side.data(), vsu_side.size());
        c->vsu_end = c->upload(vsu_end.data(), vsu_end.size());
        c->npg = (int)pg_ptr.size() - 1;
        c->pg_ptr = c->upload(pg_ptr.data(), pg_ptr.size());
        c->pg_v = c->upload(pg_v.data(), pg_v.size());
        c->nps = (int)psa.size();
        c->ps_a = c->upload(psa.data(), psa.size());
        c->ps_b = c->upload(psb.data(), psb.size());
        c->ps_o = c->upload(pso.data(), pso.size());
        c->periodic = true;
        periodic_sync(c);
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_set_edge_flux(hpg_ctx* c, int nebd, const int* types, const double* ibc) {
    if (!c || nebd < 0 || (nebd > 0 && !types) || !ibc) return -2;
    if (c->physics != HPG_PHYSICS_INS) { c->err = "edge flux types: ins only"; return -2; }
    const int nb = c->h_ebid.empty() ? 0 : *std::max_element(c->h_ebid.begin(), c->h_ebid.end()) + 1;
    if (nebd < nb) { c->err = "edge flux types: one per edge boundary"; return -2; }
    for (int b = 0; b < nebd; ++b)
        if (types[b] < HPG_EDGE_GENERIC || types[b] > HPG_EDGE_CHARACTERISTIC) { c->err = "edge flux type"; return -2; }
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(ibc[k])) { c->err = "freestream state not finite"; return -2; }
    HPG_TRY(c, {
        std::vector<unsigned char> et(c->neseg);
        bool any = false;
        for (int j = 0; j < c->neseg; ++j) {
            et[j] = (unsigned char)types[c->h_ebid[j]];
            any |= et[j] != HPG_EDGE_GENERIC;
        }
        check(hipStreamSynchronize(c->stream), "sync");
        c->etype = any ? c->upload(et.data(), et.size()) : nullptr;
        c->ibc_u[0] = ibc[0];
        c->ibc_u[1] = ibc[1];
        c->prm.bc_pressure = ibc[2];
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_set_curved(hpg_ctx* c, int nside, const int* sides, const double* crds) {
    if (!c || nside < 0 || (nside > 0 && (!sides || !crds))) return -2;
    if (c->physics != HPG_PHYSICS_INS) { c->err = "curved sides: ins only"; return -2; }
    HPG_TRY(c, {
        std::vector<int> row(c->nseg, -1);
        for (int j = 0; j < c->neseg; ++j) row[c->h_eseg[j]] = -2;  // boundary side, not curved
        for (int r = 0; r < nside; ++r) {
            const int s = sides[r];
            if (s < 0 || s >= c->nseg) throw std::runtime_error("curved side index out of range");
            if (row[s] == -1) throw std::runtime_error("curved side is not a boundary side");
            if (row[s] >= 0) throw std::runtime_error("curved side listed twice");
            row[s] = r;
        }
        for (int r = 0; r < nside * c->sm0 * 2; ++r)
            if (!std::isfinite(crds[r])) throw std::runtime_error("curved side coefficients not finite");
        check(hipStreamSynchronize(c->stream), "sync");
        std::vector<int> tseg(3 * (size_t)c->ntri), tsgn(3 * (size_t)c->ntri), stri(2 * (size_t)c->nseg);
        check(hipMemcpy(tseg.data(), c->tri_seg, tseg.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(tsgn.data(), c->tri_sgn, tsgn.size() * 4, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(stri.data(), c->st_tri, stri.size() * 4, hipMemcpyDeviceToHost), "D2H");
        // elements owning a curved side, ascending; their coordinate side modes with the
        // side sign of ugtouht (+1 on the reference's boundary sides)
        std::vector<int> cl, sl;
        std::vector<double> cv;
        const int sm0 = c->sm0;
        for (int t = 0; t < c->ntri; ++t) {
            bool curved = false;
            for (int i = 0; i < 3; ++i) curved |= row[tseg[3 * t + i]] >= 0;
            if (!curved) { sl.push_back(t); continue; }
            cl.push_back(t);
            for (int i = 0; i < 3; ++i) {
                const int r = row[tseg[3 * t + i]];
                for (int m = 0; m < sm0; ++m)
                    for (int n = 0; n < 2; ++n) {
                        const double sg = ((m & 1) && tsgn[3 * t + i] < 0) ? -1.0 : 1.0;
                        cv.push_back(r >= 0 ? sg * crds[((size_t)r * sm0 + m) * 2 + n] : 0.0);
                    }
            }
        }
        std::vector<int> ec(c->neseg);
        for (int j = 0; j < c->neseg; ++j) ec[j] = std::max(row[c->h_eseg[j]], -1);
        c->ncurv = (int)cl.size();
        if (c->ncurv > 0) {
            c->clist = c->upload(cl.data(), cl.size());
            c->slist = c->upload(sl.data(), sl.size());
            c->crv = c->upload(cv.data(), cv.size());
            c->ecrv = c->upload(ec.data(), ec.size());
            c->scrd = c->upload(crds, (size_t)nside * sm0 * 2);
            for (int l = 0; l <= c->log2pmax; ++l) {
                Level& L = c->lv[l];
                if (L.p >= 2 && !c->PRmat[l]) {
                    c->PRmat[l] = c->upload(L.ops->Pr.data(), L.ops->Pr.size());
                    c->PSmat[l] = c->upload(L.ops->Ps.data(), L.ops->Ps.size());
                }
            }
        } else {
            c->clist = c->slist = c->ecrv = nullptr;
            c->crv = c->scrd = nullptr;
        }
        check(hipStreamSynchronize(c->stream), "sync");
    });
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
    if (!c || !total_ms || !nlaunch || !nelem || full < 0 || full > 3) return -2;
    HPG_TRY(c, {
        check(hipStreamSynchronize(c->stream), "sync");
        double tot = 0.0;
        long long nl = 0, ne = 0;
        for (size_t r = 0; r < c->recs.size(); ++r) {
            if (c->recs[r].full != full) continue;
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

// MFMA f32 16x16x4 lane-layout probe (test support; A 16x4, B 4x16 row-major, out 64x4)
int hpg_selftest_mfma_f32(const float* A, const float* B, float* out) {
    float *dA = nullptr, *dB = nullptr, *dO = nullptr;
    if (hipMalloc(&dA, 64 * 4) != hipSuccess || hipMalloc(&dB, 64 * 4) != hipSuccess || hipMalloc(&dO, 256 * 4) != hipSuccess)
        return -1;
    hipMemcpy(dA, A, 64 * 4, hipMemcpyHostToDevice);
    hipMemcpy(dB, B, 64 * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_mfma_probe_f32, dim3(1), dim3(64), 0, 0, dA, dB, dO);
    const hipError_t e = hipMemcpy(out, dO, 256 * 4, hipMemcpyDeviceToHost);
    hipFree(dA); hipFree(dB); hipFree(dO);
    return e == hipSuccess ? 0 : -1;
}

// Plain device copy of a known size (test support: the reference point for HBM
// traffic counters).  `reps` launches of k_copy16 over nbytes (multiple of 16).
int hpg_selftest_copy(long long nbytes, int reps, double* ms_per_copy) {
    if (nbytes <= 0 || nbytes % 16 || reps <= 0) return -2;
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, nbytes) != hipSuccess) return -1;
    if (hipMalloc(&dst, nbytes) != hipSuccess) { hipFree(src); return -1; }
    hipMemset(src, 1, nbytes);
    const long long n16 = nbytes / 16;
    const int blocks = (int)std::min<long long>((n16 + 255) / 256, 1 << 20);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, (const uint4*)src, (uint4*)dst, n16);
    hipEventRecord(e1, 0);
    const hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(src);
    hipFree(dst);
    return e == hipSuccess ? 0 : -1;
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

// tet_hp_ins path (same translation unit: shares the host helpers above)
#include "hpg_tet.hip"
