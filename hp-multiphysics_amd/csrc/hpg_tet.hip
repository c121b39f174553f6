// Context and C-ABI of the tet_hp_ins and tet_hp_cns residual/relaxation paths
// (include/hpgpu.h, hpg_tet_*).  Compiled in the same translation unit as hpg_tri.hip (included at its
// end), whose host helpers (check, nblk8, HPG_TRY, resident_blocks, persistent_grid)
// and shared kernels (k_modes, k_absmax) it uses.
#include "tet_kernels.hip"
#include "tet_kernels_cns.hip"
#include "tet_jacobian.hip"
#include "tet_ops.h"

namespace {

struct TetLevel {
    int p = 0;
    double* ops = nullptr;   // TDims<p>::SIZE
    double* opsm = nullptr;  // TMDims<p>::SIZE, the matrix-core kernel's operators (p = 2)
    float* opsm32 = nullptr;  // the same for the FP32 matrix-core kernel (its K order)
    float* ops32 = nullptr;  // the same operators in FP32 (defect-correction inner solve)
    double* P2 = nullptr;    // face values GF x tm2d
    double* W2 = nullptr;    // face weights GF
    double* vfms = nullptr;  // 4 x 10
    double* Pg = nullptr;    // G x tm values at the Gauss points (cns setup)
    double vdiag = 0.0, ediag = 0.0, sfmv0 = 0.0, sfmv1 = 0.0;
};

template <int P>
std::vector<double> build_tet_pack(const TetOps& o) {
    using D = TDims<P>;
    std::vector<double> pk(D::SIZE, 0.0);
    const std::vector<double>* proj[4] = {&o.P, &o.Pr, &o.Ps, &o.Pt};
    for (int op = 0; op < 4; ++op)
        for (int m = 0; m < D::TM; ++m)
            for (int g = 0; g < D::G; ++g) pk[op * D::TM * 32 + m * 32 + g] = (*proj[op])[g * D::TM + m];
    const std::vector<double>* integ[3] = {&o.Ir, &o.Is, &o.It};
    const int ioff[3] = {D::OP_IR, D::OP_IS, D::OP_IT};
    for (int op = 0; op < 3; ++op)
        for (int g = 0; g < D::G; ++g)
            for (int m = 0; m < D::TM; ++m) pk[ioff[op] + g * 16 + m] = (*integ[op])[m * D::G + g];
    const std::vector<double>* der[3] = {&o.Dr, &o.Ds, &o.Dt};
    for (int op = 0; op < 3; ++op)
        for (int q = 0; q < D::G; ++q)
            for (int g = 0; g < D::G; ++g) pk[D::OP_DR + op * D::G * D::DQ + g * D::DQ + q] = (*der[op])[g * D::G + q];
    return pk;
}

// operators of k_elem_tet_mfma as lane-ready A blocks: block (tile t, K-step k) lane l
// holds A[16 t + l % 16][K index of lane l].  The projections' K index is the mode
// 4 k + l / 16; the derivatives' and integrations' K index is the Gauss point that
// accumulator register k % 4 of tile k / 4 holds in lane l: 16 (k / 4) + 4 (k % 4) +
// l / 16 for the f64 matrix cores, 16 (k / 4) + 4 (l / 16) + k % 4 for the f32 ones
// (f32 = true).  Rows and K indices beyond the G points and TM modes are zero.
template <int P>
std::vector<double> build_tet_mfma_pack(const TetOps& o, bool f32 = false) {
    using D = TDims<P>;
    using M = TMDims<P>;
    std::vector<double> pk(M::SIZE, 0.0);
    auto kpoint = [f32](int k, int l) {
        return f32 ? 16 * (k / 4) + 4 * (l / 16) + k % 4 : 16 * (k / 4) + 4 * (k % 4) + l / 16;
    };
    const std::vector<double>* proj[4] = {&o.P, &o.Pr, &o.Ps, &o.Pt};
    for (int op = 0; op < 4; ++op)
        for (int t = 0; t < M::GT; ++t)
            for (int k = 0; k < M::KT; ++k)
                for (int l = 0; l < 64; ++l) {
                    const int g = 16 * t + l % 16, m = 4 * k + l / 16;
                    if (g < D::G && m < D::TM)
                        pk[M::OFF_PROJ + ((op * M::GT + t) * M::KT + k) * 64 + l] = (*proj[op])[g * D::TM + m];
                }
    const std::vector<double>* der[3] = {&o.Dr, &o.Ds, &o.Dt};
    for (int op = 0; op < 3; ++op)
        for (int t = 0; t < M::GT; ++t)
            for (int k = 0; k < M::QK; ++k)
                for (int l = 0; l < 64; ++l) {
                    const int g = 16 * t + l % 16, h = kpoint(k, l);
                    if (g < D::G && h < D::G)
                        pk[M::OFF_DER + ((op * M::GT + t) * M::QK + k) * 64 + l] = (*der[op])[g * D::G + h];
                }
    const std::vector<double>* integ[3] = {&o.Ir, &o.Is, &o.It};
    for (int op = 0; op < 3; ++op)
        for (int k = 0; k < M::QK; ++k)
            for (int l = 0; l < 64; ++l) {
                const int m = l % 16, h = kpoint(k, l);
                if (m < D::TM && h < D::G) pk[M::OFF_INT + (op * M::QK + k) * 64 + l] = (*integ[op])[m * D::G + h];
            }
    return pk;
}

// CSR of entity -> codes, built from (entity, code) pairs in ascending code order
void csr(int nent, const std::vector<std::pair<int, int>>& pairs, std::vector<int>& ptr, std::vector<int>& code) {
    ptr.assign(nent + 1, 0);
    for (auto& pr : pairs) ptr[pr.first + 1]++;
    for (int i = 0; i < nent; ++i) ptr[i + 1] += ptr[i];
    code.assign(pairs.size(), 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (auto& pr : pairs) code[fill[pr.first]++] = pr.second;
}

}  // namespace

struct hpg_tet_ctx {
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    std::string err;
    int npnt = 0, nseg = 0, ntet = 0, nbf = 0, log2pmax = 0, log2p = 0, em0 = 0;
    int nv = TNV;            // 4 (ins: u, v, w, p) or 5 (cns: p, u, v, w, RT)
    hpg_params prm{};
    double kcond = 0.0;      // cns: R mu / prandtl * gamma / (gamma - 1)
    bool coarse_flag = false, isfrst = false;
    double *pnts = nullptr;
    int *tet_pnt = nullptr, *tet_seg = nullptr, *seg_pnt = nullptr, *bf_pnt = nullptr, *bf_seg = nullptr;
    unsigned char *vess = nullptr, *eess = nullptr;
    int *vt_ptr = nullptr, *vt_code = nullptr, *et_ptr = nullptr, *et_code = nullptr;
    int *ve_ptr = nullptr, *ve_code = nullptr, *vf_ptr = nullptr, *vf_code = nullptr, *ef_ptr = nullptr,
        *ef_code = nullptr;
    double *ugv = nullptr, *uge = nullptr, *ug0v = nullptr, *ug0e = nullptr, *resv = nullptr, *rese = nullptr;
    double *rrv = nullptr, *rre = nullptr, *r0v = nullptr, *r0e = nullptr;
    double *dresv[2] = {nullptr, nullptr}, *drese[2] = {nullptr, nullptr};
    double *vprcn = nullptr, *eprcn = nullptr, *iprcn = nullptr, *tau = nullptr, *lf = nullptr, *flf = nullptr;
    double *tpc = nullptr, *vpc = nullptr, *epc = nullptr;  // cns: 5 x 5 blocks per element / vertex / edge
    int* d_err = nullptr;
    unsigned long long* d_max = nullptr;
    TetLevel lv[2];
    // FP32 defect correction (hpg_tet_set_precision): device scale scalars and maxima,
    // the FP64 residual R64(u0) and state u0 of the outer iteration (allocated on first use)
    int precision = HPG_FP64, ninner = 1;
    double* fd_scal = nullptr;
    unsigned* fd_max = nullptr;
    bool fd_reset = true;
    double *srcv = nullptr, *srce = nullptr, *u0cv = nullptr, *u0ce = nullptr;
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
        check(hipMemsetAsync(p, 0, n * sizeof(T), stream), "memset");
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const T* h, size_t n) {
        T* d = alloc<T>(n);
        if (n) check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream), "H2D");
        return d;
    }
    ~hpg_tet_ctx() {
        if (stream) hipStreamSynchronize(stream);
        for (void* p : allocs) hipFree(p);
        for (hipEvent_t e : events) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
    int em() const { return (1 << log2p) - 1; }
    bool cns() const { return nv == CNV; }
};

namespace {

// elist: optional device list of `count` elements to evaluate (default: all ntet)
template <int P>
void launch_elem_tetc(hpg_tet_ctx* c, double beta, const int* elist, int count) {
    TetcElemArgs a;
    a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg;
    a.ntet = count; a.elist = elist; a.em0 = c->em0; a.ugv = c->ugv; a.uge = c->uge; a.tau = c->tau; a.lf = c->lf;
    a.ops = c->lv[c->log2p].ops; a.mu = c->prm.mu; a.kcond = c->kcond; a.R = c->prm.Rgas; a.gamma = c->prm.gamma;
    a.atm = 0.0; a.beta = beta;
    const int noct = (count + 7) / 8;
    if (P == 2) {
        // matrix-core form: groups of 16 elements, four groups per block
        a.ops = c->lv[c->log2p].opsm;
        const int ngrp = (count + 15) / 16;
        if (beta > 0.0) {
            static const int cap = resident_blocks((const void*)k_elem_tetc_mfma<2, true>);
            hipLaunchKernelGGL((k_elem_tetc_mfma<2, true>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256), 0,
                               c->stream, a, ngrp);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_tetc_mfma<2, false>);
            hipLaunchKernelGGL((k_elem_tetc_mfma<2, false>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256), 0,
                               c->stream, a, ngrp);
        }
    } else if (beta > 0.0) {
        static const int cap = resident_blocks((const void*)k_elem_tetc<P, true>);
        hipLaunchKernelGGL((k_elem_tetc<P, true>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
    } else {
        static const int cap = resident_blocks((const void*)k_elem_tetc<P, false>);
        hipLaunchKernelGGL((k_elem_tetc<P, false>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
    }
}

template <int P>
void launch_elem_tet(hpg_tet_ctx* c, int stage, const int* elist = nullptr, int count = -1) {
    if (count < 0) count = c->ntet;
    if (count == 0) return;  // an empty element list: no launch (a grid of zero blocks is invalid)
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    if (c->cns()) {
        const size_t r = c->recs.size();
        if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
        launch_elem_tetc<P>(c, beta, elist, count);
        if (c->timing) {
            check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
            c->recs.push_back({beta > 0.0 ? 1 : 0, count});
        }
        return;
    }
    TetElemArgs a;
    a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg; a.tet_sgn = nullptr;
    a.ntet = count; a.elist = elist; a.em0 = c->em0; a.ugv = c->ugv; a.uge = c->uge; a.tau = c->tau; a.lf = c->lf;
    a.ops = c->lv[c->log2p].ops; a.rho = c->prm.rho; a.mu = c->prm.mu; a.beta = beta;
    const int noct = (count + 7) / 8;
    const size_t r = c->recs.size();
    if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
    if (P == 2) {
        // matrix-core form: groups of 16 elements, four groups per block
        a.ops = c->lv[c->log2p].opsm;
        const int ngrp = (count + 15) / 16;
        if (beta > 0.0) {
            static const int cap = resident_blocks((const void*)k_elem_tet_mfma<2, true>);
            hipLaunchKernelGGL((k_elem_tet_mfma<2, true>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256), 0,
                               c->stream, a, ngrp);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_tet_mfma<2, false>);
            hipLaunchKernelGGL((k_elem_tet_mfma<2, false>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256), 0,
                               c->stream, a, ngrp);
        }
    } else if (beta > 0.0) {
        static const int cap = resident_blocks((const void*)k_elem_tet<P, true>);
        hipLaunchKernelGGL((k_elem_tet<P, true>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
    } else {
        static const int cap = resident_blocks((const void*)k_elem_tet<P, false>);
        hipLaunchKernelGGL((k_elem_tet<P, false>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
    }
    if (c->timing) {
        check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
        c->recs.push_back({beta > 0.0 ? 1 : 0, count});
    }
}

// FP32 element residual of the defect correction's inner solve (timing records 3 / 2)
template <int P>
void launch_elem_tet32(hpg_tet_ctx* c, int stage) {
    const int nst = c->prm.nstage, count = c->ntet, noct = (count + 7) / 8;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const size_t r = c->recs.size();
    if (c->timing) check(hipEventRecord(c->event(2 * r), c->stream), "hipEventRecord");
    if (c->cns()) {
        TetcElemArgs32 a;
        a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg;
        a.ntet = count; a.elist = nullptr; a.em0 = c->em0; a.ugv = c->ugv; a.uge = c->uge; a.tau = c->tau; a.lf = c->lf;
        a.ops = c->lv[c->log2p].ops32; a.mu = c->prm.mu; a.kcond = c->kcond; a.R = c->prm.Rgas;
        a.gamma = c->prm.gamma; a.atm = 0.0; a.beta = beta;
        if (P == 2) {
            a.ops = c->lv[c->log2p].opsm32;
            const int ngrp = (count + 15) / 16;
            if (beta > 0.0) {
                static const int cap = resident_blocks((const void*)k_elem_tetc32_mfma<2, true>);
                hipLaunchKernelGGL((k_elem_tetc32_mfma<2, true>), dim3(persistent_grid((ngrp + 3) / 4, cap)),
                                   dim3(256), 0, c->stream, a, ngrp);
            } else {
                static const int cap = resident_blocks((const void*)k_elem_tetc32_mfma<2, false>);
                hipLaunchKernelGGL((k_elem_tetc32_mfma<2, false>), dim3(persistent_grid((ngrp + 3) / 4, cap)),
                                   dim3(256), 0, c->stream, a, ngrp);
            }
        } else if (beta > 0.0) {
            static const int cap = resident_blocks((const void*)k_elem_tetc32<P, true>);
            hipLaunchKernelGGL((k_elem_tetc32<P, true>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_tetc32<P, false>);
            hipLaunchKernelGGL((k_elem_tetc32<P, false>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
        }
    } else {
        TetElemArgs32 a;
        a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg; a.tet_sgn = nullptr;
        a.ntet = count; a.elist = nullptr; a.em0 = c->em0; a.ugv = c->ugv; a.uge = c->uge; a.tau = c->tau; a.lf = c->lf;
        a.ops = c->lv[c->log2p].ops32; a.rho = c->prm.rho; a.mu = c->prm.mu; a.beta = beta;
        if (P == 2) {
            // matrix-core form: groups of 16 elements, four groups per block
            a.ops = c->lv[c->log2p].opsm32;
            const int ngrp = (count + 15) / 16;
            if (beta > 0.0) {
                static const int cap = resident_blocks((const void*)k_elem_tet32_mfma<2, true>);
                hipLaunchKernelGGL((k_elem_tet32_mfma<2, true>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256),
                                   0, c->stream, a, ngrp);
            } else {
                static const int cap = resident_blocks((const void*)k_elem_tet32_mfma<2, false>);
                hipLaunchKernelGGL((k_elem_tet32_mfma<2, false>), dim3(persistent_grid((ngrp + 3) / 4, cap)), dim3(256),
                                   0, c->stream, a, ngrp);
            }
        } else if (beta > 0.0) {
            static const int cap = resident_blocks((const void*)k_elem_tet32<P, true>);
            hipLaunchKernelGGL((k_elem_tet32<P, true>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
        } else {
            static const int cap = resident_blocks((const void*)k_elem_tet32<P, false>);
            hipLaunchKernelGGL((k_elem_tet32<P, false>), dim3(persistent_grid(noct, cap)), dim3(256), 0, c->stream, a, noct);
        }
    }
    if (c->timing) {
        check(hipEventRecord(c->event(2 * r + 1), c->stream), "hipEventRecord");
        c->recs.push_back({beta > 0.0 ? 3 : 2, count});
    }
}

template <int NVn>
void tet_modes_nv(hpg_tet_ctx* c, int op, double* xv, double* xe, double* yv, double* ye, const double* zv,
                  const double* ze, double k) {
    const long nv = (long)c->npnt * NVn;
    hipLaunchKernelGGL(k_modes<NVn>, dim3(nblk(nv)), dim3(256), 0, c->stream, op, (long)c->npnt, 1, 1, xv, yv, zv, k);
    const int em = c->em();
    if (em > 0) {
        const long ne = (long)c->nseg * em * NVn;
        hipLaunchKernelGGL(k_modes<NVn>, dim3(nblk(ne)), dim3(256), 0, c->stream, op, (long)c->nseg, em, c->em0, xe, ye,
                           ze, k);
    }
}

void tet_modes(hpg_tet_ctx* c, int op, double* xv, double* xe, double* yv, double* ye, const double* zv,
               const double* ze, double k) {
    if (c->cns()) tet_modes_nv<CNV>(c, op, xv, xe, yv, ye, zv, ze, k);
    else tet_modes_nv<TNV>(c, op, xv, xe, yv, ye, zv, ze, k);
}

// tet_hp::rsdl (tet_hp/update.cpp:9-131).  fp32: the element residual in FP32 arithmetic
// (the inner solve of the defect correction; the face fluxes and the assembly stay FP64)
void do_rsdl_tet(hpg_tet_ctx* c, int stage, bool fp32 = false) {
    const int nst = c->prm.nstage;
    const double beta = stage < nst ? c->prm.beta[stage] : 1.0;
    const TetLevel& L = c->lv[c->log2p];
    if (fp32) { if (L.p == 1) launch_elem_tet32<1>(c, stage); else launch_elem_tet32<2>(c, stage); }
    else if (L.p == 1) launch_elem_tet<1>(c, stage); else launch_elem_tet<2>(c, stage);
    if (c->nbf > 0 && c->cns()) {
        const double* b = c->prm.bc;
        if (L.p == 1)
            hipLaunchKernelGGL(k_face_tetc<1>, dim3(nblk8(c->nbf)), dim3(256), 0, c->stream, c->nbf, c->bf_pnt, c->bf_seg,
                               c->pnts, c->ugv, c->uge, c->em0, L.P2, L.W2, c->prm.gamma, 0.0, b[1], b[2], b[3], b[4],
                               c->flf);
        else
            hipLaunchKernelGGL(k_face_tetc<2>, dim3(nblk8(c->nbf)), dim3(256), 0, c->stream, c->nbf, c->bf_pnt, c->bf_seg,
                               c->pnts, c->ugv, c->uge, c->em0, L.P2, L.W2, c->prm.gamma, 0.0, b[1], b[2], b[3], b[4],
                               c->flf);
    } else if (c->nbf > 0) {
        if (L.p == 1)
            hipLaunchKernelGGL(k_face_tet<1>, dim3(nblk8(c->nbf)), dim3(256), 0, c->stream, c->nbf, c->bf_pnt, c->bf_seg,
                               c->pnts, c->ugv, c->uge, c->em0, L.P2, L.W2, c->prm.rho, c->flf);
        else
            hipLaunchKernelGGL(k_face_tet<2>, dim3(nblk8(c->nbf)), dim3(256), 0, c->stream, c->nbf, c->bf_pnt, c->bf_seg,
                               c->pnts, c->ugv, c->uge, c->em0, L.P2, L.W2, c->prm.rho, c->flf);
    }
    const int tm = L.p == 1 ? 4 : 10, tm2 = L.p == 1 ? 3 : 6;
    TetGatherArgs g;
    const int nvv = c->nv;
    g.tm = tm; g.tm2 = tm2; g.lf = c->lf; g.flf = c->flf; g.omb = 1.0 - beta; g.fnv = c->cns() ? CNV : 1;
    g.nent = c->npnt; g.mbase = 0; g.fbase = 0; g.stride = nvv; g.ptr = c->vt_ptr; g.code = c->vt_code;
    g.fptr = c->vf_ptr; g.fcode = c->vf_code; g.res = c->resv; g.rr = c->rrv;
    if (c->cns()) hipLaunchKernelGGL(k_gather_tet<CNV>, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, g);
    else hipLaunchKernelGGL(k_gather_tet<TNV>, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, g);
    if (c->em() > 0) {
        g.nent = c->nseg; g.mbase = 4; g.fbase = 3; g.stride = c->em0 * nvv; g.ptr = c->et_ptr; g.code = c->et_code;
        g.fptr = c->ef_ptr; g.fcode = c->ef_code; g.res = c->rese; g.rr = c->rre;
        if (c->cns()) hipLaunchKernelGGL(k_gather_tet<CNV>, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, g);
        else hipLaunchKernelGGL(k_gather_tet<TNV>, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, g);
    }
    if (c->coarse_flag) {
        const int l = c->log2p;
        if (c->isfrst) {
            tet_modes(c, 2, c->resv, c->rese, c->dresv[l], c->drese[l], c->r0v, c->r0e, c->prm.fadd);
            c->isfrst = false;
        } else {
            tet_modes(c, 1, c->resv, c->rese, c->dresv[l], c->drese[l], nullptr, nullptr, 0.0);
        }
    }
    check(hipGetLastError(), "tet rsdl launch");
}

TetMinvArgs tet_minv_args(hpg_tet_ctx* c, bool pre_dirichlet) {
    const TetLevel& L = c->lv[c->log2p];
    TetMinvArgs a{};
    a.npnt = c->npnt; a.nseg = c->nseg; a.em0 = c->em0; a.em = c->em(); a.pre_dirichlet = pre_dirichlet ? 1 : 0;
    a.ve_ptr = c->ve_ptr; a.ve_code = c->ve_code; a.et_ptr = c->et_ptr; a.et_code = c->et_code;
    a.tet_pnt = c->tet_pnt; a.vess = c->vess; a.eess = c->eess; a.resv = c->resv; a.rese = c->rese;
    a.vprcn = c->vprcn; a.eprcn = c->eprcn; a.iprcn = c->iprcn;
    a.vdiag = L.vdiag; a.ediag = L.ediag; a.sfmv0 = L.sfmv0; a.sfmv1 = L.sfmv1; a.vfms = L.vfms;
    return a;
}

// tet_hp_cns::minvrt with the stage update fused where the data allow it: edges in the
// edge sweep; vertices in the vertex sweep only without edge modes, else afterwards
// (the edge sweep reads the stage state of the vertices)
void do_minvrt_tetc(hpg_tet_ctx* c, double cfa) {
    const TetLevel& L = c->lv[c->log2p];
    TetcMinvArgs a{};
    a.npnt = c->npnt; a.nseg = c->nseg; a.em0 = c->em0; a.em = c->em();
    a.ve_ptr = c->ve_ptr; a.ve_code = c->ve_code; a.et_ptr = c->et_ptr; a.et_code = c->et_code;
    a.tet_pnt = c->tet_pnt; a.seg_pnt = c->seg_pnt; a.vess = c->vess; a.eess = c->eess;
    a.resv = c->resv; a.rese = c->rese; a.vprcn = c->vprcn; a.eprcn = c->eprcn; a.iprcn = c->iprcn;
    a.vpc = c->vpc; a.epc = c->epc; a.ugv = c->ugv;
    a.vdiag = L.vdiag; a.sfmv0 = L.sfmv0; a.sfmv1 = L.sfmv1; a.gm1 = c->prm.gamma - 1.0; a.atm = 0.0; a.vfms = L.vfms;
    const bool upd = cfa >= 0.0;
    a.cfa = cfa;
    if (upd && a.em == 0) { a.ugv_out = c->ugv; a.ug0v = c->ug0v; }
    hipLaunchKernelGGL(k_minv_vrtx_tetc, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, a);
    if (a.em > 0) {
        if (upd) { a.uge_out = c->uge; a.ug0e = c->ug0e; }
        hipLaunchKernelGGL(k_minv_edge_tetc, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, a);
        if (upd) {
            const long n = (long)c->npnt * CNV;
            hipLaunchKernelGGL(k_modes<CNV>, dim3(nblk(n)), dim3(256), 0, c->stream, 4, (long)c->npnt, 1, 1, c->ugv,
                               c->ug0v, c->resv, cfa);
        }
    }
    check(hipGetLastError(), "tet cns minvrt launch");
}

void do_minvrt_tet(hpg_tet_ctx* c, bool pre_dirichlet, double cfa) {
    if (c->cns()) {
        do_minvrt_tetc(c, cfa);
        return;
    }
    TetMinvArgs a = tet_minv_args(c, pre_dirichlet);
    if (cfa >= 0.0) {
        a.ugv = c->ugv; a.ug0v = c->ug0v; a.cfa = cfa;
        if (a.em > 0) { a.uge = c->uge; a.ug0e = c->ug0e; }
    }
    hipLaunchKernelGGL(k_minv_vrtx_tet, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, a);
    if (a.em > 0) hipLaunchKernelGGL(k_minv_edge_tet, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, a);
    check(hipGetLastError(), "tet minvrt launch");
}

template <int P>
void launch_setup_tetc(hpg_tet_ctx* c) {
    TetcSetupArgs a;
    a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg; a.ugv = c->ugv; a.uge = c->uge;
    a.Pg = c->lv[c->log2p].Pg; a.ntet = c->ntet; a.em0 = c->em0; a.p = P; a.prec = 1;
    a.mu = c->prm.mu; a.adis = c->prm.adis; a.gamma = c->prm.gamma; a.kcond = c->kcond; a.R = c->prm.Rgas; a.atm = 0.0;
    a.tau = c->tau; a.tpc = c->tpc; a.iprcn = c->iprcn; a.err = c->d_err;
    hipLaunchKernelGGL(k_setup_tetc<P>, dim3(nblk8(c->ntet)), dim3(256), 0, c->stream, a);
}

int do_setup_tetc(hpg_tet_ctx* c) {
    check(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream), "memset");
    if (c->lv[c->log2p].p == 1) launch_setup_tetc<1>(c); else launch_setup_tetc<2>(c);
    hipLaunchKernelGGL(k_setup_gather_tetc, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->vt_ptr,
                       c->vt_code, c->iprcn, c->tpc, c->vprcn, c->vpc);
    if (c->em() > 0)
        hipLaunchKernelGGL(k_setup_gather_tetc, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, c->nseg, c->et_ptr,
                           c->et_code, c->iprcn, c->tpc, c->eprcn, c->epc);
    check(hipGetLastError(), "tet cns setup launch");
    int h = 0;
    check(hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
    check(hipStreamSynchronize(c->stream), "sync");
    return h;
}

int do_setup_tet(hpg_tet_ctx* c) {
    if (c->cns()) return do_setup_tetc(c);
    TetSetupArgs a;
    a.pnts = c->pnts; a.tet_pnt = c->tet_pnt; a.ugv = c->ugv; a.ntet = c->ntet; a.p = c->lv[c->log2p].p;
    a.rho = c->prm.rho; a.mu = c->prm.mu; a.adis = c->prm.adis; a.iprcn = c->iprcn; a.tau = c->tau; a.err = c->d_err;
    check(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream), "memset");
    hipLaunchKernelGGL(k_setup_tet, dim3(nblk8(c->ntet)), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_setup_gather_tet, dim3(nblk8(c->npnt)), dim3(256), 0, c->stream, c->npnt, c->vt_ptr,
                       c->vt_code, c->iprcn, c->vprcn);
    if (c->em() > 0)
        hipLaunchKernelGGL(k_setup_gather_tet, dim3(nblk8(c->nseg)), dim3(256), 0, c->stream, c->nseg, c->et_ptr,
                           c->et_code, c->iprcn, c->eprcn);
    check(hipGetLastError(), "tet setup launch");
    int h = 0;
    check(hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
    check(hipStreamSynchronize(c->stream), "sync");
    return h;
}

// y += sfac[0] x (the scaled defect source of the inner solve)
__global__ void k_axpy_dev(long n, const double* __restrict__ x, const double* __restrict__ sfac, double* __restrict__ y) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] += sfac[0] * x[k];
}

// FP32 defect correction of tet_hp::update / tet_hp_cns::update with FP64 storage: the
// outer iteration of the tri cns form (mixed_update_cns).  R64(u0) is kept in src; the
// inner state v = fl32(u0) + dhat lives in ug, so minvrt and the stage update run as
// in the FP64 update; stage 0 is R64(u0)/c, the later stages evaluate the element
// residual in FP32 and add (1/c - 1) R64(u0); finally u = u0 + c (v - fl32(u0)).
void mixed_update_tet(hpg_tet_ctx* c) {
    const long nv = (long)c->npnt * c->nv, ne = (long)c->nseg * c->em0 * c->nv;
    do_rsdl_tet(c, c->prm.nstage);
    auto copy = [&](double* dst, const double* src, long n) {
        if (n) check(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, c->stream), "D2D");
    };
    auto grid = [](long n) { return dim3(std::min<unsigned>(nblk(n), RED_BLOCKS)); };
    copy(c->srcv, c->resv, nv); copy(c->srce, c->rese, ne);
    copy(c->u0cv, c->ugv, nv); copy(c->u0ce, c->uge, ne);
    hipLaunchKernelGGL(k_round32, grid(nv), dim3(256), 0, c->stream, nv, c->ugv, c->fd_max);
    if (ne) hipLaunchKernelGGL(k_round32, grid(ne), dim3(256), 0, c->stream, ne, c->uge, c->fd_max);
    hipLaunchKernelGGL(k_fd_scale, dim3(1), dim3(1), 0, c->stream, c->fd_scal, c->fd_max, c->fd_reset ? 1 : 0, FD_H);
    c->fd_reset = false;
    for (int it = 0; it < c->ninner; ++it) {
        copy(c->ug0v, c->ugv, nv); copy(c->ug0e, c->uge, ne);
        for (int s = 0; s < c->prm.nstage; ++s) {
            if (it == 0 && s == 0) {
                hipLaunchKernelGGL(k_scale64, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->resv, c->fd_scal);
                if (ne) hipLaunchKernelGGL(k_scale64, dim3(nblk(ne)), dim3(256), 0, c->stream, ne, c->rese, c->fd_scal);
            } else {
                do_rsdl_tet(c, s, true);
                hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(nv)), dim3(256), 0, c->stream, nv, c->srcv, c->fd_scal + 1, c->resv);
                if (ne) hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(ne)), dim3(256), 0, c->stream, ne, c->srce, c->fd_scal + 1, c->rese);
            }
            do_minvrt_tet(c, !c->cns(), c->prm.alpha[s] * c->prm.cfl[c->log2p]);
        }
    }
    hipLaunchKernelGGL(k_fd_final64, grid(nv), dim3(256), 0, c->stream, nv, c->ugv, c->u0cv, c->fd_scal, c->fd_max + 1);
    if (ne) hipLaunchKernelGGL(k_fd_final64, grid(ne), dim3(256), 0, c->stream, ne, c->uge, c->u0ce, c->fd_scal, c->fd_max + 1);
    check(hipGetLastError(), "tet mixed update launch");
}

double tet_absmax(hpg_tet_ctx* c, long nent, int nm, int mstride, const double* x, const unsigned char* mask,
                  double out[CNV]) {
    unsigned long long h[CNV] = {0, 0, 0, 0, 0};
    if (nent > 0 && nm > 0) {
        check(hipMemsetAsync(c->d_max, 0, CNV * sizeof(unsigned long long), c->stream), "memset");
        const long n = nent * nm;
        const unsigned nb = std::min<unsigned>(nblk(n), 2048u);
        if (c->cns())
            hipLaunchKernelGGL(k_absmax<CNV>, dim3(nb), dim3(256), 0, c->stream, nent, nm, mstride, x, mask, c->d_max);
        else
            hipLaunchKernelGGL(k_absmax<TNV>, dim3(nb), dim3(256), 0, c->stream, nent, nm, mstride, x, mask, c->d_max);
        check(hipMemcpyAsync(h, c->d_max, c->nv * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream), "memcpy");
        check(hipStreamSynchronize(c->stream), "sync");
    }
    double r = 0.0;
    for (int n = 0; n < c->nv; ++n) {
        double v;
        std::memcpy(&v, &h[n], sizeof(double));
        out[n] = v;
        r = std::max(r, v);
    }
    return r;
}

}  // namespace

extern "C" {

int hpg_tet_create(const hpg_tet_mesh* m, const hpg_params* prm, hpg_tet_ctx** out) {
    if (!m || !prm || !out) return -2;
    *out = nullptr;
    if ((prm->physics != HPG_PHYSICS_INS && prm->physics != HPG_PHYSICS_CNS) || prm->log2p < 0 || prm->log2p > 1 ||
        prm->nstage < 1 || prm->nstage > 7)
        return -2;
    if (prm->physics == HPG_PHYSICS_CNS && !(prm->gamma > 1.0 && prm->prandtl > 0.0 && prm->Rgas > 0.0 && prm->bc[4] > 0.0))
        return -2;
    if (m->npnt <= 0 || m->nseg <= 0 || m->ntet <= 0 || m->nbf < 0) return -2;
    std::unique_ptr<hpg_tet_ctx> c(new hpg_tet_ctx);
    try {
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        c->prm = *prm;
        c->nv = prm->physics == HPG_PHYSICS_CNS ? CNV : TNV;
        // tet_hp/cns/allocate.cpp:50
        c->kcond = prm->Rgas * prm->mu / prm->prandtl * prm->gamma / (prm->gamma - 1.0);
        c->prm.alpha[prm->nstage] = 1.0;
        c->prm.beta[prm->nstage] = 1.0;
        c->npnt = m->npnt; c->nseg = m->nseg; c->ntet = m->ntet; c->nbf = m->nbf;
        c->log2pmax = c->log2p = prm->log2p;
        c->em0 = (1 << prm->log2p) - 1;
        const int np = m->npnt, ns = m->nseg, nt = m->ntet, nf = m->nbf;
        // host checks of the index arrays the kernels gather through
        for (long k = 0; k < 4L * nt; ++k)
            if (m->tet_pnt[k] < 0 || m->tet_pnt[k] >= np) throw std::runtime_error("tet_pnt out of range");
        for (long k = 0; k < 6L * nt; ++k)
            if (m->tet_seg[k] < 0 || m->tet_seg[k] >= ns) throw std::runtime_error("tet_seg out of range");
        for (long k = 0; k < 2L * ns; ++k)
            if (m->seg_pnt[k] < 0 || m->seg_pnt[k] >= np) throw std::runtime_error("seg_pnt out of range");
        for (long k = 0; k < 3L * nf; ++k)
            if (m->bf_pnt[k] < 0 || m->bf_pnt[k] >= np || m->bf_seg[k] < 0 || m->bf_seg[k] >= ns)
                throw std::runtime_error("boundary face index out of range");
        if ((long)nt * 16 > 2147483647L || (long)nf * 8 > 2147483647L) throw std::runtime_error("mesh too large");
        c->pnts = c->upload(m->pnts, 3 * (size_t)np);
        c->tet_pnt = c->upload(m->tet_pnt, 4 * (size_t)nt);
        c->tet_seg = c->upload(m->tet_seg, 6 * (size_t)nt);
        c->seg_pnt = c->upload(m->seg_pnt, 2 * (size_t)ns);
        c->bf_pnt = c->upload(m->bf_pnt, 3 * (size_t)nf);
        c->bf_seg = c->upload(m->bf_seg, 3 * (size_t)nf);
        // essential variables on the vertices / edges of the inflow faces: u, v, w (ins,
        // bits 0-2) or u, v, w, RT (cns, bits 1-4)
        const unsigned char ebits = c->nv == CNV ? 30 : 7;
        std::vector<unsigned char> vess(np, 0), eess(ns, 0);
        for (long k = 0; k < 3L * nf; ++k) { vess[m->bf_pnt[k]] = ebits; eess[m->bf_seg[k]] = ebits; }
        c->vess = c->upload(vess.data(), vess.size());
        c->eess = c->upload(eess.data(), eess.size());
        // incidence in ascending element / edge / face order
        std::vector<std::pair<int, int>> vt, et, ve, vf, ef;
        vt.reserve(4L * nt); et.reserve(6L * nt);
        for (int t = 0; t < nt; ++t) {
            for (int i = 0; i < 4; ++i) vt.push_back({m->tet_pnt[4 * t + i], 16 * t + i});
            for (int i = 0; i < 6; ++i) et.push_back({m->tet_seg[6 * t + i], 16 * t + i});
        }
        for (int e = 0; e < ns; ++e)
            for (int i = 0; i < 2; ++i) ve.push_back({m->seg_pnt[2 * e + i], 2 * e + i});
        for (int f = 0; f < nf; ++f)
            for (int j = 0; j < 3; ++j) {
                vf.push_back({m->bf_pnt[3 * f + j], 8 * f + j});
                ef.push_back({m->bf_seg[3 * f + j], 8 * f + j});
            }
        std::vector<int> ptr, code;
        csr(np, vt, ptr, code); c->vt_ptr = c->upload(ptr.data(), ptr.size()); c->vt_code = c->upload(code.data(), code.size());
        csr(ns, et, ptr, code); c->et_ptr = c->upload(ptr.data(), ptr.size()); c->et_code = c->upload(code.data(), code.size());
        csr(np, ve, ptr, code); c->ve_ptr = c->upload(ptr.data(), ptr.size()); c->ve_code = c->upload(code.data(), code.size());
        csr(np, vf, ptr, code); c->vf_ptr = c->upload(ptr.data(), ptr.size()); c->vf_code = c->upload(code.data(), code.size());
        csr(ns, ef, ptr, code); c->ef_ptr = c->upload(ptr.data(), ptr.size()); c->ef_code = c->upload(code.data(), code.size());
        const int NVc = c->nv;
        const size_t nv = (size_t)np * NVc, ne = (size_t)ns * c->em0 * NVc;
        double** vv[] = {&c->ugv, &c->ug0v, &c->resv, &c->rrv, &c->r0v};
        double** ee[] = {&c->uge, &c->ug0e, &c->rese, &c->rre, &c->r0e};
        for (auto p : vv) *p = c->alloc<double>(nv);
        for (auto p : ee) *p = c->alloc<double>(ne);
        for (int l = 0; l <= c->log2pmax; ++l) {
            c->dresv[l] = c->alloc<double>(nv);
            c->drese[l] = c->alloc<double>(ne);
        }
        c->vprcn = c->alloc<double>(nv);
        c->eprcn = c->alloc<double>((size_t)ns * NVc);
        c->iprcn = c->alloc<double>((size_t)nt * NVc);
        c->tau = c->alloc<double>((size_t)nt * (c->nv == CNV ? 25 : TNV));
        c->lf = c->alloc<double>((size_t)nt * 2 * NVc * (c->log2pmax ? 10 : 4));
        c->flf = c->alloc<double>((size_t)nf * 6 * (c->nv == CNV ? CNV : 1));
        if (c->nv == CNV) {
            c->tpc = c->alloc<double>((size_t)nt * 25);
            c->vpc = c->alloc<double>((size_t)np * 25);
            c->epc = c->alloc<double>((size_t)ns * 25);
        }
        c->d_err = c->alloc<int>(1);
        c->d_max = c->alloc<unsigned long long>(CNV);
        c->fd_scal = c->alloc<double>(2);
        c->fd_max = c->alloc<unsigned>(2);
        for (int l = 0; l <= c->log2pmax; ++l) {
            const TetOps o(1 << l);
            TetLevel& L = c->lv[l];
            L.p = o.p;
            const std::vector<double> pk = o.p == 1 ? build_tet_pack<1>(o) : build_tet_pack<2>(o);
            L.ops = c->upload(pk.data(), pk.size());
            const std::vector<float> pk32(pk.begin(), pk.end());
            L.ops32 = c->upload(pk32.data(), pk32.size());
            if (o.p == 2) {
                const std::vector<double> pm = build_tet_mfma_pack<2>(o);
                L.opsm = c->upload(pm.data(), pm.size());
                const std::vector<double> pm32d = build_tet_mfma_pack<2>(o, true);
                const std::vector<float> pm32(pm32d.begin(), pm32d.end());
                L.opsm32 = c->upload(pm32.data(), pm32.size());
            }
            L.P2 = c->upload(o.P2.data(), o.P2.size());
            L.W2 = c->upload(o.W2.data(), o.W2.size());
            L.Pg = c->upload(o.P.data(), o.P.size());
            L.vdiag = o.vdiag;
            if (o.em > 0) {
                L.ediag = o.ediag[0]; L.sfmv0 = o.sfmv[0]; L.sfmv1 = o.sfmv[1];
                L.vfms = c->upload(o.vfms.data(), o.vfms.size());
            }
        }
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const std::exception&) {
        return -1;
    }
    *out = c.release();
    return 0;
}

int hpg_tet_destroy(hpg_tet_ctx* c) {
    delete c;
    return 0;
}

static int tet_copy(hpg_tet_ctx* c, bool h2d, double* dv, double* de, const double* hv, const double* he, double* ov,
                    double* oe) {
    HPG_TRY(c, {
        const size_t nv = (size_t)c->npnt * c->nv * 8, ne = (size_t)c->nseg * c->em0 * c->nv * 8;
        if (h2d) {
            check(hipMemcpyAsync(dv, hv, nv, hipMemcpyHostToDevice, c->stream), "H2D");
            if (ne) check(hipMemcpyAsync(de, he, ne, hipMemcpyHostToDevice, c->stream), "H2D");
        } else {
            check(hipMemcpyAsync(ov, dv, nv, hipMemcpyDeviceToHost, c->stream), "D2H");
            if (ne) check(hipMemcpyAsync(oe, de, ne, hipMemcpyDeviceToHost, c->stream), "D2H");
        }
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_tet_set_state(hpg_tet_ctx* c, const double* v, const double* e) {
    if (!c || !v || (c->em0 && !e)) return -2;
    c->fd_reset = true;
    return tet_copy(c, true, c->ugv, c->uge, v, e, nullptr, nullptr);
}
int hpg_tet_get_state(hpg_tet_ctx* c, double* v, double* e) {
    if (!c || !v || (c->em0 && !e)) return -2;
    return tet_copy(c, false, c->ugv, c->uge, nullptr, nullptr, v, e);
}
int hpg_tet_get_residual(hpg_tet_ctx* c, double* v, double* e) {
    if (!c || !v || (c->em0 && !e)) return -2;
    return tet_copy(c, false, c->resv, c->rese, nullptr, nullptr, v, e);
}
int hpg_tet_set_residual(hpg_tet_ctx* c, const double* v, const double* e) {
    if (!c || !v || (c->em0 && !e)) return -2;
    return tet_copy(c, true, c->resv, c->rese, v, e, nullptr, nullptr);
}
int hpg_tet_get_preconditioner(hpg_tet_ctx* c, double* vprcn, double* eprcn, double* iprcn, double* tau) {
    if (!c || !vprcn || !eprcn || !iprcn || !tau) return -2;
    HPG_TRY(c, {
        const size_t k = c->nv, kt = c->cns() ? 25 : TNV;
        check(hipMemcpyAsync(vprcn, c->vprcn, (size_t)c->npnt * k * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(eprcn, c->eprcn, (size_t)c->nseg * k * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(iprcn, c->iprcn, (size_t)c->ntet * k * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(tau, c->tau, (size_t)c->ntet * kt * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}
int hpg_tet_get_cns_blocks(hpg_tet_ctx* c, double* vpc, double* epc) {
    if (!c || !vpc || !epc) return -2;
    if (!c->cns()) { c->err = "5 x 5 preconditioner blocks exist for cns only"; return -2; }
    HPG_TRY(c, {
        check(hipMemcpyAsync(vpc, c->vpc, (size_t)c->npnt * 25 * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipMemcpyAsync(epc, c->epc, (size_t)c->nseg * 25 * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}
int hpg_tet_rsdl(hpg_tet_ctx* c, int stage) {
    if (!c || stage < 0 || stage > c->prm.nstage) return -2;
    HPG_TRY(c, { do_rsdl_tet(c, stage); });
    return 0;
}
int hpg_tet_setup_preconditioner(hpg_tet_ctx* c) {
    if (!c) return -2;
    int e = 0;
    HPG_TRY(c, { e = do_setup_tet(c); });
    return e;
}
int hpg_tet_minvrt(hpg_tet_ctx* c) {
    if (!c) return -2;
    HPG_TRY(c, { do_minvrt_tet(c, false, -1.0); });
    return 0;
}
int hpg_tet_update(hpg_tet_ctx* c) {
    if (!c) return -2;
    if (c->precision == HPG_FP32_DEFECT && !c->coarse_flag) {
        HPG_TRY(c, mixed_update_tet(c));
        return 0;
    }
    HPG_TRY(c, {
        // tet_hp::update (update.cpp:181-417) and tet_hp_cns::update (cns/update.cpp:18-143,
        // no Dirichlet pass before minvrt)
        check(hipMemcpyAsync(c->ug0v, c->ugv, (size_t)c->npnt * c->nv * 8, hipMemcpyDeviceToDevice, c->stream), "D2D");
        if (c->em0)
            check(hipMemcpyAsync(c->ug0e, c->uge, (size_t)c->nseg * c->em0 * c->nv * 8, hipMemcpyDeviceToDevice,
                                 c->stream), "D2D");
        for (int s = 0; s < c->prm.nstage; ++s) {
            do_rsdl_tet(c, s);
            do_minvrt_tet(c, !c->cns(), c->prm.alpha[s] * c->prm.cfl[c->log2p]);
        }
    });
    return 0;
}
int hpg_tet_element_jacobian(hpg_tet_ctx* c, const int* tind, int nel, double eps_r, double eps_a, double* K) {
    if (!c || nel < 0) return -2;
    if (nel == 0) return 0;
    if (!tind || !K) return -2;
    HPG_TRY(c, {
        const TetLevel& L = c->lv[c->log2p];
        const int tm = L.p == 1 ? 4 : 10, nvv = c->nv, N = tm * nvv;
        std::vector<int> tp((size_t)c->ntet * 4);
        check(hipMemcpyAsync(tp.data(), c->tet_pnt, tp.size() * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
        std::vector<char> seen(c->ntet, 0);
        for (int j = 0; j < nel; ++j) {
            if (tind[j] < 0 || tind[j] >= c->ntet) throw std::runtime_error("element index out of range");
            if (seen[tind[j]]) throw std::runtime_error("element listed twice");
            seen[tind[j]] = 1;
        }
        // split of the batch into vertex-disjoint sets: the greedy colouring of its
        // vertex lists (at p <= 2 elements sharing no vertex share no edge either)
        std::vector<int> bv((size_t)nel * 4), order, off, colr;
        for (int j = 0; j < nel; ++j)
            for (int k = 0; k < 4; ++k) bv[4 * (size_t)j + k] = tp[4 * (size_t)tind[j] + k];
        colour(nel, 4, bv.data(), c->npnt, order, off, colr);
        std::vector<int> jel(nel), jpos(nel);
        for (int i = 0; i < nel; ++i) { jel[i] = tind[order[i]]; jpos[i] = order[i]; }
        std::vector<void*> bufs;
        auto get = [&](size_t bytes) {
            void* q = nullptr;
            if (hipMalloc(&q, std::max<size_t>(bytes, 8)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
            bufs.push_back(q);
            return q;
        };
        struct Release {
            hpg_tet_ctx* c; std::vector<void*>& b;
            ~Release() { hipStreamSynchronize(c->stream); for (void* q : b) hipFree(q); }
        } release{c, bufs};
        TetJacArgs a{};
        a.jel = (const int*)get((size_t)nel * 4);
        a.jpos = (const int*)get((size_t)nel * 4);
        check(hipMemcpyAsync((void*)a.jel, jel.data(), (size_t)nel * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        check(hipMemcpyAsync((void*)a.jpos, jpos.data(), (size_t)nel * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        a.tet_pnt = c->tet_pnt; a.tet_seg = c->tet_seg; a.em0 = c->em0; a.tm = tm; a.nv = nvv;
        a.ugv = c->ugv; a.uge = c->uge; a.lf = c->lf;
        a.dw = (double*)get((size_t)nel * 8);
        a.save = (double*)get((size_t)nel * 8);
        a.base = (double*)get((size_t)nel * N * 8);
        a.K = (double*)get((size_t)nel * N * N * 8);
        a.eps_r = eps_r; a.eps_a = eps_a;
        // element residuals of batch positions [b0, b1) only
        auto elem = [&](int b0, int b1) {
            if (L.p == 1) launch_elem_tet<1>(c, 0, a.jel + b0, b1 - b0);
            else launch_elem_tet<2>(c, 0, a.jel + b0, b1 - b0);
        };
        // dw and Rbar of every batch element from the unperturbed state
        TetJacArgs all = a;
        all.begin = 0; all.end = nel; all.mode = -1; all.var = 0;
        hipLaunchKernelGGL(k_tjac_dw, dim3(nblk(nel, 64)), dim3(64), 0, c->stream, all);
        elem(0, nel);
        hipLaunchKernelGGL(k_tjac_collect, dim3(nblk((long)nel * N)), dim3(256), 0, c->stream, all);
        for (size_t k = 0; k + 1 < off.size(); ++k) {
            TetJacArgs b = a;
            b.begin = off[k]; b.end = off[k + 1];
            const int ns = b.end - b.begin;
            for (int col = 0; col < N; ++col) {
                b.mode = col / nvv; b.var = col % nvv;
                hipLaunchKernelGGL(k_tjac_perturb, dim3(nblk(ns, 64)), dim3(64), 0, c->stream, b);
                elem(b.begin, b.end);
                hipLaunchKernelGGL(k_tjac_collect, dim3(nblk((long)ns * N)), dim3(256), 0, c->stream, b);
            }
        }
        check(hipGetLastError(), "tet jacobian launch");
        check(hipMemcpyAsync(K, a.K, (size_t)nel * N * N * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
        check(hipStreamSynchronize(c->stream), "sync");
    });
    return 0;
}

int hpg_tet_maxres(hpg_tet_ctx* c, double* per_var, double* out) {
    if (!c || !out) return -2;
    HPG_TRY(c, {
        double a[CNV];
        *out = tet_absmax(c, c->npnt, 1, 1, c->resv, nullptr, a);
        if (per_var)
            for (int n = 0; n < c->nv; ++n) per_var[n] = a[n];
    });
    return 0;
}
int hpg_tet_resnorm(hpg_tet_ctx* c, double* out) {
    if (!c || !out) return -2;
    HPG_TRY(c, {
        double a[CNV], b[CNV] = {0, 0, 0, 0, 0};
        double r = tet_absmax(c, c->npnt, 1, 1, c->resv, c->vess, a);
        if (c->em() > 0) r = std::max(r, tet_absmax(c, c->nseg, c->em(), c->em0, c->rese, c->eess, b));
        *out = r;
    });
    return 0;
}
int hpg_tet_mg_restrict(hpg_tet_ctx* c) {
    if (!c || c->log2p == 0) return -2;
    HPG_TRY(c, {
        c->isfrst = true;
        c->coarse_flag = true;
        --c->log2p;
        tet_modes(c, 3, c->resv, c->rese, c->r0v, c->r0e, nullptr, nullptr, 0.0);
        check(hipGetLastError(), "tet restrict");
    });
    return 0;
}
int hpg_tet_mg_prolongate(hpg_tet_ctx* c) {
    if (!c || c->log2p >= c->log2pmax) return -2;
    ++c->log2p;
    if (c->log2p == c->log2pmax) c->coarse_flag = false;
    return 0;
}
int hpg_tet_set_precision(hpg_tet_ctx* c, int mode, int ninner) {
    if (!c || (mode != HPG_FP64 && mode != HPG_FP32_DEFECT) || ninner < 1 || ninner > 1000) return -2;
    if (mode == HPG_FP32_DEFECT && !c->srcv) {
        HPG_TRY(c, {
            const size_t nv = (size_t)c->npnt * c->nv, ne = (size_t)c->nseg * c->em0 * c->nv;
            c->srcv = c->alloc<double>(nv); c->srce = c->alloc<double>(ne);
            c->u0cv = c->alloc<double>(nv); c->u0ce = c->alloc<double>(ne);
            check(hipStreamSynchronize(c->stream), "sync");
        });
    }
    c->precision = mode;
    c->ninner = ninner;
    c->fd_reset = true;
    return 0;
}

int hpg_tet_residual_fp32(hpg_tet_ctx* c, double* resv, double* rese) {
    if (!c || !resv) return -2;
    if (c->log2p != c->log2pmax) { c->err = "finest level only"; return -2; }
    HPG_TRY(c, do_rsdl_tet(c, c->prm.nstage, true));
    return hpg_tet_get_residual(c, resv, rese);
}

int hpg_tet_get_log2p(hpg_tet_ctx* c, int* l) {
    if (!c || !l) return -2;
    *l = c->log2p;
    return 0;
}
int hpg_tet_synchronize(hpg_tet_ctx* c) {
    if (!c) return -2;
    HPG_TRY(c, { check(hipStreamSynchronize(c->stream), "sync"); });
    return 0;
}
const char* hpg_tet_last_error(hpg_tet_ctx* c) { return c ? c->err.c_str() : "null context"; }
int hpg_tet_kernel_timing(hpg_tet_ctx* c, int on) {
    if (!c) return -2;
    c->timing = on != 0;
    c->recs.clear();
    return 0;
}
int hpg_tet_kernel_timing_read(hpg_tet_ctx* c, int full, double* total_ms, long long* launches, long long* elements) {
    if (!c || !total_ms || !launches || !elements) return -2;
    HPG_TRY(c, {
        check(hipStreamSynchronize(c->stream), "sync");
        double t = 0.0;
        long long nl = 0, ne = 0;
        for (size_t r = 0; r < c->recs.size(); ++r) {
            if (c->recs[r].full != full) continue;
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, c->events[2 * r], c->events[2 * r + 1]), "hipEventElapsedTime");
            t += ms;
            ++nl;
            ne += c->recs[r].nelem;
        }
        *total_ms = t; *launches = nl; *elements = ne;
    });
    return 0;
}

}  // extern "C"
