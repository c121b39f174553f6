// HIP kernels of the tri_hp / tri_hp_ins residual and relaxation path (gfx950).
//
// Element residual: 16 elements per wave64, every basis transform a dense
// contraction on v_mfma_f64_16x16x4f64.  Rows of an MFMA tile are Gauss points
// (or modes), columns are the 16 elements of the wave.  Operands: A (16x4) lane l
// holds A[l%16][l/16], B (4x16) lane l holds B[l/16][l%16]; the f64 accumulator
// register i of lane l holds D[4*i + l/16][l%16] (checked on the device by
// hpg_selftest_mfma).  Accumulators are fed straight back as the B operand of the
// next contraction: the operator matrices are stored with their K index permuted
// so that K-step (tile t, register i) reads row 16t + 4*i + l/16, which is exactly
// what lane l holds.  No data moves between lanes from the modal gather to the
// modal scatter.
#include <hip/hip_runtime.h>
#include "tri_kernels.h"

namespace hpg {

typedef double d4 __attribute__((ext_vector_type(4)));

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// v_mfma_f32_16x16x4f32: accumulator register i of lane l holds D[4*(l/16) + i][l%16]
__device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef d4 type; };
template <> struct Vec4<float> { typedef f4 type; };

// ---------------------------------------------------------------------------------
// tri_hp::ugtouht / lftog index map for mode k of element t (tri_hp/gtol.cpp:14-55,
// 364-402): returns the offset of the NV-vector and the side sign sgn^m.
// a[i] for a lane-varying i < 3 without dynamic register indexing (which would spill
// the array to scratch)
__device__ __forceinline__ int sel3(const int* a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

template <int P, int NVn = NV>
__device__ __forceinline__ int mode_offset(int k, int t, const int* vtx, const int* seg, const int* sgn,
                                           int sm0, int im0, int& kind, double& sign) {
    using D = Dims<P>;
    sign = 1.0;
    if (k < 3) { kind = 0; return sel3(vtx, k) * NVn; }
    if (k < D::BM) {
        constexpr int SMD = D::SM > 0 ? D::SM : 1;
        const int i = (k - 3) / SMD, m = (k - 3) - i * SMD;
        kind = 1;
        sign = ((m & 1) && sel3(sgn, i) < 0) ? -1.0 : 1.0;
        return (sel3(seg, i) * sm0 + m) * NVn;
    }
    kind = 2;
    return (t * im0 + (k - D::BM)) * NVn;
}

// Per-lane gather state of one element, loaded one group ahead of its use.
// Slot s of lane l is mode 4s + l/16 of element l%16: the K-step of the projection
// that reads it and the accumulator register that returns its residual.
struct ElemConn {
    int t;  // element, -1 for a padding lane
    int vtx[3], seg[3], sgn[3];
};

template <int P>
struct ElemSlots {
    static constexpr int KT = Dims<P>::KT;
    int t;
    int off[KT];   // offset of the slot's NV-vector, bitwise-negated for a side sign of -1
    int first;     // bit s: the slot's first visit in this stage
    double x[3], y[3];
};

// kind of mode k: 0 vertex, 1 side, 2 interior (mode_offset's classification)
template <int P>
__device__ __forceinline__ int mode_kind(int k) { return k < 3 ? 0 : (k < Dims<P>::BM ? 1 : 2); }

// element of lane e in 16-element group g; a null order walks the elements in their
// own (mesh) order, which keeps neighbouring elements in one group
template <class A>
__device__ __forceinline__ int elem_of_group(const A& a, int g, int gend, int e) {
    const int idx = a.begin + g * 16 + e;
    if (!(g < gend && idx < a.end)) return -1;
    return a.order ? a.order[idx] : idx;
}

template <class A>
__device__ __forceinline__ ElemConn elem_conn(const A& a, int t) {
    ElemConn c;
    c.t = t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c.vtx[k] = t >= 0 ? a.tri_pnt[3 * t + k] : 0;
        c.seg[k] = t >= 0 ? a.tri_seg[3 * t + k] : 0;
        c.sgn[k] = t >= 0 ? a.tri_sgn[3 * t + k] : 1;
    }
    return c;
}

template <int P, class A>
__device__ __forceinline__ ElemSlots<P> elem_slots(const A& a, const ElemConn& c, int q) {
    using D = Dims<P>;
    constexpr int SMD = D::SM > 0 ? D::SM : 1;
    ElemSlots<P> S;
    S.t = c.t;
    const bool valid = c.t >= 0;
    // padding lanes compute on the reference triangle with a zero state
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S.x[k] = valid ? a.pnts[2 * c.vtx[k]] : (k == 1 ? 1.0 : 0.0);
        S.y[k] = valid ? a.pnts[2 * c.vtx[k] + 1] : (k == 2 ? 1.0 : 0.0);
    }
#pragma unroll
    for (int s = 0; s < D::KT; ++s) {
        const int k = 4 * s + q;
        int kind;
        double sg;
        const int off = mode_offset<P>(k, valid ? c.t : 0, c.vtx, c.seg, c.sgn, a.sm0, a.im0, kind, sg);
        S.off[s] = sg < 0.0 ? ~off : off;
        const bool live = valid && k < D::TM;
        const bool first = live && (kind == 2 || (kind == 0 ? a.fcv[sel3(c.vtx, k)]
                                                             : a.fcs[sel3(c.seg, (k - 3) / SMD)]) == a.colour);
        if (s == 0) S.first = 0;
        S.first |= first ? (1 << s) : 0;
    }
    return S;
}

// tri_hp_ins::element_rsdl, straight branch (tri_hp/ins/rsdl.cpp:259-433), steady,
// for the elements order[begin:end) (one colour: no two share a vertex or side).
// Persistent: each wave walks 16-element groups of its XCD's contiguous share of the
// colour and keeps a three-deep load pipeline (element ids three groups ahead,
// connectivity two ahead, modes/geometry/first-visit flags one ahead; the old
// residual values of the current group are requested before its contractions), so
// the dependent gathers overlap the MFMA work of earlier groups.  The grid is a
// multiple of 8 blocks of 4 waves.
// CURV: the curved branch (tri_hp/ins/rsdl.cpp:83-257) for the elements order[begin:end)
// that own a curved boundary side: the coordinate derivatives dcrd are projected to
// every Gauss point from the element's coordinate modes (tri_hp/metric.cpp:12-21:
// crdtocht + proj_bdry), so the metric terms, the Jacobian and tau vary per point.
template <typename T, int P, bool FULL, bool EM = false, bool CURV = false>
__device__ __forceinline__ void elem_ins_body(const ElemArgsT<T>& a, T* __restrict__ eres = nullptr,
                                              int egrid = 0) {
    using D = Dims<P>;
    using V4 = typename Vec4<T>::type;
    constexpr int GT = D::GT, KT = D::KT;
    __shared__ T sops[D::PACK];
    for (int k = threadIdx.x; k < D::PACK; k += blockDim.x) sops[k] = a.ops[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = lane & 15, q = lane >> 4;
    // XCD-aware split: hardware deals consecutive blocks round-robin to the 8 XCDs;
    // XCD x owns groups [x*gpx, (x+1)*gpx) so neighbouring elements share its L2
    // egrid: blocks of the launch that walk elements (the rest do other work)
    const int WX = ((egrid ? egrid : (int)gridDim.x) >> 3) * 4;
    const int rank = (blockIdx.x >> 3) * 4 + w;
    const int xcd = blockIdx.x & 7;
    const int ng = (a.end - a.begin + 15) >> 4;
    const int gpx = (ng + 7) >> 3;
    const int g0 = xcd * gpx + rank;
    const int gend = min(ng, (xcd + 1) * gpx);
    if (g0 >= gend) return;

    ElemSlots<P> S = elem_slots<P>(a, elem_conn(a, elem_of_group(a, g0, gend, e)), q);
    ElemConn cnext = elem_conn(a, elem_of_group(a, g0 + WX, gend, e));
    int tnext2 = elem_of_group(a, g0 + 2 * WX, gend, e);

    for (int g = g0; g < gend; g += WX) {
        // old residual values of this group, consumed by the scatter at the end and
        // requested late (register pressure): before the last contraction of the full
        // residual, after the contractions of the inviscid one
        T ro[NV][KT], io[NV][KT];
        auto load_old = [&]() {
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const int k = 4 * s + q, kind = mode_kind<P>(k);
                const bool live = S.t >= 0 && k < D::TM;
                const int off = S.off[s] < 0 ? ~S.off[s] : S.off[s];
                const T* di = kind == 0 ? a.resv : (kind == 1 ? a.ress : a.resi);
                const T* dr = kind == 0 ? a.rrv : (kind == 1 ? a.rrs : a.rri);
                // element-major output (EM) reads back only the interior modes, which
                // this element owns; the gathers read the carried parts of the others
                const bool own = !EM || kind == 2;
#pragma unroll
                for (int n = 0; n < NV; ++n) {
                    ro[n][s] = (live && own && (FULL || a.keep != T(0))) ? dr[off + n] : T(0);
                    io[n][s] = (live && !EM && !((S.first >> s) & 1)) ? di[off + n] : T(0);
                }
            }
        };
        // prefetch: next group's slots, connectivity two ahead, element ids three ahead
        const ElemSlots<P> Sn = elem_slots<P>(a, cnext, q);
        const ElemConn cn2 = elem_conn(a, tnext2);
        const int tn3 = elem_of_group(a, g + 3 * WX, gend, e);

        // modal coefficients of this group: K-step s of lane l is mode 4s + l/16
        T b[NV][KT];
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            const int k = 4 * s + q, kind = mode_kind<P>(k);
            const bool live = S.t >= 0 && k < D::TM;
            const int off = S.off[s] < 0 ? ~S.off[s] : S.off[s];
            const T sg = S.off[s] < 0 ? T(-1) : T(1);
            const T* src = kind == 0 ? a.ugv : (kind == 1 ? a.ugs : a.ugi);
#pragma unroll
            for (int n = 0; n < NV; ++n) b[n][s] = live ? sg * src[off + n] : T(0);
        }
        const double x0 = S.x[0], y0 = S.y[0], x1 = S.x[1], y1 = S.y[1], x2 = S.x[2], y2 = S.y[2];
        // ldcrd(n,0) = dx_n/dr, ldcrd(n,1) = dx_n/ds (rsdl.cpp:44-47)
        const T l00 = T(0.5 * (x2 - x1)), l01 = T(0.5 * (x0 - x1));
        const T l10 = T(0.5 * (y2 - y1)), l11 = T(0.5 * (y0 - y1));

        const V4 z = {T(0), T(0), T(0), T(0)};
        // curved: dx/dr, dx/ds, dy/dr, dy/ds at the Gauss points.  The coordinate modes
        // are taken relative to vertex 1 (the derivatives of a constant vanish), which
        // keeps the FP32 form free of the cancellation of absolute coordinates.
        V4 Xr[CURV ? GT : 1], Xs[CURV ? GT : 1], Yr[CURV ? GT : 1], Ys[CURV ? GT : 1];
        if constexpr (CURV) {
            constexpr int SMD = D::SM > 0 ? D::SM : 1;
            const int cidx = a.begin + g * 16 + e;
            T bx[KT], by[KT];
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const int k = 4 * s + q;
                T cx = T(0), cy = T(0);
                if (k == 0) { cx = T(x0 - x1); cy = T(y0 - y1); }
                else if (k == 2) { cx = T(x2 - x1); cy = T(y2 - y1); }
                else if (k >= 3 && k < D::BM && S.t >= 0) {
                    const int i = (k - 3) / SMD, m = (k - 3) - i * SMD;
                    const double* cr = a.crv + ((size_t)(cidx * 3 + i) * a.sm0 + m) * 2;
                    cx = T(cr[0]); cy = T(cr[1]);
                }
                bx[s] = cx; by[s] = cy;
            }
#pragma unroll
            for (int mt = 0; mt < GT; ++mt) {
                V4 c1 = z, c2 = z, c3 = z, c4 = z;
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const T ar = sops[D::OFF_PR + (mt * KT + s) * 64 + lane];
                    const T as = sops[D::OFF_PS + (mt * KT + s) * 64 + lane];
                    c1 = mfma(ar, bx[s], c1); c2 = mfma(as, bx[s], c2);
                    c3 = mfma(ar, by[s], c3); c4 = mfma(as, by[s], c4);
                }
                Xr[mt] = c1; Xs[mt] = c2; Yr[mt] = c3; Ys[mt] = c4;
            }
        }
        // metric terms of Gauss point (mt, i): the element constants of a straight
        // element, the projected derivatives of a curved one.  The padding rows of the
        // tiles (points beyond the G Gauss points) have zero operator rows, hence a zero
        // Jacobian; they keep the straight metric so that their values stay finite (the
        // integrations weight them by zero, and 0 x inf would reach every mode).
        auto metric = [&](int mt, int i, T& m00, T& m01, T& m10, T& m11) {
            if constexpr (CURV) {
                m00 = Xr[mt][i]; m01 = Xs[mt][i]; m10 = Yr[mt][i]; m11 = Ys[mt][i];
                if (m00 * m11 - m10 * m01 == T(0)) { m00 = l00; m01 = l01; m10 = l10; m11 = l11; }
            } else {
                m00 = l00; m01 = l01; m10 = l10; m11 = l11;
            }
        };
        V4 U[GT], V[GT], Pp[GT];
#pragma unroll
        for (int mt = 0; mt < GT; ++mt) {
            V4 cu = z, cv = z, cp = z;
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const T av = sops[D::OFF_P + (mt * KT + s) * 64 + lane];
                cu = mfma(av, b[0][s], cu);
                cv = mfma(av, b[1][s], cv);
                cp = mfma(av, b[2][s], cp);
            }
            U[mt] = cu; V[mt] = cv; Pp[mt] = cp;
        }

        // inviscid fluxes (rsdl.cpp:262-291): continuity C, momentum A
        V4 C0[GT], C1[GT], A00[GT], A01[GT], A10[GT], A11[GT];
        const T rho = a.rho;
#pragma unroll
        for (int mt = 0; mt < GT; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                T m00, m01, m10, m11;
                metric(mt, i, m00, m01, m10, m11);
                const T u = U[mt][i], v = V[mt][i], pr = Pp[mt][i];
                const T fx = rho * u, fy = rho * v;
                const T c0 = +m11 * fx - m01 * fy;
                const T c1 = -m10 * fx + m00 * fy;
                C0[mt][i] = c0; C1[mt][i] = c1;
                A00[mt][i] = u * c0 + m11 * pr;
                A01[mt][i] = u * c1 - m10 * pr;
                A10[mt][i] = v * c0 - m01 * pr;
                A11[mt][i] = v * c1 + m00 * pr;
            }

        // lf_im = intgrtrs of the inviscid fluxes (rsdl.cpp:292-294)
        V4 LI0 = z, LI1 = z, LI2 = z;
#pragma unroll
        for (int kt = 0; kt < GT; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // f64 layout: K-step (tile kt, register i) holds points 16 kt + 4 i .. + 3;
                // a step of padding points only (past the G Gauss points) adds nothing
                if (sizeof(T) == 8 && 16 * kt + 4 * i >= D::G) continue;
                const int kk = kt * 4 + i;
                const T ar = sops[D::OFF_IR + kk * 64 + lane];
                const T as = sops[D::OFF_IS + kk * 64 + lane];
                LI0 = mfma(ar, A00[kt][i], LI0); LI0 = mfma(as, A01[kt][i], LI0);
                LI1 = mfma(ar, A10[kt][i], LI1); LI1 = mfma(as, A11[kt][i], LI1);
                LI2 = mfma(ar, C0[kt][i], LI2);  LI2 = mfma(as, C1[kt][i], LI2);
            }

        V4 LR0 = z, LR1 = z, LR2 = z;
        if constexpr (FULL) {
            // derivatives of u, v at the Gauss points (proj with r/s derivatives)
            V4 Ur[GT], Us[GT], Vr[GT], Vs[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt) {
                V4 c1 = z, c2 = z, c3 = z, c4 = z;
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const T ar = sops[D::OFF_PR + (mt * KT + s) * 64 + lane];
                    const T as = sops[D::OFF_PS + (mt * KT + s) * 64 + lane];
                    c1 = mfma(ar, b[0][s], c1); c2 = mfma(as, b[0][s], c2);
                    c3 = mfma(ar, b[1][s], c3); c4 = mfma(as, b[1][s], c4);
                }
                Ur[mt] = c1; Us[mt] = c2; Vr[mt] = c3; Vs[mt] = c4;
            }
            // viscous stress tensor (rsdl.cpp:303-326, curved :139-159 per point); steady
            // so rhobd0 = 0
            const T cjcb = l00 * l11 - l10 * l01;
            const T cjcbi = a.mu / cjcb;
            T v0000 = -cjcbi * (T(2) * l11 * l11 + l01 * l01);
            T v0011 = -cjcbi * (T(2) * l10 * l10 + l00 * l00);
            T v0001 = cjcbi * (T(2) * l11 * l10 + l01 * l00);
            T v1100 = -cjcbi * (l11 * l11 + T(2) * l01 * l01);
            T v1111 = -cjcbi * (l10 * l10 + T(2) * l00 * l00);
            T v1101 = cjcbi * (l11 * l10 + T(2) * l01 * l00);
            T v0100 = cjcbi * l01 * l11;
            T v0111 = cjcbi * l00 * l10;
            T v0101 = -cjcbi * l01 * l10;
            T v0110 = -cjcbi * l00 * l11;
            V4 F00[GT], F01[GT], F10[GT], F11[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if constexpr (CURV) {
                        T m00, m01, m10, m11;
                        metric(mt, i, m00, m01, m10, m11);
                        const T cji = a.mu / (m00 * m11 - m10 * m01);
                        v0000 = -cji * (T(2) * m11 * m11 + m01 * m01);
                        v0011 = -cji * (T(2) * m10 * m10 + m00 * m00);
                        v0001 = cji * (T(2) * m11 * m10 + m01 * m00);
                        v1100 = -cji * (m11 * m11 + T(2) * m01 * m01);
                        v1111 = -cji * (m10 * m10 + T(2) * m00 * m00);
                        v1101 = cji * (m11 * m10 + T(2) * m01 * m00);
                        v0100 = cji * m01 * m11;
                        v0111 = cji * m00 * m10;
                        v0101 = -cji * m01 * m10;
                        v0110 = -cji * m00 * m11;
                    }
                    const T ur = Ur[mt][i], us = Us[mt][i], vr = Vr[mt][i], vs = Vs[mt][i];
                    const T f00 = v0000 * ur + v0100 * vr + v0001 * us + v0101 * vs;
                    const T f01 = v0001 * ur + v0110 * vr + v0011 * us + v0111 * vs;
                    const T f10 = v0100 * ur + v1100 * vr + v0110 * us + v1101 * vs;
                    const T f11 = v0101 * ur + v1101 * vr + v0111 * us + v1111 * vs;
                    F00[mt][i] = f00; F01[mt][i] = f01; F10[mt][i] = f10; F11[mt][i] = f11;
                    A00[mt][i] += f00; A01[mt][i] += f01; A10[mt][i] += f10; A11[mt][i] += f11;
                }
            // strong residual: derivr/derivs of the total fluxes (rsdl.cpp:369-375)
            V4 R0[GT], R1[GT], R2[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt) {
                V4 r0 = z, r1 = z, r2 = z;
#pragma unroll
                for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (sizeof(T) == 8 && 16 * kt + 4 * i >= D::G) continue;  // padding points only
                        const int kk = (mt * GT + kt) * 4 + i;
                        const T dr = sops[D::OFF_DR + kk * 64 + lane];
                        const T ds = sops[D::OFF_DS + kk * 64 + lane];
                        r0 = mfma(dr, A00[kt][i], r0); r0 = mfma(ds, A01[kt][i], r0);
                        r1 = mfma(dr, A10[kt][i], r1); r1 = mfma(ds, A11[kt][i], r1);
                        r2 = mfma(dr, C0[kt][i], r2);  r2 = mfma(ds, C1[kt][i], r2);
                    }
                R0[mt] = r0; R1[mt] = r1; R2[mt] = r2;
            }
            if (sizeof(T) == 8 && a.lin) {
                // linearisation data of this 16-element group, coalesced per (field, tile, reg)
                float* L = a.lin + (size_t)(a.gbase + g) * (5 * GT * 4 * 64);
#pragma unroll
                for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int o = (mt * 4 + i) * 64 + lane;
                        L[0 * GT * 256 + o] = (float)U[mt][i];
                        L[1 * GT * 256 + o] = (float)V[mt][i];
                        L[2 * GT * 256 + o] = (float)R0[mt][i];
                        L[3 * GT * 256 + o] = (float)R1[mt][i];
                        L[4 * GT * 256 + o] = (float)R2[mt][i];
                    }
            }
            // per-point tau (CALC_TAU2, rsdl.cpp:377-395) and the stabilisation (:397-421)
            const T dx1 = T(x0 - x2), dy1 = T(y0 - y2), dx2 = T(x1 - x0), dy2 = T(y1 - y0);
            const T ar2 = dx1 * dy2 - dy1 * dx2;
            const T perim = sqrt(dx1 * dx1 + dy1 * dy1) + sqrt(dx2 * dx2 + dy2 * dy2) +
                            sqrt((dx1 + dx2) * (dx1 + dx2) + (dy1 + dy2) * (dy1 + dy2));
            const T h = (ar2 / perim) / (T(0.25) * (P + 1) * (P + 1));
            const T nu = a.mu / rho;
            const T gv = T(2) * nu / h;
            const T gvis = gv * gv;
            const T tfac = a.adis * h / cjcb;
#pragma unroll
            for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    T m00, m01, m10, m11;
                    metric(mt, i, m00, m01, m10, m11);
                    const T u = U[mt][i], v = V[mt][i];
                    const T qq = u * u + v * v;
                    T gam = T(3) * qq + gvis;
                    if (a.mu == T(0)) gam = fmax(gam, T(0.1));
                    // curved: cjcb of this Gauss point (rsdl.cpp:385)
                    const T tf = CURV ? a.adis * h / (m00 * m11 - m10 * m01) : tfac;
                    const T tau0 = tf / sqrt(gam);
                    const T tau2 = sqrt(qq) * tau0;
                    const T t0 = tau0 * R0[mt][i], t1 = tau0 * R1[mt][i], t2 = tau2 * R2[mt][i];
                    F00[mt][i] -= (m11 * (T(2) * u) - m01 * v) * t0 - m01 * u * t1 + m11 * t2;
                    F01[mt][i] -= (-m10 * (T(2) * u) + m00 * v) * t0 + m00 * u * t1 - m10 * t2;
                    F10[mt][i] -= +m11 * v * t0 + (m11 * u - m01 * (T(2) * v)) * t1 - m01 * t2;
                    F11[mt][i] -= -m10 * v * t0 + (-m10 * u + m00 * (T(2) * v)) * t1 + m00 * t2;
                    C0[mt][i] = -(m11 * t0 - m01 * t1);
                    C1[mt][i] = -(-m10 * t0 + m00 * t1);
                }
            load_old();
            // lf_re = intgrtrs of viscous - stabilisation fluxes (rsdl.cpp:424-426)
#pragma unroll
            for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (sizeof(T) == 8 && 16 * kt + 4 * i >= D::G) continue;  // padding points only
                    const int kk = kt * 4 + i;
                    const T ar = sops[D::OFF_IR + kk * 64 + lane];
                    const T as = sops[D::OFF_IS + kk * 64 + lane];
                    LR0 = mfma(ar, F00[kt][i], LR0); LR0 = mfma(as, F01[kt][i], LR0);
                    LR1 = mfma(ar, F10[kt][i], LR1); LR1 = mfma(as, F11[kt][i], LR1);
                    LR2 = mfma(ar, C0[kt][i], LR2);  LR2 = mfma(as, C1[kt][i], LR2);
                }
        }


        if constexpr (!FULL) load_old();
        if constexpr (EM) {
            // element-major output of the vertex and side modes, side sign applied:
            // eres[(t*BM + k)*2*NV + n] = lf_im, [... + NV + n] = beta lf_re (full
            // residual only); the gather kernels add the carried real part and sum over
            // the elements.  Interior modes belong to this element: final here, started
            // from keep * res_r as on a first visit.
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const int k = 4 * s + q, kind = mode_kind<P>(k);
                if (S.t < 0 || k >= D::TM) continue;
                const T sg = S.off[s] < 0 ? T(-1) : T(1);
                const T li[3] = {sg * LI0[s], sg * LI1[s], sg * LI2[s]};
                T lr[3] = {T(0), T(0), T(0)};
                if constexpr (FULL) {
                    const T bt = a.beta;
                    lr[0] = sg * (LR0[s] * bt);
                    lr[1] = sg * (LR1[s] * bt);
                    lr[2] = sg * (LR2[s] * bt);
                }
                if (kind < 2) {
                    T* o = eres + ((size_t)S.t * D::BM + k) * (2 * NV);
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        o[n] = li[n];
                        if constexpr (FULL) o[NV + n] = lr[n];
                    }
                } else {
                    const int off = S.off[s];
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        const T r0 = a.keep == T(0) ? T(0) : a.keep * ro[n][s];
                        if (FULL || a.keep != T(1)) a.rri[off + n] = r0 + lr[n];
                        if (sizeof(T) == 4 && a.srci)
                            a.resi[off + n] = r0 + lr[n] + li[n] + T(a.srci[off + n] * a.sfac[0]);
                        else
                            a.resi[off + n] = r0 + lr[n] + li[n];
                    }
                }
            }
            S = Sn;
            cnext = cn2;
            tnext2 = tn3;
            continue;
        }
        // lftog: res = sum lf_im + res_r,  res_r = (1-beta) res_r + beta sum lf_re
        // (tri_hp/nstage.cpp:21-82).  The first colour that touches a coefficient
        // starts both sums from keep * res_r; interiors belong to one element.
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            const int k = 4 * s + q, kind = mode_kind<P>(k);
            if (S.t < 0 || k >= D::TM) continue;
            const int off = S.off[s] < 0 ? ~S.off[s] : S.off[s];
            const T sg = S.off[s] < 0 ? T(-1) : T(1);
            T* di = kind == 0 ? a.resv : (kind == 1 ? a.ress : a.resi);
            T* dr = kind == 0 ? a.rrv : (kind == 1 ? a.rrs : a.rri);
            const T li[3] = {sg * LI0[s], sg * LI1[s], sg * LI2[s]};
            T lr[3] = {T(0), T(0), T(0)};
            if constexpr (FULL) {
                const T bt = a.beta;
                lr[0] = sg * (LR0[s] * bt);
                lr[1] = sg * (LR1[s] * bt);
                lr[2] = sg * (LR2[s] * bt);
            }
            if ((S.first >> s) & 1) {
                // FP32 inner solve: the constant source sfac * R64(u0) enters once, on
                // the first visit (null in the FP64 kernel)
                const double* sr = kind == 0 ? a.srcv : (kind == 1 ? a.srcs : a.srci);
#pragma unroll
                for (int n = 0; n < NV; ++n) {
                    const T r0 = a.keep == T(0) ? T(0) : a.keep * ro[n][s];
                    if (FULL || a.keep != T(1)) dr[off + n] = r0 + lr[n];
                    if (sizeof(T) == 4 && sr)
                        di[off + n] = r0 + lr[n] + li[n] + T(sr[off + n] * a.sfac[0]);
                    else
                        di[off + n] = r0 + lr[n] + li[n];
                }
            } else {
#pragma unroll
                for (int n = 0; n < NV; ++n) {
                    if constexpr (FULL) dr[off + n] = ro[n][s] + lr[n];
                    di[off + n] = io[n][s] + lr[n] + li[n];
                }
            }
        }
        S = Sn;
        cnext = cn2;
        tnext2 = tn3;
    }
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_ins(ElemArgs a) {
    elem_ins_body<double, P, FULL>(a);
}

// FP32 instantiation (the inner solve of the FP32 defect correction): the same
// residual in single precision on v_mfma_f32_16x16x4f32; its operator pack permutes
// the output rows of Ir/Is so that register s of lane group q again holds mode 4s + q
template <int P, bool FULL>
__global__ __launch_bounds__(256, 2) void k_elem_ins32(ElemArgs32 a) {
    elem_ins_body<float, P, FULL>(a);
}

// ---------------------------------------------------------------------------------
// Mode-array helpers over v(nent, mstride, NV) with the first nm modes active.
// op 0: x = 0, y *= c           (rsdl start, tri_hp/nstage.cpp:21-49)
// op 1: x += y                  (rsdl end, nstage.cpp:76-82; mg_source later visits)
// op 2: y = c*z - x, x += y     (mg_source first visit, nstage.cpp:113-128)
// op 3: y = x                   (mg_restrict, movco.cpp:18-37)
// op 4: x = y - c*z             (update, nstage.cpp:387-409)
template <int NVn>
__global__ void k_modes(int op, long nent, int nm, int mstride, double* __restrict__ x, double* __restrict__ y,
                        const double* __restrict__ zz, double c) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long span = (long)nm * NVn;
    if (k >= nent * span) return;
    const long ent = k / span, r = k - ent * span;
    const long ix = ent * (long)mstride * NVn + r;
    switch (op) {
        case 0: x[ix] = 0.0; y[ix] *= c; break;
        case 1: x[ix] += y[ix]; break;
        case 2: { const double d = c * zz[ix] - x[ix]; y[ix] = d; x[ix] += d; } break;
        case 3: y[ix] = x[ix]; break;
        case 4: x[ix] = y[ix] - c * zz[ix]; break;
    }
}

// edge-boundary flux residual, tri_hp/hp_boundary.cpp:973-1049 with bdry_ins::generic::flux
// S: storage type of the state and residual (double, or float for the FP32 inner
// solve of the defect correction); the flux arithmetic is FP64 in both
template <typename S>
struct EdgeArgsT {
    const int* eseg;
    int begin, end;
    const int* seg_pnt;
    const double* pnts;
    const S* ugv;
    const S* ugs;
    S* resv;
    S* ress;
    const double* g1;   // gpx x (sm+2)
    const double* dg1;
    const double* w1;
    int gpx, sm, sm0;
    double rho, bc_pressure;
    // curved boundary sides (may be null): row of each edge idx in scrd, -1 if straight;
    // scrd rows hold the side's coordinate modes (sm0 modes, x/y) in the side's direction
    const int* ecrv;
    const double* scrd;
    // flux type of each edge idx (may be null: all generic), the freestream velocity of
    // the euler and characteristic fluxes (the freestream pressure is bc_pressure), mu
    // for the characteristic's gam
    const unsigned char* etype;
    double ub0, ub1, mu;
};
typedef EdgeArgsT<double> EdgeArgs;

enum EdgeFlux : unsigned char { EDGE_GENERIC = 0, EDGE_EULER = 1, EDGE_CHARACTERISTIC = 2 };

// boundary flux of tri_hp_ins at one Gauss point (steady, mesh velocity zero), unit
// normal (nx, ny), jcb the point's length factor |dx/dxi|:
//   generic (inflow, outflow, symmetry), bdry_ins.h:34-49: mass flux with the interior
//     velocity, the freestream pressure pb on the momentum;
//   euler, bdry_ins.h:96-113: mass flux and momentum with the freestream velocity, the
//     interior pressure;
//   characteristic, tri_hp/ins/bdry.cpp:245-320: the boundary state from the
//     characteristic swap between the interior and the freestream (its natural fluxes
//     zero), then the euler-form flux of that state.
template <int P>
__device__ __forceinline__ void edge_flux_ins(int type, const double* u, double nx, double ny, double jcb,
                                              double rho, double mu, double ub0, double ub1, double pb,
                                              double* fl) {
    if (type == EDGE_GENERIC) {
        const double f2 = rho * (u[0] * nx + u[1] * ny);
        fl[0] = f2 * u[0] + pb * nx;
        fl[1] = f2 * u[1] + pb * ny;
        fl[2] = f2;
        return;
    }
    double b0 = ub0, b1 = ub1, b2 = u[2];
    if (type == EDGE_CHARACTERISTIC) {
        const double rhoi = 1. / rho, nu = mu / rho;
        const double hmax = jcb * 2.0 / (0.25 * (P + 1) * (P + 1));
        const double qmax = u[0] * u[0] + u[1] * u[1];
        const double gv = 2. * nu / hmax;
        const double gam = 3.0 * qmax + gv * gv;
        const double ul = u[0] * nx + u[1] * ny, vl = -u[0] * ny + u[1] * nx, pl = u[2];
        const double ur = ub0 * nx + ub1 * ny, vr = -ub0 * ny + ub1 * nx, pr = pb;
        const double cl = sqrt(ul * ul + gam), cr = sqrt(ur * ur + gam);
        const double c = 0.5 * (cl + cr), sn = 0.5 * (ul + ur), vt = 0.5 * (vl + vr);
        const double den = 1. / (2 * c);
        const double lam0 = sn, lam1 = sn + c, lam2 = sn - c;
        const double w0 = ((pl - pr) * rhoi + (ul * lam1 - ur * lam2)) * den;
        const double w1 = lam0 > 0.0 ? vt * ((pr - pl) * rhoi + lam2 * (ur - ul)) * den / (lam0 - lam2) + vl
                                     : vt * ((pr - pl) * rhoi + lam1 * (ur - ul)) * den / (lam0 - lam1) + vr;
        const double w2 = (rho * (ul - ur) * gam - lam2 * pl + lam1 * pr) * den;
        b0 = w0 * nx - w1 * ny;
        b1 = w0 * ny + w1 * nx;
        b2 = w2;
    }
    const double f2 = rho * (b0 * nx + b1 * ny);
    fl[0] = f2 * b0 + b2 * nx;
    fl[1] = f2 * b1 + b2 * ny;
    fl[2] = f2;
}

// P fixes the edge's mode count (p + 1) and Gauss points (p + 1), so the per-lane
// arrays stay in registers
// EM: element-major output eout[(idx*2 + k)*NV + n] for endpoint k of boundary edge
// idx (p = 1, summed into the vertices by k_vstage1) instead of the RMW scatter
template <int P, typename S, bool EM>
__device__ __forceinline__ void edge_ins_body(const EdgeArgsT<S>& a, S* __restrict__ eout, int idx) {
    constexpr int SM = P - 1, N1 = SM + 2, GPX = P + 1;
    if (idx >= a.end) return;
    const int sind = a.eseg[idx];
    const int v0 = a.seg_pnt[2 * sind], v1 = a.seg_pnt[2 * sind + 1];
    double uh[NV][N1];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        uh[n][0] = a.ugv[v0 * NV + n];
        uh[n][1] = a.ugv[v1 * NV + n];
#pragma unroll
        for (int m = 0; m < SM; ++m) uh[n][m + 2] = a.ugs[(sind * a.sm0 + m) * NV + n];
    }
    const double X0 = a.pnts[2 * v0], Y0 = a.pnts[2 * v0 + 1], X1 = a.pnts[2 * v1], Y1 = a.pnts[2 * v1 + 1];
    // crdtocht1d (tri_hp/gtol.cpp:296-320): the side modes of a curved side's coordinates
    double cx[SM > 0 ? SM : 1], cy[SM > 0 ? SM : 1];
    const int crow = a.ecrv ? a.ecrv[idx] : -1;
    const int etype = a.etype ? a.etype[idx] : EDGE_GENERIC;
#pragma unroll
    for (int m = 0; m < SM; ++m) {
        cx[m] = crow >= 0 ? a.scrd[((size_t)crow * a.sm0 + m) * 2] : 0.0;
        cy[m] = crow >= 0 ? a.scrd[((size_t)crow * a.sm0 + m) * 2 + 1] : 0.0;
    }
    double lf[NV][N1];
#pragma unroll
    for (int n = 0; n < NV; ++n)
#pragma unroll
        for (int k = 0; k < N1; ++k) lf[n][k] = 0.0;
#pragma unroll
    for (int ip = 0; ip < GPX; ++ip) {
        const double* g = a.g1 + ip * N1;
        const double* dg = a.dg1 + ip * N1;
        double dxc = X0 * dg[0] + X1 * dg[1];
        double dyc = Y0 * dg[0] + Y1 * dg[1];
        if (crow >= 0) {
#pragma unroll
            for (int m = 0; m < SM; ++m) {
                dxc += cx[m] * dg[m + 2];
                dyc += cy[m] * dg[m + 2];
            }
        }
        double u[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N1; ++k) s += uh[n][k] * g[k];
            u[n] = s;
        }
        double nx = dyc, ny = -dxc;
        const double jcb = sqrt(nx * nx + ny * ny);
        nx /= jcb; ny /= jcb;
        double fl[NV];
        edge_flux_ins<P>(etype, u, nx, ny, jcb, a.rho, a.mu, a.ub0, a.ub1, a.bc_pressure, fl);
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const double r = fl[n] * jcb * a.w1[ip];
#pragma unroll
            for (int k = 0; k < N1; ++k) lf[n][k] += g[k] * r;
        }
    }
    if constexpr (EM) {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            eout[(2 * idx) * NV + n] = (S)lf[n][0];
            eout[(2 * idx + 1) * NV + n] = (S)lf[n][1];
        }
        return;
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        a.resv[v0 * NV + n] += (S)lf[n][0];
        a.resv[v1 * NV + n] += (S)lf[n][1];
#pragma unroll
        for (int m = 0; m < SM; ++m) a.ress[(sind * a.sm0 + m) * NV + n] += (S)lf[n][m + 2];
    }
}

template <int P, typename S = double>
__global__ __launch_bounds__(64) void k_edge_ins(EdgeArgsT<S> a) {
    edge_ins_body<P, S, false>(a, nullptr, a.begin + blockIdx.x * blockDim.x + threadIdx.x);
}

// Element-major residual (no colouring, mesh order): the element body with its EM
// output, persistent over all elements; k_gather_vs then sums the vertex and side
// modes in a fixed order.  At p = 2 the bound of 256 registers keeps two waves per
// SIMD without spilling (272 otherwise, one wave).
template <int P, bool FULL>
__global__ __launch_bounds__(256, P == 2 ? 2 : 1) void k_elem_ins_em(ElemArgs a, double* eres) {
    elem_ins_body<double, P, FULL, true>(a, eres);
}

// the FP32 inner solve's element residual in element-major form (as k_elem_ins32)
template <int P, bool FULL>
__global__ __launch_bounds__(256, 2) void k_elem_ins32_em(ElemArgs32 a, float* eres) {
    elem_ins_body<float, P, FULL, true>(a, eres);
}

// the curved elements (order[begin:end), a.crv per entry) in element-major form: a
// launch of its own after the straight elements' launch, over the few elements that
// own a curved boundary side
template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_ins_em_curved(ElemArgs a, double* eres) {
    elem_ins_body<double, P, FULL, true, true>(a, eres);
}

template <int P, bool FULL>
__global__ __launch_bounds__(256) void k_elem_ins32_em_curved(ElemArgs32 a, float* eres) {
    elem_ins_body<float, P, FULL, true, true>(a, eres);
}

// lftog of the element-major vertex and side modes (tri_hp/nstage.cpp:21-82): one lane
// per (vertex, variable), then one per (side, mode, variable).  Each sum starts from
// keep * res_r and adds the elements' (beta lf_re, lf_im) in a fixed order (ascending
// element at a vertex, the side's two elements in ascending order), as the first
// colour and the later colours of the scatter do.  FP32 (the defect correction's inner
// solve): the constant source sfac[0] * src enters with the first element, as on the
// scatter's first visit.
template <typename T>
struct GatherArgsT {
    int npnt, nseg, sm, sm0, bm;
    const int* vt_ptr;
    const int* vt_tri;
    const int* vt_loc;
    const int* st_tri;
    const int* st_loc;
    const T* eres;
    T* resv;
    T* rrv;
    T* ress;
    T* rrs;
    T keep;
    int full;
    const double* srcv;  // FP32 only, may be null
    const double* srcs;
    const double* sfac;
};

template <typename T, int NVn>
__device__ __forceinline__ void gather_sum(const GatherArgsT<T>& a, const T* e, int n, T& acc, T& accr) {
    const T li = e[n];
    if (a.full) {
        const T lr = e[NVn + n];
        accr += lr;
        acc = (acc + lr) + li;
    } else {
        acc = acc + li;
    }
}

template <typename T, int NVn>
__device__ __forceinline__ void gather_sum_neg(const GatherArgsT<T>& a, const T* e, int n, T& acc, T& accr) {
    const T li = -e[n];
    if (a.full) {
        const T lr = -e[NVn + n];
        accr += lr;
        acc = (acc + lr) + li;
    } else {
        acc = acc + li;
    }
}

template <typename T, int NVn>
__global__ __launch_bounds__(256) void k_gather_vs(GatherArgsT<T> a) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nvl = (long)a.npnt * NVn;
    const T keep = a.keep;
    if (g < nvl) {
        const int v = (int)(g / NVn), n = (int)(g - (long)v * NVn);
        const T r0 = keep == T(0) ? T(0) : keep * a.rrv[g];
        T acc = r0, accr = r0;
        for (int k = a.vt_ptr[v]; k < a.vt_ptr[v + 1]; ++k) {
            gather_sum<T, NVn>(a, a.eres + ((long)a.vt_tri[k] * a.bm + a.vt_loc[k]) * (2 * NVn), n, acc, accr);
            if (k == a.vt_ptr[v] && a.srcv) acc = acc + T(a.srcv[g] * a.sfac[0]);
        }
        if (a.full || keep != T(1)) a.rrv[g] = accr;
        a.resv[g] = acc;
        return;
    }
    const long h = g - nvl;
    const long per = (long)a.sm * NVn;
    if (h >= (long)a.nseg * per) return;
    const int s = (int)(h / per);
    const int r = (int)(h - (long)s * per);
    const int m = r / NVn, n = r - m * NVn;
    const long o = ((long)s * a.sm0 + m) * NVn + n;
    const T r0 = keep == T(0) ? T(0) : keep * a.rrs[o];
    T acc = r0, accr = r0;
    bool first = true;
    for (int slot = 0; slot < 2; ++slot) {
        const int t = a.st_tri[2 * s + slot];
        if (t < 0) continue;
        // st_loc bit 2: the element is a periodic partner's, whose side runs the other
        // way: its odd modes change sign in this side's frame
        const int sl = a.st_loc[2 * s + slot];
        const int mode = 3 + (sl & 3) * a.sm + m;
        if ((sl & 4) && (m & 1))
            gather_sum_neg<T, NVn>(a, a.eres + ((long)t * a.bm + mode) * (2 * NVn), n, acc, accr);
        else
            gather_sum<T, NVn>(a, a.eres + ((long)t * a.bm + mode) * (2 * NVn), n, acc, accr);
        if (first && a.srcs) acc = acc + T(a.srcs[o] * a.sfac[0]);
        first = false;
    }
    if (a.full || keep != T(1)) a.rrs[o] = accr;
    a.ress[o] = acc;
}

// p = 1 element residual on the vector ALU, one thread per element (mesh order), the
// element-major output of k_elem_ins_em (at p = 1, eres[(t*3 + k)*2*NV + n]).  At p = 1
// the 16 x 16 x 4 MFMA tiles are
// mostly padding (4 Gauss points, 3 modes: 80 matrix instructions per 16 elements for
// about 1.2 kflop of work per element), so the residual is written out as scalar
// loops over the dense operators (uniform across the wave: scalar loads).  The same
// sequence as elem_ins_body: proj, inviscid fluxes, intgrtrs (tri_hp/ins/rsdl.cpp:
// 262-294); derivatives, viscous tensor, strong residual, tau and stabilisation,
// intgrtrs of the real part (:296-426).
// Dense p = 1 operators: P, Pr, Ps [4][3] | Ir, Is [3][4] | Dr, Ds [4][4]
constexpr int OP1_P = 0, OP1_PR = 12, OP1_PS = 24, OP1_IR = 36, OP1_IS = 48, OP1_DR = 60, OP1_DS = 76, OP1_SIZE = 92;

template <bool FULL>
__global__ __launch_bounds__(256) void k_elem_ins1_valu(ElemArgs a, const double* __restrict__ op,
                                                        double* __restrict__ eres, EdgeArgs e,
                                                        double* __restrict__ eout, int egrid) {
    if ((int)blockIdx.x >= egrid) {
        edge_ins_body<1, double, true>(e, eout, e.begin + ((int)blockIdx.x - egrid) * (int)blockDim.x + (int)threadIdx.x);
        return;
    }
    const int t = a.begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= a.end) return;
    constexpr int G = 4, M = 3;
    int vtx[3];
    double x[3], y[3], b[NV][M];
#pragma unroll
    for (int k = 0; k < 3; ++k) vtx[k] = a.tri_pnt[3 * t + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = a.pnts[2 * vtx[k]];
        y[k] = a.pnts[2 * vtx[k] + 1];
#pragma unroll
        for (int n = 0; n < NV; ++n) b[n][k] = a.ugv[vtx[k] * NV + n];
    }
    const double l00 = 0.5 * (x[2] - x[1]), l01 = 0.5 * (x[0] - x[1]);
    const double l10 = 0.5 * (y[2] - y[1]), l11 = 0.5 * (y[0] - y[1]);
    double U[G], V[G], Pp[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        double u = 0.0, v = 0.0, pr = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const double w = op[OP1_P + g * M + m];
            u += w * b[0][m]; v += w * b[1][m]; pr += w * b[2][m];
        }
        U[g] = u; V[g] = v; Pp[g] = pr;
    }
    const double rho = a.rho;
    double C0[G], C1[G], A00[G], A01[G], A10[G], A11[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const double fx = rho * U[g], fy = rho * V[g];
        const double c0 = +l11 * fx - l01 * fy;
        const double c1 = -l10 * fx + l00 * fy;
        C0[g] = c0; C1[g] = c1;
        A00[g] = U[g] * c0 + l11 * Pp[g];
        A01[g] = U[g] * c1 - l10 * Pp[g];
        A10[g] = V[g] * c0 - l01 * Pp[g];
        A11[g] = V[g] * c1 + l00 * Pp[g];
    }
    double LI[NV][M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double ar = op[OP1_IR + m * G + g], as = op[OP1_IS + m * G + g];
            s0 += ar * A00[g] + as * A01[g];
            s1 += ar * A10[g] + as * A11[g];
            s2 += ar * C0[g] + as * C1[g];
        }
        LI[0][m] = s0; LI[1][m] = s1; LI[2][m] = s2;
    }
    double* o = eres + (size_t)t * M * (2 * NV);
    if constexpr (FULL) {
        double Ur[G], Us[G], Vr[G], Vs[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double c1 = 0.0, c2 = 0.0, c3 = 0.0, c4 = 0.0;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double ar = op[OP1_PR + g * M + m], as = op[OP1_PS + g * M + m];
                c1 += ar * b[0][m]; c2 += as * b[0][m];
                c3 += ar * b[1][m]; c4 += as * b[1][m];
            }
            Ur[g] = c1; Us[g] = c2; Vr[g] = c3; Vs[g] = c4;
        }
        const double cjcb = l00 * l11 - l10 * l01;
        const double cjcbi = a.mu / cjcb;
        const double v0000 = -cjcbi * (2.0 * l11 * l11 + l01 * l01);
        const double v0011 = -cjcbi * (2.0 * l10 * l10 + l00 * l00);
        const double v0001 = cjcbi * (2.0 * l11 * l10 + l01 * l00);
        const double v1100 = -cjcbi * (l11 * l11 + 2.0 * l01 * l01);
        const double v1111 = -cjcbi * (l10 * l10 + 2.0 * l00 * l00);
        const double v1101 = cjcbi * (l11 * l10 + 2.0 * l01 * l00);
        const double v0100 = cjcbi * l01 * l11;
        const double v0111 = cjcbi * l00 * l10;
        const double v0101 = -cjcbi * l01 * l10;
        const double v0110 = -cjcbi * l00 * l11;
        double F00[G], F01[G], F10[G], F11[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double ur = Ur[g], us = Us[g], vr = Vr[g], vs = Vs[g];
            F00[g] = v0000 * ur + v0100 * vr + v0001 * us + v0101 * vs;
            F01[g] = v0001 * ur + v0110 * vr + v0011 * us + v0111 * vs;
            F10[g] = v0100 * ur + v1100 * vr + v0110 * us + v1101 * vs;
            F11[g] = v0101 * ur + v1101 * vr + v0111 * us + v1111 * vs;
            A00[g] += F00[g]; A01[g] += F01[g]; A10[g] += F10[g]; A11[g] += F11[g];
        }
        double R0[G], R1[G], R2[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double r0 = 0.0, r1 = 0.0, r2 = 0.0;
#pragma unroll
            for (int h = 0; h < G; ++h) {
                const double dr = op[OP1_DR + g * G + h], ds = op[OP1_DS + g * G + h];
                r0 += dr * A00[h] + ds * A01[h];
                r1 += dr * A10[h] + ds * A11[h];
                r2 += dr * C0[h] + ds * C1[h];
            }
            R0[g] = r0; R1[g] = r1; R2[g] = r2;
        }
        const double dx1 = x[0] - x[2], dy1 = y[0] - y[2], dx2 = x[1] - x[0], dy2 = y[1] - y[0];
        const double ar2 = dx1 * dy2 - dy1 * dx2;
        const double perim = sqrt(dx1 * dx1 + dy1 * dy1) + sqrt(dx2 * dx2 + dy2 * dy2) +
                             sqrt((dx1 + dx2) * (dx1 + dx2) + (dy1 + dy2) * (dy1 + dy2));
        const double h = (ar2 / perim) / (0.25 * 2 * 2);
        const double nu = a.mu / rho;
        const double gv = 2.0 * nu / h;
        const double gvis = gv * gv;
        const double tfac = a.adis * h / cjcb;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double u = U[g], v = V[g];
            const double qq = u * u + v * v;
            double gam = 3.0 * qq + gvis;
            if (a.mu == 0.0) gam = fmax(gam, 0.1);
            const double tau0 = tfac / sqrt(gam);
            const double tau2 = sqrt(qq) * tau0;
            const double t0 = tau0 * R0[g], t1 = tau0 * R1[g], t2 = tau2 * R2[g];
            F00[g] -= (l11 * (2.0 * u) - l01 * v) * t0 - l01 * u * t1 + l11 * t2;
            F01[g] -= (-l10 * (2.0 * u) + l00 * v) * t0 + l00 * u * t1 - l10 * t2;
            F10[g] -= +l11 * v * t0 + (l11 * u - l01 * (2.0 * v)) * t1 - l01 * t2;
            F11[g] -= -l10 * v * t0 + (-l10 * u + l00 * (2.0 * v)) * t1 + l00 * t2;
            C0[g] = -(l11 * t0 - l01 * t1);
            C1[g] = -(-l10 * t0 + l00 * t1);
        }
        const double bt = a.beta;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const double ar = op[OP1_IR + m * G + g], as = op[OP1_IS + m * G + g];
                s0 += ar * F00[g] + as * F01[g];
                s1 += ar * F10[g] + as * F11[g];
                s2 += ar * C0[g] + as * C1[g];
            }
            double* om = o + m * (2 * NV);
            om[0] = LI[0][m]; om[1] = LI[1][m]; om[2] = LI[2][m];
            om[NV + 0] = s0 * bt; om[NV + 1] = s1 * bt; om[NV + 2] = s2 * bt;
        }
    } else {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double* om = o + m * (2 * NV);
            om[0] = LI[0][m]; om[1] = LI[1][m]; om[2] = LI[2][m];
        }
    }
}

// ---------------------------------------------------------------------------------
// XCD-aware block index: launches use a multiple of 8 blocks; consecutive hardware
// blocks go round-robin to the 8 XCDs, so XCD x is given the contiguous logical
// range [x * per, (x + 1) * per) and neighbouring entities share one L2
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// setup_preconditioner, element part (tri_hp/ins/setup_preconditioner.cpp:33-219,
// diagonal branch, REFINED_WAY, straight elements, steady)
struct SetupArgs {
    const double* pnts;
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const double* ugv;
    const double* ugs;
    const double* ugi;
    const double* Pmat;  // G x tm
    double* tprcn;
    int* err;
    int ntri, p, sm, im, tm, G, sm0, im0;
    double rho, mu, adis;
    // curved elements (k_setup_elem_curved): list, coordinate side modes per entry
    // (the layout of ElemArgsT::crv), dense d/dr and d/ds projections (G x tm)
    const int* clist;
    int ncurv;
    const double* crv;
    const double* PRmat;
    const double* PSmat;
};

template <int P>
__global__ __launch_bounds__(128) void k_setup_elem(SetupArgs a) {
    constexpr int TM = Dims<P>::TM, SM = Dims<P>::SM, G = Dims<P>::G;
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= a.ntri) return;
    int vtx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vtx[k] = a.tri_pnt[3 * t + k];
    double uh[2][TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        int off;
        double sg = 1.0;
        const double* src;
        if (k < 3) { off = vtx[k] * NV; src = a.ugv; }
        else if (k < 3 * P) {
            const int i = (k - 3) / SM, m = (k - 3) - i * SM;
            const int sgn = a.tri_sgn[3 * t + i];
            sg = ((m & 1) && sgn < 0) ? -1.0 : 1.0;
            off = (a.tri_seg[3 * t + i] * a.sm0 + m) * NV;
            src = a.ugs;
        } else { off = (t * a.im0 + (k - 3 * P)) * NV; src = a.ugi; }
        uh[0][k] = sg * src[off];
        uh[1][k] = sg * src[off + 1];
    }
    double qmax = 0.0;
    for (int g = 0; g < G; ++g) {
        double u = 0.0, v = 0.0;
#pragma unroll
        for (int k = 0; k < TM; ++k) {
            const double pv = a.Pmat[g * TM + k];
            u += pv * uh[0][k];
            v += pv * uh[1][k];
        }
        const double q = u * u + v * v;
        qmax = (qmax > q) ? qmax : q;  // MAX of sys/param.h, as the reference
    }
    const double X[3] = {a.pnts[2 * vtx[0]], a.pnts[2 * vtx[1]], a.pnts[2 * vtx[2]]};
    const double Y[3] = {a.pnts[2 * vtx[0] + 1], a.pnts[2 * vtx[1] + 1], a.pnts[2 * vtx[2] + 1]};
    const double dx1 = X[0] - X[2], dy1 = Y[0] - Y[2], dx2 = X[1] - X[0], dy2 = Y[1] - Y[0];
    double jcb = 0.25 * (dx1 * dy2 - dy1 * dx2);
    double hmax = 0.0;
    for (int j = 0; j < 3; ++j) {
        const int j1 = (j + 1) % 3;
        const double hh = (X[j] - X[j1]) * (X[j] - X[j1]) + (Y[j] - Y[j1]) * (Y[j] - Y[j1]);
        hmax = hh > hmax ? hh : hmax;
    }
    hmax = sqrt(hmax);
    const double pf = 0.25 * (a.p + 1) * (a.p + 1);
    const double h = 4. * jcb / (pf * hmax);
    hmax = hmax / pf;
    if (!(h > 0.0) || qmax != qmax) {
        atomicOr(a.err, 1);
        a.tprcn[3 * t + 0] = a.tprcn[3 * t + 1] = a.tprcn[3 * t + 2] = 1.0;
        return;
    }
    const double nu = a.mu / a.rho;
    double gam = 3.0 * qmax + (2. * nu / hmax) * (2. * nu / hmax);
    if (a.mu == 0.0) gam = fmax(gam, 0.1);
    const double qs = sqrt(qmax);
    const double lam1 = qs + sqrt(qmax + gam);
    jcb *= 2. * nu * (1. / (hmax * hmax) + 1. / (h * h)) + 3 * lam1 / h;
    a.tprcn[3 * t + 0] = a.rho * jcb;
    a.tprcn[3 * t + 1] = a.rho * jcb;
    a.tprcn[3 * t + 2] = jcb / gam;
}

// the curved branch of the same (tri_hp/ins/setup_preconditioner.cpp:44-90) for the
// elements clist[0:ncurv), after k_setup_elem: h from the smallest Jacobian over the
// Gauss points and hmax from the longest projected side tangent, both through the
// coordinate derivatives at every point; tprcn overwritten
template <int P>
__global__ __launch_bounds__(64) void k_setup_elem_curved(SetupArgs a) {
    constexpr int TM = Dims<P>::TM, SM = Dims<P>::SM, BM = Dims<P>::BM, G = Dims<P>::G;
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= a.ncurv) return;
    const int t = a.clist[ci];
    int vtx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vtx[k] = a.tri_pnt[3 * t + k];
    double uh[2][TM], cx[BM], cy[BM];
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        int off;
        double sg = 1.0;
        const double* src;
        if (k < 3) { off = vtx[k] * NV; src = a.ugv; }
        else if (k < BM) {
            const int i = (k - 3) / SM, m = (k - 3) - i * SM;
            const int sgn = a.tri_sgn[3 * t + i];
            sg = ((m & 1) && sgn < 0) ? -1.0 : 1.0;
            off = (a.tri_seg[3 * t + i] * a.sm0 + m) * NV;
            src = a.ugs;
        } else { off = (t * a.im0 + (k - BM)) * NV; src = a.ugi; }
        uh[0][k] = sg * src[off];
        uh[1][k] = sg * src[off + 1];
    }
    const double X[3] = {a.pnts[2 * vtx[0]], a.pnts[2 * vtx[1]], a.pnts[2 * vtx[2]]};
    const double Y[3] = {a.pnts[2 * vtx[0] + 1], a.pnts[2 * vtx[1] + 1], a.pnts[2 * vtx[2] + 1]};
    // coordinate modes relative to vertex 1 (crdtocht, tri_hp/gtol.cpp:204-245)
    cx[0] = X[0] - X[1]; cy[0] = Y[0] - Y[1];
    cx[1] = 0.0; cy[1] = 0.0;
    cx[2] = X[2] - X[1]; cy[2] = Y[2] - Y[1];
#pragma unroll
    for (int k = 3; k < BM; ++k) {
        const int i = (k - 3) / SM, m = (k - 3) - i * SM;
        cx[k] = a.crv[(((size_t)ci * 3 + i) * a.sm0 + m) * 2];
        cy[k] = a.crv[(((size_t)ci * 3 + i) * a.sm0 + m) * 2 + 1];
    }
    const double dx1 = X[0] - X[2], dy1 = Y[0] - Y[2], dx2 = X[1] - X[0], dy2 = Y[1] - Y[0];
    double jcb = 0.25 * (dx1 * dy2 - dy1 * dx2);
    double jcbmin = jcb, hmax = 0.0, qmax = 0.0;
    for (int g = 0; g < G; ++g) {
        double u = 0.0, v = 0.0, xr = 0.0, xs = 0.0, yr = 0.0, ys = 0.0;
#pragma unroll
        for (int k = 0; k < TM; ++k) {
            const double pv = a.Pmat[g * TM + k];
            u += pv * uh[0][k];
            v += pv * uh[1][k];
        }
#pragma unroll
        for (int k = 0; k < BM; ++k) {
            const double pr = a.PRmat[g * TM + k], ps = a.PSmat[g * TM + k];
            xr += pr * cx[k]; xs += ps * cx[k];
            yr += pr * cy[k]; ys += ps * cy[k];
        }
        jcbmin = fmin(jcbmin, xr * ys - yr * xs);
        double hh = xr * xr + yr * yr;
        hmax = hh > hmax ? hh : hmax;
        hh = xs * xs + ys * ys;
        hmax = hh > hmax ? hh : hmax;
        hh = (xs - xr) * (xs - xr) + (ys - yr) * (ys - yr);
        hmax = hh > hmax ? hh : hmax;
        const double q = u * u + v * v;
        qmax = (qmax > q) ? qmax : q;
    }
    hmax = 2. * sqrt(hmax);
    const double pf = 0.25 * (a.p + 1) * (a.p + 1);
    const double h = 4. * jcbmin / (pf * hmax);
    hmax = hmax / pf;
    if (!(h > 0.0) || qmax != qmax) {
        atomicOr(a.err, 1);
        a.tprcn[3 * t + 0] = a.tprcn[3 * t + 1] = a.tprcn[3 * t + 2] = 1.0;
        return;
    }
    const double nu = a.mu / a.rho;
    double gam = 3.0 * qmax + (2. * nu / hmax) * (2. * nu / hmax);
    if (a.mu == 0.0) gam = fmax(gam, 0.1);
    const double qs = sqrt(qmax);
    const double lam1 = qs + sqrt(qmax + gam);
    jcb *= 2. * nu * (1. / (hmax * hmax) + 1. / (h * h)) + 3 * lam1 / h;
    a.tprcn[3 * t + 0] = a.rho * jcb;
    a.tprcn[3 * t + 1] = a.rho * jcb;
    a.tprcn[3 * t + 2] = jcb / gam;
}

// vprcn = 1/(vdiag * sum of incident tprcn) (setup_preconditioner.cpp:213-219 and
// tri_hp/setup_preconditioner.cpp:68-76); incident elements in ascending order
__global__ void k_setup_vrtx(int npnt, const int* ptr, const int* tri, const double* tprcn, double* vprcn, double vdiag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= npnt) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
        const int t = tri[k];
        s0 += tprcn[3 * t]; s1 += tprcn[3 * t + 1]; s2 += tprcn[3 * t + 2];
    }
    vprcn[3 * v] = 1.0 / (vdiag * s0);
    vprcn[3 * v + 1] = 1.0 / (vdiag * s1);
    vprcn[3 * v + 2] = 1.0 / (vdiag * s2);
}

__global__ void k_setup_side(int nseg, const int* st_tri, const double* tprcn, double* sprcn) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int k = 0; k < 2; ++k) {
        const int t = st_tri[2 * s + k];
        if (t < 0) continue;
        a0 += tprcn[3 * t]; a1 += tprcn[3 * t + 1]; a2 += tprcn[3 * t + 2];
    }
    sprcn[3 * s] = 1.0 / a0; sprcn[3 * s + 1] = 1.0 / a1; sprcn[3 * s + 2] = 1.0 / a2;
}

// ---------------------------------------------------------------------------------
// periodic boundaries.  The reference keeps a partial residual on each copy of a
// periodic vertex or side and averages the preconditioned values across the pair
// (tri_hp/setup_preconditioner.cpp:24-36, 171-176, 255-258, 331-334 with the comm of
// tri_mesh/tri_boundary.h:113-148), with the preconditioner sums averaged too, which
// equals assembling the pair as one set of unknowns.  Here every copy holds the
// assembled values: the vertex residual of each copy is summed over the group after
// the residual (k_periodic_combine), the side and vertex sweeps read the union of the
// copies' adjacency, and the state copies are kept equal (k_periodic_sync).
// Vertex groups: copies pg_v[pg_ptr[g]:pg_ptr[g+1]] (ascending); one lane per (group, n).
template <typename T, int NVn>
__global__ void k_periodic_combine(int ngrp, const int* __restrict__ ptr, const int* __restrict__ pv, T* res) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrp * NVn) return;
    const int grp = g / NVn, n = g - grp * NVn;
    T sum = T(0);
    for (int k = ptr[grp]; k < ptr[grp + 1]; ++k) sum += res[pv[k] * NVn + n];
    for (int k = ptr[grp]; k < ptr[grp + 1]; ++k) res[pv[k] * NVn + n] = sum;
}

// r += fl32(sfac[0] * src): the FP32 inner solve's constant source on the vertices of a
// periodic mesh, added after the copies are summed
__global__ void k_add_src32(long n, const double* __restrict__ src, const double* __restrict__ sfac, float* r) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = r[i] + float(src[i] * sfac[0]);
}

// state copies from the first copy of each vertex group and from side a of each pair
// (side b = o^m side a); one lane per (group, n), then per (pair, mode, n)
template <typename T, int NVn>
__global__ void k_periodic_sync(int ngrp, const int* __restrict__ ptr, const int* __restrict__ pv, int npair,
                                const int* __restrict__ sa, const int* __restrict__ sb, const int* __restrict__ so,
                                int sm, int sm0, T* v, T* sd) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngrp * NVn) {
        const int grp = g / NVn, n = g - grp * NVn;
        const T x = v[pv[ptr[grp]] * NVn + n];
        for (int k = ptr[grp] + 1; k < ptr[grp + 1]; ++k) v[pv[k] * NVn + n] = x;
        return;
    }
    const int h = g - ngrp * NVn;
    if (h >= npair * sm * NVn) return;
    const int pr = h / (sm * NVn), r = h - pr * (sm * NVn), m = r / NVn, n = r - m * NVn;
    const T x = sd[((long)sa[pr] * sm0 + m) * NVn + n];
    sd[((long)sb[pr] * sm0 + m) * NVn + n] = (so[pr] < 0 && (m & 1)) ? -x : x;
}

// ---------------------------------------------------------------------------------
// minvrt, tri_hp/setup_preconditioner.cpp:102-373 (diagonal preconditioner)
template <typename T>
struct MinvArgsT {
    const int* tri_pnt;
    const int* tri_seg;
    const int* tri_sgn;
    const int* vs_ptr;
    const int* vs_side;
    const int* vs_end;
    const int* vt_ptr;
    const int* vt_tri;
    const int* vt_loc;
    const int* st_tri;
    const int* st_loc;
    const unsigned char* vess;
    const unsigned char* sess;
    T* resv;
    T* ress;
    T* resi;
    const T* vprcn;
    const T* sprcn;
    // element scalings of the side rows (side assembly couplings) and of the interior
    // rows; both are tprcn for the analytic preconditioner, the learned one scales them
    // by mode type
    const T* tprcn_s;
    const T* tprcn_i;
    const T* lump;  // sfmv | ifmb | vfms | sdiag | sfms | bfmi | idiag
    int npnt, nseg, ntri, sm, im, bm, sm0, im0;
    int o_sfmv, o_ifmb, o_vfms, o_sdiag, o_sfms, o_bfmi, o_idiag;
    // fused stage update ug = ug0 - cfa * res (tri_hp/nstage.cpp:387-409) when ug != null
    T* ugv;
    T* ugs;
    T* ugi;
    const T* ug0v;
    const T* ug0s;
    const T* ug0i;
    T cfa;
};


// The minvrt sweeps run one lane per (entity, variable): the diagonal preconditioner
// never couples the variables, so lane (e, n) carries variable n of entity e through
// every subtraction in the same order as the per-entity loop.  Consecutive lanes then
// read consecutive doubles of the (entity, mode, variable) records, and the gathers of
// neighbouring records touch a third of the cache lines a lane-per-entity sweep does.
template <typename T, int NVn>
__device__ __forceinline__ bool minv_lane(int count, int& e, int& n) {
    const long g = (long)xcd_block() * blockDim.x + threadIdx.x;
    e = (int)(g / NVn);
    n = (int)(g - (long)e * NVn);
    return e < count;
}

// vertices: subtract side and interior couplings (:110-135), scale (:151-152),
// vertex Dirichlet (:179-183)
template <typename T, int P, int NVn>
__global__ __launch_bounds__(256) void k_minv_vrtx(MinvArgsT<T> a) {
    constexpr int SM = Dims<P>::SM, IM = Dims<P>::IM;
    int v, n;
    if (!minv_lane<T, NVn>(a.npnt, v, n)) return;
    T r = a.resv[v * NVn + n];
    if constexpr (SM > 0) {
        for (int k = a.vs_ptr[v]; k < a.vs_ptr[v + 1]; ++k) {
            const int s = a.vs_side[k], i = a.vs_end[k];
#pragma unroll
            for (int m = 0; m < SM; ++m) r -= a.lump[a.o_sfmv + i * SM + m] * a.ress[(s * a.sm0 + m) * NVn + n];
        }
        if constexpr (IM > 0)
            for (int k = a.vt_ptr[v]; k < a.vt_ptr[v + 1]; ++k) {
                const int t = a.vt_tri[k], i = a.vt_loc[k];
#pragma unroll
                for (int m = 0; m < IM; ++m) r -= a.lump[a.o_ifmb + i * IM + m] * a.resi[(t * a.im0 + m) * NVn + n];
            }
    }
    // every load before the first store: the stores could alias later loads
    const T x = (a.vess[v] >> n) & 1 ? T(0) : r * a.vprcn[v * NVn + n];
    const T u0 = a.ugv ? a.ug0v[v * NVn + n] : T(0);
    a.resv[v * NVn + n] = x;
    if (a.ugv) a.ugv[v * NVn + n] = u0 - a.cfa * x;
}

// p = 1 stage on the vertices, one lane per (vertex, variable): the lftog sums of the
// element-major element and edge results in a fixed order (tri_hp/nstage.cpp:21-82,
// hp_boundary.cpp:973-1049), the multigrid source (nstage.cpp:76-82, 113-128) and,
// with `ug` set, the vertex minvrt (setup_preconditioner.cpp:151-152, 179-183) and the
// stage update ug = ug0 - cfa * res (nstage.cpp:387-409).  At p = 1 every mode is a
// vertex mode, so this is the whole stage after the element residual.
struct VStageArgs {
    int npnt;
    const int* vt_ptr;
    const int* vt_tri;
    const int* vt_loc;
    const int* vb_ptr;  // boundary-edge endpoints at each vertex: entries 2*idx + k of eout
    const int* vb_ent;
    const double* eres;  // (ntri, 3, 2, NV)
    const double* eout;  // (neseg, 2, NV)
    double* resv;
    double* rrv;
    double keep;
    int full;           // the stage evaluates the real part (beta > 0)
    int mg;             // 0 none, 1 res += dres, 2 dres = fadd r0 - res, res += dres
    double* dres;
    const double* r0;
    double fadd;
    const unsigned char* vess;
    const double* vprcn;
    double* ug;         // null: residual only
    double* ug0;
    double cfa;
    int save_u0;        // first stage: ug0 = ug
};

template <int NVn>
__global__ __launch_bounds__(256) void k_vstage1(VStageArgs a) {
    int v, n;
    if (!minv_lane<double, NVn>(a.npnt, v, n)) return;
    const long o = (long)v * NVn + n;
    const double keep = a.keep;
    const double r0 = (keep == 0.0 || !(a.full || keep != 0.0)) ? 0.0 : keep * a.rrv[o];
    double acc = r0, accr = r0;
    for (int k = a.vt_ptr[v]; k < a.vt_ptr[v + 1]; ++k) {
        const double* e = a.eres + ((long)a.vt_tri[k] * 3 + a.vt_loc[k]) * (2 * NVn);
        const double li = e[n];
        if (a.full) {
            const double lr = e[NVn + n];
            accr += lr;
            acc = (acc + lr) + li;
        } else {
            acc = acc + li;
        }
    }
    for (int k = a.vb_ptr[v]; k < a.vb_ptr[v + 1]; ++k) acc += a.eout[(long)a.vb_ent[k] * NVn + n];
    if (a.mg == 1) {
        acc += a.dres[o];
    } else if (a.mg == 2) {
        const double d = a.fadd * a.r0[o] - acc;
        a.dres[o] = d;
        acc += d;
    }
    if (a.full || keep != 1.0) a.rrv[o] = accr;
    if (a.ug) {
        const double x = (a.vess[v] >> n) & 1 ? 0.0 : acc * a.vprcn[o];
        const double u = a.save_u0 ? a.ug[o] : a.ug0[o];
        if (a.save_u0) a.ug0[o] = u;
        a.resv[o] = x;
        a.ug[o] = u - a.cfa * x;
    } else {
        a.resv[o] = acc;
    }
}

// sides: subtract interior couplings (:137-148), then the vertex solution (:192-214)
template <typename T, int P, int NVn>
__global__ __launch_bounds__(256) void k_minv_side(MinvArgsT<T> a) {
    constexpr int SM = Dims<P>::SM, IM = Dims<P>::IM;
    int s, n;
    if (!minv_lane<T, NVn>(a.nseg, s, n)) return;
    T r[SM];
#pragma unroll
    for (int m = 0; m < SM; ++m) r[m] = a.ress[(s * a.sm0 + m) * NVn + n];
    if constexpr (IM > 0)
        for (int k = 0; k < 2; ++k) {
            const int t = a.st_tri[2 * s + k];
            if (t < 0) continue;
            // a periodic partner's element (st_loc bit 2) sees this side reversed
            const int sl = a.st_loc[2 * s + k], i = sl & 3;
            const int sgn = (sl & 4) ? -a.tri_sgn[3 * t + i] : a.tri_sgn[3 * t + i];
            T bi[IM];
#pragma unroll
            for (int m = 0; m < IM; ++m) bi[m] = a.resi[(t * a.im0 + m) * NVn + n];
            T ms = T(1);
#pragma unroll
            for (int j = 0; j < SM; ++j) {
#pragma unroll
                for (int m = 0; m < IM; ++m) {
                    const T c = ms * a.lump[a.o_ifmb + (3 + i * SM + j) * IM + m];
                    r[j] -= c * bi[m];
                }
                ms *= sgn;
            }
        }
    for (int k = 0; k < 2; ++k) {
        const int t = a.st_tri[2 * s + k];
        if (t < 0) continue;
        const int sl = a.st_loc[2 * s + k], i = sl & 3;
        const int sgn = (sl & 4) ? -a.tri_sgn[3 * t + i] : a.tri_sgn[3 * t + i];
        const T tp = a.tprcn_s[NVn * t + n];
        T uh[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) uh[j] = a.resv[a.tri_pnt[3 * t + (i + j) % 3] * NVn + n] * tp;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T ms = T(1);
#pragma unroll
            for (int m = 0; m < SM; ++m) {
                r[m] -= ms * a.lump[a.o_vfms + j * SM + m] * uh[j];
                ms *= sgn;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < SM; ++m) a.ress[(s * a.sm0 + m) * NVn + n] = r[m];
}

template <typename T, int NVn>
__device__ __forceinline__ T side_scaled(const MinvArgsT<T>& a, int s, int mode, int n) {
    if ((a.sess[s] >> n) & 1) return T(0);
    return a.ress[(s * a.sm0 + mode) * NVn + n] * (a.sprcn[s * NVn + n] * a.lump[a.o_sdiag + mode]);
}

// remove solved side mode `MODE` from the higher modes (:244-328); the solved value
// sprcn*sdiag*res with sdirichlet is formed on the fly, stored later by k_minv_sscale
template <typename T, int P, int MODE, int NVn>
__global__ __launch_bounds__(256) void k_minv_sub(MinvArgsT<T> a) {
    constexpr int SM = Dims<P>::SM;
    int s, n;
    if (!minv_lane<T, NVn>(a.nseg, s, n)) return;
    T r[SM];
#pragma unroll
    for (int m = MODE + 1; m < SM; ++m) r[m] = a.ress[(s * a.sm0 + m) * NVn + n];
    for (int k = 0; k < 2; ++k) {
        const int t = a.st_tri[2 * s + k];
        if (t < 0) continue;
        const int sl = a.st_loc[2 * s + k], i = sl & 3;
        const T tp = a.tprcn_s[NVn * t + n];
        T uh[3];
        int sign[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int jj = (i + j) % 3;
            const int sj = a.tri_seg[3 * t + jj];
            sign[j] = a.tri_sgn[3 * t + jj];
            const T sg = (MODE % 2) ? T(sign[j]) : T(1);
            uh[j] = sg * side_scaled<T, NVn>(a, sj, MODE, n) * tp;
        }
        // into this side's frame: a periodic partner's element (st_loc bit 2) has its
        // own copy of the side running the other way
        const int s0 = (sl & 4) ? -sign[0] : sign[0];
        T ms = ((MODE + 1) % 2) ? T(s0) : T(1);
#pragma unroll
        for (int m = MODE + 1; m < SM; ++m) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const T c = ms * a.lump[a.o_sfms + (MODE * SM + m) * 3 + j];
                r[m] -= c * uh[j];
            }
            ms *= s0;
        }
    }
#pragma unroll
    for (int m = MODE + 1; m < SM; ++m) a.ress[(s * a.sm0 + m) * NVn + n] = r[m];
}

// store the solved side modes: res.s *= sprcn*sdiag(mode), sdirichlet (:245-267, :331-343);
// one lane per (side, mode, variable), so on the finest level the sweep is a plain
// contiguous stream
template <typename T, int P, int NVn>
__global__ __launch_bounds__(256) void k_minv_sscale(MinvArgsT<T> a) {
    constexpr int SM = Dims<P>::SM;
    const long g = (long)xcd_block() * blockDim.x + threadIdx.x;
    const int s = (int)(g / (SM * NVn));
    if (s >= a.nseg) return;
    const int mn = (int)(g - (long)s * (SM * NVn)), m = mn / NVn, n = mn - m * NVn;
    const int ix = (s * a.sm0 + m) * NVn + n;
    // every load before the first store: the stores could alias later loads
    const T x = side_scaled<T, NVn>(a, s, m, n);
    const T u0 = a.ugs ? a.ug0s[ix] : T(0);
    a.ress[ix] = x;
    if (a.ugs) a.ugs[ix] = u0 - a.cfa * x;
}

// interiors: banded solve with the literal DPBTRSNU2 indexing (utilities/myblas.c:91-122),
// divide by tprcn, subtract the boundary solution through bfmi (:347-370)
template <typename T, int P, int NVn>
__global__ __launch_bounds__(256) void k_minv_int(MinvArgsT<T> a) {
    constexpr int SM = Dims<P>::SM, IM = Dims<P>::IM, BM = Dims<P>::BM;
    int t, n;
    if (!minv_lane<T, NVn>(a.ntri, t, n)) return;
    T b[IM];
#pragma unroll
    for (int k = 0; k < IM; ++k) b[k] = a.resi[(t * a.im0 + k) * NVn + n];
    const T* ab = a.lump + a.o_idiag;  // stride 2, ofdg 1
#pragma unroll
    for (int i = 0; i < IM; ++i) {
        const int frst = (1 - i > 0 ? 1 - i : 0), ba = -frst;
#pragma unroll
        for (int j = frst; j < 1; ++j) b[i] -= ab[i * 2 + j] * b[ba + j];
        b[i] /= ab[i * 2 + 1];
    }
#pragma unroll
    for (int i = IM - 1; i >= 0; --i) {
        b[i] /= ab[i * 2 + 1];
        const int frst = (1 - i > 0 ? 1 - i : 0), ba = -frst;
#pragma unroll
        for (int j = frst; j < 1; ++j) b[ba + j] -= ab[i * 2 + j] * b[i];
    }
    T uh[BM];
#pragma unroll
    for (int i = 0; i < 3; ++i) uh[i] = a.resv[a.tri_pnt[3 * t + i] * NVn + n];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int sd = a.tri_seg[3 * t + i], sgn = a.tri_sgn[3 * t + i];
        T ms = T(1);
#pragma unroll
        for (int m = 0; m < SM; ++m) {
            uh[3 + i * SM + m] = ms * a.ress[(sd * a.sm0 + m) * NVn + n];
            ms *= sgn;
        }
    }
    const T tp = a.tprcn_i[NVn * t + n];
#pragma unroll
    for (int k = 0; k < IM; ++k) {
        b[k] /= tp;
#pragma unroll
        for (int i = 0; i < BM; ++i) b[k] -= a.lump[a.o_bfmi + i * IM + k] * uh[i];
    }
    T u0[IM];
#pragma unroll
    for (int k = 0; k < IM; ++k) u0[k] = a.ugi ? a.ug0i[(t * a.im0 + k) * NVn + n] : T(0);
#pragma unroll
    for (int k = 0; k < IM; ++k) {
        const int ix = (t * a.im0 + k) * NVn + n;
        a.resi[ix] = b[k];
        if (a.ugi) a.ugi[ix] = u0[k] - a.cfa * b[k];
    }
}

// ---------------------------------------------------------------------------------
// max |x| over (nent, nm active modes of mstride, NV), per variable, into out[NV] as
// ordered bit patterns of non-negative doubles; mask (optional) skips essential entries
template <int NVn>
__global__ void k_absmax(long nent, int nm, int mstride, const double* __restrict__ x,
                         const unsigned char* __restrict__ mask, unsigned long long* out) {
    __shared__ unsigned long long red[NVn][256];
    unsigned long long mx[NVn] = {};
    const long span = (long)nm;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < nent * span; k += (long)gridDim.x * blockDim.x) {
        const long ent = k / span, m = k - ent * span;
        const unsigned msk = mask ? mask[ent] : 0u;
        for (int n = 0; n < NVn; ++n) {
            if ((msk >> n) & 1) continue;
            const double v = fabs(x[(ent * mstride + m) * NVn + n]);
            const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
            mx[n] = bits > mx[n] ? bits : mx[n];
        }
    }
    for (int n = 0; n < NVn; ++n) red[n][threadIdx.x] = mx[n];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int n = 0; n < NVn; ++n) {
                const unsigned long long o = red[n][threadIdx.x + st];
                red[n][threadIdx.x] = o > red[n][threadIdx.x] ? o : red[n][threadIdx.x];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int n = 0; n < NVn; ++n) atomicMax(&out[n], red[n][0]);
}

// ---------------------------------------------------------------------------------
// h-multigrid transfers between a mesh and its coarse neighbour (tri_hp/movco.cpp:41-102,
// tri_hp/movfin.cpp:16-82), vertex values only (coarse levels run at p = 1)

// res0(coarse v) = sum over the fine vertices held by triangles of v, in ascending fine
// order, of weight x fine residual (ccnnct transposed into a gather)
template <int NVn>
__global__ void k_hrestrict(int nc, const int* __restrict__ ptr, const int* __restrict__ fidx,
                            const double* __restrict__ wt, const double* __restrict__ fres, double* __restrict__ res0) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nc) return;
    double s[NVn];
#pragma unroll
    for (int n = 0; n < NVn; ++n) s[n] = 0.0;
    for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
        const int i = fidx[k];
        const double w = wt[k];
#pragma unroll
        for (int n = 0; n < NVn; ++n) s[n] += w * fres[i * NVn + n];
    }
#pragma unroll
    for (int n = 0; n < NVn; ++n) res0[v * NVn + n] = s[n];
}

// coarse ug = fine ug interpolated at the coarse vertices (fcnnct); vug_frst = ug
template <int NVn>
__global__ void k_hinterp(int nc, const int* __restrict__ fv, const double* __restrict__ w,
                          const double* __restrict__ fug, double* __restrict__ ug, double* __restrict__ vug) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nc) return;
    double s[NVn];
#pragma unroll
    for (int n = 0; n < NVn; ++n) s[n] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int f = fv[3 * v + j];
        const double ww = w[3 * v + j];
#pragma unroll
        for (int n = 0; n < NVn; ++n) s[n] += ww * fug[f * NVn + n];
    }
#pragma unroll
    for (int n = 0; n < NVn; ++n) { ug[v * NVn + n] = s[n]; vug[v * NVn + n] = s[n]; }
}

// vug_frst -= ug: minus the coarse correction
__global__ void k_hcorr(long n, double* __restrict__ vug, const double* __restrict__ ug) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) vug[k] -= ug[k];
}

// fine res = interpolated coarse correction (ccnnct), fine ug += res
template <int NVn>
__global__ void k_hprolong(int nf, const int* __restrict__ cv, const double* __restrict__ w,
                           const double* __restrict__ vug, double* __restrict__ fres, double* __restrict__ fug) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    double r[NVn];
#pragma unroll
    for (int n = 0; n < NVn; ++n) r[n] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int c = cv[3 * i + j];
        const double ww = w[3 * i + j];
#pragma unroll
        for (int n = 0; n < NVn; ++n) r[n] -= ww * vug[c * NVn + n];
    }
#pragma unroll
    for (int n = 0; n < NVn; ++n) {
        fres[i * NVn + n] = r[n];
        fug[i * NVn + n] += r[n];
    }
}

// MFMA lane-layout self test: D = A(16x4) * B(4x16) with the layout the element
// kernel assumes; out[lane*4+i] = D element held by (lane, register i)
__global__ void k_mfma_probe(const double* A, const double* B, double* out) {
    const int l = threadIdx.x;
    const double av = A[(l & 15) * 4 + (l >> 4)];
    const double bv = B[(l >> 4) * 16 + (l & 15)];
    d4 c = {0.0, 0.0, 0.0, 0.0};
    c = mfma(av, bv, c);
    for (int i = 0; i < 4; ++i) out[l * 4 + i] = c[i];
}

// same probe for v_mfma_f32_16x16x4f32 (FP32 inputs, FP32 accumulation)
__global__ void k_mfma_probe_f32(const float* A, const float* B, float* out) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int l = threadIdx.x;
    const float av = A[(l & 15) * 4 + (l >> 4)];
    const float bv = B[(l >> 4) * 16 + (l & 15)];
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) out[l * 4 + i] = c[i];
}

// explicit instantiations used by the host context
template __global__ void k_elem_ins<1, true>(ElemArgs);
template __global__ void k_elem_ins<1, false>(ElemArgs);
template __global__ void k_elem_ins<2, true>(ElemArgs);
template __global__ void k_elem_ins<2, false>(ElemArgs);
template __global__ void k_elem_ins<4, true>(ElemArgs);
template __global__ void k_elem_ins<4, false>(ElemArgs);
template __global__ void k_elem_ins32<1, true>(ElemArgs32);
template __global__ void k_elem_ins32<1, false>(ElemArgs32);
template __global__ void k_elem_ins32<2, true>(ElemArgs32);
template __global__ void k_elem_ins32<2, false>(ElemArgs32);
template __global__ void k_elem_ins32<4, true>(ElemArgs32);
template __global__ void k_elem_ins32<4, false>(ElemArgs32);


// grid-stride copy of n 4-byte words (device-to-device copies inside the solver)
__global__ void k_copy_words(long n, const unsigned* __restrict__ src, unsigned* __restrict__ dst) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) dst[k] = src[k];
}

// grid-stride 16-byte copy (reference kernel for traffic counters)
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n16) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

}  // namespace hpg
