// FP32 kernels of the mixed-precision (defect-correction) path.
//
// The FP64 residual kernel stores, per Gauss point, u, v and the strong residual
// of the three equations at the outer state u0.  The inner solve then iterates on
// the linear problem  R64(u0) + J(u0) delta = 0  with the reference's multistage
// scheme, where J(u0) delta is the exact derivative of the tri_hp_ins element
// residual (tri_hp/ins/rsdl.cpp:259-433, including tau's dependence on |u|^2)
// evaluated here in FP32 on v_mfma_f32_16x16x4f32 with FP32 accumulation.  The
// FP64 residual stays the judge of convergence.
#include <hip/hip_runtime.h>
#include "tri_kernels.h"

namespace hpg {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Accumulator row held by register i of lane group q for v_mfma_f32_16x16x4f32:
// 4q + i (checked on the device by hpg_selftest_mfma_f32).
__host__ __device__ constexpr int f32_row(int i, int q) { return 4 * q + i; }

template <int P, bool FULL, bool PROBE = false>
__global__ __launch_bounds__(256) void k_elem_lin(LinArgs a) {
    using D = Dims<P>;
    constexpr int GT = D::GT, KT = D::KT;
    constexpr int SMD = D::SM > 0 ? D::SM : 1;
    __shared__ float sops[D::PACK];
    for (int k = threadIdx.x; k < D::PACK; k += blockDim.x) sops[k] = a.ops[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = lane & 15, q = lane >> 4;
    // Persistent and XCD-aware like k_elem_ins: XCD x walks the groups
    // [x*gpx, (x+1)*gpx) of the colour, each wave every WX-th one, with element ids
    // three groups ahead, connectivity two ahead and the mode offsets, geometry and
    // delta coefficients one group ahead of their use.
    const int WX = (gridDim.x >> 3) * 4;
    const int rank = (blockIdx.x >> 3) * 4 + w;
    const int xcd = blockIdx.x & 7;
    const int ng = (a.end - a.begin + 15) >> 4;
    const int gpx = (ng + 7) >> 3;
    const int g0 = xcd * gpx + rank;
    const int gend = min(ng, (xcd + 1) * gpx);
    if (g0 >= gend) return;

    // K-step s of lane l is mode 4s + l/16: the delta coefficients of a group
    auto gather = [&](const ElemSlots<P>& S, float (&bb)[NV][KT]) {
#pragma unroll
        for (int s = 0; s < KT; ++s) {
            const int k = 4 * s + q, kind = mode_kind<P>(k);
            const bool live = S.t >= 0 && k < D::TM;
            const int off = S.off[s] < 0 ? ~S.off[s] : S.off[s];
            const float sg = S.off[s] < 0 ? -1.f : 1.f;
            const float* src = kind == 0 ? a.dv : (kind == 1 ? a.ds : a.di);
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                if constexpr (PROBE)
                    bb[n][s] = (k == a.probe_m && n == a.probe_n) ? 1.f : 0.f;
                else
                    bb[n][s] = live ? sg * src[off + n] : 0.f;
            }
        }
    };

    ElemConn cur = elem_conn(a, elem_of_group(a, g0, gend, e));
    ElemSlots<P> S = elem_slots<P>(a, cur, q);
    float b[NV][KT];
    gather(S, b);
    ElemConn cnext = elem_conn(a, elem_of_group(a, g0 + WX, gend, e));
    int tnext2 = elem_of_group(a, g0 + 2 * WX, gend, e);

    for (int g = g0; g < gend; g += WX) {
        const ElemSlots<P> Sn = elem_slots<P>(a, cnext, q);
        float bn[NV][KT];
        gather(Sn, bn);
        const ElemConn cn2 = elem_conn(a, tnext2);
        const int tn3 = elem_of_group(a, g + 3 * WX, gend, e);

        const bool valid = cur.t >= 0;
        const int t = cur.t;
        const int* vtx = cur.vtx;
        const int* seg = cur.seg;
        const int* sgn = cur.sgn;
        // padding lanes run on the reference triangle with a zero delta
        const double X0 = S.x[0], Y0 = S.y[0], X1 = S.x[1], Y1 = S.y[1], X2 = S.x[2], Y2 = S.y[2];
        const float l00 = (float)(0.5 * (X2 - X1)), l01 = (float)(0.5 * (X0 - X1));
        const float l10 = (float)(0.5 * (Y2 - Y1)), l11 = (float)(0.5 * (Y0 - Y1));

        // frozen data of this group: the FP64 kernel stored row 16mt + 4i' + q' at
        // (mt*4 + i')*64 + e + 16q'
        const float* L = a.lin + (size_t)(a.gbase + g) * (5 * GT * 4 * 64);
        auto frozen = [&](int f, int mt, int i) {
            const int r = f32_row(i, q);
            return L[f * GT * 256 + (mt * 4 + (r >> 2)) * 64 + e + 16 * (r & 3)];
        };

        const f4 z = {0.f, 0.f, 0.f, 0.f};
        f4 dU[GT], dV[GT], dP[GT];
#pragma unroll
        for (int mt = 0; mt < GT; ++mt) {
            f4 cu = z, cv = z, cp = z;
#pragma unroll
            for (int s = 0; s < KT; ++s) {
                const float av = sops[D::OFF_P + (mt * KT + s) * 64 + lane];
                cu = mfma32(av, b[0][s], cu);
                cv = mfma32(av, b[1][s], cv);
                cp = mfma32(av, b[2][s], cp);
            }
            dU[mt] = cu; dV[mt] = cv; dP[mt] = cp;
        }

        const float rho = a.rho;
        f4 C0[GT], C1[GT], A00[GT], A01[GT], A10[GT], A11[GT];
        f4 U0[GT], V0[GT];
#pragma unroll
        for (int mt = 0; mt < GT; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float u = frozen(0, mt, i), v = frozen(1, mt, i);
                U0[mt][i] = u; V0[mt][i] = v;
                const float du = dU[mt][i], dv = dV[mt][i], dp = dP[mt][i];
                const float c0 = rho * (l11 * u - l01 * v), c1 = rho * (-l10 * u + l00 * v);
                const float dc0 = rho * (l11 * du - l01 * dv), dc1 = rho * (-l10 * du + l00 * dv);
                C0[mt][i] = dc0; C1[mt][i] = dc1;
                A00[mt][i] = du * c0 + u * dc0 + l11 * dp;
                A01[mt][i] = du * c1 + u * dc1 - l10 * dp;
                A10[mt][i] = dv * c0 + v * dc0 - l01 * dp;
                A11[mt][i] = dv * c1 + v * dc1 + l00 * dp;
            }
        f4 LI0 = z, LI1 = z, LI2 = z;
#pragma unroll
        for (int kt = 0; kt < GT; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kk = kt * 4 + i;
                const float ar = sops[D::OFF_IR + kk * 64 + lane];
                const float as = sops[D::OFF_IS + kk * 64 + lane];
                LI0 = mfma32(ar, A00[kt][i], LI0); LI0 = mfma32(as, A01[kt][i], LI0);
                LI1 = mfma32(ar, A10[kt][i], LI1); LI1 = mfma32(as, A11[kt][i], LI1);
                LI2 = mfma32(ar, C0[kt][i], LI2);  LI2 = mfma32(as, C1[kt][i], LI2);
            }

        f4 LR0 = z, LR1 = z, LR2 = z;
        if constexpr (FULL) {
            f4 Ur[GT], Us[GT], Vr[GT], Vs[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt) {
                f4 c1 = z, c2 = z, c3 = z, c4 = z;
#pragma unroll
                for (int s = 0; s < KT; ++s) {
                    const float ar = sops[D::OFF_PR + (mt * KT + s) * 64 + lane];
                    const float as = sops[D::OFF_PS + (mt * KT + s) * 64 + lane];
                    c1 = mfma32(ar, b[0][s], c1); c2 = mfma32(as, b[0][s], c2);
                    c3 = mfma32(ar, b[1][s], c3); c4 = mfma32(as, b[1][s], c4);
                }
                Ur[mt] = c1; Us[mt] = c2; Vr[mt] = c3; Vs[mt] = c4;
            }
            const double cjcbd = (0.5 * (X2 - X1)) * (0.5 * (Y0 - Y1)) - (0.5 * (Y2 - Y1)) * (0.5 * (X0 - X1));
            const float cjcb = (float)cjcbd;
            const float cjcbi = a.mu / cjcb;
            const float v0000 = -cjcbi * (2.f * l11 * l11 + l01 * l01);
            const float v0011 = -cjcbi * (2.f * l10 * l10 + l00 * l00);
            const float v0001 = cjcbi * (2.f * l11 * l10 + l01 * l00);
            const float v1100 = -cjcbi * (l11 * l11 + 2.f * l01 * l01);
            const float v1111 = -cjcbi * (l10 * l10 + 2.f * l00 * l00);
            const float v1101 = cjcbi * (l11 * l10 + 2.f * l01 * l00);
            const float v0100 = cjcbi * l01 * l11;
            const float v0111 = cjcbi * l00 * l10;
            const float v0101 = -cjcbi * l01 * l10;
            const float v0110 = -cjcbi * l00 * l11;
            f4 F00[GT], F01[GT], F10[GT], F11[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float ur = Ur[mt][i], us = Us[mt][i], vr = Vr[mt][i], vs = Vs[mt][i];
                    const float f00 = v0000 * ur + v0100 * vr + v0001 * us + v0101 * vs;
                    const float f01 = v0001 * ur + v0110 * vr + v0011 * us + v0111 * vs;
                    const float f10 = v0100 * ur + v1100 * vr + v0110 * us + v1101 * vs;
                    const float f11 = v0101 * ur + v1101 * vr + v0111 * us + v1111 * vs;
                    F00[mt][i] = f00; F01[mt][i] = f01; F10[mt][i] = f10; F11[mt][i] = f11;
                    A00[mt][i] += f00; A01[mt][i] += f01; A10[mt][i] += f10; A11[mt][i] += f11;
                }
            f4 R0[GT], R1[GT], R2[GT];
#pragma unroll
            for (int mt = 0; mt < GT; ++mt) {
                f4 r0 = z, r1 = z, r2 = z;
#pragma unroll
                for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int kk = (mt * GT + kt) * 4 + i;
                        const float dr = sops[D::OFF_DR + kk * 64 + lane];
                        const float ds = sops[D::OFF_DS + kk * 64 + lane];
                        r0 = mfma32(dr, A00[kt][i], r0); r0 = mfma32(ds, A01[kt][i], r0);
                        r1 = mfma32(dr, A10[kt][i], r1); r1 = mfma32(ds, A11[kt][i], r1);
                        r2 = mfma32(dr, C0[kt][i], r2);  r2 = mfma32(ds, C1[kt][i], r2);
                    }
                R0[mt] = r0; R1[mt] = r1; R2[mt] = r2;
            }
            const double dx1 = X0 - X2, dy1 = Y0 - Y2, dx2 = X1 - X0, dy2 = Y1 - Y0;
            const double ar2 = dx1 * dy2 - dy1 * dx2;
            const double perim = sqrt(dx1 * dx1 + dy1 * dy1) + sqrt(dx2 * dx2 + dy2 * dy2) +
                                 sqrt((dx1 + dx2) * (dx1 + dx2) + (dy1 + dy2) * (dy1 + dy2));
            const double hd = (ar2 / perim) / (0.25 * (P + 1) * (P + 1));
            const float gv = (float)(2.0 * ((double)a.mu / (double)rho) / hd);
            const float gvis = gv * gv;
            const float tfac = (float)(a.adis * hd / cjcbd);
#pragma unroll
            for (int mt = 0; mt < GT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float u = U0[mt][i], v = V0[mt][i];
                    const float du = dU[mt][i], dv = dV[mt][i];
                    const float s0 = frozen(2, mt, i), s1 = frozen(3, mt, i), s2 = frozen(4, mt, i);
                    const float qq = u * u + v * v;
                    float gam = 3.f * qq + gvis;
                    float dq = 2.f * (u * du + v * dv);
                    float dgam = 3.f * dq;
                    if (a.mu == 0.f && gam < 0.1f) { gam = 0.1f; dgam = 0.f; }
                    const float tau0 = tfac / sqrtf(gam);
                    const float dtau0 = -0.5f * tau0 * dgam / gam;
                    const float sq = sqrtf(qq);
                    const float tau2 = sq * tau0;
                    const float dtau2 = (qq > 0.f ? dq / (2.f * sq) : 0.f) * tau0 + sq * dtau0;
                    const float t0 = tau0 * s0, t1 = tau0 * s1, t2 = tau2 * s2;
                    const float dt0 = dtau0 * s0 + tau0 * R0[mt][i];
                    const float dt1 = dtau0 * s1 + tau0 * R1[mt][i];
                    const float dt2 = dtau2 * s2 + tau2 * R2[mt][i];
                    F00[mt][i] -= (2.f * l11 * du - l01 * dv) * t0 + (2.f * l11 * u - l01 * v) * dt0 -
                                  l01 * (du * t1 + u * dt1) + l11 * dt2;
                    F01[mt][i] -= (-2.f * l10 * du + l00 * dv) * t0 + (-2.f * l10 * u + l00 * v) * dt0 +
                                  l00 * (du * t1 + u * dt1) - l10 * dt2;
                    F10[mt][i] -= l11 * (dv * t0 + v * dt0) + (l11 * du - 2.f * l01 * dv) * t1 +
                                  (l11 * u - 2.f * l01 * v) * dt1 - l01 * dt2;
                    F11[mt][i] -= -l10 * (dv * t0 + v * dt0) + (-l10 * du + 2.f * l00 * dv) * t1 +
                                  (-l10 * u + 2.f * l00 * v) * dt1 + l00 * dt2;
                    C0[mt][i] = -(l11 * dt0 - l01 * dt1);
                    C1[mt][i] = -(-l10 * dt0 + l00 * dt1);
                }
#pragma unroll
            for (int kt = 0; kt < GT; ++kt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int kk = kt * 4 + i;
                    const float ar = sops[D::OFF_IR + kk * 64 + lane];
                    const float as = sops[D::OFF_IS + kk * 64 + lane];
                    LR0 = mfma32(ar, F00[kt][i], LR0); LR0 = mfma32(as, F01[kt][i], LR0);
                    LR1 = mfma32(ar, F10[kt][i], LR1); LR1 = mfma32(as, F11[kt][i], LR1);
                    LR2 = mfma32(ar, C0[kt][i], LR2);  LR2 = mfma32(as, C1[kt][i], LR2);
                }
        }

        if (valid && PROBE) {
            // element-local diagonal entry J_e(m n, m n) of the full residual (beta = 1)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (f32_row(i, q) != a.probe_m) continue;
                const float li[3] = {LI0[i], LI1[i], LI2[i]};
                const float lr[3] = {LR0[i], LR1[i], LR2[i]};
                // summed per mode type (0 vertex, 1 side, 2 interior): diag[t][type][n]
                a.diag[((size_t)t * 3 + mode_kind<P>(a.probe_m)) * NV + a.probe_n] +=
                    fabs((double)li[a.probe_n] + (double)lr[a.probe_n]);
            }
        } else if (valid) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = f32_row(i, q);
                if (k >= D::TM) continue;
                int kind;
                double sgd;
                const int off = mode_offset<P>(k, t, vtx, seg, sgn, a.sm0, a.im0, kind, sgd);
                const float sg = (float)sgd;
                float* di = kind == 0 ? a.resv : (kind == 1 ? a.ress : a.resi);
                float* dr = kind == 0 ? a.rrv : (kind == 1 ? a.rrs : a.rri);
                const double* sr = kind == 0 ? a.srcv : (kind == 1 ? a.srcs : a.srci);
                const bool first =
                    kind == 2 || (kind == 0 ? a.fcv[sel3(vtx, k)] : a.fcs[sel3(seg, (k - 3) / SMD)]) == a.colour;
                const float li[3] = {sg * LI0[i], sg * LI1[i], sg * LI2[i]};
                float lr[3] = {0.f, 0.f, 0.f};
                if constexpr (FULL) {
                    lr[0] = sg * (LR0[i] * a.beta);
                    lr[1] = sg * (LR1[i] * a.beta);
                    lr[2] = sg * (LR2[i] * a.beta);
                }
                if (first) {
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        const float r0 = a.keep == 0.f ? 0.f : a.keep * dr[off + n];
                        if (FULL || a.keep != 1.f) dr[off + n] = r0 + lr[n];
                        const float src = sr ? (float)sr[off + n] : 0.f;
                        di[off + n] = r0 + src + lr[n] + li[n];
                    }
                } else {
#pragma unroll
                    for (int n = 0; n < NV; ++n) {
                        if constexpr (FULL) dr[off + n] += lr[n];
                        di[off + n] += lr[n] + li[n];
                    }
                }
            }
        }
        S = Sn;
        cur = cnext;
        cnext = cn2;
        tnext2 = tn3;
#pragma unroll
        for (int n = 0; n < NV; ++n)
#pragma unroll
            for (int s = 0; s < KT; ++s) b[n][s] = bn[n][s];
    }
}

// linearised edge flux: d(flx) for the bdry_ins::generic flux (tri_hp/ins/bdry_ins.h:34-49)
struct EdgeLinArgs {
    const int* eseg;
    int begin, end;
    const int* seg_pnt;
    const double* pnts;
    const double* ugv;  // outer state u0
    const double* ugs;
    const float* dv;    // delta
    const float* ds;
    float* resv;
    float* ress;
    const double* g1;
    const double* dg1;
    const double* w1;
    int gpx, sm, sm0;
    double rho;
};

template <int P>
__global__ __launch_bounds__(64) void k_edge_lin(EdgeLinArgs a) {
    constexpr int SM = P - 1, N1 = SM + 2, GPX = P + 1;
    const int idx = a.begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.end) return;
    const int sind = a.eseg[idx];
    const int v0 = a.seg_pnt[2 * sind], v1 = a.seg_pnt[2 * sind + 1];
    double uh[2][N1], dh[2][N1];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        uh[n][0] = a.ugv[v0 * NV + n]; uh[n][1] = a.ugv[v1 * NV + n];
        dh[n][0] = a.dv[v0 * NV + n]; dh[n][1] = a.dv[v1 * NV + n];
#pragma unroll
        for (int m = 0; m < SM; ++m) {
            uh[n][m + 2] = a.ugs[(sind * a.sm0 + m) * NV + n];
            dh[n][m + 2] = a.ds[(sind * a.sm0 + m) * NV + n];
        }
    }
    const double X0 = a.pnts[2 * v0], Y0 = a.pnts[2 * v0 + 1], X1 = a.pnts[2 * v1], Y1 = a.pnts[2 * v1 + 1];
    double lf[NV][N1];
#pragma unroll
    for (int n = 0; n < NV; ++n)
#pragma unroll
        for (int k = 0; k < N1; ++k) lf[n][k] = 0.0;
#pragma unroll
    for (int ip = 0; ip < GPX; ++ip) {
        const double* g = a.g1 + ip * N1;
        const double* dg = a.dg1 + ip * N1;
        const double dxc = X0 * dg[0] + X1 * dg[1];
        const double dyc = Y0 * dg[0] + Y1 * dg[1];
        double u[2], du[2];
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            double s = 0.0, ds = 0.0;
#pragma unroll
            for (int k = 0; k < N1; ++k) { s += uh[n][k] * g[k]; ds += dh[n][k] * g[k]; }
            u[n] = s; du[n] = ds;
        }
        double nx = dyc, ny = -dxc;
        const double jcb = sqrt(nx * nx + ny * ny);
        nx /= jcb; ny /= jcb;
        const double f2 = a.rho * (u[0] * nx + u[1] * ny);
        const double d2 = a.rho * (du[0] * nx + du[1] * ny);
        const double fl[NV] = {d2 * u[0] + f2 * du[0], d2 * u[1] + f2 * du[1], d2};
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const double r = fl[n] * jcb * a.w1[ip];
#pragma unroll
            for (int k = 0; k < N1; ++k) lf[n][k] += g[k] * r;
        }
    }
#pragma unroll
    for (int n = 0; n < NV; ++n) {
        a.resv[v0 * NV + n] += (float)lf[n][0];
        a.resv[v1 * NV + n] += (float)lf[n][1];
#pragma unroll
        for (int m = 0; m < SM; ++m) a.ress[(sind * a.sm0 + m) * NV + n] += (float)lf[n][m + 2];
    }
}

// ---- FP32 defect correction with the nonlinear FP32 residual ----------------------
// The inner solve runs the multistage scheme on F(v) = R32(v) + (1/c - 1) R64(u0) from
// v = b = fl32(u0): R32(b + dhat) - R64(u0) ~ J(u0) dhat, so dhat solves the scaled
// defect equation and u0 += c dhat.  The scale c keeps |dhat| near `h` |u0| (see
// k_fd_scale), where FP32 differences of the residual resolve J(u0) dhat to about
// 1e-3 relative; without it corrections below FP32 resolution of u0 would be lost.

// max of |x| over a block, one atomicMax of the float bits (non-negative floats order
// like their bit patterns)
__device__ __forceinline__ void block_absmax(float x, unsigned* out) {
    __shared__ float red[4];
    for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_down(x, off, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmaxf(m, red[k]);
        atomicMax(out, __float_as_uint(m));
    }
}

// The reductions below are grid-stride loops over a grid capped at RED_BLOCKS blocks, so
// at most that many atomics meet on the one result word.
constexpr int RED_BLOCKS = 2048;

// b = v = fl32(u), umax = max |b| (blocks of 256)
__global__ __launch_bounds__(256) void k_base32(long n, const double* __restrict__ u, float* __restrict__ b,
                                                float* __restrict__ v, unsigned* umax) {
    float m = 0.f;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const float x = (float)u[k];
        b[k] = x;
        v[k] = x;
        m = fmaxf(m, fabsf(x));
    }
    block_absmax(m, umax);
}

// dmax = max |v - b| (blocks of 256)
__global__ __launch_bounds__(256) void k_diffmax32(long n, const float* __restrict__ v, const float* __restrict__ b,
                                                   unsigned* dmax) {
    float m = 0.f;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(v[k] - b[k]));
    block_absmax(m, dmax);
}

// scale of this outer iteration from the last one: c = |delta_prev| / (h |u0|) with
// |delta_prev| = c_prev max|dhat_prev|, capped at 1 (then the inner solve is the plain
// FP32 nonlinear smoother); c = 1 after a reset.  Writes c, 1/c - 1 and clears the
// |dhat| maximum for this iteration.
__global__ void k_fd_scale(double* scal, unsigned* mx, int reset, double h) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double un = (double)__uint_as_float(mx[0]), dn = (double)__uint_as_float(mx[1]);
    double c = 1.0;
    if (!reset && dn > 0.0 && un > 0.0) c = fmin(1.0, scal[0] * dn / (h * un));
    if (!(c > 1e-30)) c = 1.0;
    scal[0] = c;
    scal[1] = 1.0 / c - 1.0;
    mx[0] = 0u;
    mx[1] = 0u;
}

// y = fl32(x / c) with 1/c = scal[1] + 1: the inner residual at v = b, F(b) = R64(u0) / c
__global__ void k_scale_f32(long n, const double* __restrict__ x, const double* __restrict__ scal,
                            float* __restrict__ y) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = (float)(x[k] * (scal[1] + 1.0));
}

// ug += c (v - b), FP64
__global__ void k_add_scaled(long n, double* __restrict__ ug, const float* __restrict__ v, const float* __restrict__ b,
                             const double* __restrict__ scal) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) ug[k] += scal[0] * ((double)v[k] - (double)b[k]);
}

// ---- FP32 defect correction with FP64 storage (cns, tet): the inner state v lives in
// the FP64 state arrays, so the FP64 preconditioner, boundary and projection steps run
// on it unchanged; only the element residual is evaluated in FP32

// u = fl64(fl32(u)) in place (the base b = fl32(u0) of the inner solve), umax = max |b|
__global__ __launch_bounds__(256) void k_round32(long n, double* __restrict__ u, unsigned* umax) {
    float m = 0.f;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const float x = (float)u[k];
        u[k] = (double)x;
        m = fmaxf(m, fabsf(x));
    }
    block_absmax(m, umax);
}

// x *= 1/c (1/c = scal[1] + 1): F(b) = R64(u0) / c
__global__ void k_scale64(long n, double* __restrict__ x, const double* __restrict__ scal) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] *= scal[1] + 1.0;
}

// dmax = max |v - fl32(u0)|, then v = u0 + c (v - fl32(u0)): the FP64 update
__global__ __launch_bounds__(256) void k_fd_final64(long n, double* __restrict__ v, const double* __restrict__ u0,
                                                    const double* __restrict__ scal, unsigned* dmax) {
    float m = 0.f;
    const double c = scal[0];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const double b = (double)(float)u0[k];
        const double dh = v[k] - b;
        m = fmaxf(m, fabsf((float)dh));
        v[k] = u0[k] + c * dh;
    }
    block_absmax(m, dmax);
}

// clear n unsigned values (a fill kernel rather than a memset node)
__global__ void k_clear_u32(int n, unsigned* x) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = 0u;
}

// FP64 -> FP32 copy
__global__ void k_to_f32(long n, const double* __restrict__ x, float* __restrict__ y) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = (float)x[k];
}

template __global__ void k_elem_lin<1, true>(LinArgs);
template __global__ void k_elem_lin<1, false>(LinArgs);
template __global__ void k_elem_lin<2, true>(LinArgs);
template __global__ void k_elem_lin<2, false>(LinArgs);
template __global__ void k_elem_lin<4, true>(LinArgs);
template __global__ void k_elem_lin<4, false>(LinArgs);
template __global__ void k_elem_lin<1, true, true>(LinArgs);
template __global__ void k_elem_lin<2, true, true>(LinArgs);
template __global__ void k_elem_lin<4, true, true>(LinArgs);

// learned-preconditioner inputs per element (11): vertex offsets from the centroid
// scaled by sqrt(area) (6), vertex-mean u, v, p (3), log(mu / (rho sqrt(area))),
// log(area / tprcn_u) (the analytic local pseudo-time step)
__global__ void k_mlp_features(int ntri, const double* pnts, const int* tri_pnt, const double* ugv,
                               const double* tprcn, double mu, double rho, float* X) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const int v0 = tri_pnt[3 * t], v1 = tri_pnt[3 * t + 1], v2 = tri_pnt[3 * t + 2];
    const double x[3] = {pnts[2 * v0], pnts[2 * v1], pnts[2 * v2]};
    const double y[3] = {pnts[2 * v0 + 1], pnts[2 * v1 + 1], pnts[2 * v2 + 1]};
    const double area = 0.5 * ((x[0] - x[2]) * (y[1] - y[0]) - (y[0] - y[2]) * (x[1] - x[0]));
    const double L = sqrt(fabs(area));
    const double xc = (x[0] + x[1] + x[2]) / 3.0, yc = (y[0] + y[1] + y[2]) / 3.0;
    float* o = X + (size_t)t * 11;
    for (int k = 0; k < 3; ++k) {
        o[2 * k] = (float)((x[k] - xc) / L);
        o[2 * k + 1] = (float)((y[k] - yc) / L);
    }
    for (int n = 0; n < NV; ++n) o[6 + n] = (float)((ugv[v0 * NV + n] + ugv[v1 * NV + n] + ugv[v2 * NV + n]) / 3.0);
    o[9] = (float)log(1e-12 + mu / (rho * L));
    o[10] = (float)log(fabs(area) / fmax(tprcn[3 * t], 1e-300));
}

// learned element scalings of the vertex, side and interior rows of the preconditioner:
// t_type = tprcn * exp(clamp(y[type][n], -2, 2)), Y = ntri x 9 ([type][var], type 0
// vertex, 1 side, 2 interior)
__global__ void k_mlp_apply(int ntri, const float* Y, const double* tprcn, double* tv, double* ts, double* ti) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntri * NV) return;
    const int t = k / NV, n = k - t * NV;
    const float* y = Y + (size_t)t * 3 * NV;
    const double tp = tprcn[k];
    tv[k] = tp * exp((double)fminf(2.f, fmaxf(-2.f, y[n])));
    ts[k] = tp * exp((double)fminf(2.f, fmaxf(-2.f, y[NV + n])));
    ti[k] = tp * exp((double)fminf(2.f, fmaxf(-2.f, y[2 * NV + n])));
}

}  // namespace hpg
