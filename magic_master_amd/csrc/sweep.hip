// Fused directional sweep of the MAGIC unsplit f-wave step (FP64).
//
// One workgroup owns one pencil; lane i owns interface i (between cells i-1
// and i, Fortran numbering 0..n+2).  Per interface it does, in registers:
//   rpn3   Roe averages, f-wave decomposition with well-balanced gravity,
//          entropy fix, amdq/apdq                   (rpn3euX.f:138-411)
//   limiters: tracer CFL-Superbee/minmod and Lax-Liu SuperPower/CFL-Superbee
//          on the f-waves, with the upwind jump read from the neighbour
//          lane through LDS                           (rpn3euX.f:417-676)
//   flux3  second-order correction cqxx, 14 transverse splits (the four
//          cqxx splits of flux3fw.f:333-346 are two distinct ones) and the
//          double-transverse corner terms              (flux3fw.f:239-625)
//   step3  the 9-point increment of its two cells (i-1 and i) into the
//          pencil and its 8 transverse neighbours       (step3.f:189-226)
// The 9-point scatter is race-free and deterministic: a launch only runs
// pencils of one colour (a mod 3, b mod 3), and within a pencil the
// contributions to cell i-1 (from interface i) and to cell i (from
// interface i) are applied in two phases separated by a barrier.
#include "magic_dev.h"
#include "magic_solver.h"

namespace magic {

__device__ __forceinline__ void report(const SweepArgs &A, int code, int i, int pa, int pb)
{
    if (atomicCAS(A.status, 0, code) == 0) {
        // Fortran indices (i along the pencil, pa along g, pb along h)
        int fi, fj, fk;
        if (A.sweep == 1) { fi = i; fj = pa; fk = pb; }
        else if (A.sweep == 2) { fi = pb; fj = i; fk = pa; }
        else { fi = pa; fj = pb; fk = i; }
        A.status[1] = A.sweep; A.status[2] = fi; A.status[3] = fj; A.status[4] = fk;
    }
}

template <int NTR>
__global__ __launch_bounds__(512) void sweep_kernel(SweepArgs A)
{
    constexpr int NC = 5 + NTR;
    constexpr double beta = 0.666666666666666;
    extern __shared__ double smem[];
    const int n = A.n;
    const int NI = n + 3;
    double *s_del = smem;                  // [5][NI] jumps delta(1:5, i)
    double *s_s2 = s_del + 5 * NI;         // [NI]    s(2, i)
    double *s_trw = s_s2 + NI;             // [NTR][NI] unlimited tracer f-waves
    double *s_cfo = s_trw + NTR * NI;      // [NI]    CFLup value left by interface i
    int *s_flag = (int *)(s_cfo + NI);     // [NI]    interface i assigns CFLup

    const int i = threadIdx.x;             // interface 0..n+2
    const int pa = A.cA + 3 * (int)blockIdx.x;
    const int pb = A.cB + 3 * (int)blockIdx.y;
    const long long pbase = (long long)(pa + A.mbc - 1) * A.sA + (long long)(pb + A.mbc - 1) * A.sB;
    const bool act = i <= n + 2;
    const double dtdn = A.dtdn;

    Roe R = {0, 0, 0, 0, 1, 0, 0, 0};
    double fw1[NC], fw2[NC], fw3[NC], amdq[NC], apdq[NC];
    double s1 = 0, s2 = 0, s3 = 0, c1 = 0, c2 = 0, c3 = 0;
    bool f1 = false, f2 = false, f5 = false;

    if (act) {
        const long long ca = pbase + (long long)(i - 1 + A.mbc - 1) * A.sN;   // cell i-1
        const long long cb = ca + A.sN;                                      // cell i
        double qa[NC], qb[NC];
#pragma unroll
        for (int m = 0; m < NC; m++) {
            qa[m] = A.qold[(long long)A.comp[m] * A.P + ca];
            qb[m] = A.qold[(long long)A.comp[m] * A.P + cb];
        }
        const double ga = A.gam[ca], gb = A.gam[cb];
        if (!(qa[0] > 0.0) || !(qb[0] > 0.0)) report(A, 1, i, pa, pb);
        // Roe averages: rpn3euX.f:159-188
        const double rhsqrtl = sqrt(qa[0]), rhsqrtr = sqrt(qb[0]);
        const double rhsq2 = rhsqrtl + rhsqrtr;
        const double gamma = (rhsqrtr * gb + rhsqrtl * ga) / rhsq2;
        const double gamma1 = gamma - 1.0;
        const double pl = (ga - 1.0) * (qa[4] - 0.5 * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) / qa[0]);
        const double pr = (gb - 1.0) * (qb[4] - 0.5 * (qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]) / qb[0]);
        R.u = (qa[1] / rhsqrtl + qb[1] / rhsqrtr) / rhsq2;
        R.v = (qa[2] / rhsqrtl + qb[2] / rhsqrtr) / rhsq2;
        R.w = (qa[3] / rhsqrtl + qb[3] / rhsqrtr) / rhsq2;
        R.enth = ((qa[4] + pl) / rhsqrtl + (qb[4] + pr) / rhsqrtr) / rhsq2;
        R.u2v2w2 = R.u * R.u + R.v * R.v + R.w * R.w;
        const double a2 = gamma1 * (R.enth - 0.5 * R.u2v2w2);
        if (!(a2 > 0.0)) report(A, 2, i, pa, pb);
        R.a = sqrt(a2);
        R.i2a = 1.0 / (2.0 * R.a);
        R.g1a2 = gamma1 / a2;
        R.euv = R.enth - R.u2v2w2;
        // jumps and eigen-coefficients: rpn3euX.f:194-216
        const double ur = qb[1] / qb[0], ul = qa[1] / qa[0];
        const double d1 = qb[1] - qa[1];
        double d2 = (qb[1] * ur + pr) - (qa[1] * ul + pl);
        const double d3 = ur * qb[2] - ul * qa[2];
        const double d4 = ur * qb[3] - ul * qa[3];
        double d5 = (qb[4] + pr) * ur - (qa[4] + pl) * ul;
        if (A.gravity) {
            const double vg = A.vgrav[cb];
            d2 = d2 - 0.5 * (qb[0] + qa[0]) * vg;
            d5 = d5 - 0.5 * (qb[1] + qa[1]) * vg;
        }
        const double b4 = R.g1a2 * (R.euv * d1 + R.u * d2 + R.v * d3 + R.w * d4 - d5);
        const double b2 = d3 - R.v * d1;
        const double b3 = d4 - R.w * d1;
        const double b5 = (d2 + (R.a - R.u) * d1 - R.a * b4) / (2.0 * R.a);
        const double b1 = d1 - b4 - b5;
        // f-waves: rpn3euX.f:226-269
        fw1[0] = b1; fw1[1] = b1 * (R.u - R.a); fw1[2] = b1 * R.v; fw1[3] = b1 * R.w;
        fw1[4] = b1 * (R.enth - R.u * R.a);
        s1 = R.u - R.a;
        fw2[0] = b4; fw2[1] = b4 * R.u; fw2[2] = b2 + b4 * R.v; fw2[3] = b3 + b4 * R.w;
        fw2[4] = b2 * R.v + b3 * R.w + b4 * 0.5 * R.u2v2w2;
        s2 = R.u;
        fw3[0] = b5; fw3[1] = b5 * (R.u + R.a); fw3[2] = b5 * R.v; fw3[3] = b5 * R.w;
        fw3[4] = b5 * (R.enth + R.u * R.a);
        s3 = R.u + R.a;
#pragma unroll
        for (int m = 5; m < NC; m++) {
            fw1[m] = 0.0;
            fw3[m] = 0.0;
            fw2[m] = (m - 4 <= A.nspecadv) ? s2 * (qb[m] - qa[m]) : ur * qb[m] - ul * qa[m];
        }
        // entropy fix and amdq: rpn3euX.f:316-398
#pragma unroll
        for (int m = 0; m < NC; m++) amdq[m] = 0.0;
        {
            const double gmm = ga, gm1 = ga - 1.0;
            const double rhoim1 = qa[0];
            const double pim1 = gm1 * (qa[4] - 0.5 * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) / rhoim1);
            const double cim1 = sqrt(gmm * pim1 / rhoim1);
            const double s0 = qa[1] / rhoim1 - cim1;
            if (!(s0 >= 0.0 && s1 > 0.0)) {
                const double rho1 = qa[0] + fw1[0] / s1;
                const double rhou1 = qa[1] + fw1[1] / s1;
                const double rhov1 = qa[2] + fw1[2] / s1;
                const double rhow1 = qa[3] + fw1[3] / s1;
                const double en1 = qa[4] + fw1[4] / s1;
                const double p1 = gm1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1 + rhow1 * rhow1) / rho1);
                const double cc1 = sqrt(gmm * p1 / rho1);
                const double sa = rhou1 / rho1 - cc1;
                double sfract;
                if (s0 < 0.0 && sa > 0.0) sfract = s0 * (sa - s1) / (sa - s0);
                else if (s1 < 0.0) sfract = s1;
                else sfract = 0.0;
#pragma unroll
                for (int m = 0; m < NC; m++) amdq[m] = sfract * fw1[m] / s1;
                if (s2 < 0.0) {
#pragma unroll
                    for (int m = 0; m < NC; m++) amdq[m] = amdq[m] + fw2[m];
                    const double gm = gb, g1 = gb - 1.0;
                    const double rhoi = qb[0];
                    const double pi_ = g1 * (qb[4] - 0.5 * (qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]) / rhoi);
                    const double ci = sqrt(gm * pi_ / rhoi);
                    const double s3r = qb[1] / rhoi + ci;
                    const double rho2 = qb[0] - fw3[0] / s3;
                    const double rhou2 = qb[1] - fw3[1] / s3;
                    const double rhov2 = qb[2] - fw3[2] / s3;
                    const double rhow2 = qb[3] - fw3[3] / s3;
                    const double en2 = qb[4] - fw3[4] / s3;
                    const double p2 = g1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2 + rhow2 * rhow2) / rho2);
                    const double cc2 = sqrt(gm * p2 / rho2);
                    const double sb = rhou2 / rho2 + cc2;
                    bool add = true;
                    if (sb < 0.0 && s3r > 0.0) sfract = sb * (s3r - s3) / (s3r - sb);
                    else if (s3 < 0.0) sfract = s3;
                    else add = false;
                    if (add) {
#pragma unroll
                        for (int m = 0; m < NC; m++) amdq[m] = amdq[m] + sfract * fw3[m] / s3;
                    }
                }
            }
        }
        // apdq = sum of f-waves - amdq: rpn3euX.f:404-411
#pragma unroll
        for (int m = 0; m < NC; m++) apdq[m] = ((0.0 + fw1[m]) + fw2[m]) + fw3[m] - amdq[m];
        // local Courant numbers: rpn3euX.f:417-421 (uniform dt/dn)
        c1 = fabs(s1 * dtdn); c2 = fabs(s2 * dtdn); c3 = fabs(s3 * dtdn);
        // which CFLup assignments this interface makes (rpn3euX.f:536-649)
        const double b2r = fw2[2] - b4 * R.v;
        f1 = b1 != 0.0; f2 = b2r != 0.0; f5 = b5 != 0.0;
        s_del[0 * NI + i] = d1; s_del[1 * NI + i] = d2; s_del[2 * NI + i] = d3;
        s_del[3 * NI + i] = d4; s_del[4 * NI + i] = d5;
        s_s2[i] = s2;
#pragma unroll
        for (int t = 0; t < NTR; t++) s_trw[t * NI + i] = fw2[5 + t];
        s_flag[i] = (f1 || f2 || f5) ? 1 : 0;
        s_cfo[i] = f5 ? c3 : (f2 ? c2 : c1);
    }
    __syncthreads();

    const bool lim = act && i >= 1 && i <= n + 1;
    double cqxx[NC];
    double cflmax = 0.0;
    if (lim) {
        // tracer limiters: rpn3euX.f:434-467
#pragma unroll
        for (int t = 0; t < NTR; t++) {
            const double fc = fw2[5 + t];
            if (fc != 0.0) {
                const double r = (s2 > 0.0) ? s_trw[t * NI + i - 1] / fc : s_trw[t * NI + i + 1] / fc;
                double wl;
                if (t + 1 <= A.nspecadv) {
                    const int id2 = (int)(i - copysign(1.0, s2));
                    wl = cflsupbee(0.5, beta, r, c2, fabs(s_s2[id2] * dtdn));
                } else {
                    wl = fmaxF(0.0, fminF(1.0, r));
                }
                fw2[5 + t] = wl * fc;
            }
        }
        // Lax-Liu limiting of waves 1-5: rpn3euX.f:492-676
        const double b1 = fw1[0], b4 = fw2[0];
        const double b2 = fw2[2] - b4 * R.v, b3 = fw2[3] - b4 * R.w, b5 = fw3[0];
        double wl1 = 1.0, wl2 = 1.0, wl3 = 1.0, wl4 = 1.0, wl5 = 1.0;
        auto del = [&](int k, int id) { return s_del[k * NI + id]; };
        if (b1 != 0.0) {
            const int id = (int)(i - copysign(1.0, s1));
            const double b4Id = R.g1a2 * (R.euv * del(0, id) + R.u * del(1, id) + R.v * del(2, id) +
                                          R.w * del(3, id) - del(4, id));
            const double b5Id = (del(1, id) + (R.a - R.u) * del(0, id) - R.a * b4Id) / (2.0 * R.a);
            const double b1Id = del(0, id) - b4Id - b5Id;
            wl1 = suppow(b1Id / b1, c1, c1);
        }
        if (b2 != 0.0) {
            const int id = (int)(i - copysign(1.0, s2));
            const double b2Id = del(2, id) - R.v * del(0, id);
            wl2 = cflsupbee(1.0, beta, b2Id / b2, c2, c2);
        }
        if (b3 != 0.0 || b4 != 0.0) {
            double cflup;
            if (f2) cflup = c2;
            else if (f1) cflup = c1;
            else {
                cflup = 0.0;
                for (int j = i - 1; j >= 1; j--)
                    if (s_flag[j]) { cflup = s_cfo[j]; break; }
            }
            const int id = (int)(i - copysign(1.0, s2));
            if (b3 != 0.0) {
                const double b3Id = del(3, id) - R.w * del(0, id);
                wl3 = cflsupbee(1.0, beta, b3Id / b3, c2, cflup);
            }
            if (b4 != 0.0) {
                const double b4Id = R.g1a2 * (R.euv * del(0, id) + R.u * del(1, id) + R.v * del(2, id) +
                                              R.w * del(3, id) - del(4, id));
                wl4 = cflsupbee(1.0, beta, b4Id / b4, c2, cflup);
            }
        }
        if (b5 != 0.0) {
            const int id = (int)(i - copysign(1.0, s3));
            const double b4Id = R.g1a2 * (R.euv * del(0, id) + R.u * del(1, id) + R.v * del(2, id) +
                                          R.w * del(3, id) - del(4, id));
            const double b5Id = (del(1, id) + (R.a - R.u) * del(0, id) - R.a * b4Id) / (2.0 * R.a);
            wl5 = suppow(b5Id / b5, c3, c3);
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            fw1[m] = fw1[m] * wl1;
            fw3[m] = fw3[m] * wl5;
        }
        const double b2l = wl2 * b2, b3l = wl3 * b3, b4l = wl4 * b4;
        fw2[0] = b4l;
        fw2[1] = b4l * R.u;
        fw2[2] = b2l + b4l * R.v;
        fw2[3] = b3l + b4l * R.w;
        fw2[4] = b2l * R.v + b3l * R.w + b4l * 0.5 * R.u2v2w2;
        // second-order correction: flux3fw.f:270-284
        const double k1 = 0.5 * copysign(1.0, s1) * (1.0 - fabs(s1) * dtdn);
        const double k2 = 0.5 * copysign(1.0, s2) * (1.0 - fabs(s2) * dtdn);
        const double k3 = 0.5 * copysign(1.0, s3) * (1.0 - fabs(s3) * dtdn);
#pragma unroll
        for (int m = 0; m < NC; m++) cqxx[m] = k1 * fw1[m] + k2 * fw2[m] + k3 * fw3[m];
        // Courant number over interfaces 1..n+1: flux3fw.f:246-254
        cflmax = fmaxF(fmaxF(c1, c2), c3);
    } else {
#pragma unroll
        for (int m = 0; m < NC; m++) cqxx[m] = 0.0;
    }
    // CFL: wavefront max, one atomic per wave (non-negative doubles order
    // like their bit patterns).
    for (int o = 32; o > 0; o >>= 1) cflmax = fmaxF(cflmax, __shfl_xor(cflmax, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(A.cfl_bits, (unsigned long long)__double_as_longlong(cflmax));

    const double d = dtdn;
    const double k6g = (1.0 / 6.0) * d * A.fh;  // corner terms of gadd (flux3fw.f:443-483)
    const double k6h = (1.0 / 6.0) * d * A.fg;  // corner terms of hadd (flux3fw.f:568-608)
    double Bmq[NC], Bpq[NC], Cmq[NC], Cpq[NC];
    if (lim) {
        tsplit<NC, 2, false>(R, cqxx, Bmq, Bpq);
        tsplit<NC, 3, false>(R, cqxx, Cmq, Cpq);
    }

    // ---- phase L: interface i -> cell i-1 (the "i-1" terms of flux3) ----
    const double cg = A.fg, ch = A.fh;     // dt/dg and dt/dh of the 9-point update
    if (lim && i >= 2) {
        Acc<NC> acc;
        acc.zero();
        double Bm[NC], Bp[NC], Cm[NC], Cp[NC], t1[NC], t2[NC], x1[NC], x2[NC];
        tsplit<NC, 2, false>(R, amdq, Bm, Bp);
        tsplit<NC, 3, false>(R, amdq, Cm, Cp);
        // G: 2-D part and correction (flux3fw.f:426-429, 495-498)
        acc.g(1, 0, -0.5 * d * cg, Bm); acc.g(2, 0, -0.5 * d * cg, Bp);
        acc.g(1, 0, -d * cg, Bmq);      acc.g(2, 0, -d * cg, Bpq);
        // corner part: rptt3 of c*amdq2 in the g direction (flux3fw.f:388-412, 464-483)
        axpy3(t1, Cp, 3.0, Cpq);   // cpamdq2
        axpy3(t2, Cm, 3.0, Cmq);   // cmamdq2
        tsplit<NC, 2, true>(R, t1, x1, x2);   // bmcpamdq, bpcpamdq
        acc.g(2, 0, k6g * cg, x2);  acc.g(1, 0, k6g * cg, x1);
        acc.g(2, 1, -k6g * cg, x2); acc.g(1, 1, -k6g * cg, x1);
        tsplit<NC, 2, true>(R, t2, x1, x2);   // bmcmamdq, bpcmamdq
        acc.g(2, 0, -k6g * cg, x2); acc.g(1, 0, -k6g * cg, x1);
        acc.g(2, -1, k6g * cg, x2); acc.g(1, -1, k6g * cg, x1);
        // H: 2-D part and correction (flux3fw.f:552-555, 620-623)
        acc.h(1, 0, -0.5 * d * ch, Cm); acc.h(2, 0, -0.5 * d * ch, Cp);
        acc.h(1, 0, -d * ch, Cmq);      acc.h(2, 0, -d * ch, Cpq);
        // corner part: rptt3 of b*amdq + 3 b*cqxx in the h direction
        // (flux3fw.f:510-538, 589-608)
        axpy3(t1, Bp, 3.0, Bpq);   // bpamdq
        axpy3(t2, Bm, 3.0, Bmq);   // bmamdq
        tsplit<NC, 3, true>(R, t1, x1, x2);   // C-, C+ of bpamdq
        acc.h(2, 0, k6h * ch, x2);  acc.h(1, 0, k6h * ch, x1);
        acc.h(2, 1, -k6h * ch, x2); acc.h(1, 1, -k6h * ch, x1);
        tsplit<NC, 3, true>(R, t2, x1, x2);   // C-, C+ of bmamdq
        acc.h(2, 0, -k6h * ch, x2); acc.h(1, 0, -k6h * ch, x1);
        acc.h(2, -1, k6h * ch, x2); acc.h(1, -1, k6h * ch, x1);
        // F and Godunov terms of cell i-1: -dtdn*amdq(i) (flux3fw.f:242) and
        // -dtdn*fadd(i) (step3.f:194)
#pragma unroll
        for (int m = 0; m < NC; m++) acc.D[1][1][m] += -d * amdq[m] - d * cqxx[m];
        scatter(A, pbase + (long long)(i - 1 + A.mbc - 1) * A.sN, acc);
    }
    __syncthreads();
    // ---- phase R: interface i -> cell i (the "i" terms of flux3) ----
    if (lim && i <= n) {
        Acc<NC> acc;
        acc.zero();
        double Bm[NC], Bp[NC], Cm[NC], Cp[NC], t1[NC], t2[NC], x1[NC], x2[NC];
        tsplit<NC, 2, false>(R, apdq, Bm, Bp);
        tsplit<NC, 3, false>(R, apdq, Cm, Cp);
        acc.g(1, 0, -0.5 * d * cg, Bm); acc.g(2, 0, -0.5 * d * cg, Bp);
        acc.g(1, 0, d * cg, Bmq);       acc.g(2, 0, d * cg, Bpq);
        axpy3(t1, Cp, -3.0, Cpq);  // cpapdq2
        axpy3(t2, Cm, -3.0, Cmq);  // cmapdq2
        tsplit<NC, 2, true>(R, t1, x1, x2);   // bmcpapdq, bpcpapdq
        acc.g(2, 0, k6g * cg, x2);  acc.g(1, 0, k6g * cg, x1);
        acc.g(2, 1, -k6g * cg, x2); acc.g(1, 1, -k6g * cg, x1);
        tsplit<NC, 2, true>(R, t2, x1, x2);   // bmcmapdq, bpcmapdq
        acc.g(2, 0, -k6g * cg, x2); acc.g(1, 0, -k6g * cg, x1);
        acc.g(2, -1, k6g * cg, x2); acc.g(1, -1, k6g * cg, x1);
        acc.h(1, 0, -0.5 * d * ch, Cm); acc.h(2, 0, -0.5 * d * ch, Cp);
        acc.h(1, 0, d * ch, Cmq);       acc.h(2, 0, d * ch, Cpq);
        axpy3(t1, Bp, -3.0, Bpq);  // bpapdq
        axpy3(t2, Bm, -3.0, Bmq);  // bmapdq
        tsplit<NC, 3, true>(R, t1, x1, x2);
        acc.h(2, 0, k6h * ch, x2);  acc.h(1, 0, k6h * ch, x1);
        acc.h(2, 1, -k6h * ch, x2); acc.h(1, 1, -k6h * ch, x1);
        tsplit<NC, 3, true>(R, t2, x1, x2);
        acc.h(2, 0, -k6h * ch, x2); acc.h(1, 0, -k6h * ch, x1);
        acc.h(2, -1, k6h * ch, x2); acc.h(1, -1, k6h * ch, x1);
        // cell i gets -dtdn*apdq(i) (flux3fw.f:241) and +dtdn*fadd(i) (step3.f:194)
#pragma unroll
        for (int m = 0; m < NC; m++) acc.D[1][1][m] += -d * apdq[m] + d * cqxx[m];
        scatter(A, pbase + (long long)(i + A.mbc - 1) * A.sN, acc);
    }
}

template __global__ void sweep_kernel<0>(SweepArgs);
template __global__ void sweep_kernel<1>(SweepArgs);
template __global__ void sweep_kernel<2>(SweepArgs);
template __global__ void sweep_kernel<3>(SweepArgs);
template __global__ void sweep_kernel<4>(SweepArgs);
template __global__ void sweep_kernel<5>(SweepArgs);

}  // namespace magic
