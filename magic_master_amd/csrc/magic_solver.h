// Per-interface solver math shared by the sweep kernels (FP64): Roe state,
// limiters, transverse splits and the 9-point increment accumulator.
#pragma once
#include "magic_dev.h"

namespace magic {

struct Roe {
    double u, v, w, enth, a, g1a2, euv, u2v2w2;
};

__device__ __forceinline__ double fminF(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double fmaxF(double a, double b) { return a > b ? a : b; }

// rpn3euX.f:686-705
__device__ __forceinline__ double cflsupbee(double theta, double beta, double r, double cfl,
                                            double cflup)
{
    double cfle = fmaxF(0.001, fminF(0.999, cfl));
    double cflupe = fmaxF(0.001, fminF(0.999, cflup));
    double weight = (1.0 - cflupe) / (1.0 - cfle);
    double s1 = theta * weight * 2.0 / cfle;
    double s2 = (1.0 + cfle) / 3.0;
    double phimax = theta * 2.0 / (1.0 - cfle);
    double ultra = fmaxF(0.0, fminF(s1 * r, phimax));
    double c1 = 1.0 + (s2 - beta / 2.0) * (r - 1.0);
    double c2 = 1.0 + (s2 + beta / 2.0) * (r - 1.0);
    return fmaxF(0.0, fminF(ultra, fmaxF(c1, c2)));
}

// rpn3euX.f:709-731
__device__ __forceinline__ double suppow(double r, double cfl, double cflup)
{
    double cfle = fmaxF(0.001, fminF(0.999, cfl));
    double cflupe = fmaxF(0.001, fminF(0.999, cflup));
    double s2 = (1.0 + cfle) / 3.0;
    double s3 = 1.0 - s2;
    double weight = (1.0 - cflupe) / (1.0 - cfle);
    double pp = (r <= 1.0) ? weight * (2.0 / cfle) * 2.0 * s3 : fabs(2.0 / (1.0 - cfle)) * 2.0 * s2;
    double rabs = fabs(r);
    double rfrac = fabs((1.0 - rabs) / (1.0 + rabs));
    double signum = fmaxF(0.0, copysign(1.0, r));
    return signum * (s3 + s2 * r) * (1.0 - pow(rfrac, pp));
}

// Transverse split of a flux difference by the Roe eigensystem of the
// interface: rpt3euX.f:106-214; QUIRK = rptt3's icoor = 3 coefficient
// (rptt3euX.f:169).  Local components: 0 rho, 1 normal momentum, 2/3 the
// transverse momenta (mv, mw of rpn3euX.f:109-121), 4 energy, 5.. tracers.
template <int NC, int ICOOR, bool QUIRK>
__device__ __forceinline__ void tsplit(const Roe &R, const double (&in)[NC], double (&bm)[NC],
                                       double (&bp)[NC])
{
    const double q1 = in[0], qu = in[1], qv = in[2], qw = in[3], q5 = in[4];
    const double a4 = R.g1a2 * (R.euv * q1 + R.u * qu + R.v * qv + R.w * qw - q5);
    double w1[5], w2[5], w3[5], sb1, sb2, sb3;
    if (ICOOR == 2) {
        const double a2 = qu - R.u * q1, a3 = qw - R.w * q1;
        const double a5 = (qv + (R.a - R.v) * q1 - R.a * a4) / (2.0 * R.a);
        const double a1 = q1 - a4 - a5;
        w1[0] = a1; w1[1] = a1 * R.u; w1[2] = a1 * (R.v - R.a); w1[3] = a1 * R.w;
        w1[4] = a1 * (R.enth - R.v * R.a); sb1 = R.v - R.a;
        w2[0] = a4; w2[1] = a2 + R.u * a4; w2[2] = R.v * a4; w2[3] = a3 + R.w * a4;
        w2[4] = a4 * 0.5 * R.u2v2w2 + a2 * R.u + a3 * R.w; sb2 = R.v;
        w3[0] = a5; w3[1] = a5 * R.u; w3[2] = a5 * (R.v + R.a); w3[3] = a5 * R.w;
        w3[4] = a5 * (R.enth + R.v * R.a); sb3 = R.v + R.a;
    } else {
        const double a2 = (QUIRK ? qw : qu) - R.u * q1, a3 = qv - R.v * q1;
        const double a5 = (qw + (R.a - R.w) * q1 - R.a * a4) / (2.0 * R.a);
        const double a1 = q1 - a4 - a5;
        w1[0] = a1; w1[1] = a1 * R.u; w1[2] = a1 * R.v; w1[3] = a1 * (R.w - R.a);
        w1[4] = a1 * (R.enth - R.w * R.a); sb1 = R.w - R.a;
        w2[0] = a4; w2[1] = a2 + R.u * a4; w2[2] = a3 + R.v * a4; w2[3] = R.w * a4;
        w2[4] = a4 * 0.5 * R.u2v2w2 + a2 * R.u + a3 * R.v; sb2 = R.w;
        w3[0] = a5; w3[1] = a5 * R.u; w3[2] = a5 * R.v; w3[3] = a5 * (R.w + R.a);
        w3[4] = a5 * (R.enth + R.w * R.a); sb3 = R.w + R.a;
    }
    const double m1 = fminF(sb1, 0.0), m2 = fminF(sb2, 0.0), m3 = fminF(sb3, 0.0);
    const double p1 = fmaxF(sb1, 0.0), p2 = fmaxF(sb2, 0.0), p3 = fmaxF(sb3, 0.0);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        bm[m] = m1 * w1[m] + m2 * w2[m] + m3 * w3[m];
        bp[m] = p1 * w1[m] + p2 * w2[m] + p3 * w3[m];
    }
#pragma unroll
    for (int m = 5; m < NC; m++) {
        bm[m] = m2 * in[m];
        bp[m] = p2 * in[m];
    }
}

// Accumulators D[da+1][db+1][m] of the increments a cell receives, for the
// target cell offset (da, db) in the (g, h) plane (step3.f:189-226).
template <int NC>
struct Acc {
    double D[3][3][NC];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++)
#pragma unroll
                for (int m = 0; m < NC; m++) D[x][y][m] = 0.0;
    }
    // gadd(side, off) += c*v  folded with factor fg: side 1 -> da = -1,
    // side 2 -> da = +1, each also entering the own column with the
    // opposite sign (flux difference).
    __device__ __forceinline__ void g(int side, int off, double c, const double (&v)[NC])
    {
#pragma unroll
        for (int m = 0; m < NC; m++) {
            const double x = c * v[m];
            if (side == 1) { D[0][off + 1][m] -= x; D[1][off + 1][m] += x; }
            else           { D[2][off + 1][m] += x; D[1][off + 1][m] -= x; }
        }
    }
    __device__ __forceinline__ void h(int side, int off, double c, const double (&v)[NC])
    {
#pragma unroll
        for (int m = 0; m < NC; m++) {
            const double x = c * v[m];
            if (side == 1) { D[off + 1][0][m] -= x; D[off + 1][1][m] += x; }
            else           { D[off + 1][2][m] += x; D[off + 1][1][m] -= x; }
        }
    }
};

template <int NC>
__device__ __forceinline__ void axpy3(double (&out)[NC], const double (&a)[NC], double c,
                                      const double (&b)[NC])
{
#pragma unroll
    for (int m = 0; m < NC; m++) out[m] = a[m] + c * b[m];
}

template <int NC>
__device__ __forceinline__ void scatter(const SweepArgs &A, long long base, const Acc<NC> &acc)
{
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const long long off = base + (long long)(x - 1) * A.sA + (long long)(y - 1) * A.sB;
#pragma unroll
            for (int m = 0; m < NC; m++) {
                double *p = A.qnew + (long long)A.comp[m] * A.P + off;
                *p = *p + acc.D[x][y][m];
            }
        }
}


}  // namespace magic
