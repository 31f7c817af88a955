// Per-interface solver math shared by the sweep kernels: Roe state,
// limiters, transverse splits and the 9-point increment accumulator, for a
// floating-point type T (double: the FP64 form; float: the FP32 perturbation
// form).  Constants are written as T(...) so that the double instantiation
// performs exactly the reference's operations.
#pragma once
#include "magic_dev.h"

namespace magic {

template <typename T>
struct RoeT {
    T u, v, w, enth, a, g1a2, euv, u2v2w2;
    T i2a;   // 1 / (2 a): the wave-strength divisor of rpn3 and of every transverse split
};
using Roe = RoeT<double>;

// The well-balanced gravity jump (rpn3euX.f:194-208): across a hydrostatic
// column p_b - p_a - .5 (rho_a + rho_b) aux17 cancels to rounding, so a
// single fused multiply-add in it, or in the pressures it is made of, moves
// the jump by a unit of the pressure.  These pieces are rounded as the
// reference rounds them (no contraction) whatever the surrounding code's
// contraction setting.
template <typename T>
__device__ __forceinline__ T pressure_nc(T gam, T E, T m1, T m2, T m3, T irho)
{
#pragma clang fp contract(off)
    return (gam - T(1.0)) * (E - T(0.5) * (m1 * m1 + m2 * m2 + m3 * m3) * irho);
}
template <typename T>
__device__ __forceinline__ T flux_jump_nc(T mb, T ub, T pb, T ma, T ua, T pa)
{
#pragma clang fp contract(off)
    return (mb * ub + pb) - (ma * ua + pa);
}
template <typename T>
__device__ __forceinline__ T gravity_jump_nc(T d, T ra, T rb, T vg)
{
#pragma clang fp contract(off)
    return d - T(0.5) * (rb + ra) * vg;
}

// Reciprocal and square root of the Riemann solver's densities, sound
// speeds and wave speeds: correctly rounded in FP64 (as the reference); in
// the FP32 form, whose error bar is a fraction of the field peak, the
// hardware's single-instruction approximations (about one unit in the last
// place).  Never used for a ratio of two quantities that can both be tiny.
template <typename T>
__device__ __forceinline__ T rcpT(T x) { return T(1.0) / x; }
template <>
__device__ __forceinline__ float rcpT<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ __forceinline__ T sqrtT(T x) { return sqrt(x); }
template <>
__device__ __forceinline__ float sqrtT<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// The FP32 form solves the normal Riemann problem in FP64 (the sweep's
// part 1): its jumps and wave strengths subtract background-sized products
// (enthalpy times velocity, enthalpy times the density jump), whose FP32
// rounding the flow carries up into the thin upper layers; in FP32 the
// form's temperature error there grew to 3e-4 of the field peak in 300
// steps, in FP64 to 3e-5-6e-5.  Its reciprocals and square roots need not
// be correctly rounded there: the hardware's FP64 approximations refined by
// one Newton step (relative error near 1e-15, far inside that form's error
// bar).  RF selects them (true: the FP32 form) over the correctly rounded
// operations of the FP64 form.
template <bool RF>
__device__ __forceinline__ double rcp1(double x)
{
    if constexpr (RF) {
        const double r = __builtin_amdgcn_rcp(x);
        return fma(r, fma(-x, r, 1.0), r);
    } else {
        return 1.0 / x;
    }
}
template <bool RF>
__device__ __forceinline__ double sqrt1(double x)
{
    if constexpr (RF) {
        // Goldschmidt step from the reciprocal square root; 0 and negative
        // (or NaN) arguments take the plain hardware square root
        const double y = __builtin_amdgcn_rsq(x);
        const double s0 = x * y, h = 0.5 * y;
        const double r = fma(-s0, h, 0.5);
        const double s1 = fma(s0, r, s0);
        return x > 0.0 ? s1 : __builtin_amdgcn_sqrt(x);
    } else {
        return sqrt(x);
    }
}

// flang's MIN/MAX: a < b ? a : b, a > b ? a : b
template <typename T>
__device__ __forceinline__ T fminF(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T fmaxF(T a, T b) { return a > b ? a : b; }

// x**y for 0 <= x < 1, y > 0 (suppow's rfrac**pp): the library pow in FP64;
// in the FP32 form 2**(y log2 x) from the hardware's exp2 and log2
// (log2 0 = -inf gives 0, as pow does).
__device__ __forceinline__ double powT(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float powT(float x, float y)
{
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// rpn3euX.f:686-705
template <typename T>
__device__ __forceinline__ T cflsupbee(T theta, T beta, T r, T cfl, T cflup)
{
    T cfle = fmaxF(T(0.001), fminF(T(0.999), cfl));
    T cflupe = fmaxF(T(0.001), fminF(T(0.999), cflup));
    // cfle is clamped to [0.001, 0.999]: its reciprocals are taken once
    const T i1mc = rcpT(T(1.0) - cfle), icfle = rcpT(cfle);
    T weight = (T(1.0) - cflupe) * i1mc;
    T s1 = theta * weight * T(2.0) * icfle;
    T s2 = (T(1.0) + cfle) * (T(1.0) / T(3.0));
    T phimax = theta * T(2.0) * i1mc;
    T ultra = fmaxF(T(0.0), fminF(s1 * r, phimax));
    T c1 = T(1.0) + (s2 - beta / T(2.0)) * (r - T(1.0));
    T c2 = T(1.0) + (s2 + beta / T(2.0)) * (r - T(1.0));
    return fmaxF(T(0.0), fminF(ultra, fmaxF(c1, c2)));
}

// rpn3euX.f:709-731
template <typename T>
__device__ __forceinline__ T suppow(T r, T cfl, T cflup)
{
    if constexpr (sizeof(T) == 4) {
        // FP32: a denormal wave strength can make the ratio r overflow to
        // +-inf, and (1-|r|)/(1+|r|) is then NaN.  For |r| >= 1e30 the
        // result is already 0 in both precisions (rfrac rounds to 1).
        r = fminF(fmaxF(r, T(-1e30)), T(1e30));
    }
    T cfle = fmaxF(T(0.001), fminF(T(0.999), cfl));
    T cflupe = fmaxF(T(0.001), fminF(T(0.999), cflup));
    const T i1mc = rcpT(T(1.0) - cfle), icfle = rcpT(cfle);
    T s2 = (T(1.0) + cfle) * (T(1.0) / T(3.0));
    T s3 = T(1.0) - s2;
    T weight = (T(1.0) - cflupe) * i1mc;
    T pp = (r <= T(1.0)) ? weight * (T(2.0) * icfle) * T(2.0) * s3 : fabs(T(2.0) * i1mc) * T(2.0) * s2;
    T rabs = fabs(r);
    T rfrac = fabs((T(1.0) - rabs) * rcpT(T(1.0) + rabs));
    T signum = fmaxF(T(0.0), copysign(T(1.0), r));
    return signum * (s3 + s2 * r) * (T(1.0) - powT(rfrac, pp));
}

// Transverse split of a flux difference by the Roe eigensystem of the
// interface: rpt3euX.f:106-214; QUIRK = rptt3's icoor = 3 coefficient
// (rptt3euX.f:169).  Local components: 0 rho, 1 normal momentum, 2/3 the
// transverse momenta (mv, mw of rpn3euX.f:109-121), 4 energy, 5.. tracers.
template <int NC, int ICOOR, bool QUIRK, typename T>
__device__ __forceinline__ void tsplit(const RoeT<T> &R, const T (&in)[NC], T (&bm)[NC], T (&bp)[NC])
{
    const T q1 = in[0], qu = in[1], qv = in[2], qw = in[3], q5 = in[4];
    const T a4 = R.g1a2 * (R.euv * q1 + R.u * qu + R.v * qv + R.w * qw - q5);
    T w1[5], w2[5], w3[5], sb1, sb2, sb3;
    if (ICOOR == 2) {
        const T a2 = qu - R.u * q1, a3 = qw - R.w * q1;
        const T a5 = (qv + (R.a - R.v) * q1 - R.a * a4) * R.i2a;
        const T a1 = q1 - a4 - a5;
        w1[0] = a1; w1[1] = a1 * R.u; w1[2] = a1 * (R.v - R.a); w1[3] = a1 * R.w;
        w1[4] = a1 * (R.enth - R.v * R.a); sb1 = R.v - R.a;
        w2[0] = a4; w2[1] = a2 + R.u * a4; w2[2] = R.v * a4; w2[3] = a3 + R.w * a4;
        w2[4] = a4 * T(0.5) * R.u2v2w2 + a2 * R.u + a3 * R.w; sb2 = R.v;
        w3[0] = a5; w3[1] = a5 * R.u; w3[2] = a5 * (R.v + R.a); w3[3] = a5 * R.w;
        w3[4] = a5 * (R.enth + R.v * R.a); sb3 = R.v + R.a;
    } else {
        const T a2 = (QUIRK ? qw : qu) - R.u * q1, a3 = qv - R.v * q1;
        const T a5 = (qw + (R.a - R.w) * q1 - R.a * a4) * R.i2a;
        const T a1 = q1 - a4 - a5;
        w1[0] = a1; w1[1] = a1 * R.u; w1[2] = a1 * R.v; w1[3] = a1 * (R.w - R.a);
        w1[4] = a1 * (R.enth - R.w * R.a); sb1 = R.w - R.a;
        w2[0] = a4; w2[1] = a2 + R.u * a4; w2[2] = a3 + R.v * a4; w2[3] = R.w * a4;
        w2[4] = a4 * T(0.5) * R.u2v2w2 + a2 * R.u + a3 * R.v; sb2 = R.w;
        w3[0] = a5; w3[1] = a5 * R.u; w3[2] = a5 * R.v; w3[3] = a5 * (R.w + R.a);
        w3[4] = a5 * (R.enth + R.w * R.a); sb3 = R.w + R.a;
    }
    const T m1 = fminF(sb1, T(0.0)), m2 = fminF(sb2, T(0.0)), m3 = fminF(sb3, T(0.0));
    const T p1 = fmaxF(sb1, T(0.0)), p2 = fmaxF(sb2, T(0.0)), p3 = fmaxF(sb3, T(0.0));
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

// The split of components M0 .. M0+G-1 alone.  A tracer's split involves
// only its own component and the contact speed (the v or w of the
// interface), so a group of tracers (M0 >= 5) is split without the Euler
// components; M0 = 0 is the split above.
template <int M0, int G, int ICOOR, bool QUIRK, typename T>
__device__ __forceinline__ void tsplit_group(const RoeT<T> &R, const T (&in)[G], T (&bm)[G], T (&bp)[G])
{
    if constexpr (M0 == 0) {
        tsplit<G, ICOOR, QUIRK>(R, in, bm, bp);
    } else {
        static_assert(M0 >= 5, "a group without the Euler components holds tracers only");
        const T s = ICOOR == 2 ? R.v : R.w;
        const T m2 = fminF(s, T(0.0)), p2 = fmaxF(s, T(0.0));
#pragma unroll
        for (int m = 0; m < G; m++) {
            bm[m] = m2 * in[m];
            bp[m] = p2 * in[m];
        }
    }
}

// Accumulators D[da+1][db+1][m] of the increments a cell receives, for the
// target cell offset (da, db) in the (g, h) plane (step3.f:189-226).
template <int NC, typename T = double>
struct Acc {
    T D[3][3][NC];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++)
#pragma unroll
                for (int m = 0; m < NC; m++) D[x][y][m] = T(0.0);
    }
    // gadd(side, off) += c*v  folded with factor fg: side 1 -> da = -1,
    // side 2 -> da = +1, each also entering the own column with the
    // opposite sign (flux difference).
    __device__ __forceinline__ void g(int side, int off, T c, const T (&v)[NC])
    {
#pragma unroll
        for (int m = 0; m < NC; m++) {
            const T x = c * v[m];
            if (side == 1) { D[0][off + 1][m] -= x; D[1][off + 1][m] += x; }
            else           { D[2][off + 1][m] += x; D[1][off + 1][m] -= x; }
        }
    }
    __device__ __forceinline__ void h(int side, int off, T c, const T (&v)[NC])
    {
#pragma unroll
        for (int m = 0; m < NC; m++) {
            const T x = c * v[m];
            if (side == 1) { D[off + 1][0][m] -= x; D[off + 1][1][m] += x; }
            else           { D[off + 1][2][m] += x; D[off + 1][1][m] -= x; }
        }
    }
    __device__ __forceinline__ void c11(int m, T x) { D[1][1][m] += x; }
};

template <int NC, typename T>
__device__ __forceinline__ void axpy3(T (&out)[NC], const T (&a)[NC], T c, const T (&b)[NC])
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
