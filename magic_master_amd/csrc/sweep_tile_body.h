// Tiled directional sweep of the MAGIC unsplit f-wave step, for a float
// type T and two state forms:
//   T = double, PERT = false: q is the conserved state (the FP64 form);
//   T = float,  PERT = true : q is the perturbation q - q_bar from a
//     background column at rest (the FP32 form); the f-wave jumps are formed
//     from perturbation pressures, so the hydrostatic balance cancels
//     analytically, and the Courant number is reduced in FP64.
// This header holds the kernel; sweep_tile.hip and sweep_tile32.hip
// instantiate it under their own floating-point contraction settings.
//
// Same per-interface algebra as sweep.hip (rpn3euX.f, rpt3euX.f, rptt3euX.f,
// flux3fw.f, step3.f); different data movement:
//   * a workgroup owns TA pencils that are neighbours along the tile axis "a"
//     (x-sweep: a = y; y- and z-sweeps: a = x, so that tile-axis-major lanes
//     read rows of x) and marches along the pencils in chunks of CH
//     interfaces (overlap 2); TA, CH and the lane order form the tile shape;
//   * the 3x3 increments of each interface (its cells i-1 and i into the
//     pencil and its 8 transverse neighbours, step3.f:189-226) are summed in
//     an LDS tile LT[b offset][a][cell] in sub-steps (one per phase and a
//     offset, so no two lanes add to the same entry at once; separated by the
//     wave's program order or by barriers, see the sub-step loop), and the
//     tile is flushed to q once per chunk (FP64: one atomic add per entry;
//     FP32: one coalesced read-modify-write per row);
//   * tiles are coloured 2 ways along a and 3 ways along the other transverse
//     axis b, so no two workgroups of a launch flush to the same rows.
// The result is deterministic.
#pragma once
#include "magic_dev.h"
#include "magic_solver.h"

namespace magic {

// Tile shapes: TA pencils x CH interfaces per chunk (TA*CH = 256 lanes),
// lanes either tile-axis-major (consecutive lanes on neighbouring pencils)
// or pencil-major (consecutive lanes along one pencil).  The host picks one
// per sweep direction (a fixed table or the autotuner); every shape computes
// the same increments, only the grouping of their sums differs.
template <int SH>
struct TileShape;
template <>
struct TileShape<0> { static constexpr int TA = 16; static constexpr bool PM = false; };
template <>
struct TileShape<1> { static constexpr int TA = 8; static constexpr bool PM = false; };
template <>
struct TileShape<2> { static constexpr int TA = 8; static constexpr bool PM = true; };
template <>
struct TileShape<3> { static constexpr int TA = 4; static constexpr bool PM = true; };
static_assert(TileShape<TILE_NSHAPE - 1>::TA > 0, "one TileShape per shape index");

// Pointer into global memory.  The work-list kernel reads its state pointers
// from an argument block in memory, where the compiler cannot tell their
// address space and would use flat loads and stores; flat accesses count on
// the LDS counter too, so every LDS wait would also wait for them.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) { return (gptr<T>)p; }

// The FP64 tile flush adds each tile entry to q with one atomic add per
// element and discards the result, so no wave waits for q to arrive (one
// wave per SIMD has nothing else to run meanwhile); the colouring keeps
// every element of q to one adding workgroup per launch, so the sum is the
// same as a read-modify-write.  Measured on the full-domain workload: FP64
// step +3 %; in FP32 (two waves per SIMD) the atomics cost 17 % (one grid)
// and 27 % (ensemble), so FP32 keeps the read-modify-write.
template <typename T>
constexpr bool tile_flush_atomic() { return sizeof(T) == 8; }
__device__ __forceinline__ void add_noret(gptr<double> p, double v) { (void)__builtin_amdgcn_global_atomic_fadd_f64(p, v); }
__device__ __forceinline__ void add_noret(gptr<float> p, float v) { (void)__builtin_amdgcn_global_atomic_fadd_f32(p, v); }

// The Roe state the limiters use: the FP64 form's own; the FP32 form's FP64
// copy from its solve.
template <typename T>
__device__ __forceinline__ const RoeT<double> &limiter_roe(const RoeT<T> &R, const RoeT<double> &RD)
{
    if constexpr (sizeof(T) == 8) return R;
    else return RD;
}

template <typename AT>
__device__ __forceinline__ void report_t(const AT &A, int code, int i, int pg, int ph)
{
    if (atomicCAS(A.status, 0, code) == 0) {
        int fi, fj, fk;
        if (A.sweep == 1) { fi = i; fj = pg; fk = ph; }
        else if (A.sweep == 2) { fi = ph; fj = i; fk = pg; }
        else { fi = pg; fj = ph; fk = i; }
        A.status[1] = A.sweep; A.status[2] = fi; A.status[3] = fj; A.status[4] = fk;
    }
}

// One tile (bx along the tile axis, by along the coloured axis) of one
// colour launch; called by every lane of a 256-lane workgroup.
template <typename T, bool PERT, int NTR, int SH>
__device__ __forceinline__ void tile_sweep_body(const SweepArgsT<T> &A, int a_is_g, int bx, int by)
{
    constexpr int NC = 5 + NTR;
    constexpr int NT = 256, TA = TileShape<SH>::TA, CH = NT / TA;
    constexpr bool PM = TileShape<SH>::PM;
    constexpr int NCELL = CH - 1;               // cells S .. S+CH-2 of a chunk
    constexpr int TROW = TA + 2;                // a offsets A0-1 .. A0+TA
    constexpr T beta = T(0.666666666666666);
    static_assert(TA * CH == NT && TA == tile_ta(SH), "tile is 256 lanes");

    __shared__ double s_del[5][NT];   // FP64 in both forms: the limiters' upwind wave strengths
    __shared__ T s_s2[NT];
    __shared__ T s_trw[NTR > 0 ? NTR : 1][NT];
    __shared__ T s_cfo[NT];
    __shared__ int s_flag[NT];
    __shared__ T s_cprev[TA];
    __shared__ T LT[NC * 3 * TROW * NCELL + 1];   // + a slot that only ever holds 0

    const int t = threadIdx.x;
    const int pa = PM ? (t / CH) : (t % TA);    // pencil within the tile
    const int c = PM ? (t % CH) : (t / TA);     // interface offset within the chunk
    const int nb = PM ? 1 : TA;                 // lane distance between neighbouring interfaces
    const int n = A.n;
    const int A0 = (2 * bx + A.cA) * TA;
    const int a = A0 + pa;                      // tile-axis coordinate of this lane's pencil
    const int b = A.cB + 3 * by;                // other transverse coordinate
    const bool pen = a <= A.nA + 1;
    const int pg = a_is_g ? a : b, ph = a_is_g ? b : a;   // (g, h) coordinates
    const long long pbase = (long long)(a + A.mbc - 1) * A.sA + (long long)(b + A.mbc - 1) * A.sB;
    const T dtdn = T(A.dtdn);
    const T cg = T(A.fg), ch = T(A.fh);
    const T d = dtdn;
    const gptr<const T> qold = to_global(A.qold);
    const gptr<T> qnew = to_global(A.qnew);
    const gptr<const double> gam = to_global(A.gam), vgrav = to_global(A.vgrav);
    const gptr<const double> crho = to_global(A.crho), cE = to_global(A.cE), cgam = to_global(A.cgam),
                             cvg = to_global(A.cvg), cres = to_global(A.cres);
    const T k6g = (T(1.0) / T(6.0)) * d * ch;
    const T k6h = (T(1.0) / T(6.0)) * d * cg;

    // T index: (m, db, a offset ai, cell ci); the fastest index follows memory
    auto tix = [&](int m, int db, int ai, int ci) {
        return PM ? ((m * 3 + db) * TROW + ai) * NCELL + ci : ((m * 3 + db) * NCELL + ci) * TROW + ai;
    };
    for (int e = t; e <= NC * 3 * TROW * NCELL; e += NT) LT[e] = T(0.0);
    if (t < TA) s_cprev[t] = T(0.0);
    double cflmax = 0.0;            // Courant number, reduced in FP64
    const int nchunk = (n + 1 + (CH - 2) - 1) / (CH - 2);

    // The states of a lane's interface in the next chunk are loaded while
    // the current chunk's transverse terms are computed (registers), so the
    // HBM latency of part 1 is hidden behind part 3.  Not in the FP64
    // pencil-major kernels, where the registers would spill.
    constexpr bool PREFETCH = !(PM && sizeof(T) == 8) && !(sizeof(T) == 8 && NTR > 2);
    T pqa[NC], pqb[NC], pga = T(0.0), pgb = T(0.0), pvg = T(0.0);
    auto fetch = [&](int Iq) {
        if (pen && Iq <= n + 2) {
            const long long ca = pbase + (long long)(Iq - 1 + A.mbc - 1) * A.sN;
            const long long cb = ca + A.sN;
#pragma unroll
            for (int m = 0; m < NC; m++) {
                pqa[m] = qold[(long long)A.comp[m] * A.P + ca];
                pqb[m] = qold[(long long)A.comp[m] * A.P + cb];
            }
            if constexpr (!PERT) {
                pga = gam[ca];
                pgb = gam[cb];
                if (A.gravity) pvg = vgrav[cb];
            }
        }
    };
    if constexpr (PREFETCH) fetch(c);

    for (int chunk = 0; chunk < nchunk; chunk++) {
        const int S = chunk * (CH - 2);
        const int I = S + c;                    // Fortran interface index
        const bool act1 = pen && I <= n + 2;
        const bool act2 = act1 && c >= 1 && c <= CH - 2 && I >= 1 && I <= n + 1;
        // No barrier here: the previous chunk's last readers of the s_*
        // arrays (part 2 and the CFLup carry) are behind its sub-step
        // barriers, and its flush's reads of the tile are behind this
        // chunk's barrier after part 1.

        // ---------------- part 1: Roe solve of interface I (rpn3euX.f:138-421)
        RoeT<T> R = {0, 0, 0, 0, 1, 0, 0, 0, 0};
        T fw1[NC], fw2[NC], fw3[NC], amdq[NC], apdq[NC];
        T s1 = 0, s2 = 0, s3 = 0, c1 = 0, c2 = 0, c3 = 0;
        bool f1 = false, f2 = false, f5 = false;
#pragma unroll
        for (int m = 0; m < NC; m++) { fw1[m] = fw2[m] = fw3[m] = amdq[m] = apdq[m] = T(0.0); }
        // The solve runs in FP64 in both forms (T1 = double): the FP64 form
        // writes its results in place; the FP32 form solves into FP64
        // arrays and rounds the results (see rcp1 in magic_solver.h).
        auto riemann = [&](auto &RR, auto &F1, auto &F2, auto &F3, auto &AM, auto &AP, auto &W) {
            using T1 = double;
            constexpr bool RF = sizeof(T) == 4;
            T1 qa[NC], qb[NC], vgc = T1(0.0);
            T1 ga, gb;
            if constexpr (PREFETCH) {
#pragma unroll
                for (int m = 0; m < NC; m++) {
                    qa[m] = pqa[m];
                    qb[m] = pqb[m];
                }
                ga = pga;
                gb = pgb;
                vgc = pvg;
            } else {
                const long long ca = pbase + (long long)(I - 1 + A.mbc - 1) * A.sN;
                const long long cb = ca + A.sN;
#pragma unroll
                for (int m = 0; m < NC; m++) {
                    qa[m] = qold[(long long)A.comp[m] * A.P + ca];
                    qb[m] = qold[(long long)A.comp[m] * A.P + cb];
                }
                if constexpr (!PERT) {
                    ga = gam[ca];
                    gb = gam[cb];
                    if (A.gravity) vgc = vgrav[cb];
                }
            }
            // perturbation form: q holds q - background; the background is a
            // column at rest (rho_bar(k), E_bar(k), gamma(k)), so the full
            // states differ from q only in rho and E
            T1 rpa = 0, rpb = 0, ppa = 0, ppb = 0;
            int kib = 0;
            if constexpr (PERT) {
                const int ka = A.sweep == 3 ? I - 1 : b, kb = A.sweep == 3 ? I : b;
                const int kia = ka + A.mbc - 1;
                kib = kb + A.mbc - 1;
                ga = cgam[kia];
                gb = cgam[kib];
                rpa = qa[0];
                rpb = qb[0];
                const T1 epa = qa[4], epb = qb[4];
                qa[0] = crho[kia] + rpa;
                qb[0] = crho[kib] + rpb;
                qa[4] = cE[kia] + epa;
                qb[4] = cE[kib] + epb;
                // p - p_bar = (gamma-1) (E' - kinetic energy), p_bar = (gamma-1) E_bar
                ppa = (ga - T1(1.0)) * (epa - T1(0.5) * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) / qa[0]);
                ppb = (gb - T1(1.0)) * (epb - T1(0.5) * (qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]) / qb[0]);
            }
            if (!(qa[0] > T1(0.0)) || !(qb[0] > T1(0.0))) report_t(A, 1, I, pg, ph);
            const T1 rhsqrtl = sqrt1<RF>(qa[0]), rhsqrtr = sqrt1<RF>(qb[0]);
            const T1 rhsq2 = rhsqrtl + rhsqrtr;
            // reciprocals taken once where rpn3euX.f:159-216 divides by the
            // same value several times (the products differ from the
            // quotients by at most one rounding)
            const T1 il = rcp1<RF>(rhsqrtl), ir = rcp1<RF>(rhsqrtr), i2 = rcp1<RF>(rhsq2);
            const T1 iqa = rcp1<RF>(qa[0]), iqb = rcp1<RF>(qb[0]);
            const T1 gamma = (rhsqrtr * gb + rhsqrtl * ga) * i2;
            const T1 gamma1 = gamma - T1(1.0);
            const T1 pl = pressure_nc(ga, qa[4], qa[1], qa[2], qa[3], iqa);
            const T1 pr = pressure_nc(gb, qb[4], qb[1], qb[2], qb[3], iqb);
            RR.u = (qa[1] * il + qb[1] * ir) * i2;
            RR.v = (qa[2] * il + qb[2] * ir) * i2;
            RR.w = (qa[3] * il + qb[3] * ir) * i2;
            RR.enth = ((qa[4] + pl) * il + (qb[4] + pr) * ir) * i2;
            RR.u2v2w2 = RR.u * RR.u + RR.v * RR.v + RR.w * RR.w;
            const T1 a2 = gamma1 * (RR.enth - T1(0.5) * RR.u2v2w2);
            if (!(a2 > T1(0.0))) report_t(A, 2, I, pg, ph);
            RR.a = sqrt1<RF>(a2);
            RR.i2a = rcp1<RF>(T1(2.0) * RR.a);
            RR.g1a2 = gamma1 / a2;
            RR.euv = RR.enth - RR.u2v2w2;
            const T1 ur = qb[1] * iqb, ul = qa[1] * iqa;
            const T1 d1 = qb[1] - qa[1];
            T1 d2;
            if constexpr (PERT) d2 = (qb[1] * ur + ppb) - (qa[1] * ul + ppa);   // background part below
            else d2 = flux_jump_nc(qb[1], ur, pr, qa[1], ul, pl);
            const T1 d3 = ur * qb[2] - ul * qa[2];
            const T1 d4 = ur * qb[3] - ul * qa[3];
            T1 d5 = (qb[4] + pr) * ur - (qa[4] + pl) * ul;
            if (A.gravity) {
                if constexpr (PERT) {
                    // the background's own jump p_bar_b - p_bar_a - .5 (rho_bar_a +
                    // rho_bar_b) aux17 is the column residual (zero to FP64 rounding)
                    const T1 vg = cvg[kib];
                    d2 = d2 - T1(0.5) * (rpb + rpa) * vg + cres[kib];
                    d5 = d5 - T1(0.5) * (qb[1] + qa[1]) * vg;
                } else {
                    const T1 vg = vgc;
                    d2 = gravity_jump_nc(d2, qa[0], qb[0], vg);
                    d5 = d5 - T1(0.5) * (qb[1] + qa[1]) * vg;
                }
            }
            const T1 b4 = RR.g1a2 * (RR.euv * d1 + RR.u * d2 + RR.v * d3 + RR.w * d4 - d5);
            const T1 b2 = d3 - RR.v * d1;
            const T1 b3 = d4 - RR.w * d1;
            const T1 b5 = (d2 + (RR.a - RR.u) * d1 - RR.a * b4) * RR.i2a;
            const T1 b1 = d1 - b4 - b5;
            F1[0] = b1; F1[1] = b1 * (RR.u - RR.a); F1[2] = b1 * RR.v; F1[3] = b1 * RR.w;
            F1[4] = b1 * (RR.enth - RR.u * RR.a);
            s1 = RR.u - RR.a;
            F2[0] = b4; F2[1] = b4 * RR.u; F2[2] = b2 + b4 * RR.v; F2[3] = b3 + b4 * RR.w;
            F2[4] = b2 * RR.v + b3 * RR.w + b4 * T1(0.5) * RR.u2v2w2;
            s2 = RR.u;
            F3[0] = b5; F3[1] = b5 * (RR.u + RR.a); F3[2] = b5 * RR.v; F3[3] = b5 * RR.w;
            F3[4] = b5 * (RR.enth + RR.u * RR.a);
            s3 = RR.u + RR.a;
#pragma unroll
            for (int m = 5; m < NC; m++)
                F2[m] = (m - 4 <= A.nspecadv) ? s2 * (qb[m] - qa[m]) : ur * qb[m] - ul * qa[m];
            // entropy fix and amdq: rpn3euX.f:316-398
            {
                const T1 gmm = ga, gm1 = ga - T1(1.0);
                const T1 pim1 = gm1 * (qa[4] - T1(0.5) * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) * iqa);
                const T1 cim1 = sqrt1<RF>(gmm * pim1 * iqa);
                const T1 s0 = qa[1] * iqa - cim1;
                if (!(s0 >= T1(0.0) && s1 > T1(0.0))) {
                    const T1 is1 = rcp1<RF>(s1);
                    const T1 rho1 = qa[0] + F1[0] * is1;
                    const T1 rhou1 = qa[1] + F1[1] * is1;
                    const T1 rhov1 = qa[2] + F1[2] * is1;
                    const T1 rhow1 = qa[3] + F1[3] * is1;
                    const T1 en1 = qa[4] + F1[4] * is1;
                    const T1 ir1 = rcp1<RF>(rho1);
                    const T1 p1 = gm1 * (en1 - T1(0.5) * (rhou1 * rhou1 + rhov1 * rhov1 + rhow1 * rhow1) * ir1);
                    const T1 cc1 = sqrt1<RF>(gmm * p1 * ir1);
                    const T1 sa = rhou1 * ir1 - cc1;
                    T1 sfract;
                    if (s0 < T1(0.0) && sa > T1(0.0)) sfract = s0 * (sa - s1) / (sa - s0);
                    else if (s1 < T1(0.0)) sfract = s1;
                    else sfract = T1(0.0);
#pragma unroll
                    for (int m = 0; m < NC; m++) AM[m] = sfract * F1[m] * is1;
                    if (s2 < T1(0.0)) {
#pragma unroll
                        for (int m = 0; m < NC; m++) AM[m] = AM[m] + F2[m];
                        const T1 gm = gb, g1 = gb - T1(1.0);
                        const T1 pi_ = g1 * (qb[4] - T1(0.5) * (qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]) * iqb);
                        const T1 ci = sqrt1<RF>(gm * pi_ * iqb);
                        const T1 s3r = qb[1] * iqb + ci;
                        const T1 is3 = rcp1<RF>(s3);
                        const T1 rho2 = qb[0] - F3[0] * is3;
                        const T1 rhou2 = qb[1] - F3[1] * is3;
                        const T1 rhov2 = qb[2] - F3[2] * is3;
                        const T1 rhow2 = qb[3] - F3[3] * is3;
                        const T1 en2 = qb[4] - F3[4] * is3;
                        const T1 ir2 = rcp1<RF>(rho2);
                        const T1 p2 = g1 * (en2 - T1(0.5) * (rhou2 * rhou2 + rhov2 * rhov2 + rhow2 * rhow2) * ir2);
                        const T1 cc2 = sqrt1<RF>(gm * p2 * ir2);
                        const T1 sb = rhou2 * ir2 + cc2;
                        bool add = true;
                        if (sb < T1(0.0) && s3r > T1(0.0)) sfract = sb * (s3r - s3) / (s3r - sb);
                        else if (s3 < T1(0.0)) sfract = s3;
                        else add = false;
                        if (add) {
#pragma unroll
                            for (int m = 0; m < NC; m++) AM[m] = AM[m] + sfract * F3[m] * is3;
                        }
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < NC; m++) AP[m] = ((T1(0.0) + F1[m]) + F2[m]) + F3[m] - AM[m];
            c1 = fabs(s1 * dtdn); c2 = fabs(s2 * dtdn); c3 = fabs(s3 * dtdn);
            const T1 b2r = F2[2] - b4 * RR.v;
            f1 = b1 != T1(0.0); f2 = b2r != T1(0.0); f5 = b5 != T1(0.0);
            W[0] = b1; W[1] = b2r; W[2] = F2[3] - b4 * RR.w; W[3] = b4; W[4] = b5;
            s_del[0][t] = d1; s_del[1][t] = d2; s_del[2][t] = d3; s_del[3][t] = d4; s_del[4][t] = d5;
            s_s2[t] = s2;
#pragma unroll
            for (int k = 0; k < NTR; k++) s_trw[k][t] = F2[5 + k];
            s_flag[t] = (act2 && (f1 || f2 || f5)) ? 1 : 0;
            s_cfo[t] = f5 ? c3 : (f2 ? c2 : c1);
        };
        // FP32 form: the solve's FP64 Roe state and wave strengths (b1 .. b5),
        // kept for the limiters' Lax-Liu ratios, whose upwind wave strengths
        // cancel as the solve's do
        RoeT<double> RD = {0, 0, 0, 0, 1, 0, 0, 0, 0};
        double WB[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (act1) {
            if constexpr (sizeof(T) == 8) {
                riemann(R, fw1, fw2, fw3, amdq, apdq, WB);
            } else {
                RoeT<double> &RR = RD;
                // zero as the FP32 arrays above: the solve leaves the tracer
                // entries of waves 1 and 3 at zero
                double F1[NC], F2[NC], F3[NC], AM[NC], AP[NC];
#pragma unroll
                for (int m = 0; m < NC; m++) F1[m] = F2[m] = F3[m] = AM[m] = AP[m] = 0.0;
                riemann(RR, F1, F2, F3, AM, AP, WB);
#pragma unroll
                for (int m = 0; m < NC; m++) {
                    fw1[m] = (T)F1[m]; fw2[m] = (T)F2[m]; fw3[m] = (T)F3[m];
                    amdq[m] = (T)AM[m]; apdq[m] = (T)AP[m];
                }
                R = {(T)RR.u, (T)RR.v, (T)RR.w, (T)RR.enth, (T)RR.a, (T)RR.g1a2, (T)RR.euv, (T)RR.u2v2w2, (T)RR.i2a};
            }
        } else {
            s_flag[t] = 0;
        }
        __syncthreads();

        // ---------------- part 2: limiters and cqxx (rpn3euX.f:434-676, flux3fw.f:270-284)
        T cqxx[NC];
#pragma unroll
        for (int m = 0; m < NC; m++) cqxx[m] = T(0.0);
        if (act2) {
#pragma unroll
            for (int k = 0; k < NTR; k++) {
                const T fc = fw2[5 + k];
                if (fc != T(0.0)) {
                    const T r = (s2 > T(0.0)) ? s_trw[k][t - nb] / fc : s_trw[k][t + nb] / fc;
                    T wl;
                    if (k + 1 <= A.nspecadv) {
                        const int up = (int)(-copysign(T(1.0), s2));      // Id(2,i) - i
                        wl = cflsupbee(T(0.5), beta, r, c2, fabs(s_s2[t + up * nb] * dtdn));
                    } else {
                        wl = fmaxF(T(0.0), fminF(T(1.0), r));
                    }
                    fw2[5 + k] = wl * fc;
                }
            }
            // wave strengths and ratios in FP64 (TL); the limiter functions
            // in the form's type
            using TL = double;
            const RoeT<double> &RL = limiter_roe(R, RD);
            TL b1, b2, b3, b4, b5;
            if constexpr (sizeof(T) == 8) {
                b1 = fw1[0];
                b4 = fw2[0];
                b2 = fw2[2] - b4 * R.v;
                b3 = fw2[3] - b4 * R.w;
                b5 = fw3[0];
            } else {
                b1 = WB[0]; b2 = WB[1]; b3 = WB[2]; b4 = WB[3]; b5 = WB[4];
            }
            T wl1 = T(1.0), wl2 = T(1.0), wl3 = T(1.0), wl4 = T(1.0), wl5 = T(1.0);
            auto del = [&](int k, int up) { return s_del[k][t + up * nb]; };
            if (b1 != TL(0.0)) {
                const int up = (int)(-copysign(T(1.0), s1));
                const TL b4Id = RL.g1a2 * (RL.euv * del(0, up) + RL.u * del(1, up) + RL.v * del(2, up) +
                                           RL.w * del(3, up) - del(4, up));
                const TL b5Id = (del(1, up) + (RL.a - RL.u) * del(0, up) - RL.a * b4Id) * RL.i2a;
                const TL b1Id = del(0, up) - b4Id - b5Id;
                wl1 = suppow((T)(b1Id / b1), c1, c1);
            }
            if (b2 != TL(0.0)) {
                const int up = (int)(-copysign(T(1.0), s2));
                const TL b2Id = del(2, up) - RL.v * del(0, up);
                wl2 = cflsupbee(T(1.0), beta, (T)(b2Id / b2), c2, c2);
            }
            if (b3 != TL(0.0) || b4 != TL(0.0)) {
                T cflup;
                if (f2) cflup = c2;
                else if (f1) cflup = c1;
                else {
                    cflup = s_cprev[pa];
                    for (int j = c - 1; j >= 1; j--)
                        if (s_flag[t - (c - j) * nb]) { cflup = s_cfo[t - (c - j) * nb]; break; }
                }
                const int up = (int)(-copysign(T(1.0), s2));
                if (b3 != TL(0.0)) {
                    const TL b3Id = del(3, up) - RL.w * del(0, up);
                    wl3 = cflsupbee(T(1.0), beta, (T)(b3Id / b3), c2, cflup);
                }
                if (b4 != TL(0.0)) {
                    const TL b4Id = RL.g1a2 * (RL.euv * del(0, up) + RL.u * del(1, up) + RL.v * del(2, up) +
                                               RL.w * del(3, up) - del(4, up));
                    wl4 = cflsupbee(T(1.0), beta, (T)(b4Id / b4), c2, cflup);
                }
            }
            if (b5 != TL(0.0)) {
                const int up = (int)(-copysign(T(1.0), s3));
                const TL b4Id = RL.g1a2 * (RL.euv * del(0, up) + RL.u * del(1, up) + RL.v * del(2, up) +
                                           RL.w * del(3, up) - del(4, up));
                const TL b5Id = (del(1, up) + (RL.a - RL.u) * del(0, up) - RL.a * b4Id) * RL.i2a;
                wl5 = suppow((T)(b5Id / b5), c3, c3);
            }
#pragma unroll
            for (int m = 0; m < 5; m++) {
                fw1[m] = fw1[m] * wl1;
                fw3[m] = fw3[m] * wl5;
            }
            const TL b2l = wl2 * b2, b3l = wl3 * b3, b4l = wl4 * b4;
            fw2[0] = (T)b4l;
            fw2[1] = (T)(b4l * RL.u);
            fw2[2] = (T)(b2l + b4l * RL.v);
            fw2[3] = (T)(b3l + b4l * RL.w);
            fw2[4] = (T)(b2l * RL.v + b3l * RL.w + b4l * TL(0.5) * RL.u2v2w2);
            const T k1 = T(0.5) * copysign(T(1.0), s1) * (T(1.0) - fabs(s1) * dtdn);
            const T k2 = T(0.5) * copysign(T(1.0), s2) * (T(1.0) - fabs(s2) * dtdn);
            const T k3 = T(0.5) * copysign(T(1.0), s3) * (T(1.0) - fabs(s3) * dtdn);
#pragma unroll
            for (int m = 0; m < NC; m++) cqxx[m] = k1 * fw1[m] + k2 * fw2[m] + k3 * fw3[m];
            if constexpr (PERT)
                cflmax = fmaxF(cflmax, fmaxF(fmaxF(fabs((double)s1), fabs((double)s2)), fabs((double)s3)) * A.dtdn);
            else
                cflmax = fmaxF(cflmax, fmaxF(fmaxF(c1, c2), c3));
        }
        if (PREFETCH && chunk + 1 < nchunk) fetch(S + (CH - 2) + c);

        // ---------------- part 3: transverse terms, two phases
        // Each phase recomputes the two cqxx splits it needs, in an order that
        // keeps at most four split vectors alive next to the accumulator.
        // Each lane folds its increments into a register accumulator, added
        // to the tile in three barrier-separated sub-steps (one per a offset,
        // so no two lanes add to one entry at once): deterministic.  (Adding
        // each increment to the tile with shared-memory atomics instead, with
        // or without a two-waves-per-SIMD register budget, measured 3-5 %
        // slower in FP64 and 7x slower in FP32 on the full-domain workload.)
        // In FP64 with tracers the components go in groups, the Euler
        // components and then the tracers, at most three at a time, with one
        // accumulator alive at a time (at ten components it spilled a
        // kilobyte per lane, and the full-domain step with five tracers ran
        // 2.8x slower; with one tracer the groups hold 19 fewer AGPRs, and
        // the full-domain step runs 4 % faster, bit for bit the same).
        // Groups add to disjoint tile entries, so a group's first sub-step
        // needs no barrier after the previous group's last one.  FP32
        // accumulators fit without groups, and the extra barriers cost 3 %
        // there.
        if constexpr (NTR == 0 || sizeof(T) == 4) {
            for (int phase = 0; phase < 2; phase++) {
                // phase 0 (L): interface I -> cell I-1; phase 1 (R): -> cell I
                const bool on = act2 && (phase == 0 ? I >= 2 : I <= n);
                const int ci = phase == 0 ? c - 1 : c;  // cell index within the chunk's tile
                auto transverse = [&](auto &acc) {
                    const T sgn = phase == 0 ? T(1.0) : -T(1.0);   // +3 cqxx mixing for amdq, -3 for apdq
                    const T cc = phase == 0 ? -d : d;        // sign of the correction-wave term
                    T asd[NC];
#pragma unroll
                    for (int m = 0; m < NC; m++) asd[m] = phase == 0 ? amdq[m] : apdq[m];
                    {
                        // C splits (h direction): 2-D part, correction, then the
                        // g-direction corner terms of c*asdq +- 3 c*cqxx (flux3fw.f:388-412)
                        T Cm[NC], Cp[NC], Cmq[NC], Cpq[NC];
                        tsplit<NC, 3, false>(R, cqxx, Cmq, Cpq);
                        tsplit<NC, 3, false>(R, asd, Cm, Cp);
                        acc.h(1, 0, -T(0.5) * d * ch, Cm); acc.h(2, 0, -T(0.5) * d * ch, Cp);
                        acc.h(1, 0, cc * ch, Cmq);      acc.h(2, 0, cc * ch, Cpq);
                        axpy3(Cp, Cp, T(3.0) * sgn, Cpq);
                        axpy3(Cm, Cm, T(3.0) * sgn, Cmq);
                        tsplit<NC, 2, true>(R, Cp, Cmq, Cpq);      // B-, B+ of c+asdq2
                        acc.g(2, 0, k6g * cg, Cpq);  acc.g(1, 0, k6g * cg, Cmq);
                        acc.g(2, 1, -k6g * cg, Cpq); acc.g(1, 1, -k6g * cg, Cmq);
                        tsplit<NC, 2, true>(R, Cm, Cmq, Cpq);      // B-, B+ of c-asdq2
                        acc.g(2, 0, -k6g * cg, Cpq); acc.g(1, 0, -k6g * cg, Cmq);
                        acc.g(2, -1, k6g * cg, Cpq); acc.g(1, -1, k6g * cg, Cmq);
                    }
                    {
                        // B splits (g direction): 2-D part, correction, then the
                        // h-direction corner terms of b*asdq +- 3 b*cqxx (flux3fw.f:510-538)
                        T Bm[NC], Bp[NC], Bmq[NC], Bpq[NC];
                        tsplit<NC, 2, false>(R, cqxx, Bmq, Bpq);
                        tsplit<NC, 2, false>(R, asd, Bm, Bp);
                        acc.g(1, 0, -T(0.5) * d * cg, Bm); acc.g(2, 0, -T(0.5) * d * cg, Bp);
                        acc.g(1, 0, cc * cg, Bmq);      acc.g(2, 0, cc * cg, Bpq);
                        axpy3(Bp, Bp, T(3.0) * sgn, Bpq);
                        axpy3(Bm, Bm, T(3.0) * sgn, Bmq);
                        tsplit<NC, 3, true>(R, Bp, Bmq, Bpq);      // C-, C+ of b+asdq
                        acc.h(2, 0, k6h * ch, Bpq);  acc.h(1, 0, k6h * ch, Bmq);
                        acc.h(2, 1, -k6h * ch, Bpq); acc.h(1, 1, -k6h * ch, Bmq);
                        tsplit<NC, 3, true>(R, Bm, Bmq, Bpq);      // C-, C+ of b-asdq
                        acc.h(2, 0, -k6h * ch, Bpq); acc.h(1, 0, -k6h * ch, Bmq);
                        acc.h(2, -1, k6h * ch, Bpq); acc.h(1, -1, k6h * ch, Bmq);
                    }
                    if constexpr (sizeof(T) == 4) {
                        // The FP32 form forms the cell's own normal update in
                        // FP64 with the unrounded dt/dn and rounds it once: the
                        // second-order correction nearly cancels the fluctuation
                        // in smooth flow, and rounding the two products apart
                        // (with dt/dn in FP32) grew, over some 300 steps, into
                        // an error of T' of 1e-4 of its peak in the thin upper
                        // layers of weak plumes.
                        const double dD = A.dtdn;
#pragma unroll
                        for (int m = 0; m < NC; m++)
                            acc.c11(m, (T)(phase == 0 ? -dD * (double)amdq[m] - dD * (double)cqxx[m]
                                                      : -dD * (double)apdq[m] + dD * (double)cqxx[m]));
                    } else {
#pragma unroll
                        for (int m = 0; m < NC; m++)
                            acc.c11(m, phase == 0 ? -d * amdq[m] - d * cqxx[m] : -d * apdq[m] + d * cqxx[m]);
                    }
                };
                Acc<NC, T> acc;
                acc.zero();
                if (on) transverse(acc);
                // one sub-step per offset along a: lanes then add to distinct T entries
#pragma unroll
                for (int da = 0; da < 3; da++) {
                    if (on) {
                        // the accumulator is indexed [g offset][h offset]; both
                        // branches keep every index a compile-time constant so
                        // it stays in registers
                        if (a_is_g) {
#pragma unroll
                            for (int db = 0; db < 3; db++)
#pragma unroll
                                for (int m = 0; m < NC; m++) LT[tix(m, db, pa + da, ci)] += acc.D[da][db][m];
                        } else {
#pragma unroll
                            for (int db = 0; db < 3; db++)
#pragma unroll
                                for (int m = 0; m < NC; m++) LT[tix(m, db, pa + da, ci)] += acc.D[db][da][m];
                        }
                    }
                    // With tile-axis-major lanes a wave holds every pencil of
                    // its chunk cells, so the entries one phase's sub-steps add
                    // to are the wave's own (cells c-1 or c of its lanes): only
                    // the wave's program order separates its sub-steps (the
                    // fence keeps the compiler from moving one sub-step's tile
                    // reads above the previous one's writes; it emits nothing),
                    // and the workgroup waits once per phase.  Pencil-major
                    // waves share entries: a barrier after every sub-step.
                    if (PM || da == 2) __syncthreads();
                    else __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    if (phase == 0 && da == (PM ? 0 : 2) && t < TA) {
                        // CFLup carried into the next chunk: the last assignment
                        // among this chunk's interfaces (every lane is past part
                        // 2, which reads s_cprev, once this barrier is passed)
                        for (int j = CH - 2; j >= 1; j--) {
                            const int tt = PM ? t * CH + j : j * TA + t;
                            if (s_flag[tt]) { s_cprev[t] = s_cfo[tt]; break; }
                        }
                    }
                }
            }
        } else {
            for (int phase = 0; phase < 2; phase++) {
                // phase 0 (L): interface I -> cell I-1; phase 1 (R): -> cell I
                const bool on = act2 && (phase == 0 ? I >= 2 : I <= n);
                const int ci = phase == 0 ? c - 1 : c;  // cell index within the chunk's tile
                const T sgn = phase == 0 ? T(1.0) : -T(1.0);   // +3 cqxx mixing for amdq, -3 for apdq
                const T cc = phase == 0 ? -d : d;        // sign of the correction-wave term
                // components M0 .. M0+GW-1; `final`: the phase's last group
                auto group = [&](auto m0c, auto gwc, bool final) {
                    constexpr int M0 = decltype(m0c)::value, GW = decltype(gwc)::value;
                    Acc<GW, T> acc;
                    acc.zero();
                    if (on) {
                        T asd[GW];
#pragma unroll
                        for (int m = 0; m < GW; m++) asd[m] = phase == 0 ? amdq[M0 + m] : apdq[M0 + m];
                        const T(&cq)[GW] = *reinterpret_cast<const T(*)[GW]>(cqxx + M0);
                        {
                            // C splits (h direction): 2-D part, correction, then the
                            // g-direction corner terms of c*asdq +- 3 c*cqxx (flux3fw.f:388-412)
                            T Cm[GW], Cp[GW], Cmq[GW], Cpq[GW];
                            tsplit_group<M0, GW, 3, false>(R, cq, Cmq, Cpq);
                            tsplit_group<M0, GW, 3, false>(R, asd, Cm, Cp);
                            acc.h(1, 0, -T(0.5) * d * ch, Cm); acc.h(2, 0, -T(0.5) * d * ch, Cp);
                            acc.h(1, 0, cc * ch, Cmq);      acc.h(2, 0, cc * ch, Cpq);
                            axpy3(Cp, Cp, T(3.0) * sgn, Cpq);
                            axpy3(Cm, Cm, T(3.0) * sgn, Cmq);
                            tsplit_group<M0, GW, 2, true>(R, Cp, Cmq, Cpq);      // B-, B+ of c+asdq2
                            acc.g(2, 0, k6g * cg, Cpq);  acc.g(1, 0, k6g * cg, Cmq);
                            acc.g(2, 1, -k6g * cg, Cpq); acc.g(1, 1, -k6g * cg, Cmq);
                            tsplit_group<M0, GW, 2, true>(R, Cm, Cmq, Cpq);      // B-, B+ of c-asdq2
                            acc.g(2, 0, -k6g * cg, Cpq); acc.g(1, 0, -k6g * cg, Cmq);
                            acc.g(2, -1, k6g * cg, Cpq); acc.g(1, -1, k6g * cg, Cmq);
                        }
                        {
                            // B splits (g direction): 2-D part, correction, then the
                            // h-direction corner terms of b*asdq +- 3 b*cqxx (flux3fw.f:510-538)
                            T Bm[GW], Bp[GW], Bmq[GW], Bpq[GW];
                            tsplit_group<M0, GW, 2, false>(R, cq, Bmq, Bpq);
                            tsplit_group<M0, GW, 2, false>(R, asd, Bm, Bp);
                            acc.g(1, 0, -T(0.5) * d * cg, Bm); acc.g(2, 0, -T(0.5) * d * cg, Bp);
                            acc.g(1, 0, cc * cg, Bmq);      acc.g(2, 0, cc * cg, Bpq);
                            axpy3(Bp, Bp, T(3.0) * sgn, Bpq);
                            axpy3(Bm, Bm, T(3.0) * sgn, Bmq);
                            tsplit_group<M0, GW, 3, true>(R, Bp, Bmq, Bpq);      // C-, C+ of b+asdq
                            acc.h(2, 0, k6h * ch, Bpq);  acc.h(1, 0, k6h * ch, Bmq);
                            acc.h(2, 1, -k6h * ch, Bpq); acc.h(1, 1, -k6h * ch, Bmq);
                            tsplit_group<M0, GW, 3, true>(R, Bm, Bmq, Bpq);      // C-, C+ of b-asdq
                            acc.h(2, 0, -k6h * ch, Bpq); acc.h(1, 0, -k6h * ch, Bmq);
                            acc.h(2, -1, k6h * ch, Bpq); acc.h(1, -1, k6h * ch, Bmq);
                        }
#pragma unroll
                        for (int m = 0; m < GW; m++)
                            acc.c11(m, phase == 0 ? -d * amdq[M0 + m] - d * cqxx[M0 + m] : -d * apdq[M0 + m] + d * cqxx[M0 + m]);
                    }
                    // one sub-step per offset along a: lanes then add to distinct T entries
#pragma unroll
                    for (int da = 0; da < 3; da++) {
                        if (on) {
                            // the accumulator is indexed [g offset][h offset]; both
                            // branches keep every index a compile-time constant so
                            // it stays in registers
                            if (a_is_g) {
#pragma unroll
                                for (int db = 0; db < 3; db++)
#pragma unroll
                                    for (int m = 0; m < GW; m++) LT[tix(M0 + m, db, pa + da, ci)] += acc.D[da][db][m];
                            } else {
#pragma unroll
                                for (int db = 0; db < 3; db++)
#pragma unroll
                                    for (int m = 0; m < GW; m++) LT[tix(M0 + m, db, pa + da, ci)] += acc.D[db][da][m];
                            }
                        }
                        // as in the branch above: tile-axis-major waves wait
                        // once per phase (after its final group), pencil-major
                        // ones after every sub-step but a group's last
                        if (PM ? (da < 2 || (final && phase == 1)) : (da == 2 && final)) __syncthreads();
                        else if (!PM) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        if (phase == 0 && (PM ? (M0 == 0 && da == 0) : (final && da == 2)) && t < TA) {
                            // CFLup carried into the next chunk: the last assignment
                            // among this chunk's interfaces (every lane is past part
                            // 2, which reads s_cprev, once this barrier is passed)
                            for (int j = CH - 2; j >= 1; j--) {
                                const int tt = PM ? t * CH + j : j * TA + t;
                                if (s_flag[tt]) { s_cprev[t] = s_cfo[tt]; break; }
                            }
                        }
                    }
                };
                if constexpr (NTR <= 3) {
                    group(IConst<0>{}, IConst<5>{}, false);
                    group(IConst<5>{}, IConst<NTR>{}, true);
                } else {
                    group(IConst<0>{}, IConst<5>{}, false);
                    group(IConst<5>{}, IConst<3>{}, false);
                    group(IConst<8>{}, IConst<NTR - 3>{}, true);
                }
            }
        }

        // ---------------- flush the tile to q (one read-modify-write per row)
        // Each lane owns PPT (a offset, cell) positions of every (m, db) plane
        // of the tile; m and db are compile-time so the plane bases come from
        // scalar registers.  The flush is branch-free so that all loads are
        // in flight together and no store waits for another: a position
        // outside the grid reads and writes the lane's slot of A.sink, and
        // its tile entry is the shared zero slot.
        {
            constexpr int NPOS = TROW * NCELL;
            constexpr int PPT = (NPOS + NT - 1) / NT;
            constexpr int ZERO = NC * 3 * NPOS;
            long long poff[PPT];
            bool pok[PPT];
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const int e = t + k * NT;
                const int ai = PM ? e / NCELL : e % TROW, ci = PM ? e % NCELL : e / TROW;
                const int cell = S + ci;                // Fortran index along the pencil
                const int aa = A0 - 1 + ai;             // tile-axis coordinate of the target row
                pok[k] = e < NPOS && cell >= 1 && cell <= n && aa <= A.nA + 2;
                poff[k] = (long long)(cell + A.mbc - 1) * A.sN + (long long)(aa + A.mbc - 1) * A.sA;
            }
            if constexpr (tile_flush_atomic<T>()) {
#pragma unroll
                for (int m = 0; m < NC; m++)
#pragma unroll
                    for (int db = 0; db < 3; db++) {
                        const gptr<T> base =
                            qnew + (long long)A.comp[m] * A.P + (long long)(b + db - 1 + A.mbc - 1) * A.sB;
#pragma unroll
                        for (int k = 0; k < PPT; k++) {
                            const int e = t + k * NT;
                            const int li = e < NPOS ? (m * 3 + db) * NPOS + e : ZERO;
                            const T v = LT[li];
                            LT[li] = T(0.0);
                            if (pok[k]) add_noret(base + poff[k], v);
                        }
                    }
            } else {
            const gptr<T> sink = to_global(A.sink) + t;
            T val[NC][3][PPT];
#pragma unroll
            for (int m = 0; m < NC; m++)
#pragma unroll
                for (int db = 0; db < 3; db++) {
                    const gptr<T> base = qnew + (long long)A.comp[m] * A.P + (long long)(b + db - 1 + A.mbc - 1) * A.sB;
#pragma unroll
                    for (int k = 0; k < PPT; k++) val[m][db][k] = *(pok[k] ? base + poff[k] : sink);
                }
#pragma unroll
            for (int m = 0; m < NC; m++)
#pragma unroll
                for (int db = 0; db < 3; db++)
#pragma unroll
                    for (int k = 0; k < PPT; k++) {
                        const int e = t + k * NT;
                        const int li = e < NPOS ? (m * 3 + db) * NPOS + e : ZERO;   // = tix(m, db, ai, ci)
                        val[m][db][k] = val[m][db][k] + LT[li];
                        LT[li] = T(0.0);
                    }
#pragma unroll
            for (int m = 0; m < NC; m++)
#pragma unroll
                for (int db = 0; db < 3; db++) {
                    const gptr<T> base = qnew + (long long)A.comp[m] * A.P + (long long)(b + db - 1 + A.mbc - 1) * A.sB;
#pragma unroll
                    for (int k = 0; k < PPT; k++) *(pok[k] ? base + poff[k] : sink) = val[m][db][k];
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) cflmax = fmaxF(cflmax, __shfl_xor(cflmax, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(A.cfl_bits, (unsigned long long)__double_as_longlong(cflmax));
}


// Waves per SIMD the register allocation aims at: two for the FP32 form
// with tile-axis-major lanes and at most one tracer (left to themselves
// these kernels miss 256 registers by a few, since the Riemann solve runs in
// FP64, and halve the occupancy: 50 % slower on the full-domain FP32 step),
// one otherwise.
template <typename T, int NTR, int SH>
constexpr int tile_waves()
{
    if (sizeof(T) != 4 || TileShape<SH>::PM) return 1;
    return NTR <= 1 ? 2 : 1;
}

// The sweep kernel, for one grid and for a batch of grids alike (one
// compiled body with one call site, so both round alike), in two modes:
//  * work list (grid.x = 0): a persistent grid whose workgroups take tiles
//    from a work list (item = {argument block, a_is_g, bx, by}) through the
//    atomic counter next[0] until it runs out; tiles of grids whose control
//    flag (*run) is 0 in this attempt are skipped.  next[1] counts the
//    workgroups that are done; the last one sets both back to 0 for the
//    next launch (every other workgroup has taken its last item by then);
//  * one tile per workgroup (grid.x > 0 tiles along a): workgroup b takes
//    tile {argument block grid.y, a_is_g grid.z, b mod grid.x, b / grid.x},
//    as the hardware dispatches them, so that the launches of two streams
//    share the compute units tile by tile (one grid's sweep overlap).
template <typename T, bool PERT, int NTR, int SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(tile_waves<T, NTR, SH>())))
void tile_sweep_list(const SweepArgsT<T> *__restrict__ args, const int4 *items, int nitems,
                                                        const int *const *run, int *next, int4 grid)
{
    __shared__ int s_item;
    bool first = true;
    for (;;) {
        int4 w;
        if (grid.x > 0) {
            if (!first) break;
            first = false;
            w = make_int4(grid.y, grid.z, (int)blockIdx.x % grid.x, (int)blockIdx.x / grid.x);
        } else {
            __syncthreads();   // the previous tile's readers of s_item and of the tile arrays are done
            if (threadIdx.x == 0) s_item = atomicAdd(next, 1);
            __syncthreads();
            const int it = s_item;
            if (it >= nitems) break;
            w = items[it];
        }
        // the item is the same for the whole workgroup: scalar, so that the
        // argument block is read with scalar loads
        const int ai = __builtin_amdgcn_readfirstlane(w.x);
        if (*run[ai] == 0) continue;
        tile_sweep_body<T, PERT, NTR, SH>(args[ai], __builtin_amdgcn_readfirstlane(w.y),
                                          __builtin_amdgcn_readfirstlane(w.z), __builtin_amdgcn_readfirstlane(w.w));
    }
    if (grid.x == 0 && threadIdx.x == 0) {
        __threadfence();   // this workgroup's last take precedes its count
        if (atomicAdd(next + 1, 1) == (int)gridDim.x - 1) {
            __threadfence();
            atomicExch(next, 0);
            atomicExch(next + 1, 0);
        }
    }
}

}  // namespace magic
