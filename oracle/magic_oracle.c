/*
 * magic_oracle.c -- CPU restatement of the MAGIC Clawpack 3-D f-wave step
 * (Euler + gravity + tracers), written from the reference's documented
 * algorithm with the same floating-point operation order, so that it matches
 * the compiled reference bit for bit when both are built without FMA
 * contraction (gcc -O2 -ffp-contract=off here, flang -O2 -ffp-contract=off for
 * the reference).
 *
 * TEST INFRASTRUCTURE ONLY: this is the parity checker for the HIP library and
 * the CPU baseline leg of bench.py.  The product path never calls it.
 *
 * Every function cites the reference file:line it restates.  Fortran MIN/MAX
 * are restated as (a<b?a:b) / (a>b?a:b), which is what flang emits (checked on
 * signed zeros and NaN); DSIGN(1,x) is copysign(1,x).
 *
 * Deliberate, documented differences from the reference (each affects only
 * values the reference leaves undefined):
 *   - rpn3's CFLup (rpn3euX.f:536-649) is read before any assignment when the
 *     first limited interface has b1 = b2 = 0 and b3 or b4 != 0; it starts at
 *     0.0 here.
 *   - rpn3's CFL(2,mx+2) (used by the tracer limiter, rpn3euX.f:453-454, when
 *     s(2,mx+1) <= 0) is never assigned by the reference (loop 240 stops at
 *     mx+1); it is computed here like the other CFL entries.
 *   - conduct's z-ghost loop runs i up to mz+mbc (conduct.f:86-91).  When
 *     mz > mx the reference writes past the end of its local array u; here u
 *     carries slack so the same linear-address writes stay in bounds.
 *   - bc3's y-lower wall (bc3_mpi.f:366-370) indexes q out of order and the
 *     z-periodic branch copies back an unfilled buffer (bc3_mpi.f:583-588 vs
 *     782-794); here both do what their comments state.
 */
#include "magic_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double fminF(double a, double b) { return a < b ? a : b; }
static inline double fmaxF(double a, double b) { return a > b ? a : b; }
static inline double dsign1(double x) { return copysign(1.0, x); }

/* ------------------------------------------------------------------------ */
/* Pencil helpers: Fortran index i in [1-mbc, n+mbc] -> 0-based i+mbc-1.     */
#define PI_(i) ((i) + mbc - 1)
#define Q1(m, i) q1d[(size_t)PI_(i) * meqn + ((m) - 1)]
#define FW(m, w, i) fwave[((size_t)PI_(i) * 3 + ((w) - 1)) * meqn + ((m) - 1)]
#define S(w, i) s[(size_t)PI_(i) * 3 + ((w) - 1)]
#define V1(arr, m, i) arr[(size_t)PI_(i) * meqn + ((m) - 1)]

enum { R_U2V2W2 = 0, R_U, R_V, R_W, R_ENTH, R_A, R_G1A2, R_EUV };
#define ROE(f, i) roe[(size_t)(f) * N + PI_(i)]

/* ------------------------------------------------------------------------ */
/* Limiters: rpn3euX.f:686-705 (cflsupbee) and 709-731 (suppow).            */
double orc_cflsupbee(double theta, double beta, double r, double cfl, double cflup)
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

double orc_suppow(double r, double cfl, double cflup)
{
    double cfle = fmaxF(0.001, fminF(0.999, cfl));
    double cflupe = fmaxF(0.001, fminF(0.999, cflup));
    double s2 = (1.0 + cfle) / 3.0;
    double s3 = 1.0 - s2;
    double weight = (1.0 - cflupe) / (1.0 - cfle);
    double pp;
    if (r <= 1.0)
        pp = weight * (2.0 / cfle) * 2.0 * s3;
    else
        pp = fabs(2.0 / (1.0 - cfle)) * 2.0 * s2;
    double rabs = fabs(r);
    double rfrac = fabs((1.0 - rabs) / (1.0 + rabs));
    double signum = fmaxF(0.0, dsign1(r));
    return signum * (s3 + s2 * r) * (1.0 - pow(rfrac, pp));
}

static void comp_map(int ixyz, int *mu, int *mv, int *mw)
{
    /* rpn3euX.f:109-121, rpt3euX.f:90-102 */
    if (ixyz == 1) { *mu = 2; *mv = 3; *mw = 4; }
    else if (ixyz == 2) { *mu = 3; *mv = 4; *mw = 2; }
    else { *mu = 4; *mv = 2; *mw = 3; }
}

/* ------------------------------------------------------------------------ */
/* rpn3: rpn3euX.f:3-681 with efix = laxliu = thirdorder = .true.,
 * nonlocal = .false. (rpn3euX.f:76-79), ql = qr = q1d, auxl = auxr.        */
int orc_rpn3(int ixyz, int meqn, int mbc, int n, int nspecadv,
             const double *q1d, const double *aux19, const double *aux17,
             double dtcom, double dxcom,
             double *fwave, double *s, double *amdq, double *apdq, double *roe)
{
    const int N = n + 2 * mbc;
    const double theta = 1.0, beta = 0.666666666666666;
    int mu, mv, mw;
    comp_map(ixyz, &mu, &mv, &mw);

    double *delta = (double *)malloc(sizeof(double) * 5 * N);
    double *dtdx = (double *)malloc(sizeof(double) * N);
    int *Id = (int *)malloc(sizeof(int) * 3 * N);
    double *CFL = (double *)malloc(sizeof(double) * 3 * N);
#define DEL(k, i) delta[(size_t)PI_(i) * 5 + ((k) - 1)]
#define IDX(w, i) Id[(size_t)PI_(i) * 3 + ((w) - 1)]
#define CF(w, i) CFL[(size_t)PI_(i) * 3 + ((w) - 1)]
#define A19(i) aux19[PI_(i)]
#define A17(i) aux17[PI_(i)]
    int status = ORC_OK;

    /* Roe averages, jumps, wave decomposition: rpn3euX.f:138-271 */
    for (int i = 2 - mbc; i <= n + mbc; i++) {
        if (Q1(1, i - 1) <= 0.0 || Q1(1, i) <= 0.0) { status = ORC_RHO_NONPOS; goto done; }
        dtdx[PI_(i)] = dtcom / dxcom;
        double rhsqrtl = sqrt(Q1(1, i - 1));
        double rhsqrtr = sqrt(Q1(1, i));
        double rhsq2 = rhsqrtl + rhsqrtr;
        double gamma = (rhsqrtr * A19(i) + rhsqrtl * A19(i - 1)) / rhsq2;
        double gamma1 = gamma - 1.0;
        double pl = (A19(i - 1) - 1.0) * (Q1(5, i - 1) - 0.5 * (Q1(mu, i - 1) * Q1(mu, i - 1) +
                     Q1(mv, i - 1) * Q1(mv, i - 1) + Q1(mw, i - 1) * Q1(mw, i - 1)) / Q1(1, i - 1));
        double pr = (A19(i) - 1.0) * (Q1(5, i) - 0.5 * (Q1(mu, i) * Q1(mu, i) +
                     Q1(mv, i) * Q1(mv, i) + Q1(mw, i) * Q1(mw, i)) / Q1(1, i));
        double u = (Q1(mu, i - 1) / rhsqrtl + Q1(mu, i) / rhsqrtr) / rhsq2;
        double v = (Q1(mv, i - 1) / rhsqrtl + Q1(mv, i) / rhsqrtr) / rhsq2;
        double w = (Q1(mw, i - 1) / rhsqrtl + Q1(mw, i) / rhsqrtr) / rhsq2;
        double enth = (((Q1(5, i - 1) + pl) / rhsqrtl + (Q1(5, i) + pr) / rhsqrtr)) / rhsq2;
        double u2v2w2 = u * u + v * v + w * w;
        double a2 = gamma1 * (enth - 0.5 * u2v2w2);
        if (a2 <= 0.0) { status = ORC_A2_NONPOS; goto done; }
        double a = sqrt(a2);
        double g1a2 = gamma1 / a2;
        double euv = enth - u2v2w2;
        ROE(R_U2V2W2, i) = u2v2w2; ROE(R_U, i) = u; ROE(R_V, i) = v; ROE(R_W, i) = w;
        ROE(R_ENTH, i) = enth; ROE(R_A, i) = a; ROE(R_G1A2, i) = g1a2; ROE(R_EUV, i) = euv;

        double ur = Q1(mu, i) / Q1(1, i);
        double ul = Q1(mu, i - 1) / Q1(1, i - 1);
        double d1 = Q1(mu, i) - Q1(mu, i - 1);
        double d2 = (Q1(mu, i) * ur + pr) - (Q1(mu, i - 1) * ul + pl);
        double d3 = ur * Q1(mv, i) - ul * Q1(mv, i - 1);
        double d4 = ur * Q1(mw, i) - ul * Q1(mw, i - 1);
        double d5 = (Q1(5, i) + pr) * ur - (Q1(5, i - 1) + pl) * ul;
        if (ixyz == 3) { /* well-balanced gravity, rpn3euX.f:202-208 */
            d2 = d2 - 0.5 * (Q1(1, i) + Q1(1, i - 1)) * A17(i);
            d5 = d5 - 0.5 * (Q1(mu, i) + Q1(mu, i - 1)) * A17(i);
        }
        DEL(1, i) = d1; DEL(2, i) = d2; DEL(3, i) = d3; DEL(4, i) = d4; DEL(5, i) = d5;
        double b4 = g1a2 * (euv * d1 + u * d2 + v * d3 + w * d4 - d5);
        double b2 = d3 - v * d1;
        double b3 = d4 - w * d1;
        double b5 = (d2 + (a - u) * d1 - a * b4) / (2.0 * a);
        double b1 = d1 - b4 - b5;

        FW(1, 1, i) = b1;
        FW(mu, 1, i) = b1 * (u - a);
        FW(mv, 1, i) = b1 * v;
        FW(mw, 1, i) = b1 * w;
        FW(5, 1, i) = b1 * (enth - u * a);
        S(1, i) = u - a;

        FW(1, 2, i) = b4;
        FW(mu, 2, i) = b4 * u;
        FW(mv, 2, i) = b2 + b4 * v;
        FW(mw, 2, i) = b3 + b4 * w;
        FW(5, 2, i) = b2 * v + b3 * w + b4 * 0.5 * u2v2w2;
        S(2, i) = u;

        FW(1, 3, i) = b5;
        FW(mu, 3, i) = b5 * (u + a);
        FW(mv, 3, i) = b5 * v;
        FW(mw, 3, i) = b5 * w;
        FW(5, 3, i) = b5 * (enth + u * a);
        S(3, i) = u + a;

        for (int m = 6; m <= meqn; m++) {
            FW(m, 1, i) = 0.0;
            FW(m, 3, i) = 0.0;
            if ((m - 5) <= nspecadv)
                FW(m, 2, i) = S(2, i) * (Q1(m, i) - Q1(m, i - 1));
            else
                FW(m, 2, i) = ur * Q1(m, i) - ul * Q1(m, i - 1);
        }
    }

    /* Entropy fix and amdq: rpn3euX.f:311-399 */
    for (int i = 2 - mbc; i <= n + mbc; i++) {
        double gamma = A19(i - 1);
        double gamma1 = gamma - 1.0;
        double rhoim1 = Q1(1, i - 1);
        double pim1 = gamma1 * (Q1(5, i - 1) - 0.5 * (Q1(mu, i - 1) * Q1(mu, i - 1) +
                      Q1(mv, i - 1) * Q1(mv, i - 1) + Q1(mw, i - 1) * Q1(mw, i - 1)) / rhoim1);
        double cim1 = sqrt(gamma * pim1 / rhoim1);
        double s0 = Q1(mu, i - 1) / rhoim1 - cim1;
        if (s0 >= 0.0 && S(1, i) > 0.0) {
            for (int m = 1; m <= meqn; m++) V1(amdq, m, i) = 0.0;
            continue;
        }
        double rho1 = Q1(1, i - 1) + FW(1, 1, i) / S(1, i);
        double rhou1 = Q1(mu, i - 1) + FW(mu, 1, i) / S(1, i);
        double rhov1 = Q1(mv, i - 1) + FW(mv, 1, i) / S(1, i);
        double rhow1 = Q1(mw, i - 1) + FW(mw, 1, i) / S(1, i);
        double en1 = Q1(5, i - 1) + FW(5, 1, i) / S(1, i);
        double p1 = gamma1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1 + rhow1 * rhow1) / rho1);
        double c1 = sqrt(gamma * p1 / rho1);
        double s1 = rhou1 / rho1 - c1;
        double sfract;
        if (s0 < 0.0 && s1 > 0.0)
            sfract = s0 * (s1 - S(1, i)) / (s1 - s0);
        else if (S(1, i) < 0.0)
            sfract = S(1, i);
        else
            sfract = 0.0;
        for (int m = 1; m <= meqn; m++) V1(amdq, m, i) = sfract * FW(m, 1, i) / S(1, i);
        if (S(2, i) >= 0.0) continue;
        for (int m = 1; m <= meqn; m++) V1(amdq, m, i) = V1(amdq, m, i) + FW(m, 2, i);
        gamma = A19(i);
        gamma1 = gamma - 1.0;
        double rhoi = Q1(1, i);
        double pi = gamma1 * (Q1(5, i) - 0.5 * (Q1(mu, i) * Q1(mu, i) +
                    Q1(mv, i) * Q1(mv, i) + Q1(mw, i) * Q1(mw, i)) / rhoi);
        double ci = sqrt(gamma * pi / rhoi);
        double s3 = Q1(mu, i) / rhoi + ci;
        double rho2 = Q1(1, i) - FW(1, 3, i) / S(3, i);
        double rhou2 = Q1(mu, i) - FW(mu, 3, i) / S(3, i);
        double rhov2 = Q1(mv, i) - FW(mv, 3, i) / S(3, i);
        double rhow2 = Q1(mw, i) - FW(mw, 3, i) / S(3, i);
        double en2 = Q1(5, i) - FW(5, 3, i) / S(3, i);
        double p2 = gamma1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2 + rhow2 * rhow2) / rho2);
        double c2 = sqrt(gamma * p2 / rho2);
        double s2 = rhou2 / rho2 + c2;
        if (s2 < 0.0 && s3 > 0.0)
            sfract = s2 * (s3 - S(3, i)) / (s3 - s2);
        else if (S(3, i) < 0.0)
            sfract = S(3, i);
        else
            continue;
        for (int m = 1; m <= meqn; m++)
            V1(amdq, m, i) = V1(amdq, m, i) + sfract * FW(m, 3, i) / S(3, i);
    }
    /* apdq = sum of f-waves - amdq: rpn3euX.f:404-411 */
    for (int m = 1; m <= meqn; m++)
        for (int i = 2 - mbc; i <= n + mbc; i++) {
            double df = 0.0;
            for (int mws = 1; mws <= 3; mws++) df = df + FW(m, mws, i);
            V1(apdq, m, i) = df - V1(amdq, m, i);
        }
    /* Upwind index and local Courant numbers: rpn3euX.f:417-421 (extended to
     * n+2 so that CFL(2,n+2) is defined, see the file header). */
    for (int i = 0; i <= n + 2; i++)
        for (int mws = 1; mws <= 3; mws++) {
            if (i <= n + 1) IDX(mws, i) = (int)(i - dsign1(S(mws, i)));
            int iu = (i <= n + 1) ? IDX(mws, i) : i;
            int ic = i > iu ? i : iu;
            if (ic > n + mbc) ic = n + mbc;
            CF(mws, i) = fabs(S(mws, i) * dtdx[PI_(ic)]);
        }

    /* Tracer limiters: rpn3euX.f:434-467 */
    if (meqn > 5) {
        double wlimitr = 0.0;
        for (int m = 6; m <= meqn; m++) {
            double fwavec = 0.0;
            for (int i = 0; i <= n + 1; i++) {
                double fwavel = fwavec;
                fwavec = FW(m, 2, i);
                double fwaver = FW(m, 2, i + 1);
                if (i == 0) continue;
                if (fwavec == 0.0) continue;
                double r;
                if (S(2, i) > 0.0) r = fwavel / fwavec;
                else r = fwaver / fwavec;
                if ((m - 5) <= nspecadv)
                    wlimitr = orc_cflsupbee(0.5, beta, r, CF(2, i), CF(2, IDX(2, i)));
                else
                    wlimitr = fmaxF(0.0, fminF(1.0, r));
                FW(m, 2, i) = wlimitr * FW(m, 2, i);
            }
        }
    }

    /* Lax-Liu limiting of waves 1-5 and recombination: rpn3euX.f:482-676 */
    {
        double wl1 = 0.0, wl2 = 0.0, wl3 = 0.0, wl4 = 0.0, wl5 = 0.0;
        double b1, b2, b3, b4, b5;
        double cflup = 0.0;
        for (int i = 0; i <= n + 1; i++) {
            double u = ROE(R_U, i), v = ROE(R_V, i), w = ROE(R_W, i), a = ROE(R_A, i);
            double g1a2 = ROE(R_G1A2, i), euv = ROE(R_EUV, i), u2v2w2 = ROE(R_U2V2W2, i);
            b1 = FW(1, 1, i);
            b4 = FW(1, 2, i);
            b2 = FW(mv, 2, i) - b4 * v;
            b3 = FW(mw, 2, i) - b4 * w;
            b5 = FW(1, 3, i);
            if (i == 0) continue;
            if (b1 != 0.0) {
                int id = IDX(1, i);
                double b4Id = g1a2 * (euv * DEL(1, id) + u * DEL(2, id) + v * DEL(3, id) +
                                      w * DEL(4, id) - DEL(5, id));
                double b5Id = (DEL(2, id) + (a - u) * DEL(1, id) - a * b4Id) / (2.0 * a);
                double b1Id = DEL(1, id) - b4Id - b5Id;
                double r1 = b1Id / b1;
                cflup = CF(1, i);
                wl1 = orc_suppow(r1, CF(1, i), cflup);
            }
            if (b2 != 0.0) {
                int id = IDX(2, i);
                double b2Id = DEL(3, id) - v * DEL(1, id);
                double r2 = b2Id / b2;
                cflup = CF(2, i);
                wl2 = orc_cflsupbee(theta, beta, r2, CF(2, i), cflup);
            }
            if (b3 != 0.0) {
                int id = IDX(2, i);
                double b3Id = DEL(4, id) - w * DEL(1, id);
                double r3 = b3Id / b3;
                wl3 = orc_cflsupbee(theta, beta, r3, CF(2, i), cflup);
            }
            if (b4 != 0.0) {
                int id = IDX(2, i);
                double b4Id = g1a2 * (euv * DEL(1, id) + u * DEL(2, id) + v * DEL(3, id) +
                                      w * DEL(4, id) - DEL(5, id));
                double r4 = b4Id / b4;
                wl4 = orc_cflsupbee(theta, beta, r4, CF(2, i), cflup);
            }
            if (b5 != 0.0) {
                int id = IDX(3, i);
                double b4Id = g1a2 * (euv * DEL(1, id) + u * DEL(2, id) + v * DEL(3, id) +
                                      w * DEL(4, id) - DEL(5, id));
                double b5Id = (DEL(2, id) + (a - u) * DEL(1, id) - a * b4Id) / (2.0 * a);
                double r5 = b5Id / b5;
                cflup = CF(3, i);
                wl5 = orc_suppow(r5, CF(3, i), cflup);
            }
            for (int m = 1; m <= 5; m++) {
                FW(m, 1, i) = FW(m, 1, i) * wl1;
                FW(m, 3, i) = FW(m, 3, i) * wl5;
            }
            double b2ltd = wl2 * b2, b3ltd = wl3 * b3, b4ltd = wl4 * b4;
            FW(1, 2, i) = b4ltd;
            FW(mu, 2, i) = b4ltd * u;
            FW(mv, 2, i) = b2ltd + b4ltd * v;
            FW(mw, 2, i) = b3ltd + b4ltd * w;
            FW(5, 2, i) = b2ltd * v + b3ltd * w + b4ltd * 0.5 * u2v2w2;
        }
    }
done:
    free(delta); free(dtdx); free(Id); free(CFL);
    return status;
#undef DEL
#undef IDX
#undef CF
#undef A19
#undef A17
}

/* ------------------------------------------------------------------------ */
/* rpt3 (rpt3euX.f:106-214) and rptt3 (rptt3euX.f:108-214).  They differ only
 * in the icoor = 3 'a2' coefficient (rptt3euX.f:169 uses component mw).     */
static void transverse(int ixyz, int icoor, int quirk, int meqn, int mbc, int n,
                       const double *roe, const double *asdq, double *bm, double *bp)
{
    const int N = n + 2 * mbc;
    int mu, mv, mw;
    comp_map(ixyz, &mu, &mv, &mw);
    double waveb[3][64];
    double sb[3];
    for (int i = 2 - mbc; i <= n + mbc; i++) {
        double u = ROE(R_U, i), v = ROE(R_V, i), w = ROE(R_W, i), a = ROE(R_A, i);
        double enth = ROE(R_ENTH, i), g1a2 = ROE(R_G1A2, i), euv = ROE(R_EUV, i);
        double u2v2w2 = ROE(R_U2V2W2, i);
        double q1 = V1(asdq, 1, i), qmu = V1(asdq, mu, i), qmv = V1(asdq, mv, i);
        double qmw = V1(asdq, mw, i), q5 = V1(asdq, 5, i);
        double a4 = g1a2 * (euv * q1 + u * qmu + v * qmv + w * qmw - q5);
        if (icoor == 2) {
            double a2 = qmu - u * q1;
            double a3 = qmw - w * q1;
            double a5 = (qmv + (a - v) * q1 - a * a4) / (2.0 * a);
            double a1 = q1 - a4 - a5;
            waveb[0][0] = a1; waveb[0][mu - 1] = a1 * u; waveb[0][mv - 1] = a1 * (v - a);
            waveb[0][mw - 1] = a1 * w; waveb[0][4] = a1 * (enth - v * a);
            sb[0] = v - a;
            waveb[1][0] = a4; waveb[1][mu - 1] = a2 + u * a4; waveb[1][mv - 1] = v * a4;
            waveb[1][mw - 1] = a3 + w * a4; waveb[1][4] = a4 * 0.5 * u2v2w2 + a2 * u + a3 * w;
            sb[1] = v;
            waveb[2][0] = a5; waveb[2][mu - 1] = a5 * u; waveb[2][mv - 1] = a5 * (v + a);
            waveb[2][mw - 1] = a5 * w; waveb[2][4] = a5 * (enth + v * a);
            sb[2] = v + a;
        } else {
            double a2 = (quirk ? qmw : qmu) - u * q1;
            double a3 = qmv - v * q1;
            double a5 = (qmw + (a - w) * q1 - a * a4) / (2.0 * a);
            double a1 = q1 - a4 - a5;
            waveb[0][0] = a1; waveb[0][mu - 1] = a1 * u; waveb[0][mv - 1] = a1 * v;
            waveb[0][mw - 1] = a1 * (w - a); waveb[0][4] = a1 * (enth - w * a);
            sb[0] = w - a;
            waveb[1][0] = a4; waveb[1][mu - 1] = a2 + u * a4; waveb[1][mv - 1] = a3 + v * a4;
            waveb[1][mw - 1] = w * a4; waveb[1][4] = a4 * 0.5 * u2v2w2 + a2 * u + a3 * v;
            sb[1] = w;
            waveb[2][0] = a5; waveb[2][mu - 1] = a5 * u; waveb[2][mv - 1] = a5 * v;
            waveb[2][mw - 1] = a5 * (w + a); waveb[2][4] = a5 * (enth + w * a);
            sb[2] = w + a;
        }
        for (int m = 6; m <= meqn; m++) {
            waveb[0][m - 1] = 0.0;
            waveb[1][m - 1] = V1(asdq, m, i);
            waveb[2][m - 1] = 0.0;
        }
        for (int m = 1; m <= meqn; m++) {
            double vm = 0.0, vp = 0.0;
            for (int mws = 0; mws < 3; mws++) {
                vm = vm + fminF(sb[mws], 0.0) * waveb[mws][m - 1];
                vp = vp + fmaxF(sb[mws], 0.0) * waveb[mws][m - 1];
            }
            V1(bm, m, i) = vm;
            V1(bp, m, i) = vp;
        }
    }
}

void orc_rpt3(int ixyz, int icoor, int meqn, int mbc, int n, const double *roe,
              const double *asdq, double *bmasdq, double *bpasdq)
{
    transverse(ixyz, icoor, 0, meqn, mbc, n, roe, asdq, bmasdq, bpasdq);
}

void orc_rptt3(int ixyz, int icoor, int meqn, int mbc, int n, const double *roe,
               const double *bsasdq, double *cmbsasdq, double *cpbsasdq)
{
    transverse(ixyz, icoor, 1, meqn, mbc, n, roe, bsasdq, cmbsasdq, cpbsasdq);
}

/* ------------------------------------------------------------------------ */
/* flux3: lib/flux3fw.f:4-628 with method(2) = 2, method(3) = 22, maux > 0,
 * mthlim = 0.  gadd/hadd are (meqn, pencil, 2, -1:1).                      */
typedef struct {
    int N, meqn;
    double *q1d, *a19, *a17, *dtdx1d, *fwave, *s, *amdq, *apdq, *cqxx, *roe;
    double *qadd, *fadd, *gadd, *hadd;
    double *t[28];
} flux_ws;

static void ws_alloc(flux_ws *W, int N, int meqn)
{
    W->N = N; W->meqn = meqn;
    size_t v = (size_t)N * meqn;
    W->q1d = calloc(v, sizeof(double));
    W->a19 = calloc(N, sizeof(double));
    W->a17 = calloc(N, sizeof(double));
    W->dtdx1d = calloc(N, sizeof(double));
    W->fwave = calloc(v * 3, sizeof(double));
    W->s = calloc((size_t)N * 3, sizeof(double));
    W->amdq = calloc(v, sizeof(double));
    W->apdq = calloc(v, sizeof(double));
    W->cqxx = calloc(v, sizeof(double));
    W->roe = calloc((size_t)N * 8, sizeof(double));
    W->qadd = calloc(v, sizeof(double));
    W->fadd = calloc(v, sizeof(double));
    W->gadd = calloc(v * 6, sizeof(double));
    W->hadd = calloc(v * 6, sizeof(double));
    for (int k = 0; k < 28; k++) W->t[k] = calloc(v, sizeof(double));
}

static void ws_free(flux_ws *W)
{
    free(W->q1d); free(W->a19); free(W->a17); free(W->dtdx1d); free(W->fwave); free(W->s);
    free(W->amdq); free(W->apdq); free(W->cqxx); free(W->roe); free(W->qadd);
    free(W->fadd); free(W->gadd); free(W->hadd);
    for (int k = 0; k < 28; k++) free(W->t[k]);
}

#define GADD(arr, m, i, side, off) arr[(((size_t)((off) + 1) * 2 + ((side) - 1)) * N + PI_(i)) * meqn + ((m) - 1)]

static int flux3(int ixyz, int meqn, int mbc, int n, int nspecadv, flux_ws *W,
                 double dtdy, double dtdz, double dtcom, double dxcom, double *cfl1d)
{
    const int N = n + 2 * mbc;
    double *q1d = W->q1d, *dtdx1d = W->dtdx1d, *fwave = W->fwave, *s = W->s;
    double *amdq = W->amdq, *apdq = W->apdq, *cqxx = W->cqxx, *roe = W->roe;
    double *qadd = W->qadd, *fadd = W->fadd, *gadd = W->gadd, *hadd = W->hadd;
    double *bmamdq = W->t[0], *bmapdq = W->t[1], *bpamdq = W->t[2], *bpapdq = W->t[3];
    double *cmamdq = W->t[4], *cmapdq = W->t[5], *cpamdq = W->t[6], *cpapdq = W->t[7];
    double *cmamdq2 = W->t[8], *cmapdq2 = W->t[9], *cpamdq2 = W->t[10], *cpapdq2 = W->t[11];
    double *bmcqxxp = W->t[12], *bpcqxxp = W->t[13], *bmcqxxm = W->t[14], *bpcqxxm = W->t[15];
    double *cmcqxxp = W->t[16], *cpcqxxp = W->t[17], *cmcqxxm = W->t[18], *cpcqxxm = W->t[19];
    double *bmcmamdq = W->t[20], *bmcmapdq = W->t[21], *bpcmamdq = W->t[22], *bpcmapdq = W->t[23];
    double *bmcpamdq = W->t[24], *bmcpapdq = W->t[25], *bpcpamdq = W->t[26], *bpcpapdq = W->t[27];
#define DT1(i) dtdx1d[PI_(i)]

    /* flux3fw.f:198-214 */
    for (int i = 1 - mbc; i <= n + mbc; i++)
        for (int m = 1; m <= meqn; m++) {
            V1(qadd, m, i) = 0.0;
            V1(fadd, m, i) = 0.0;
            for (int side = 1; side <= 2; side++)
                for (int off = -1; off <= 1; off++) {
                    GADD(gadd, m, i, side, off) = 0.0;
                    GADD(hadd, m, i, side, off) = 0.0;
                }
        }
    /* flux3fw.f:233-235 */
    int st = orc_rpn3(ixyz, meqn, mbc, n, nspecadv, q1d, W->a19, W->a17, dtcom, dxcom,
                      fwave, s, amdq, apdq, roe);
    if (st) return st;
    /* Godunov update: flux3fw.f:239-243 */
    for (int i = 1; i <= n + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            V1(qadd, m, i) = V1(qadd, m, i) - DT1(i) * V1(apdq, m, i);
            V1(qadd, m, i - 1) = V1(qadd, m, i - 1) - DT1(i - 1) * V1(amdq, m, i);
        }
    /* Courant number: flux3fw.f:246-254 */
    double c = 0.0;
    for (int i = 1; i <= n + 1; i++)
        for (int mw = 1; mw <= 3; mw++)
            c = fmaxF(fmaxF(c, DT1(i) * S(mw, i)), -DT1(i - 1) * S(mw, i));
    *cfl1d = c;
    /* Second-order correction cqxx: flux3fw.f:270-284 */
    for (int i = 2 - mbc; i <= n + mbc; i++) {
        DT1(i - 1) = 0.5 * (DT1(i - 1) + DT1(i));
        for (int m = 1; m <= meqn; m++) {
            V1(cqxx, m, i) = 0.0;
            for (int mw = 1; mw <= 3; mw++)
                V1(cqxx, m, i) = V1(cqxx, m, i) + 0.5 * dsign1(S(mw, i)) *
                    (1.0 - fabs(S(mw, i)) * DT1(i - 1)) * FW(m, mw, i);
            V1(fadd, m, i) = V1(fadd, m, i) + V1(cqxx, m, i);
        }
    }
    /* Transverse splits: flux3fw.f:297-346 (maux > 0 branch) */
    orc_rpt3(ixyz, 2, meqn, mbc, n, roe, amdq, bmamdq, bpamdq);
    orc_rpt3(ixyz, 2, meqn, mbc, n, roe, apdq, bmapdq, bpapdq);
    orc_rpt3(ixyz, 3, meqn, mbc, n, roe, amdq, cmamdq, cpamdq);
    orc_rpt3(ixyz, 3, meqn, mbc, n, roe, apdq, cmapdq, cpapdq);
    orc_rpt3(ixyz, 2, meqn, mbc, n, roe, cqxx, bmcqxxm, bpcqxxm);
    orc_rpt3(ixyz, 2, meqn, mbc, n, roe, cqxx, bmcqxxp, bpcqxxp);
    orc_rpt3(ixyz, 3, meqn, mbc, n, roe, cqxx, cmcqxxm, cpcqxxm);
    orc_rpt3(ixyz, 3, meqn, mbc, n, roe, cqxx, cmcqxxp, cpcqxxp);
    /* m4 = 2 mixing: flux3fw.f:388-396 */
    for (int i = 0; i <= n + 2; i++)
        for (int m = 1; m <= meqn; m++) {
            V1(cpapdq2, m, i) = V1(cpapdq, m, i) - 3.0 * V1(cpcqxxp, m, i);
            V1(cpamdq2, m, i) = V1(cpamdq, m, i) + 3.0 * V1(cpcqxxm, m, i);
            V1(cmapdq2, m, i) = V1(cmapdq, m, i) - 3.0 * V1(cmcqxxp, m, i);
            V1(cmamdq2, m, i) = V1(cmamdq, m, i) + 3.0 * V1(cmcqxxm, m, i);
        }
    /* flux3fw.f:405-412 */
    orc_rptt3(ixyz, 2, meqn, mbc, n, roe, cpapdq2, bmcpapdq, bpcpapdq);
    orc_rptt3(ixyz, 2, meqn, mbc, n, roe, cpamdq2, bmcpamdq, bpcpamdq);
    orc_rptt3(ixyz, 2, meqn, mbc, n, roe, cmapdq2, bmcmapdq, bpcmapdq);
    orc_rptt3(ixyz, 2, meqn, mbc, n, roe, cmamdq2, bmcmamdq, bpcmamdq);
    /* G updates: flux3fw.f:419-500 */
    for (int i = 1; i <= n + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            double d = DT1(i - 1);
            GADD(gadd, m, i - 1, 1, 0) = GADD(gadd, m, i - 1, 1, 0) - 0.5 * d * V1(bmamdq, m, i);
            GADD(gadd, m, i - 1, 2, 0) = GADD(gadd, m, i - 1, 2, 0) - 0.5 * d * V1(bpamdq, m, i);
            GADD(gadd, m, i, 1, 0) = GADD(gadd, m, i, 1, 0) - 0.5 * d * V1(bmapdq, m, i);
            GADD(gadd, m, i, 2, 0) = GADD(gadd, m, i, 2, 0) - 0.5 * d * V1(bpapdq, m, i);
            GADD(gadd, m, i, 2, 0) = GADD(gadd, m, i, 2, 0) + (1.0 / 6.0) * d * dtdz *
                (V1(bpcpapdq, m, i) - V1(bpcmapdq, m, i));
            GADD(gadd, m, i, 1, 0) = GADD(gadd, m, i, 1, 0) + (1.0 / 6.0) * d * dtdz *
                (V1(bmcpapdq, m, i) - V1(bmcmapdq, m, i));
            GADD(gadd, m, i, 2, 1) = GADD(gadd, m, i, 2, 1) - (1.0 / 6.0) * d * dtdz * V1(bpcpapdq, m, i);
            GADD(gadd, m, i, 1, 1) = GADD(gadd, m, i, 1, 1) - (1.0 / 6.0) * d * dtdz * V1(bmcpapdq, m, i);
            GADD(gadd, m, i, 2, -1) = GADD(gadd, m, i, 2, -1) + (1.0 / 6.0) * d * dtdz * V1(bpcmapdq, m, i);
            GADD(gadd, m, i, 1, -1) = GADD(gadd, m, i, 1, -1) + (1.0 / 6.0) * d * dtdz * V1(bmcmapdq, m, i);
            GADD(gadd, m, i - 1, 2, 0) = GADD(gadd, m, i - 1, 2, 0) + (1.0 / 6.0) * d * dtdz *
                (V1(bpcpamdq, m, i) - V1(bpcmamdq, m, i));
            GADD(gadd, m, i - 1, 1, 0) = GADD(gadd, m, i - 1, 1, 0) + (1.0 / 6.0) * d * dtdz *
                (V1(bmcpamdq, m, i) - V1(bmcmamdq, m, i));
            GADD(gadd, m, i - 1, 2, 1) = GADD(gadd, m, i - 1, 2, 1) - (1.0 / 6.0) * d * dtdz * V1(bpcpamdq, m, i);
            GADD(gadd, m, i - 1, 1, 1) = GADD(gadd, m, i - 1, 1, 1) - (1.0 / 6.0) * d * dtdz * V1(bmcpamdq, m, i);
            GADD(gadd, m, i - 1, 2, -1) = GADD(gadd, m, i - 1, 2, -1) + (1.0 / 6.0) * d * dtdz * V1(bpcmamdq, m, i);
            GADD(gadd, m, i - 1, 1, -1) = GADD(gadd, m, i - 1, 1, -1) + (1.0 / 6.0) * d * dtdz * V1(bmcmamdq, m, i);
            GADD(gadd, m, i, 2, 0) = GADD(gadd, m, i, 2, 0) + d * V1(bpcqxxp, m, i);
            GADD(gadd, m, i, 1, 0) = GADD(gadd, m, i, 1, 0) + d * V1(bmcqxxp, m, i);
            GADD(gadd, m, i - 1, 2, 0) = GADD(gadd, m, i - 1, 2, 0) - d * V1(bpcqxxm, m, i);
            GADD(gadd, m, i - 1, 1, 0) = GADD(gadd, m, i - 1, 1, 0) - d * V1(bmcqxxm, m, i);
        }
    /* m4 = 2 mixing into the B splits: flux3fw.f:510-518 */
    for (int i = 0; i <= n + 2; i++)
        for (int m = 1; m <= meqn; m++) {
            V1(bpapdq, m, i) = V1(bpapdq, m, i) - 3.0 * V1(bpcqxxp, m, i);
            V1(bpamdq, m, i) = V1(bpamdq, m, i) + 3.0 * V1(bpcqxxm, m, i);
            V1(bmapdq, m, i) = V1(bmapdq, m, i) - 3.0 * V1(bmcqxxp, m, i);
            V1(bmamdq, m, i) = V1(bmamdq, m, i) + 3.0 * V1(bmcqxxm, m, i);
        }
    /* flux3fw.f:531-538 (outputs reuse the bXcY arrays) */
    orc_rptt3(ixyz, 3, meqn, mbc, n, roe, bpapdq, bmcpapdq, bpcpapdq);
    orc_rptt3(ixyz, 3, meqn, mbc, n, roe, bpamdq, bmcpamdq, bpcpamdq);
    orc_rptt3(ixyz, 3, meqn, mbc, n, roe, bmapdq, bmcmapdq, bpcmapdq);
    orc_rptt3(ixyz, 3, meqn, mbc, n, roe, bmamdq, bmcmamdq, bpcmamdq);
    /* H updates: flux3fw.f:545-625 */
    for (int i = 1; i <= n + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            double d = DT1(i - 1);
            GADD(hadd, m, i - 1, 1, 0) = GADD(hadd, m, i - 1, 1, 0) - 0.5 * d * V1(cmamdq, m, i);
            GADD(hadd, m, i - 1, 2, 0) = GADD(hadd, m, i - 1, 2, 0) - 0.5 * d * V1(cpamdq, m, i);
            GADD(hadd, m, i, 1, 0) = GADD(hadd, m, i, 1, 0) - 0.5 * d * V1(cmapdq, m, i);
            GADD(hadd, m, i, 2, 0) = GADD(hadd, m, i, 2, 0) - 0.5 * d * V1(cpapdq, m, i);
            GADD(hadd, m, i, 2, 0) = GADD(hadd, m, i, 2, 0) + (1.0 / 6.0) * d * dtdy *
                (V1(bpcpapdq, m, i) - V1(bpcmapdq, m, i));
            GADD(hadd, m, i, 1, 0) = GADD(hadd, m, i, 1, 0) + (1.0 / 6.0) * d * dtdy *
                (V1(bmcpapdq, m, i) - V1(bmcmapdq, m, i));
            GADD(hadd, m, i, 2, 1) = GADD(hadd, m, i, 2, 1) - (1.0 / 6.0) * d * dtdy * V1(bpcpapdq, m, i);
            GADD(hadd, m, i, 1, 1) = GADD(hadd, m, i, 1, 1) - (1.0 / 6.0) * d * dtdy * V1(bmcpapdq, m, i);
            GADD(hadd, m, i, 2, -1) = GADD(hadd, m, i, 2, -1) + (1.0 / 6.0) * d * dtdy * V1(bpcmapdq, m, i);
            GADD(hadd, m, i, 1, -1) = GADD(hadd, m, i, 1, -1) + (1.0 / 6.0) * d * dtdy * V1(bmcmapdq, m, i);
            GADD(hadd, m, i - 1, 2, 0) = GADD(hadd, m, i - 1, 2, 0) + (1.0 / 6.0) * d * dtdy *
                (V1(bpcpamdq, m, i) - V1(bpcmamdq, m, i));
            GADD(hadd, m, i - 1, 1, 0) = GADD(hadd, m, i - 1, 1, 0) + (1.0 / 6.0) * d * dtdy *
                (V1(bmcpamdq, m, i) - V1(bmcmamdq, m, i));
            GADD(hadd, m, i - 1, 2, 1) = GADD(hadd, m, i - 1, 2, 1) - (1.0 / 6.0) * d * dtdy * V1(bpcpamdq, m, i);
            GADD(hadd, m, i - 1, 1, 1) = GADD(hadd, m, i - 1, 1, 1) - (1.0 / 6.0) * d * dtdy * V1(bmcpamdq, m, i);
            GADD(hadd, m, i - 1, 2, -1) = GADD(hadd, m, i - 1, 2, -1) + (1.0 / 6.0) * d * dtdy * V1(bpcmamdq, m, i);
            GADD(hadd, m, i - 1, 1, -1) = GADD(hadd, m, i - 1, 1, -1) + (1.0 / 6.0) * d * dtdy * V1(bmcmamdq, m, i);
            GADD(hadd, m, i, 2, 0) = GADD(hadd, m, i, 2, 0) + d * V1(cpcqxxp, m, i);
            GADD(hadd, m, i, 1, 0) = GADD(hadd, m, i, 1, 0) + d * V1(cmcqxxp, m, i);
            GADD(hadd, m, i - 1, 2, 0) = GADD(hadd, m, i - 1, 2, 0) - d * V1(cpcqxxm, m, i);
            GADD(hadd, m, i - 1, 1, 0) = GADD(hadd, m, i - 1, 1, 0) - d * V1(cmcqxxm, m, i);
        }
    return ORC_OK;
#undef DT1
}

/* ------------------------------------------------------------------------ */
/* Grid helpers.                                                              */
#define NXP (g->mx + 2 * g->mbc)
#define NYP (g->my + 2 * g->mbc)
#define NZP (g->mz + 2 * g->mbc)
static inline size_t cell(const orc_grid *g, int i, int j, int k)
{
    return ((size_t)(k + g->mbc - 1) * NYP + (size_t)(j + g->mbc - 1)) * NXP + (size_t)(i + g->mbc - 1);
}
#define QG(arr, m, i, j, k) arr[cell(g, i, j, k) * g->meqn + ((m) - 1)]
#define AG(ma, i, j, k) aux[cell(g, i, j, k) * g->maux + ((ma) - 1)]

/* ------------------------------------------------------------------------ */
/* step3: lib/step3.f:3-626 (mcapa = 0).                                     */
int orc_step3(const orc_grid *g, int nspecadv, const double *qold, double *qnew,
              const double *aux, double dt, double *cflout)
{
    const int meqn = g->meqn, mbc = g->mbc, mx = g->mx, my = g->my, mz = g->mz;
    const int maxm = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
    flux_ws W;
    ws_alloc(&W, maxm + 2 * mbc, meqn);
    double cfl = 0.0, cfl1d;
    const double dtdx = dt / g->dx, dtdy = dt / g->dy, dtdz = dt / g->dz;
    int st = ORC_OK;

    /* x-sweeps: lib/step3.f:121-280 */
    for (int k = 0; k <= mz + 1 && !st; k++)
        for (int j = 0; j <= my + 1; j++) {
            int n = mx, N = mx + 2 * mbc;
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                for (int m = 1; m <= meqn; m++) W.q1d[(size_t)PI_(i) * meqn + m - 1] = QG(qold, m, i, j, k);
                W.a19[PI_(i)] = AG(19, i, j, k);
                W.a17[PI_(i)] = AG(17, i, j, k);
            }
            for (int i = 0; i < N; i++) W.dtdx1d[i] = dtdx;
            st = flux3(1, meqn, mbc, n, nspecadv, &W, dtdy, dtdz, dt, g->dx, &cfl1d);
            if (st) break;
            cfl = fmaxF(cfl, cfl1d);
            double *qadd = W.qadd, *fadd = W.fadd, *gadd = W.gadd, *hadd = W.hadd;
            for (int i = 1; i <= mx; i++)
                for (int m = 1; m <= meqn; m++) {
                    QG(qnew, m, i, j, k) = QG(qnew, m, i, j, k) + V1(qadd, m, i)
                        - dtdx * (V1(fadd, m, i + 1) - V1(fadd, m, i))
                        - dtdy * (GADD(gadd, m, i, 2, 0) - GADD(gadd, m, i, 1, 0))
                        - dtdz * (GADD(hadd, m, i, 2, 0) - GADD(hadd, m, i, 1, 0));
                    QG(qnew, m, i, j - 1, k) = QG(qnew, m, i, j - 1, k)
                        - dtdy * GADD(gadd, m, i, 1, 0)
                        - dtdz * (GADD(hadd, m, i, 2, -1) - GADD(hadd, m, i, 1, -1));
                    QG(qnew, m, i, j - 1, k - 1) = QG(qnew, m, i, j - 1, k - 1)
                        - dtdy * GADD(gadd, m, i, 1, -1)
                        - dtdz * GADD(hadd, m, i, 1, -1);
                    QG(qnew, m, i, j, k - 1) = QG(qnew, m, i, j, k - 1)
                        - dtdy * (GADD(gadd, m, i, 2, -1) - GADD(gadd, m, i, 1, -1))
                        - dtdz * GADD(hadd, m, i, 1, 0);
                    QG(qnew, m, i, j + 1, k - 1) = QG(qnew, m, i, j + 1, k - 1)
                        + dtdy * GADD(gadd, m, i, 2, -1)
                        - dtdz * GADD(hadd, m, i, 1, 1);
                    QG(qnew, m, i, j + 1, k) = QG(qnew, m, i, j + 1, k)
                        + dtdy * GADD(gadd, m, i, 2, 0)
                        - dtdz * (GADD(hadd, m, i, 2, 1) - GADD(hadd, m, i, 1, 1));
                    QG(qnew, m, i, j + 1, k + 1) = QG(qnew, m, i, j + 1, k + 1)
                        + dtdy * GADD(gadd, m, i, 2, 1)
                        + dtdz * GADD(hadd, m, i, 2, 1);
                    QG(qnew, m, i, j, k + 1) = QG(qnew, m, i, j, k + 1)
                        - dtdy * (GADD(gadd, m, i, 2, 1) - GADD(gadd, m, i, 1, 1))
                        + dtdz * GADD(hadd, m, i, 2, 0);
                    QG(qnew, m, i, j - 1, k + 1) = QG(qnew, m, i, j - 1, k + 1)
                        - dtdy * GADD(gadd, m, i, 1, 1)
                        + dtdz * GADD(hadd, m, i, 2, -1);
                }
        }

    /* y-sweeps: lib/step3.f:289-450 */
    for (int k = 0; k <= mz + 1 && !st; k++)
        for (int i = 0; i <= mx + 1; i++) {
            int n = my, N = my + 2 * mbc;
            for (int j = 1 - mbc; j <= my + mbc; j++) {
                for (int m = 1; m <= meqn; m++) W.q1d[(size_t)PI_(j) * meqn + m - 1] = QG(qold, m, i, j, k);
                W.a19[PI_(j)] = AG(19, i, j, k);
                W.a17[PI_(j)] = AG(17, i, j, k);
            }
            for (int q = 0; q < N; q++) W.dtdx1d[q] = dtdy;
            st = flux3(2, meqn, mbc, n, nspecadv, &W, dtdz, dtdx, dt, g->dy, &cfl1d);
            if (st) break;
            cfl = fmaxF(cfl, cfl1d);
            double *qadd = W.qadd, *fadd = W.fadd, *gadd = W.gadd, *hadd = W.hadd;
            for (int j = 1; j <= my; j++)
                for (int m = 1; m <= meqn; m++) {
                    QG(qnew, m, i, j, k) = QG(qnew, m, i, j, k) + V1(qadd, m, j)
                        - dtdy * (V1(fadd, m, j + 1) - V1(fadd, m, j))
                        - dtdz * (GADD(gadd, m, j, 2, 0) - GADD(gadd, m, j, 1, 0))
                        - dtdx * (GADD(hadd, m, j, 2, 0) - GADD(hadd, m, j, 1, 0));
                    QG(qnew, m, i, j, k + 1) = QG(qnew, m, i, j, k + 1)
                        + dtdz * GADD(gadd, m, j, 2, 0)
                        - dtdx * (GADD(hadd, m, j, 2, 1) - GADD(hadd, m, j, 1, 1));
                    QG(qnew, m, i + 1, j, k + 1) = QG(qnew, m, i + 1, j, k + 1)
                        + dtdz * GADD(gadd, m, j, 2, 1)
                        + dtdx * GADD(hadd, m, j, 2, 1);
                    QG(qnew, m, i + 1, j, k) = QG(qnew, m, i + 1, j, k)
                        - dtdz * (GADD(gadd, m, j, 2, 1) - GADD(gadd, m, j, 1, 1))
                        + dtdx * GADD(hadd, m, j, 2, 0);
                    QG(qnew, m, i + 1, j, k - 1) = QG(qnew, m, i + 1, j, k - 1)
                        - dtdz * GADD(gadd, m, j, 1, 1)
                        + dtdx * GADD(hadd, m, j, 2, -1);
                    QG(qnew, m, i, j, k - 1) = QG(qnew, m, i, j, k - 1)
                        - dtdz * GADD(gadd, m, j, 1, 0)
                        - dtdx * (GADD(hadd, m, j, 2, -1) - GADD(hadd, m, j, 1, -1));
                    QG(qnew, m, i - 1, j, k - 1) = QG(qnew, m, i - 1, j, k - 1)
                        - dtdz * GADD(gadd, m, j, 1, -1)
                        - dtdx * GADD(hadd, m, j, 1, -1);
                    QG(qnew, m, i - 1, j, k) = QG(qnew, m, i - 1, j, k)
                        - dtdz * (GADD(gadd, m, j, 2, -1) - GADD(gadd, m, j, 1, -1))
                        - dtdx * GADD(hadd, m, j, 1, 0);
                    QG(qnew, m, i - 1, j, k + 1) = QG(qnew, m, i - 1, j, k + 1)
                        + dtdz * GADD(gadd, m, j, 2, -1)
                        - dtdx * GADD(hadd, m, j, 1, 1);
                }
        }

    /* z-sweeps: lib/step3.f:459-621 */
    for (int j = 0; j <= my + 1 && !st; j++)
        for (int i = 0; i <= mx + 1; i++) {
            int n = mz, N = mz + 2 * mbc;
            for (int k = 1 - mbc; k <= mz + mbc; k++) {
                for (int m = 1; m <= meqn; m++) W.q1d[(size_t)PI_(k) * meqn + m - 1] = QG(qold, m, i, j, k);
                W.a19[PI_(k)] = AG(19, i, j, k);
                W.a17[PI_(k)] = AG(17, i, j, k);
            }
            for (int q = 0; q < N; q++) W.dtdx1d[q] = dtdz;
            st = flux3(3, meqn, mbc, n, nspecadv, &W, dtdx, dtdy, dt, g->dz, &cfl1d);
            if (st) break;
            cfl = fmaxF(cfl, cfl1d);
            double *qadd = W.qadd, *fadd = W.fadd, *gadd = W.gadd, *hadd = W.hadd;
            for (int k = 1; k <= mz; k++)
                for (int m = 1; m <= meqn; m++) {
                    QG(qnew, m, i, j, k) = QG(qnew, m, i, j, k) + V1(qadd, m, k)
                        - dtdz * (V1(fadd, m, k + 1) - V1(fadd, m, k))
                        - dtdx * (GADD(gadd, m, k, 2, 0) - GADD(gadd, m, k, 1, 0))
                        - dtdy * (GADD(hadd, m, k, 2, 0) - GADD(hadd, m, k, 1, 0));
                    QG(qnew, m, i, j + 1, k) = QG(qnew, m, i, j + 1, k)
                        - dtdx * (GADD(gadd, m, k, 2, 1) - GADD(gadd, m, k, 1, 1))
                        + dtdy * GADD(hadd, m, k, 2, 0);
                    QG(qnew, m, i + 1, j + 1, k) = QG(qnew, m, i + 1, j + 1, k)
                        + dtdx * GADD(gadd, m, k, 2, 1)
                        + dtdy * GADD(hadd, m, k, 2, 1);
                    QG(qnew, m, i + 1, j, k) = QG(qnew, m, i + 1, j, k)
                        + dtdx * GADD(gadd, m, k, 2, 0)
                        - dtdy * (GADD(hadd, m, k, 2, 1) - GADD(hadd, m, k, 1, 1));
                    QG(qnew, m, i + 1, j - 1, k) = QG(qnew, m, i + 1, j - 1, k)
                        + dtdx * GADD(gadd, m, k, 2, -1)
                        - dtdy * GADD(hadd, m, k, 1, 1);
                    QG(qnew, m, i, j - 1, k) = QG(qnew, m, i, j - 1, k)
                        - dtdx * (GADD(gadd, m, k, 2, -1) - GADD(gadd, m, k, 1, -1))
                        - dtdy * GADD(hadd, m, k, 1, 0);
                    QG(qnew, m, i - 1, j - 1, k) = QG(qnew, m, i - 1, j - 1, k)
                        - dtdx * GADD(gadd, m, k, 1, -1)
                        - dtdy * GADD(hadd, m, k, 1, -1);
                    QG(qnew, m, i - 1, j, k) = QG(qnew, m, i - 1, j, k)
                        - dtdx * GADD(gadd, m, k, 1, 0)
                        - dtdy * (GADD(hadd, m, k, 2, -1) - GADD(hadd, m, k, 1, -1));
                    QG(qnew, m, i - 1, j + 1, k) = QG(qnew, m, i - 1, j + 1, k)
                        - dtdx * GADD(gadd, m, k, 1, 1)
                        + dtdy * GADD(hadd, m, k, 2, -1);
                }
        }
    ws_free(&W);
    *cflout = cfl;
    return st;
}

/* ------------------------------------------------------------------------ */
/* bc3: bc3_mpi.f:5-812 for one rank (lx = ly = lz = 1).                      */
int orc_bc3(const orc_grid *g, double *q, const double *aux, const int *mthbc)
{
    const int meqn = g->meqn, mbc = g->mbc, mx = g->mx, my = g->my, mz = g->mz;
    for (int f = 0; f < 6; f++)
        if (mthbc[f] == 0 || mthbc[f] >= 4) return ORC_BAD_BC;
    /* x lower / upper: bc3_mpi.f:97-312 */
    if (mthbc[0] == 1) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, 1 - ibc, j, k) = QG(q, m, 1, j, k);
    } else if (mthbc[0] == 3) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, 1 - ibc, j, k) = QG(q, m, ibc, j, k);
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++) QG(q, 2, 1 - ibc, j, k) = -QG(q, 2, ibc, j, k);
    }
    if (mthbc[1] == 1) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, mx + ibc, j, k) = QG(q, m, mx, j, k);
    } else if (mthbc[1] == 3) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, mx + ibc, j, k) = QG(q, m, mx + 1 - ibc, j, k);
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ibc = 1; ibc <= mbc; ibc++) QG(q, 2, mx + ibc, j, k) = -QG(q, 2, mx + 1 - ibc, j, k);
    }
    /* periodic in x (lx = 1): copy interior j,k only (bc3_mpi.f:120-131, 208-220, 285-312) */
    if (mthbc[0] == 2 || mthbc[1] == 2) {
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mbc; i++)
                    for (int m = 1; m <= meqn; m++) {
                        if (mthbc[0] == 2) QG(q, m, i - mbc, j, k) = QG(q, m, mx - mbc + i, j, k);
                        if (mthbc[1] == 2) QG(q, m, mx + i, j, k) = QG(q, m, i, j, k);
                    }
    }
    /* y lower / upper: bc3_mpi.f:320-536 */
    if (mthbc[2] == 1) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, i, 1 - jbc, k) = QG(q, m, i, 1, k);
    } else if (mthbc[2] == 3) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, i, 1 - jbc, k) = QG(q, m, i, jbc, k);
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int k = 1 - mbc; k <= mz + mbc; k++) QG(q, 3, i, 1 - jbc, k) = -QG(q, 3, i, jbc, k);
    }
    if (mthbc[3] == 1) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, i, my + jbc, k) = QG(q, m, i, my, k);
    } else if (mthbc[3] == 3) {
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) QG(q, m, i, my + jbc, k) = QG(q, m, i, my + 1 - jbc, k);
        for (int k = 1 - mbc; k <= mz + mbc; k++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 1 - mbc; i <= mx + mbc; i++) QG(q, 3, i, my + jbc, k) = -QG(q, 3, i, my + 1 - jbc, k);
    }
    if (mthbc[2] == 2 || mthbc[3] == 2) {
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= mbc; j++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) {
                        if (mthbc[2] == 2) QG(q, m, i, j - mbc, k) = QG(q, m, i, my - mbc + j, k);
                        if (mthbc[3] == 2) QG(q, m, i, my + j, k) = QG(q, m, i, j, k);
                    }
    }
    /* z lower / upper: bc3_mpi.f:544-809.  Extrapolation and wall scale the
     * perturbation from the background by (rho_bar_ghost/rho_bar_1)**0.5. */
    for (int side = 0; side < 2; side++) {
        int mth = mthbc[4 + side];
        if (mth != 1 && mth != 3) continue;
        for (int kbc = 1; kbc <= mbc; kbc++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int i = 1 - mbc; i <= mx + mbc; i++) {
                    int kg = side == 0 ? 1 - kbc : mz + kbc;
                    int kb = side == 0 ? 1 : mz;
                    double sc = pow(AG(1, i, j, kg) / AG(1, i, j, kb), 0.5);
                    QG(q, 1, i, j, kg) = AG(1, i, j, kg) + (QG(q, 1, i, j, kb) - AG(1, i, j, kb)) * sc;
                    QG(q, 2, i, j, kg) = AG(2, i, j, kg) + (QG(q, 2, i, j, kb) - AG(2, i, j, kb)) * sc;
                    QG(q, 3, i, j, kg) = AG(3, i, j, kg) + (QG(q, 3, i, j, kb) - AG(3, i, j, kb)) * sc;
                    QG(q, 4, i, j, kg) = QG(q, 4, i, j, kb) * sc;
                    QG(q, 5, i, j, kg) = AG(5, i, j, kg) + (QG(q, 5, i, j, kb) - AG(5, i, j, kb)) * sc;
                    for (int m = 6; m <= meqn; m++) QG(q, m, i, j, kg) = QG(q, m, i, j, kb);
                }
        if (mth == 3)
            for (int kbc = 1; kbc <= mbc; kbc++)
                for (int j = 1 - mbc; j <= my + mbc; j++)
                    for (int i = 1 - mbc; i <= mx + mbc; i++) {
                        if (side == 0) QG(q, 4, i, j, 1 - kbc) = -QG(q, 4, i, j, kbc);
                        else QG(q, 4, i, j, mz + kbc) = -QG(q, 4, i, j, mz + 1 - kbc);
                    }
    }
    if (mthbc[4] == 2 || mthbc[5] == 2) {
        for (int k = 1; k <= mbc; k++)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int i = 1 - mbc; i <= mx + mbc; i++)
                    for (int m = 1; m <= meqn; m++) {
                        if (mthbc[4] == 2) QG(q, m, i, j, k - mbc) = QG(q, m, i, j, mz - mbc + k);
                        if (mthbc[5] == 2) QG(q, m, i, j, mz + k) = QG(q, m, i, j, k);
                    }
    }
    return ORC_OK;
}

/* b4step3.f:35-45 */
void orc_b4step3(const orc_grid *g, double *q)
{
    const int meqn = g->meqn;
    size_t ncell = (size_t)NXP * NYP * NZP;
    if (meqn <= 5) return;
    for (size_t c = 0; c < ncell; c++)
        for (int m = 6; m <= meqn; m++) {
            double v = q[c * meqn + m - 1];
            q[c * meqn + m - 1] = fmaxF(0.0, fminF(v, 1.0));
        }
}

/* ------------------------------------------------------------------------ */
/* diffuse: diffuse.f:2-214.  Works on q(2:4) holding velocity perturbations. */
static void diffuse(const orc_grid *g, double *q, const double *aux, int mcoefaux, double dtr)
{
    const int meqn = g->meqn, mbc = g->mbc, mx = g->mx, my = g->my, mz = g->mz;
    const size_t nc = (size_t)NXP * NYP * NZP;
    double *qt = calloc(nc, sizeof(double));
    double *u = calloc(nc * 3, sizeof(double));
    double *coefs = calloc(nc, sizeof(double));
#define U3(i, j, k, c) u[((size_t)(c) - 1) * nc + cell(g, i, j, k)]
#define QT(i, j, k) qt[cell(g, i, j, k)]
#define CO(i, j, k) coefs[cell(g, i, j, k)]
    const double fourthirds = 4.0 / 3.0;
    double dtdx2 = (dtr) / (g->dx * g->dx);
    double dtdy2 = (dtr) / (g->dy * g->dy);
    double dtdz2 = (dtr) / (g->dz * g->dz);
    /* z ghosts of q(2:4): diffuse.f:48-54 */
    for (int kbc = 1; kbc <= mbc; kbc++)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int m = 2; m <= 4; m++) {
                    QG(q, m, i, j, 1 - kbc) = QG(q, m, i, j, 1);
                    QG(q, m, i, j, mz + kbc) = QG(q, m, i, j, mz);
                }
    for (int k = 1 - mbc; k <= mz + mbc; k++)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                CO(i, j, k) = AG(mcoefaux, i, j, k);
                U3(i, j, k, 1) = QG(q, 2, i, j, k);
                U3(i, j, k, 2) = QG(q, 3, i, j, k);
                U3(i, j, k, 3) = QG(q, 4, i, j, k);
            }
    /* cross terms: diffuse.f:70-118 */
    double dtr32 = dtr / 6.0;
    double dt2dxdy = dtr32 / (g->dx * g->dy);
    double dt2dxdz = dtr32 / (g->dx * g->dz);
    double dt2dydz = dtr32 / (g->dy * g->dz);
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++) {
                double ctxxy = U3(i + 1, j + 1, k, 2) - U3(i - 1, j + 1, k, 2) + U3(i - 1, j - 1, k, 2) - U3(i + 1, j - 1, k, 2);
                double ctxxz = U3(i + 1, j, k + 1, 3) - U3(i - 1, j, k + 1, 3) + U3(i - 1, j, k - 1, 3) - U3(i + 1, j, k - 1, 3);
                double ctyxy = U3(i + 1, j + 1, k, 1) - U3(i - 1, j + 1, k, 1) + U3(i - 1, j - 1, k, 1) - U3(i + 1, j - 1, k, 1);
                double ctyyz = U3(i, j + 1, k + 1, 3) - U3(i, j - 1, k + 1, 3) + U3(i, j - 1, k - 1, 3) - U3(i, j + 1, k - 1, 3);
                double ctzxz = U3(i + 1, j, k + 1, 1) - U3(i - 1, j, k + 1, 1) + U3(i - 1, j, k - 1, 1) - U3(i + 1, j, k - 1, 1);
                double ctzyz = U3(i, j + 1, k + 1, 2) - U3(i, j - 1, k + 1, 2) + U3(i, j - 1, k - 1, 2) - U3(i, j + 1, k - 1, 2);
                QG(q, 2, i, j, k) = U3(i, j, k, 1) + CO(i, j, k) * (dt2dxdy * ctxxy + dt2dxdz * ctxxz);
                QG(q, 3, i, j, k) = U3(i, j, k, 2) + CO(i, j, k) * (dt2dxdy * ctyxy + dt2dydz * ctyyz);
                QG(q, 4, i, j, k) = U3(i, j, k, 3) + CO(i, j, k) * (dt2dxdz * ctzxz + dt2dydz * ctzyz);
            }
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++)
                for (int c = 1; c <= 3; c++) U3(i, j, k, c) = QG(q, c + 1, i, j, k);
    /* split heat equation per component and direction: diffuse.f:135-210 */
    for (int c = 1; c <= 3; c++) {
        if (c == 1) {
            dtdx2 = fourthirds * dtdx2;
        } else if (c == 2) {
            dtdy2 = fourthirds * dtdy2;
            dtdx2 = dtdx2 / fourthirds;
        } else {
            dtdz2 = fourthirds * dtdz2;
            dtdy2 = dtdy2 / fourthirds;
        }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double am = U3(i - 1, j, k, c), aq = U3(i, j, k, c), ap = U3(i + 1, j, k, c);
                    QT(i, j, k) = (ap - aq - aq + am);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    U3(i, j, k, c) = U3(i, j, k, c) + dtdx2 * CO(i, j, k) * QT(i, j, k);
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double ab = U3(i, j - 1, k, c), aq = U3(i, j, k, c), af = U3(i, j + 1, k, c);
                    QT(i, j, k) = (af - aq - aq + ab);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    U3(i, j, k, c) = U3(i, j, k, c) + dtdy2 * CO(i, j, k) * QT(i, j, k);
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double au = U3(i, j, k + 1, c), aq = U3(i, j, k, c), ad = U3(i, j, k - 1, c);
                    QT(i, j, k) = (ad - aq - aq + au);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    QG(q, c + 1, i, j, k) = U3(i, j, k, c) + dtdz2 * CO(i, j, k) * QT(i, j, k);
    }
    free(qt); free(u); free(coefs);
#undef U3
#undef QT
#undef CO
}

/* conduct: conduct.f:2-121 on q(mdiffq) holding the temperature perturbation. */
static void conduct(const orc_grid *g, double *q, int mdiffq, const double *aux, int mcoefaux,
                    double dtr, int ndtrc)
{
    const int meqn = g->meqn, mbc = g->mbc, mx = g->mx, my = g->my, mz = g->mz;
    const size_t nc = (size_t)NXP * NYP * NZP;
    /* slack for the linear-address spill of conduct.f:86-91 when mz > mx */
    const size_t slack = (size_t)(mz > mx ? mz - mx : 0) + 2 * (size_t)mbc + 8;
    double *qt = calloc(nc, sizeof(double));
    double *u = calloc(nc + slack, sizeof(double));
    double *coefs = calloc(nc, sizeof(double));
#define UU(i, j, k) u[cell(g, i, j, k)]
#define QT(i, j, k) qt[cell(g, i, j, k)]
#define CO(i, j, k) coefs[cell(g, i, j, k)]
    double dtdx2 = (dtr) / (g->dx * g->dx);
    double dtdy2 = (dtr) / (g->dy * g->dy);
    double dtdz2 = (dtr) / (g->dz * g->dz);
    for (int k = 1 - mbc; k <= mz + mbc; k++)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                CO(i, j, k) = AG(mcoefaux, i, j, k);
                UU(i, j, k) = QG(q, mdiffq, i, j, k);
            }
    for (int nn = 1; nn <= ndtrc; nn++) {
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double am = UU(i - 1, j, k), aq = UU(i, j, k), ap = UU(i + 1, j, k);
                    QT(i, j, k) = (ap - aq - aq + am);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    UU(i, j, k) = UU(i, j, k) + dtdx2 * CO(i, j, k) * QT(i, j, k);
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double ab = UU(i, j - 1, k), aq = UU(i, j, k), af = UU(i, j + 1, k);
                    QT(i, j, k) = (af - aq - aq + ab);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    UU(i, j, k) = UU(i, j, k) + dtdy2 * CO(i, j, k) * QT(i, j, k);
        /* z ghosts with the reference's i-bound mz+mbc (conduct.f:86-91):
         * linear addressing, as Fortran does for an out-of-row index. */
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int i = 1 - mbc; i <= mz + mbc; i++)
                for (int kbc = 1; kbc <= mbc; kbc++) {
                    long base_j = (long)(j + mbc - 1) * NXP + (i + mbc - 1);
                    long a_lo = ((long)(1 - kbc + mbc - 1) * NYP) * NXP + base_j;
                    long a_1 = ((long)(1 + mbc - 1) * NYP) * NXP + base_j;
                    long a_hi = ((long)(mz + kbc + mbc - 1) * NYP) * NXP + base_j;
                    long a_mz = ((long)(mz + mbc - 1) * NYP) * NXP + base_j;
                    u[a_lo] = u[a_1];
                    u[a_hi] = u[a_mz];
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++) {
                    double au = UU(i, j, k + 1), aq = UU(i, j, k), ad = UU(i, j, k - 1);
                    QT(i, j, k) = (ad - aq - aq + au);
                }
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    UU(i, j, k) = UU(i, j, k) + dtdz2 * CO(i, j, k) * QT(i, j, k);
    }
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++) QG(q, mdiffq, i, j, k) = UU(i, j, k);
    free(qt); free(u); free(coefs);
#undef UU
#undef QT
#undef CO
}

/* ------------------------------------------------------------------------ */
/* src3: src3.f90:203-634.  forcemth 0, 1 and 3; omega1 must be 0.          */
int orc_src3(const orc_grid *g, const orc_params *p, double *q, const double *aux,
             double t, double dt, const int *mthbc_in, int *ndtr_out, int *ndtrc_out)
{
    const int meqn = g->meqn, mbc = g->mbc, mx = g->mx, my = g->my, mz = g->mz;
    if (p->omega1 != 0.0) return ORC_UNSUPPORTED;
    if (!(p->forcemth == 0.0 || p->forcemth == 1.0 || p->forcemth == 3.0)) return ORC_UNSUPPORTED;
    /* conserved -> (rho, u', v', w', T'): src3.f90:242-304 */
    for (int k = 1 - mbc; k <= mz + mbc; k++)
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                double gamma = AG(19, i, j, k);
                double gamma1 = gamma - 1.0;
                double q1 = QG(q, 1, i, j, k), q2 = QG(q, 2, i, j, k), q3 = QG(q, 3, i, j, k);
                double q4 = QG(q, 4, i, j, k);
                double qtemp = gamma1 * (QG(q, 5, i, j, k) - 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4))
                               / (q1 * AG(18, i, j, k));
                double a1 = AG(1, i, j, k), a2 = AG(2, i, j, k), a3 = AG(3, i, j, k), a4 = AG(4, i, j, k);
                double auxtemp = gamma1 * (AG(5, i, j, k) - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4))
                                 / (a1 * AG(18, i, j, k));
                QG(q, 5, i, j, k) = qtemp - auxtemp;
                QG(q, 2, i, j, k) = q2 / q1 - a2 / a1;
                QG(q, 3, i, j, k) = q3 / q1 - a3 / a1;
                QG(q, 4, i, j, k) = q4 / q1 - a4 / a1;
            }
    /* substep counts: src3.f90:314-322 */
    double dmin = g->dx < g->dy ? g->dx : g->dy;
    dmin = dmin < g->dz ? dmin : g->dz;
    double dxyzmin2 = dmin * dmin;
    double difmax = (4.0 / 3.0) * AG(11, 1, 1, mz);
    double difcfl = 0.45;
    int ndtr = (int)ceil((dt * difmax / (dxyzmin2)) / difcfl);
    double dtr = dt / ndtr;
    double conmax = AG(12, 1, 1, mz);
    double concfl = 0.45;
    /* With no viscosity at the top (aux(11) = 0) ndtr is 0, dtr is
       infinite and the reference's ceiling of it is undefined; its substep
       loop (src3.f90:341) then runs no pass at all, so the count is 0 here. */
    int ndtrc = ndtr > 0 ? (int)ceil((dtr * conmax / (dxyzmin2)) / concfl) : 0;
    double dtrc = dtr / ndtrc;
    *ndtr_out = ndtr;
    *ndtrc_out = ndtrc;
    int mthbc[6];
    memcpy(mthbc, mthbc_in, sizeof(mthbc));
    mthbc[4] = -1;
    mthbc[5] = -1;
    for (int nn = 1; nn <= ndtr; nn++) {
        int st = orc_bc3(g, q, aux, mthbc);
        if (st) return st;
        diffuse(g, q, aux, 11, dtr);
        conduct(g, q, 5, aux, 12, dtrc, ndtrc);
    }
    /* reassemble the interior: src3.f90:380-418 */
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++) {
                double gamma = AG(19, i, j, k);
                double gamma1 = gamma - 1.0;
                double a1 = AG(1, i, j, k), a2 = AG(2, i, j, k), a3 = AG(3, i, j, k), a4 = AG(4, i, j, k);
                double q1 = QG(q, 1, i, j, k);
                QG(q, 2, i, j, k) = QG(q, 2, i, j, k) * q1 + (a2 / a1) * q1;
                QG(q, 3, i, j, k) = QG(q, 3, i, j, k) * q1 + (a3 / a1) * q1;
                QG(q, 4, i, j, k) = QG(q, 4, i, j, k) * q1 + (a4 / a1) * q1;
                double auxtemp = gamma1 * (AG(5, i, j, k) - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4))
                                 / (a1 * AG(18, i, j, k));
                double q2 = QG(q, 2, i, j, k), q3 = QG(q, 3, i, j, k), q4 = QG(q, 4, i, j, k);
                QG(q, 5, i, j, k) = (auxtemp + QG(q, 5, i, j, k)) * q1 * AG(18, i, j, k) / gamma1 +
                                    0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4);
            }
    /* forcing */
    if (p->forcemth == 1.0) { /* src3.f90:430-445 with fcns.f gaussian/forcefnz */
        for (int i = 1; i <= mx; i++) {
            double xcell = g->xlower + (i - 0.5) * g->dx - t * p->vsrcx;
            for (int j = 1; j <= my; j++) {
                double ycell = g->ylower + (j - 0.5) * g->dy;
                for (int k = 1; k <= mz; k++) {
                    double zcell = g->zlower + (k - 0.5) * g->dz;
                    double gau = exp(-.5 * ((xcell - p->xpos) * (xcell - p->xpos) / (p->xwidth * p->xwidth))) *
                                 exp(-.5 * ((ycell - p->ypos) * (ycell - p->ypos) / (p->ywidth * p->ywidth))) *
                                 exp(-.5 * ((zcell - p->zpos) * (zcell - p->zpos) / (p->zwidth * p->zwidth))) *
                                 exp(-.5 * ((t - p->tcenter) * (t - p->tcenter) / (p->twidth * p->twidth)));
                    double ffz = sin(p->omega * (t - p->tcenter) - p->propx * (xcell - p->xpos) -
                                     p->propz * (ycell - p->ypos));
                    double bf = dt * p->amplitude * gau * ffz;
                    QG(q, 4, i, j, k) = QG(q, 4, i, j, k) + bf * QG(q, 1, i, j, k);
                }
            }
        }
    } else if (p->forcemth == 3.0) { /* src3.f90:514-607 */
        const double pi = 4.0 * atan(1.0);
        for (int i = 1; i <= mx; i++) {
            double xcell = g->xlower + (i - 0.5) * g->dx - t * p->vsrcx;
            for (int j = 1; j <= my; j++) {
                double ycell = g->ylower + (j - 0.5) * g->dy;
                for (int k = 1; k <= mz; k++) {
                    double zcell = g->zlower + (k - 0.5) * g->dz;
                    if (zcell > 15000.0) break;
                    double gaussian2 = exp(-.5 * ((xcell - p->xpos) * (xcell - p->xpos) / (p->xwidth * p->xwidth))) *
                                       exp(-.5 * ((ycell - p->ypos) * (ycell - p->ypos) / (p->ywidth * p->ywidth))) *
                                       exp(-.5 * ((t - p->tcenter) * (t - p->tcenter) / (p->twidth * p->twidth)));
                    double precipw = p->amplitude * gaussian2;
                    if (!(precipw > 0.1)) continue;
                    double th = 10.79 + 0.05 * precipw;
                    double ph = 6.93 - 0.03 * precipw;
                    double hamp = -0.0001 + 0.0025 * precipw;
                    double tc = 1.12 + 0.10 * precipw;
                    double pc = 0.63 + 0.02 * precipw;
                    double bc = 0.37 + 0.00 * precipw;
                    double camp = -0.0003 - 0.0005 * precipw;
                    double z2 = zcell / 1000.0;
                    double heat;
                    if ((bc <= z2) && (z2 <= pc))
                        heat = camp * sin((2.0 * pi / (4.0 * (pc - bc))) * (z2 - bc));
                    else if ((pc <= z2) && (z2 <= tc))
                        heat = -camp * sin((2.0 * pi / (4.0 * (tc - pc))) * (z2 - tc));
                    else if ((pc <= z2) && (z2 <= ph))
                        heat = hamp * sin((2.0 * pi / (4.0 * (ph - tc))) * (z2 - tc));
                    else if ((ph <= z2) && (z2 <= th))
                        heat = -hamp * sin((2.0 * pi / (4.0 * (th - ph))) * (z2 - th));
                    else
                        heat = 0.0;
                    double gamma = AG(19, i, j, k);
                    double gamma1 = gamma - 1.0;
                    double q1 = QG(q, 1, i, j, k), q2 = QG(q, 2, i, j, k), q3 = QG(q, 3, i, j, k);
                    double q4 = QG(q, 4, i, j, k);
                    double a1 = AG(1, i, j, k), a2 = AG(2, i, j, k), a3 = AG(3, i, j, k), a4 = AG(4, i, j, k);
                    double qtemp = gamma1 * (QG(q, 5, i, j, k) - 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4))
                                   / (q1 * AG(18, i, j, k));
                    double auxtemp = gamma1 * (AG(5, i, j, k) - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4))
                                     / (a1 * AG(18, i, j, k));
                    double t5 = qtemp - auxtemp;
                    t5 = t5 + dt * heat;
                    auxtemp = gamma1 * (AG(5, i, j, k) - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4))
                              / (a1 * AG(18, i, j, k));
                    QG(q, 5, i, j, k) = (auxtemp + t5) * q1 * AG(18, i, j, k) / gamma1 +
                                        0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4);
                }
            }
        }
    }
    /* clamp tracers: src3.f90:616-626 */
    if (meqn > 5)
        for (int k = 1; k <= mz; k++)
            for (int j = 1; j <= my; j++)
                for (int i = 1; i <= mx; i++)
                    for (int m = 6; m <= meqn; m++)
                        QG(q, m, i, j, k) = fmaxF(0.0, fminF(QG(q, m, i, j, k), 1.0));
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* claw3: mpilib/claw3_mpi.f:631-951, method(1) = 1 (variable dt),
 * method(5) = 1 (Godunov source splitting), one rank.                       */
int orc_claw3(const orc_grid *g, const orc_params *p, double *q, const double *aux,
              double tstart, double *tend, double *dtv, double *cflv, int *nv,
              const int *method, const int *mthbc,
              double *log_dt, double *log_cfl, int *log_acc, int maxlog, int *nlog)
{
    const size_t nq = (size_t)NXP * NYP * NZP * g->meqn;
    if (method[0] != 1 || method[4] != 1 || method[1] != 2 || method[2] != 22 || method[5] != 0)
        return ORC_UNSUPPORTED;
    double *qwork = malloc(nq * sizeof(double));
    double t = tstart, dt = dtv[0], cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    int maxn = nv[0], st = ORC_OK;
    *nlog = 0;
    nv[1] = 0;
    for (int n = 1; n <= maxn; n++) {
        double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;
        for (;;) {
            st = orc_bc3(g, q, aux, mthbc);
            if (st) goto out;
            orc_b4step3(g, q);
            memcpy(qwork, q, nq * sizeof(double));
            st = orc_step3(g, p->nspecadv, qwork, q, aux, dt, &cfl);
            if (st) goto out;
            t = told + dt;
            if (*nlog < maxlog) {
                log_dt[*nlog] = dt; log_cfl[*nlog] = cfl; log_acc[*nlog] = cfl <= cflv[0];
                (*nlog)++;
            }
            if (cfl <= cflv[0]) {
                cflmax = fmaxF(cfl, cflmax);
                break;
            }
            t = told;
            memcpy(q, qwork, nq * sizeof(double));
            dt = fminF(dtv[1], dt * cflv[1] / cfl);
        }
        {
            int ndtr, ndtrc;
            st = orc_src3(g, p, q, aux, told, dt, mthbc, &ndtr, &ndtrc);
            if (st) goto out;
        }
        if (cfl == 0.0) dt = dtv[1];
        else dt = fminF(dtv[1], dt * cflv[1] / cfl);
        dtmin = fminF(dt, dtmin);
        dtmax = fmaxF(dt, dtmax);
        nv[1] = nv[1] + 1;
        if (t >= *tend) break;
    }
out:
    *tend = t;
    cflv[2] = cflmax;
    cflv[3] = cfl;
    dtv[2] = dtmin;
    dtv[3] = dtmax;
    dtv[4] = dt;
    free(qwork);
    return st;
}
