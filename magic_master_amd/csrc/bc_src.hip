// Boundary (bc3), pre-step (b4step3) and source-pass (src3, diffuse,
// conduct) kernels of the MAGIC step, FP64, on component planes.
//
// Index convention: a cell with Fortran indices (i, j, k) in
// [1-mbc, mx+mbc] x [1-mbc, my+mbc] x [1-mbc, mz+mbc] lives at
// ((k+mbc-1)*NY + (j+mbc-1))*NX + (i+mbc-1) of every plane.
#include "magic_dev.h"
#include "bc_src.h"

namespace magic {

__device__ __forceinline__ long long cid(const Dims &d, int i, int j, int k)
{
    return ((long long)(k + d.mbc - 1) * d.NY + (j + d.mbc - 1)) * d.NX + (i + d.mbc - 1);
}

// ---------------------------------------------------------------- bc3 ----
// x faces: bc3_mpi.f:97-312.  One thread per (j, k) of the padded face.
__global__ void k_bc_x(Dims d, double *q, int nc, int lo, int hi)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int k = (int)blockIdx.y + 1 - d.mbc;
    if (j > d.my + d.mbc || k > d.mz + d.mbc) return;
    const bool inner = j >= 1 && j <= d.my && k >= 1 && k <= d.mz;
    for (int m = 0; m < nc; m++) {
        double *p = q + (long long)m * d.P;
        for (int ibc = 1; ibc <= d.mbc; ibc++) {
            if (lo == 1) p[cid(d, 1 - ibc, j, k)] = p[cid(d, 1, j, k)];
            else if (lo == 3) p[cid(d, 1 - ibc, j, k)] = (m == 1 ? -1.0 : 1.0) * p[cid(d, ibc, j, k)];
            else if (lo == 2 && inner) p[cid(d, 1 - ibc, j, k)] = p[cid(d, d.mx - d.mbc + (d.mbc + 1 - ibc), j, k)];
            if (hi == 1) p[cid(d, d.mx + ibc, j, k)] = p[cid(d, d.mx, j, k)];
            else if (hi == 3) p[cid(d, d.mx + ibc, j, k)] = (m == 1 ? -1.0 : 1.0) * p[cid(d, d.mx + 1 - ibc, j, k)];
            else if (hi == 2 && inner) p[cid(d, d.mx + ibc, j, k)] = p[cid(d, ibc, j, k)];
        }
    }
}

// y faces: bc3_mpi.f:320-536.  One thread per (i, k).
__global__ void k_bc_y(Dims d, double *q, int nc, int lo, int hi)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int k = (int)blockIdx.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || k > d.mz + d.mbc) return;
    const bool inner = k >= 1 && k <= d.mz;
    for (int m = 0; m < nc; m++) {
        double *p = q + (long long)m * d.P;
        for (int jbc = 1; jbc <= d.mbc; jbc++) {
            if (lo == 1) p[cid(d, i, 1 - jbc, k)] = p[cid(d, i, 1, k)];
            else if (lo == 3) p[cid(d, i, 1 - jbc, k)] = (m == 2 ? -1.0 : 1.0) * p[cid(d, i, jbc, k)];
            else if (lo == 2 && inner) p[cid(d, i, 1 - jbc, k)] = p[cid(d, i, d.my - d.mbc + (d.mbc + 1 - jbc), k)];
            if (hi == 1) p[cid(d, i, d.my + jbc, k)] = p[cid(d, i, d.my, k)];
            else if (hi == 3) p[cid(d, i, d.my + jbc, k)] = (m == 2 ? -1.0 : 1.0) * p[cid(d, i, d.my + 1 - jbc, k)];
            else if (hi == 2 && inner) p[cid(d, i, d.my + jbc, k)] = p[cid(d, i, jbc, k)];
        }
    }
}

// z faces: bc3_mpi.f:544-809.  Extrapolation and wall keep the perturbation
// from the background scaled by sc = (rho_bar_ghost/rho_bar_1)**0.5, which
// the host precomputes (zratio[side][kbc-1][plane index of the (i,j) column]).
__global__ void k_bc_z(Dims d, double *q, int nc, int lo, int hi, const double *aux1,
                       const double *aux2, const double *aux3, const double *aux5,
                       const double *zratio)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)blockIdx.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || j > d.my + d.mbc) return;
    const long long col = (long long)(j + d.mbc - 1) * d.NX + (i + d.mbc - 1);
    const long long NXY = (long long)d.NX * d.NY;
    for (int side = 0; side < 2; side++) {
        const int mth = side == 0 ? lo : hi;
        if (mth == 1 || mth == 3) {
            const int kb = side == 0 ? 1 : d.mz;
            const long long cb = cid(d, i, j, kb);
            for (int kbc = 1; kbc <= d.mbc; kbc++) {
                const int kg = side == 0 ? 1 - kbc : d.mz + kbc;
                const long long cg = cid(d, i, j, kg);
                const double sc = zratio[((long long)side * d.mbc + (kbc - 1)) * NXY + col];
                q[0 * d.P + cg] = aux1[cg] + (q[0 * d.P + cb] - aux1[cb]) * sc;
                q[1 * d.P + cg] = aux2[cg] + (q[1 * d.P + cb] - aux2[cb]) * sc;
                q[2 * d.P + cg] = aux3[cg] + (q[2 * d.P + cb] - aux3[cb]) * sc;
                q[3 * d.P + cg] = q[3 * d.P + cb] * sc;
                q[4 * d.P + cg] = aux5[cg] + (q[4 * d.P + cb] - aux5[cb]) * sc;
                for (int m = 5; m < nc; m++) q[m * d.P + cg] = q[m * d.P + cb];
            }
            if (mth == 3)
                for (int kbc = 1; kbc <= d.mbc; kbc++) {
                    const int kg = side == 0 ? 1 - kbc : d.mz + kbc;
                    const int km = side == 0 ? kbc : d.mz + 1 - kbc;
                    q[3 * d.P + cid(d, i, j, kg)] = -q[3 * d.P + cid(d, i, j, km)];
                }
        } else if (mth == 2) {
            for (int kbc = 1; kbc <= d.mbc; kbc++)
                for (int m = 0; m < nc; m++) {
                    double *p = q + (long long)m * d.P;
                    if (side == 0) p[cid(d, i, j, kbc - d.mbc)] = p[cid(d, i, j, d.mz - d.mbc + kbc)];
                    else p[cid(d, i, j, d.mz + kbc)] = p[cid(d, i, j, kbc)];
                }
        }
    }
}

// b4step3.f:35-45: clamp tracers to [0, 1] on every cell.
__global__ void k_clamp(double *q, long long P, int nc)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    for (int m = 5; m < nc; m++) {
        const double v = q[m * P + c];
        q[m * P + c] = fmaxD(0.0, fminD(v, 1.0));
    }
}

// ---------------------------------------------------------------- src3 ---
// conserved -> (rho, u', v', w', T') on every cell: src3.f90:242-304
__global__ void k_src_convert(long long P, double *q, const Aux a)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    const double gamma1 = a.gam[c] - 1.0;
    const double q1 = q[c], q2 = q[P + c], q3 = q[2 * P + c], q4 = q[3 * P + c];
    const double a1 = a.rho[c], a2 = a.mx[c], a3 = a.my[c], a4 = a.mz[c];
    const double qtemp = gamma1 * (q[4 * P + c] - 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4)) /
                         (q1 * a.R[c]);
    const double auxtemp = gamma1 * (a.E[c] - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4)) /
                           (a1 * a.R[c]);
    q[4 * P + c] = qtemp - auxtemp;
    q[P + c] = q2 / q1 - a2 / a1;
    q[2 * P + c] = q3 / q1 - a3 / a1;
    q[3 * P + c] = q4 / q1 - a4 / a1;
}

// diffuse.f:48-54: z ghosts of q(2:4) from the first/last interior layer.
__global__ void k_zghost_uvw(Dims d, double *q)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)blockIdx.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || j > d.my + d.mbc) return;
    for (int m = 1; m <= 3; m++) {
        double *p = q + (long long)m * d.P;
        const double lo = p[cid(d, i, j, 1)], hi = p[cid(d, i, j, d.mz)];
        for (int kbc = 1; kbc <= d.mbc; kbc++) {
            p[cid(d, i, j, 1 - kbc)] = lo;
            p[cid(d, i, j, d.mz + kbc)] = hi;
        }
    }
}

// Interior thread mapping: x fastest.
#define INTERIOR_IJK                                                          \
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;           \
    const int j = (int)blockIdx.y + 1;                                        \
    const int k = (int)blockIdx.z + 1;                                        \
    if (i > d.mx) return;                                                     \
    const long long c = cid(d, i, j, k);

// diffuse.f:70-118: cross-derivative terms from u (copy of q(2:4) taken
// after the z-ghost fill); writes q(2:4) on the interior.
__global__ void k_diff_cross(Dims d, double *q, const double *u, const double *coef,
                             double dxy, double dxz, double dyz)
{
    INTERIOR_IJK
    const long long sx = 1, sy = d.NX, sz = (long long)d.NX * d.NY;
    const double *u1 = u, *u2 = u + d.P, *u3 = u + 2 * d.P;
    auto X = [&](const double *f, long long a, long long b) {
        return f[c + a + b] - f[c - a + b] + f[c - a - b] - f[c + a - b];
    };
    const double ctxxy = X(u2, sx, sy), ctxxz = X(u3, sx, sz);
    const double ctyxy = X(u1, sx, sy), ctyyz = X(u3, sy, sz);
    const double ctzxz = X(u1, sx, sz), ctzyz = X(u2, sy, sz);
    const double co = coef[c];
    q[d.P + c] = u1[c] + co * (dxy * ctxxy + dxz * ctxxz);
    q[2 * d.P + c] = u2[c] + co * (dxy * ctyxy + dyz * ctyyz);
    q[3 * d.P + c] = u3[c] + co * (dxz * ctzxz + dyz * ctzyz);
}

// interior copy of planes src -> dst (n planes with stride P)
__global__ void k_copy_interior(Dims d, double *dst, const double *src, int nplanes)
{
    INTERIOR_IJK
    for (int m = 0; m < nplanes; m++) dst[m * d.P + c] = src[m * d.P + c];
}

// One explicit directional diffusion pass (diffuse.f:153-208, conduct.f:262-323):
//   out = u + f*coef*(second difference of u along axis),
// computed on the interior.  The interior values of u come from src and the
// values on the ghost layers (which the reference copies once per call and
// never refreshes between passes) from ghost; src may equal ghost.  The
// second difference keeps the reference's operand order: (up-mid-mid+down)
// along x and y, (down-mid-mid+up) along z.
__global__ void k_pass(Dims d, double *out, const double *src, const double *ghost,
                       const double *coef, double f, int axis)
{
    INTERIOR_IJK
    long long s;
    bool lo_in, hi_in;
    if (axis == 0) { s = 1; lo_in = i > 1; hi_in = i < d.mx; }
    else if (axis == 1) { s = d.NX; lo_in = j > 1; hi_in = j < d.my; }
    else { s = (long long)d.NX * d.NY; lo_in = k > 1; hi_in = k < d.mz; }
    const double aq = src[c];
    const double am = lo_in ? src[c - s] : ghost[c - s];
    const double ap = hi_in ? src[c + s] : ghost[c + s];
    const double qt = axis == 2 ? (am - aq - aq + ap) : (ap - aq - aq + am);
    out[c] = aq + f * coef[c] * qt;
}

// conduct.f:86-91: z ghosts of u, for j in [1-mbc, my+mbc] and i in
// [1-mbc, mz+mbc] -- the reference bounds i by mz, not mx.  When mx > mz the
// columns i > mz+mbc keep their stale z ghosts (which then enter the z
// stencil of the interior cells of those columns).  When mz > mx the
// out-of-row i addresses the array linearly: within a layer the writes land
// on later rows of the same layer with those rows' own values (what the
// in-row fill writes anyway), so this kernel fills whole rows and
// k_conduct_zspill adds the writes that cross into the next layer.
__global__ void k_conduct_zghost(Dims d, double *u)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)blockIdx.y + 1 - d.mbc;
    const int ilast = (d.mz < d.mx ? d.mz : d.mx) + d.mbc;
    if (i > ilast || j > d.my + d.mbc) return;
    const double lo = u[cid(d, i, j, 1)], hi = u[cid(d, i, j, d.mz)];
    for (int kbc = 1; kbc <= d.mbc; kbc++) {
        u[cid(d, i, j, 1 - kbc)] = lo;
        u[cid(d, i, j, d.mz + kbc)] = hi;
    }
}

// The layer-crossing writes of conduct.f:86-91 when mz > mx.  The rows at
// the end of a layer address i up to mz+mbc, S = mz-mx elements past the
// layer's last cell, so at in-layer offsets r < S the loop writes
//   layer 2-kbc <- layer 2          (u(i,j,1-kbc) = u(i,j,1), kbc = 1..mbc)
//   layer mz+1+kbc <- layer mz+1    (u(i,j,mz+kbc) = u(i,j,mz))
// after the in-row fills of the rows those offsets fall in (the host checks
// that the rows written lie before the rows that write them).  Writes past
// the last layer leave the reference's array; they are dropped here.
// Layers are Fortran k indices; r counts from cell (1-mbc, 1-mbc).
__global__ void k_conduct_zspill(Dims d, double *u, int S)
{
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= S) return;
    const long long NXY = (long long)d.NX * d.NY;
    auto L = [&](int k) { return (long long)(k + d.mbc - 1) * NXY + r; };
    const double lo = u[L(2)], hi = u[L(d.mz + 1)];
    for (int kbc = 1; kbc <= d.mbc; kbc++) {
        u[L(2 - kbc)] = lo;
        if (d.mz + 1 + kbc <= d.mz + d.mbc) u[L(d.mz + 1 + kbc)] = hi;
    }
}

// src3.f90:380-418 (reassembly), 430-445 / 514-607 (forcing), 616-626
// (tracer clamp), on the interior.
__global__ void k_src_finish(Dims d, double *q, int nc, const Aux a, Forcing f)
{
    INTERIOR_IJK
    const long long P = d.P;
    const double gamma1 = a.gam[c] - 1.0;
    const double a1 = a.rho[c], a2 = a.mx[c], a3 = a.my[c], a4 = a.mz[c], R = a.R[c];
    const double q1 = q[c];
    const double q2 = q[P + c] * q1 + (a2 / a1) * q1;
    const double q3 = q[2 * P + c] * q1 + (a3 / a1) * q1;
    double q4 = q[3 * P + c] * q1 + (a4 / a1) * q1;
    const double auxtemp = gamma1 * (a.E[c] - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4)) / (a1 * R);
    double q5 = (auxtemp + q[4 * P + c]) * q1 * R / gamma1 + 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4);
    const double xcell = f.xlower + (i - 0.5) * f.dx - f.t * f.vsrcx;
    const double ycell = f.ylower + (j - 0.5) * f.dy;
    const double zcell = f.zlower + (k - 0.5) * f.dz;
    if (f.forcemth == 1) {
        const double gau = exp(-.5 * ((xcell - f.xpos) * (xcell - f.xpos) / (f.xwidth * f.xwidth))) *
                           exp(-.5 * ((ycell - f.ypos) * (ycell - f.ypos) / (f.ywidth * f.ywidth))) *
                           exp(-.5 * ((zcell - f.zpos) * (zcell - f.zpos) / (f.zwidth * f.zwidth))) *
                           exp(-.5 * ((f.t - f.tcenter) * (f.t - f.tcenter) / (f.twidth * f.twidth)));
        const double ffz = sin(f.omega * (f.t - f.tcenter) - f.propx * (xcell - f.xpos) -
                               f.propz * (ycell - f.ypos));
        const double bf = f.dt * f.amplitude * gau * ffz;
        q4 = q4 + bf * q1;
    } else if (f.forcemth == 3 && !(zcell > 15000.0)) {
        const double g2 = exp(-.5 * ((xcell - f.xpos) * (xcell - f.xpos) / (f.xwidth * f.xwidth))) *
                          exp(-.5 * ((ycell - f.ypos) * (ycell - f.ypos) / (f.ywidth * f.ywidth))) *
                          f.tgauss;
        const double precipw = f.amplitude * g2;
        if (precipw > 0.1) {
            // Stephan & Alexander heating profile (src3.f90:534-572)
            const double pi = 3.141592653589793;
            const double th = 10.79 + 0.05 * precipw, ph = 6.93 - 0.03 * precipw;
            const double hamp = -0.0001 + 0.0025 * precipw;
            const double tc = 1.12 + 0.10 * precipw, pc = 0.63 + 0.02 * precipw;
            const double bc = 0.37 + 0.00 * precipw, camp = -0.0003 - 0.0005 * precipw;
            const double z2 = zcell / 1000.0;
            double heat;
            if ((bc <= z2) && (z2 <= pc)) heat = camp * sin((2.0 * pi / (4.0 * (pc - bc))) * (z2 - bc));
            else if ((pc <= z2) && (z2 <= tc)) heat = -camp * sin((2.0 * pi / (4.0 * (tc - pc))) * (z2 - tc));
            else if ((pc <= z2) && (z2 <= ph)) heat = hamp * sin((2.0 * pi / (4.0 * (ph - tc))) * (z2 - tc));
            else if ((ph <= z2) && (z2 <= th)) heat = -hamp * sin((2.0 * pi / (4.0 * (th - ph))) * (z2 - th));
            else heat = 0.0;
            const double qtemp = gamma1 * (q5 - 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4)) / (q1 * R);
            double t5 = qtemp - auxtemp;
            t5 = t5 + f.dt * heat;
            q5 = (auxtemp + t5) * q1 * R / gamma1 + 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4);
        }
    }
    q[P + c] = q2;
    q[2 * P + c] = q3;
    q[3 * P + c] = q4;
    q[4 * P + c] = q5;
    for (int m = 5; m < nc; m++) q[m * P + c] = fmaxD(0.0, fminD(q[m * P + c], 1.0));
}

// ---------------------------------------------------------- layout I/O ---
// reference AoS q(meqn, i, j, k) <-> component planes
__global__ void k_aos_to_planes(double *planes, const double *aos, long long P, int nc)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    for (int m = 0; m < nc; m++) planes[m * P + c] = aos[c * nc + m];
}

__global__ void k_planes_to_aos(double *aos, const double *planes, long long P, int nc)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    for (int m = 0; m < nc; m++) aos[c * nc + m] = planes[m * P + c];
}

}  // namespace magic
