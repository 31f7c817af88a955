// Boundary (bc3), pre-step (b4step3) and source-pass (src3, diffuse,
// conduct) kernels of the MAGIC step, FP64 form, on component planes.  The
// kernels shared with the FP32 form live in stencil_body.h and are
// instantiated for double at the end of this file.
//
// Index convention: a cell with Fortran indices (i, j, k) in
// [1-mbc, mx+mbc] x [1-mbc, my+mbc] x [1-mbc, mz+mbc] lives at
// ((k+mbc-1)*NY + (j+mbc-1))*NX + (i+mbc-1) of every plane.
#include "magic_dev.h"
#include "stencil_body.h"
#include "batch_ops.h"

namespace magic {

// ---------------------------------------------------------------- bc3 ----
// x, y faces and the tracer clamp: stencil_body.h.
// z faces: bc3_mpi.f:544-809.  Extrapolation and wall keep the perturbation
// from the background scaled by sc = (rho_bar_ghost/rho_bar_1)**0.5, which
// the host precomputes (zratio[side][kbc-1][plane index of the (i,j) column]).
template <class G>
__device__ __forceinline__ void b_bc_z(const uint3 BI, Dims d, double *q, int nc, int lo, int hi, const double *aux1, const double *aux2, const double *aux3, const double *aux5, const double *zratio, const G &gate)
{
    if (!gate.on()) return;
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)BI.y + 1 - d.mbc;
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
template <class G>
__global__ void k_bc_z(Dims d, double *q, int nc, int lo, int hi, const double *aux1,
                       const double *aux2, const double *aux3, const double *aux5,
                       const double *zratio, G gate)
{
    b_bc_z(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, lo, hi, aux1, aux2, aux3, aux5, zratio, gate);
}

// ---------------------------------------------------------------- src3 ---
// conserved -> (rho, u', v', w', T') on every cell: src3.f90:242-304.
// COL: the background comes as columns (index kk), else as planes.
// dq (one grid's sweeps on two streams): the second stream's increments,
// added to every component of the cell first, as a separate q += dq pass
// would add them (nc components; tracers are only added).
template <bool COL, class G>
__device__ __forceinline__ void b_src_convert(const uint3 BI, Dims d, double *q, const Aux a, const G &gate,
                                              const double *dq = nullptr, int nc = 0)
{
    if (!gate.on()) return;
    PADDED_IJK
    const long long P = d.P;
    if (dq) {
        q[c] = q[c] + dq[c];
        for (int m = 1; m < nc; m++) q[m * P + c] = q[m * P + c] + dq[m * P + c];
    }
    const long long x = COL ? (long long)kk : c;
    const double gamma1 = a.gam[x] - 1.0;
    const double q1 = q[c], q2 = q[P + c], q3 = q[2 * P + c], q4 = q[3 * P + c];
    const double a1 = a.rho[x], a2 = a.mx[x], a3 = a.my[x], a4 = a.mz[x];
    const double qtemp = gamma1 * (q[4 * P + c] - 0.5 * (1.0 / q1) * (q2 * q2 + q3 * q3 + q4 * q4)) /
                         (q1 * a.R[x]);
    const double auxtemp = gamma1 * (a.E[x] - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4)) /
                           (a1 * a.R[x]);
    q[4 * P + c] = qtemp - auxtemp;
    q[P + c] = q2 / q1 - a2 / a1;
    q[2 * P + c] = q3 / q1 - a3 / a1;
    q[3 * P + c] = q4 / q1 - a4 / a1;
}
template <bool COL, class G>
__global__ void k_src_convert(Dims d, double *q, const Aux a, G gate, const double *dq, int nc)
{
    b_src_convert<COL>(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, a, gate, dq, nc);
}

// src3.f90:380-418 (reassembly), 430-445 / 514-607 (forcing), 616-626
// (tracer clamp), on the interior.
template <bool COL, class G>
__device__ __forceinline__ void b_src_finish(const uint3 BI, Dims d, double *q, int nc, const Aux a, Forcing f, const G &gate)
{
    if (!gate.on()) return;
    gate.forcing(f.t, f.dt, f.tgauss);
    INTERIOR_IJK
    const long long P = d.P;
    const long long x = COL ? (long long)(k + d.mbc - 1) : c;
    const double gamma1 = a.gam[x] - 1.0;
    const double a1 = a.rho[x], a2 = a.mx[x], a3 = a.my[x], a4 = a.mz[x], R = a.R[x];
    const double q1 = q[c];
    const double q2 = q[P + c] * q1 + (a2 / a1) * q1;
    const double q3 = q[2 * P + c] * q1 + (a3 / a1) * q1;
    double q4 = q[3 * P + c] * q1 + (a4 / a1) * q1;
    const double auxtemp = gamma1 * (a.E[x] - 0.5 * (1.0 / a1) * (a2 * a2 + a3 * a3 + a4 * a4)) / (a1 * R);
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
            const double heat = sa_heating(precipw, zcell);
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
template <bool COL, class G>
__global__ void k_src_finish(Dims d, double *q, int nc, const Aux a, Forcing f, G gate)
{
    b_src_finish<COL>(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, a, f, gate);
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

// The FP64 form's own bodies for the batched work-list kernels
// (batch_ops.h); its scenarios keep background columns
struct Form64 {
    __device__ static void bc_z(uint3 bi, const BatchOps<double> &o, cptr<BatchOps<double>>, const StepGate &g)
    {
        b_bc_z(bi, o.d, o.q, o.nc, lohi(o.bc[4]), lohi(o.bc[5]), o.aux1, o.aux2, o.aux3, o.aux5, o.zratio, g);
    }
    __device__ static void convert(uint3 bi, const BatchOps<double> &o, const StepGate &g)
    {
        b_src_convert<true>(bi, o.d, o.q, o.ax, g);
    }
    __device__ static void finish(uint3 bi, const BatchOps<double> &o, const StepGate &g)
    {
        b_src_finish<true>(bi, o.d, o.q, o.nc, o.ax, o.f, g);
    }
    __device__ static const double *nu(const BatchOps<double> &o) { return o.nu64; }
    __device__ static const double *kap(const BatchOps<double> &o) { return o.kap64; }
};

template <>
const void *batch_op_fn<double>(int op)
{
    switch (op) {
#define MAGIC_OP(X) case X: return reinterpret_cast<const void *>(&k_batch_op<double, X, Form64>);
    MAGIC_OP(OP_BCX) MAGIC_OP(OP_BCY) MAGIC_OP(OP_BCZ) MAGIC_OP(OP_CLAMP) MAGIC_OP(OP_BACKUP) MAGIC_OP(OP_RESTORE)
    MAGIC_OP(OP_CONVERT) MAGIC_OP(OP_ZGHOST_UVW) MAGIC_OP(OP_DIFF_CROSS) MAGIC_OP(OP_DIFF_XYZ) MAGIC_OP(OP_COPY_Q5)
    MAGIC_OP(OP_PASS) MAGIC_OP(OP_CONDUCT_ZGHOST) MAGIC_OP(OP_ZSPILL) MAGIC_OP(OP_COPY_INTERIOR) MAGIC_OP(OP_FINISH)
    MAGIC_OP(OP_MERGE)
#undef MAGIC_OP
    default: return nullptr;
    }
}

// FP64 instantiations of the shared kernels
template __global__ void k_bc_x<double, NoGate>(Dims, double *, int, int, int, NoGate);
template __global__ void k_bc_y<double, NoGate>(Dims, double *, int, int, int, NoGate);
template __global__ void k_clamp<double, NoGate>(double *, long long, int, NoGate);
template __global__ void k_zghost_uvw<double, NoGate>(Dims, double *, NoGate);
template __global__ void k_bc_z<NoGate>(Dims, double *, int, int, int, const double *, const double *, const double *,
                                        const double *, const double *, NoGate);
template __global__ void k_src_convert<false, NoGate>(Dims, double *, const Aux, NoGate, const double *, int);
template __global__ void k_src_convert<true, NoGate>(Dims, double *, const Aux, NoGate, const double *, int);
template __global__ void k_src_finish<false, NoGate>(Dims, double *, int, const Aux, Forcing, NoGate);
template __global__ void k_src_finish<true, NoGate>(Dims, double *, int, const Aux, Forcing, NoGate);
template __global__ void k_diff_xyz<double, true, NoGate>(Dims, double *, const double *, const double *, double, double,
                                                         double, int, NoGate);
template __global__ void k_diff_cross<double, true, NoGate>(Dims, double *, const double *, const double *, double, double,
                                                           double, int, NoGate);
template __global__ void k_pass<double, true, NoGate>(Dims, double *, const double *, const double *, const double *, double,
                                                     int, NoGate);
template __global__ void k_diff_xyz<double, false, NoGate>(Dims, double *, const double *, const double *, double, double,
                                                          double, int, NoGate);
template __global__ void k_diff_cross<double, false, NoGate>(Dims, double *, const double *, const double *, double, double, double,
                                                        int, NoGate);
template __global__ void k_copy_interior<double, NoGate>(Dims, double *, const double *, int, NoGate);
template __global__ void k_pass<double, false, NoGate>(Dims, double *, const double *, const double *, const double *, double, int,
                                                  NoGate);
template __global__ void k_conduct_zghost<double, NoGate>(Dims, double *, NoGate);
template __global__ void k_conduct_zspill<double, NoGate>(Dims, double *, int, NoGate);

}  // namespace magic
