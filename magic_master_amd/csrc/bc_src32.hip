// Boundary and source-pass kernels of the FP32 perturbation form.
//
// The state is q' = q - q_bar on component planes (float), with a
// background column at rest: q_bar = (rho_bar(k), 0, 0, 0, E_bar(k)) and
// p_bar = (gamma-1) E_bar.  Tracers are stored as they are (their background
// is 0).  Every formula below is the FP64 form's operator (bc_src.hip, the
// reference's bc3_mpi.f / src3.f90) rewritten so that no FP32 operation
// subtracts two background-sized numbers.  Products and sums may contract
// into fused multiply-adds: this form's error bar is a fraction of the field
// peak against the FP64 form.
#pragma clang fp contract(fast)
#include "batch.h"
#include "magic_dev.h"
#include "stencil_body.h"
#include "batch_ops.h"

namespace magic {

// FP32 instantiations of the shared kernels (coefficients as columns)
template __global__ void k_bc_x<float, NoGate>(Dims, float *, int, int, int, NoGate);
template __global__ void k_bc_y<float, NoGate>(Dims, float *, int, int, int, NoGate);
template __global__ void k_clamp<float, NoGate>(float *, long long, int, NoGate);
template __global__ void k_zghost_uvw<float, NoGate>(Dims, float *, NoGate);
template __global__ void k_diff_xyz<float, true, NoGate>(Dims, float *, const float *, const float *, float, float, float,
                                                        int, NoGate);
template __global__ void k_diff_cross<float, true, NoGate>(Dims, float *, const float *, const float *, float, float, float,
                                                        int, NoGate);
template __global__ void k_copy_interior<float, NoGate>(Dims, float *, const float *, int, NoGate);
template __global__ void k_pass<float, true, NoGate>(Dims, float *, const float *, const float *, const float *, float, int,
                                                  NoGate);
template __global__ void k_conduct_zghost<float, NoGate>(Dims, float *, NoGate);
template __global__ void k_conduct_zspill<float, NoGate>(Dims, float *, int, NoGate);

// z faces (bc3_mpi.f:544-809).  The FP64 form writes
//   q_ghost = aux_ghost + (q_boundary - aux_boundary) * sc
// for rho, the horizontal momenta and E, and q_ghost = q_boundary * sc for
// the vertical momentum; with a background at rest every one of them is
// q'_ghost = q'_boundary * sc.  sc depends on the layer only.
// zf(side, kbc - 1): the factor sc of ghost layer kbc on side 0 (bottom) or 1
// (top), Zr32's s[side][kbc - 1] (FP64)
template <class G, class ZF>
__device__ __forceinline__ void b32_bc_z(const uint3 BI, Dims d, float *q, int nc, int lo, int hi, ZF zf,
                                         const G &gate)
{
    if (!gate.on()) return;
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)BI.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || j > d.my + d.mbc) return;
    for (int side = 0; side < 2; side++) {
        const int mth = side == 0 ? lo : hi;
        if (mth == 1 || mth == 3) {
            const int kb = side == 0 ? 1 : d.mz;
            const long long cb = cid(d, i, j, kb);
            for (int kbc = 1; kbc <= d.mbc; kbc++) {
                const int kg = side == 0 ? 1 - kbc : d.mz + kbc;
                const long long cg = cid(d, i, j, kg);
                const double sc = zf(side, kbc - 1);   // FP64 factor, one rounding per ghost value
                for (int m = 0; m < 5; m++) q[m * d.P + cg] = (float)((double)q[m * d.P + cb] * sc);
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
                    float *p = q + (long long)m * d.P;
                    if (side == 0) p[cid(d, i, j, kbc - d.mbc)] = p[cid(d, i, j, d.mz - d.mbc + kbc)];
                    else p[cid(d, i, j, d.mz + kbc)] = p[cid(d, i, j, kbc)];
                }
        }
    }
}
template <class G>
__global__ void k32_bc_z(Dims d, float *q, int nc, int lo, int hi, Zr32 zr, G gate)
{
    b32_bc_z(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, lo, hi,
             [&](int side, int k) { return zr.s[side][k]; }, gate);
}

// conserved perturbation -> (rho', u, v, w, T') on every cell
// (src3.f90:242-304 with u_bar = 0):
//   T' = T - T_bar = (p' - rho' (gamma-1) E_bar / rho_bar) / (rho R),
//   p' = (gamma-1) (E' - |m|^2 / (2 rho)),  rho = rho_bar + rho'.
//
// The arithmetic of the conversion and of the reassembly below is FP64 (the
// stored values are FP32): T' is a small difference of the pressure and
// density perturbations' background-weighted parts, and E' a sum of such
// parts, whose FP32 rounding the flow carries into the thin upper layers.
// dq: as the FP64 form's conversion (bc_src.hip), the second sweep
// stream's increments added first, each rounded to FP32 as q += dq rounds
template <class G>
__device__ __forceinline__ void b32_src_convert(const uint3 BI, Dims d, float *q, Col32 col, const G &gate,
                                                const float *dq = nullptr, int nc = 0)
{
    if (!gate.on()) return;
    PADDED_IJK
    const long long P = d.P;
    if (dq) {
        q[c] = q[c] + dq[c];
        for (int m = 1; m < nc; m++) q[m * P + c] = q[m * P + c] + dq[m * P + c];
    }
    const double rb = col.rho[kk], Eb = col.E[kk], g1 = col.gam[kk] - 1.0, R = col.R[kk];
    const double rp = q[c], rho = rb + rp;
    const double m2 = q[P + c], m3 = q[2 * P + c], m4 = q[3 * P + c], ep = q[4 * P + c];
    const double irho = 1.0 / rho;
    const double pp = g1 * (ep - 0.5 * (m2 * m2 + m3 * m3 + m4 * m4) * irho);
    q[4 * P + c] = (float)((pp - rp * (g1 * Eb / rb)) * irho / R);
    q[P + c] = (float)(m2 * irho);
    q[2 * P + c] = (float)(m3 * irho);
    q[3 * P + c] = (float)(m4 * irho);
}
template <class G>
__global__ void k32_src_convert(Dims d, float *q, Col32 col, G gate, const float *dq, int nc)
{
    b32_src_convert(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, col, gate, dq, nc);
}

// Reassembly, plume forcing and tracer clamp on the interior
// (src3.f90:380-418, 430-445, 514-607, 616-626):
//   m = u rho,  E' = E_bar rho'/rho_bar + T' rho R/(gamma-1) + |m|^2/(2 rho),
// the heating adding dt*heat to T' first.  Positions and the forcing are
// evaluated in FP64.
template <class G>
__device__ __forceinline__ void b32_src_finish(const uint3 BI, Dims d, float *q, int nc, Col32 col, Forcing f, const G &gate)
{
    if (!gate.on()) return;
    gate.forcing(f.t, f.dt, f.tgauss);
    INTERIOR_IJK
    const long long P = d.P;
    const int kk = k + d.mbc - 1;
    const double rb = col.rho[kk], Eb = col.E[kk], g1 = col.gam[kk] - 1.0, R = col.R[kk];
    const double rp = q[c], rho = rb + rp;
    const double m2 = q[P + c] * rho, m3 = q[2 * P + c] * rho;
    double m4 = q[3 * P + c] * rho;
    double tp = q[4 * P + c];
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
        m4 = m4 + (f.dt * f.amplitude * gau * ffz) * rho;
    } else if (f.forcemth == 3 && !(zcell > 15000.0)) {
        const double g2 = exp(-.5 * ((xcell - f.xpos) * (xcell - f.xpos) / (f.xwidth * f.xwidth))) *
                          exp(-.5 * ((ycell - f.ypos) * (ycell - f.ypos) / (f.ywidth * f.ywidth))) *
                          f.tgauss;
        const double precipw = f.amplitude * g2;
        if (precipw > 0.1) tp = tp + f.dt * sa_heating(precipw, zcell);
    }
    q[P + c] = (float)m2;
    q[2 * P + c] = (float)m3;
    q[3 * P + c] = (float)m4;
    q[4 * P + c] = (float)(Eb * (rp / rb) + tp * rho * R / g1 + 0.5 * (m2 * m2 + m3 * m3 + m4 * m4) / rho);
    for (int m = 5; m < nc; m++) q[m * P + c] = fmaxD(0.0f, fminD(q[m * P + c], 1.0f));
}
template <class G>
__global__ void k32_src_finish(Dims d, float *q, int nc, Col32 col, Forcing f, G gate)
{
    b32_src_finish(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, col, f, gate);
}


// The FP32 form's own bodies for the batched work-list kernels (batch_ops.h)
struct Form32 {
    // po: the operand block in memory (the factors are indexed by layer)
    __device__ static void bc_z(uint3 bi, const BatchOps<float> &o, cptr<BatchOps<float>> po, const StepGate &g)
    {
        b32_bc_z(bi, o.d, o.q, o.nc, lohi(o.bc[4]), lohi(o.bc[5]),
                 [=](int side, int k) { return po->zr.s[side][k]; }, g);
    }
    __device__ static void convert(uint3 bi, const BatchOps<float> &o, const StepGate &g)
    {
        b32_src_convert(bi, o.d, o.q, o.col, g);
    }
    __device__ static void finish(uint3 bi, const BatchOps<float> &o, const StepGate &g)
    {
        b32_src_finish(bi, o.d, o.q, o.nc, o.col, o.f, g);
    }
    __device__ static const float *nu(const BatchOps<float> &o) { return o.nu32; }
    __device__ static const float *kap(const BatchOps<float> &o) { return o.kap32; }
};

template <>
const void *batch_op_fn<float>(int op)
{
    switch (op) {
#define MAGIC_OP(X) case X: return reinterpret_cast<const void *>(&k_batch_op<float, X, Form32>);
    MAGIC_OP(OP_BCX) MAGIC_OP(OP_BCY) MAGIC_OP(OP_BCZ) MAGIC_OP(OP_CLAMP) MAGIC_OP(OP_BACKUP) MAGIC_OP(OP_RESTORE)
    MAGIC_OP(OP_CONVERT) MAGIC_OP(OP_ZGHOST_UVW) MAGIC_OP(OP_DIFF_CROSS) MAGIC_OP(OP_DIFF_XYZ) MAGIC_OP(OP_COPY_Q5)
    MAGIC_OP(OP_PASS) MAGIC_OP(OP_CONDUCT_ZGHOST) MAGIC_OP(OP_ZSPILL) MAGIC_OP(OP_COPY_INTERIOR) MAGIC_OP(OP_FINISH)
    MAGIC_OP(OP_MERGE)
#undef MAGIC_OP
    default: return nullptr;
    }
}

// instantiations of the single-grid kernels (the batched driver's are
// k_batch_op, above)
template __global__ void k32_bc_z<NoGate>(Dims, float *, int, int, int, Zr32, NoGate);
template __global__ void k32_src_convert<NoGate>(Dims, float *, Col32, NoGate, const float *, int);
template __global__ void k32_src_finish<NoGate>(Dims, float *, int, Col32, Forcing, NoGate);

// reference AoS q(meqn, i, j, k) (FP64, full state) <-> perturbation planes:
// the background is subtracted / added in FP64, then rounded.
__global__ void k32_aos_to_pert(Dims d, float *q, const double *aos, int nc, const double *bgrho,
                                const double *bgE)
{
    const uint3 BI = {blockIdx.x, blockIdx.y, blockIdx.z};
    PADDED_IJK
    for (int m = 0; m < nc; m++) {
        const double bg = m == 0 ? bgrho[kk] : (m == 4 ? bgE[kk] : 0.0);
        q[(long long)m * d.P + c] = (float)(aos[c * nc + m] - bg);
    }
}

__global__ void k32_pert_to_aos(Dims d, double *aos, const float *q, int nc, const double *bgrho,
                                const double *bgE)
{
    const uint3 BI = {blockIdx.x, blockIdx.y, blockIdx.z};
    PADDED_IJK
    for (int m = 0; m < nc; m++) {
        const double bg = m == 0 ? bgrho[kk] : (m == 4 ? bgE[kk] : 0.0);
        aos[c * nc + m] = bg + (double)q[(long long)m * d.P + c];
    }
}

}  // namespace magic
