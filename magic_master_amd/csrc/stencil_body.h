// Boundary, pre-step and diffusion/conduction stencil kernels shared by the
// FP64 form (T = double, coefficients as 3-D planes) and the FP32
// perturbation form (T = float, coefficients as background columns, COLC).
// bc_src.hip and bc_src32.hip instantiate them under their own
// floating-point contraction settings.
//
// Index convention: a cell with Fortran indices (i, j, k) in
// [1-mbc, mx+mbc] x [1-mbc, my+mbc] x [1-mbc, mz+mbc] lives at
// ((k+mbc-1)*NY + (j+mbc-1))*NX + (i+mbc-1) of every plane.
#pragma once
#include "bc_src.h"

// Every kernel below takes a gate G: NoGate (always on, factors as passed)
// for one grid; the batched driver passes a gate that reads its scenario's
// step-control block on the device (skip when the scenario is idle in this
// stage, factors that depend on the scenario's dt).

namespace magic {

__device__ __forceinline__ long long cid(const Dims &d, int i, int j, int k)
{
    return ((long long)(k + d.mbc - 1) * d.NY + (j + d.mbc - 1)) * d.NX + (i + d.mbc - 1);
}

// x faces: bc3_mpi.f:97-312.  One thread per (j, k) of the padded face.
template <typename T, class G>
__device__ __forceinline__ void b_bc_x(const uint3 BI, Dims d, T *q, int nc, int lo, int hi, const G &gate)
{
    if (!gate.on()) return;
    const int j = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int k = (int)BI.y + 1 - d.mbc;
    if (j > d.my + d.mbc || k > d.mz + d.mbc) return;
    const bool inner = j >= 1 && j <= d.my && k >= 1 && k <= d.mz;
    for (int m = 0; m < nc; m++) {
        T *p = q + (long long)m * d.P;
        for (int ibc = 1; ibc <= d.mbc; ibc++) {
            if (lo == 1) p[cid(d, 1 - ibc, j, k)] = p[cid(d, 1, j, k)];
            else if (lo == 3) p[cid(d, 1 - ibc, j, k)] = (m == 1 ? T(-1.0) : T(1.0)) * p[cid(d, ibc, j, k)];
            else if (lo == 2 && inner) p[cid(d, 1 - ibc, j, k)] = p[cid(d, d.mx - d.mbc + (d.mbc + 1 - ibc), j, k)];
            if (hi == 1) p[cid(d, d.mx + ibc, j, k)] = p[cid(d, d.mx, j, k)];
            else if (hi == 3) p[cid(d, d.mx + ibc, j, k)] = (m == 1 ? T(-1.0) : T(1.0)) * p[cid(d, d.mx + 1 - ibc, j, k)];
            else if (hi == 2 && inner) p[cid(d, d.mx + ibc, j, k)] = p[cid(d, ibc, j, k)];
        }
    }
}
template <typename T, class G>
__global__ void k_bc_x(Dims d, T *q, int nc, int lo, int hi, G gate)
{
    b_bc_x(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, lo, hi, gate);
}

// y faces: bc3_mpi.f:320-536.  One thread per (i, k).
template <typename T, class G>
__device__ __forceinline__ void b_bc_y(const uint3 BI, Dims d, T *q, int nc, int lo, int hi, const G &gate)
{
    if (!gate.on()) return;
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int k = (int)BI.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || k > d.mz + d.mbc) return;
    const bool inner = k >= 1 && k <= d.mz;
    for (int m = 0; m < nc; m++) {
        T *p = q + (long long)m * d.P;
        for (int jbc = 1; jbc <= d.mbc; jbc++) {
            if (lo == 1) p[cid(d, i, 1 - jbc, k)] = p[cid(d, i, 1, k)];
            else if (lo == 3) p[cid(d, i, 1 - jbc, k)] = (m == 2 ? T(-1.0) : T(1.0)) * p[cid(d, i, jbc, k)];
            else if (lo == 2 && inner) p[cid(d, i, 1 - jbc, k)] = p[cid(d, i, d.my - d.mbc + (d.mbc + 1 - jbc), k)];
            if (hi == 1) p[cid(d, i, d.my + jbc, k)] = p[cid(d, i, d.my, k)];
            else if (hi == 3) p[cid(d, i, d.my + jbc, k)] = (m == 2 ? T(-1.0) : T(1.0)) * p[cid(d, i, d.my + 1 - jbc, k)];
            else if (hi == 2 && inner) p[cid(d, i, d.my + jbc, k)] = p[cid(d, i, jbc, k)];
        }
    }
}
template <typename T, class G>
__global__ void k_bc_y(Dims d, T *q, int nc, int lo, int hi, G gate)
{
    b_bc_y(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, nc, lo, hi, gate);
}

// b4step3.f:35-45: clamp tracers to [0, 1] on every cell.
template <typename T, class G>
__device__ __forceinline__ void b_clamp(const uint3 BI, T *q, long long P, int nc, const G &gate)
{
    if (!gate.on()) return;
    const long long c = (long long)BI.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    for (int m = 5; m < nc; m++) {
        const T v = q[m * P + c];
        q[m * P + c] = fmaxD(T(0.0), fminD(v, T(1.0)));
    }
}
template <typename T, class G>
__global__ void k_clamp(T *q, long long P, int nc, G gate)
{
    b_clamp(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, q, P, nc, gate);
}

// diffuse.f:48-54: z ghosts of q(2:4) from the first/last interior layer.
template <typename T, class G>
__device__ __forceinline__ void b_zghost_uvw(const uint3 BI, Dims d, T *q, const G &gate)
{
    if (!gate.on()) return;
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)BI.y + 1 - d.mbc;
    if (i > d.mx + d.mbc || j > d.my + d.mbc) return;
    for (int m = 1; m <= 3; m++) {
        T *p = q + (long long)m * d.P;
        const T lo = p[cid(d, i, j, 1)], hi = p[cid(d, i, j, d.mz)];
        for (int kbc = 1; kbc <= d.mbc; kbc++) {
            p[cid(d, i, j, 1 - kbc)] = lo;
            p[cid(d, i, j, d.mz + kbc)] = hi;
        }
    }
}
template <typename T, class G>
__global__ void k_zghost_uvw(Dims d, T *q, G gate)
{
    b_zghost_uvw(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, gate);
}

// Interior thread mapping: x fastest.
#define INTERIOR_IJK                                                          \
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1;                \
    const int j = (int)(BI.y * blockDim.y + threadIdx.y) + 1;                \
    const int k = (int)BI.z + 1;                                             \
    if (i > d.mx || j > d.my) return;                                         \
    const long long c = cid(d, i, j, k);

// All-cells thread mapping over the padded grid: x fastest; kk is the
// padded layer index (k + mbc - 1).
#define PADDED_IJK                                                            \
    const int ip = (int)(BI.x * blockDim.x + threadIdx.x);                   \
    const int jp = (int)(BI.y * blockDim.y + threadIdx.y);                   \
    if (ip >= d.NX || jp >= d.NY) return;                                     \
    const int kk = (int)BI.z;                                                \
    const long long c = ((long long)kk * d.NY + jp) * d.NX + ip;

// diffuse.f:70-118: cross-derivative terms of q(2:4) (after the z-ghost
// fill), written to the interior of the three planes of u; q is not
// changed (its planes keep the values the reference's copy holds, whose
// ghost layers the directional passes read).  A workgroup covers 64 x-cells
// by 4 y-rows and marches over kb layers, holding q(2:4) of three
// consecutive layers with a one-cell halo in x and y in LDS, so each value
// is read from memory once per workgroup.  Launch: block (64, 4), grid
// (ceil(mx/64), ceil(my/4), ceil(mz/kb)).
template <typename T, bool COLC, class G>
__device__ __forceinline__ void b_diff_cross(const uint3 BI, Dims d, T *u, const T *q, const T *coef, T dxy, T dxz, T dyz, int kb, const G &gate)
{
    if (!gate.on()) return;
    dxy = gate.f(dxy, 0);
    dxz = gate.f(dxz, 1);
    dyz = gate.f(dyz, 2);
    constexpr int TX = 66, TL = TX * 6;     // (64 + 2) x (4 + 2) cells per layer
    __shared__ T s[3][3][TL];               // [layer k mod 3][component][cell]
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, t = ty * 64 + tx;
    const int i0 = (int)BI.x * 64 + 1, j0 = (int)BI.y * 4 + 1;
    const int k0 = (int)BI.z * kb + 1;
    const int k1 = k0 + kb - 1 < d.mz ? k0 + kb - 1 : d.mz;
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i <= d.mx && j <= d.my;
    const T *u1 = q + d.P, *u2 = q + 2 * d.P, *u3 = q + 3 * d.P;
    // layer kk (0 <= kk <= mz + 1) of the tile and its halo; cells past
    // mx + 1 or my + 1 are not read by any interior cell
    auto load = [&](int kk) {
        const int sl = kk % 3;
        for (int e = t; e < TL; e += 256) {
            const int ii = i0 - 1 + e % TX, jj = j0 - 1 + e / TX;
            if (ii <= d.mx + 1 && jj <= d.my + 1) {
                const long long c = cid(d, ii, jj, kk);
                s[sl][0][e] = u1[c];
                s[sl][1][e] = u2[c];
                s[sl][2][e] = u3[c];
            }
        }
    };
    load(k0 - 1);
    load(k0);
    const int e = (ty + 1) * TX + tx + 1;   // this lane's cell in the tile
    for (int k = k0; k <= k1; k++) {
        load(k + 1);   // into the slot of layer k - 2, whose readers passed the barrier below
        __syncthreads();
        if (own) {
            const int lm = (k - 1) % 3, l0 = k % 3, lp = (k + 1) % 3;
            // f[c+a+b] - f[c-a+b] + f[c-a-b] - f[c+a-b] for the pairs of axes
            auto Xxy = [&](int m) {
                const T *f = s[l0][m];
                return f[e + 1 + TX] - f[e - 1 + TX] + f[e - 1 - TX] - f[e + 1 - TX];
            };
            auto Xxz = [&](int m) {
                return s[lp][m][e + 1] - s[lp][m][e - 1] + s[lm][m][e - 1] - s[lm][m][e + 1];
            };
            auto Xyz = [&](int m) {
                return s[lp][m][e + TX] - s[lp][m][e - TX] + s[lm][m][e - TX] - s[lm][m][e + TX];
            };
            const T ctxxy = Xxy(1), ctxxz = Xxz(2);
            const T ctyxy = Xxy(0), ctyyz = Xyz(2);
            const T ctzxz = Xxz(0), ctzyz = Xyz(1);
            const long long c = cid(d, i, j, k);
            const T co = coef[COLC ? (long long)(k + d.mbc - 1) : c];
            u[c] = s[l0][0][e] + co * (dxy * ctxxy + dxz * ctxxz);
            u[d.P + c] = s[l0][1][e] + co * (dxy * ctyxy + dyz * ctyyz);
            u[2 * d.P + c] = s[l0][2][e] + co * (dxz * ctzxz + dyz * ctzyz);
        }
        __syncthreads();
    }
}
template <typename T, bool COLC, class G>
__global__ __launch_bounds__(256) void k_diff_cross(Dims d, T *u, const T *q, const T *coef, T dxy, T dxz, T dyz,
                                                    int kb, G gate)
{
    b_diff_cross<T, COLC>(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, u, q, coef, dxy, dxz, dyz, kb, gate);
}

// diffuse.f:153-208 for one velocity component: the x, y and z passes of
// one cycle of the component loop, fused.  u holds the cross-term result on
// the interior; the plane q holds the values before the cross terms, whose
// ghost layers are the reference's stale ghosts of every pass; the result
// replaces q's interior.  Each pass is the reference's (k_pass below, same
// operand order and roundings):
//   x: X = U + fx*co*((U[i+1]-U)-U+U[i-1])   on the interior rows
//   y: Y = X + fy*co*((X[j+1]-X)-X+X[j-1])   (q's ghosts at j = 0, my+1)
//   z: q = Y + fz*co*((Y[k-1]-Y)-Y+Y[k+1])   (q's ghosts at k = 0, mz+1)
// A workgroup covers 64 x-cells by 4 y-rows and kb layers: per layer it
// forms X on 6 rows (its rows and one on either side) in LDS, Y on its 4
// rows, and keeps Y of three consecutive layers in registers for the z pass
// (the layer below its first is recomputed).  Launch: block (64, 4), grid
// (ceil(mx/64), ceil(my/4), ceil(mz/kb)).
template <typename T, bool COLC, class G>
__device__ __forceinline__ void b_diff_xyz(const uint3 BI, Dims d, T *q, const T *u, const T *coef, T fx, T fy, T fz, int kb, const G &gate)
{
    if (!gate.on()) return;
    fx = gate.f(fx, 0);
    fy = gate.f(fy, 1);
    fz = gate.f(fz, 2);
    __shared__ T sx[6][64];
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int i = (int)BI.x * 64 + tx + 1;
    const int j0 = (int)BI.y * 4 + 1;
    const int j = j0 + ty;
    const int k0 = (int)BI.z * kb + 1;
    const int k1 = k0 + kb - 1 < d.mz ? k0 + kb - 1 : d.mz;
    const bool own = i <= d.mx && j <= d.my;
    const long long sy = d.NX, sz = (long long)d.NX * d.NY;
    auto co_at = [&](long long c, int k) { return coef[COLC ? (long long)(k + d.mbc - 1) : c]; };
    // Y at (i, j, kk) for an interior layer kk; every lane of the block calls it
    auto yres = [&](int kk) {
        __syncthreads();   // the previous layer's readers of sx are done
        for (int r = ty; r < 6; r += 4) {
            const int jr = j0 - 1 + r;
            if (i <= d.mx && jr >= 1 && jr <= d.my) {
                const long long c = cid(d, i, jr, kk);
                const T aq = u[c];
                const T am = i > 1 ? u[c - 1] : q[c - 1];
                const T ap = i < d.mx ? u[c + 1] : q[c + 1];
                const T qt = ap - aq - aq + am;
                sx[r][tx] = aq + fx * co_at(c, kk) * qt;
            }
        }
        __syncthreads();
        T y = T(0.0);
        if (own) {
            const long long c = cid(d, i, j, kk);
            const T aq = sx[ty + 1][tx];
            const T am = j > 1 ? sx[ty][tx] : q[c - sy];
            const T ap = j < d.my ? sx[ty + 2][tx] : q[c + sy];
            const T qt = ap - aq - aq + am;
            y = aq + fy * co_at(c, kk) * qt;
        }
        return y;
    };
    const long long cb = own ? cid(d, i, j, k0) : 0;
    T ym = k0 > 1 ? yres(k0 - 1) : (own ? q[cb - sz] : T(0.0));
    T y0 = yres(k0);
    for (int k = k0; k <= k1; k++) {
        const long long c = own ? cid(d, i, j, k) : 0;
        const T yp = k < d.mz ? yres(k + 1) : (own ? q[c + sz] : T(0.0));
        if (own) {
            const T qt = ym - y0 - y0 + yp;
            q[c] = y0 + fz * co_at(c, k) * qt;
        }
        ym = y0;
        y0 = yp;
    }
}
template <typename T, bool COLC, class G>
__global__ __launch_bounds__(256) void k_diff_xyz(Dims d, T *q, const T *u, const T *coef, T fx, T fy, T fz, int kb,
                                                  G gate)
{
    b_diff_xyz<T, COLC>(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, q, u, coef, fx, fy, fz, kb, gate);
}

// interior copy of planes src -> dst (n planes with stride P)
template <typename T, class G>
__device__ __forceinline__ void b_copy_interior(const uint3 BI, Dims d, T *dst, const T *src, int nplanes, const G &gate)
{
    if (!gate.on()) return;
    INTERIOR_IJK
    for (int m = 0; m < nplanes; m++) dst[m * d.P + c] = src[m * d.P + c];
}
template <typename T, class G>
__global__ void k_copy_interior(Dims d, T *dst, const T *src, int nplanes, G gate)
{
    b_copy_interior(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, dst, src, nplanes, gate);
}

// One explicit directional diffusion pass (diffuse.f:153-208, conduct.f:262-323):
//   out = u + f*coef*(second difference of u along axis),
// computed on the interior.  The interior values of u come from src and the
// values on the ghost layers (which the reference copies once per call and
// never refreshes between passes) from ghost; src may equal ghost.  The
// second difference keeps the reference's operand order: (up-mid-mid+down)
// along x and y, (down-mid-mid+up) along z.
template <typename T, bool COLC, class G>
__device__ __forceinline__ void b_pass(const uint3 BI, Dims d, T *out, const T *src, const T *ghost, const T *coef, T f, int axis, const G &gate)
{
    if (!gate.on()) return;
    f = gate.f(f, 0);
    INTERIOR_IJK
    long long s;
    bool lo_in, hi_in;
    if (axis == 0) { s = 1; lo_in = i > 1; hi_in = i < d.mx; }
    else if (axis == 1) { s = d.NX; lo_in = j > 1; hi_in = j < d.my; }
    else { s = (long long)d.NX * d.NY; lo_in = k > 1; hi_in = k < d.mz; }
    const T aq = src[c];
    const T am = lo_in ? src[c - s] : ghost[c - s];
    const T ap = hi_in ? src[c + s] : ghost[c + s];
    const T qt = axis == 2 ? (am - aq - aq + ap) : (ap - aq - aq + am);
    out[c] = aq + f * coef[COLC ? (long long)(k + d.mbc - 1) : c] * qt;
}
template <typename T, bool COLC, class G>
__global__ void k_pass(Dims d, T *out, const T *src, const T *ghost, const T *coef, T f, int axis, G gate)
{
    b_pass<T, COLC>(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, out, src, ghost, coef, f, axis, gate);
}

// conduct.f:86-91: z ghosts of u, for j in [1-mbc, my+mbc] and i in
// [1-mbc, mz+mbc] -- the reference bounds i by mz, not mx.  When mx > mz the
// columns i > mz+mbc keep their stale z ghosts (which then enter the z
// stencil of the interior cells of those columns).  When mz > mx the
// out-of-row i addresses the array linearly: within a layer the writes land
// on later rows of the same layer with those rows' own values (what the
// in-row fill writes anyway), so this kernel fills whole rows and
// k_conduct_zspill adds the writes that cross into the next layer.
template <typename T, class G>
__device__ __forceinline__ void b_conduct_zghost(const uint3 BI, Dims d, T *u, const G &gate)
{
    if (!gate.on()) return;
    const int i = (int)(BI.x * blockDim.x + threadIdx.x) + 1 - d.mbc;
    const int j = (int)BI.y + 1 - d.mbc;
    const int ilast = (d.mz < d.mx ? d.mz : d.mx) + d.mbc;
    if (i > ilast || j > d.my + d.mbc) return;
    const T lo = u[cid(d, i, j, 1)], hi = u[cid(d, i, j, d.mz)];
    for (int kbc = 1; kbc <= d.mbc; kbc++) {
        u[cid(d, i, j, 1 - kbc)] = lo;
        u[cid(d, i, j, d.mz + kbc)] = hi;
    }
}
template <typename T, class G>
__global__ void k_conduct_zghost(Dims d, T *u, G gate)
{
    b_conduct_zghost(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, u, gate);
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
template <typename T, class G>
__device__ __forceinline__ void b_conduct_zspill(const uint3 BI, Dims d, T *u, int S, const G &gate)
{
    if (!gate.on()) return;
    const int r = (int)(BI.x * blockDim.x + threadIdx.x);
    if (r >= S) return;
    const long long NXY = (long long)d.NX * d.NY;
    auto L = [&](int k) { return (long long)(k + d.mbc - 1) * NXY + r; };
    const T lo = u[L(2)], hi = u[L(d.mz + 1)];
    for (int kbc = 1; kbc <= d.mbc; kbc++) {
        u[L(2 - kbc)] = lo;
        if (d.mz + 1 + kbc <= d.mz + d.mbc) u[L(d.mz + 1 + kbc)] = hi;
    }
}
template <typename T, class G>
__global__ void k_conduct_zspill(Dims d, T *u, int S, G gate)
{
    b_conduct_zspill(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, d, u, S, gate);
}

// plain copy of n elements (copyq3 and the working copies of the source pass)
template <typename T, class G>
__device__ __forceinline__ void b_copy(const uint3 BI, unsigned nblk, T *dst, const T *src, long long n, const G &gate)
{
    if (!gate.on()) return;
    for (long long e = (long long)BI.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)nblk * blockDim.x)
        dst[e] = src[e];
}
template <typename T, class G>
__global__ void k_copy(T *dst, const T *src, long long n, G gate)
{
    b_copy(uint3{blockIdx.x, blockIdx.y, blockIdx.z}, gridDim.x, dst, src, n, gate);
}

}  // namespace magic
