// Shared declarations of the MAGIC device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace magic {

// A compile-time int passed as a value (to generic lambdas).
template <int V>
struct IConst { static constexpr int value = V; };

// One directional sweep of step3 (lib/step3.f:121-621): pencils along the
// normal direction n, with the "g" transverse direction (gadd, flux3fw.f:419)
// and the "h" transverse direction (hadd, flux3fw.f:545).
//   x-sweep: n = x, g = y, h = z      y-sweep: n = y, g = z, h = x
//   z-sweep: n = z, g = x, h = y      (step3.f:161, 329, 499 argument order)
template <typename T>
struct SweepArgsT {
    const T *__restrict__ qold;        // component planes, stride P
    T *__restrict__ qnew;
    const double *__restrict__ gam;    // aux(19) plane
    const double *__restrict__ vgrav;  // aux(17) plane
    long long P;                       // cells per plane
    int comp[10];                      // local component -> plane (rho, m_n, m_g', m_h', E, tracers)
    int n;                             // interior cells along the pencil
    int nA, nB;                        // interior extents along g and h
    int mbc;
    long long sN, sA, sB;              // strides (cells) along n, g, h
    int cA, cB;                        // pencil colour (a mod 3, b mod 3)
    int gravity;                       // 1 on the z-sweep (rpn3euX.f:202-208)
    int nspecadv;
    int sweep;                         // 1, 2, 3 (for error reports)
    double dtdn, fg, fh;               // dt/dn, dt/dg, dt/dh
    unsigned long long *cfl_bits;      // running max of the Courant number
    int *status;                       // [0] code, [1..4] sweep,i,j,k of first failure
    int i0, j0, k0;                    // Fortran index of (cell 0 along n, a = 0, b = 0) mapping helpers
    // perturbation form only: background columns indexed by k + mbc - 1
    // (rho_bar, E_bar, gamma, aux17, and the background's own z jump
    // p_bar(k) - p_bar(k-1) - .5 (rho_bar(k) + rho_bar(k-1)) aux17(k))
    const double *crho, *cE, *cgam, *cvg, *cres;
    T *sink;                           // >= 256 elements of scratch the tile flush may write
};
using SweepArgs = SweepArgsT<double>;

template <int NTR>
__global__ void sweep_kernel(SweepArgs a);

// Tiled sweep, one instantiation per tile shape SH (0 .. TILE_NSHAPE-1):
// TA = tile_ta(SH) pencils along the tile axis a per workgroup, 256/TA
// interfaces per chunk along the pencil.  a_is_g says whether the tile axis
// is the sweep's g direction; sA/sB/nA/nB then describe the tile axis a and
// the coloured axis b.  (T, PERT) = (double, false): FP64 form; (float,
// true): FP32 perturbation form.  A persistent kernel over a work list of
// tiles of one or several grids (items: {argument block, a_is_g, bx, by};
// run[argument block] points at the grid's "runs in this attempt" flag;
// next: the work counter and the count of finished workgroups, both 0 at
// launch and set back to 0 by the kernel; grid: {tiles along a, argument
// block, a_is_g, 0} for one tile per workgroup, x = 0 for the work list).
constexpr int TILE_NSHAPE = 4;
__host__ __device__ constexpr int tile_ta(int sh) { return sh == 0 ? 16 : sh == 3 ? 4 : 8; }
template <typename T, bool PERT, int NTR, int SH>
__global__ void tile_sweep_list(const SweepArgsT<T> *args, const int4 *items, int nitems, const int *const *run,
                                int *next, int4 grid);

}  // namespace magic
