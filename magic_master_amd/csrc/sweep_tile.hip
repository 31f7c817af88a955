// FP64 instantiations of the tiled sweep (one grid and the batched work-list
// form), for the tile shape MAGIC_TILE_SHAPE: this file is compiled once per
// shape.  A product and a sum of one expression may contract into a fused
// multiply-add, except in the pressures and the gravity jump of the Riemann
// problem (pressure_nc, flux_jump_nc, gravity_jump_nc), which are rounded as
// the reference rounds them.  Contraction stays within an expression
// (contract(on), not fast): fast lets the backend fuse across statements
// depending on how many uses a product has after inlining, and the grid and
// work-list kernels then fused different pairs (results one unit in the last
// place apart), whereas a batched scenario has to equal the same grid
// stepped alone bit for bit.
#pragma clang fp contract(on)
#include "sweep_tile_body.h"

#ifndef MAGIC_TILE_SHAPE
#error "compile once per tile shape with -DMAGIC_TILE_SHAPE=0..TILE_NSHAPE-1"
#endif

namespace magic {

#define MAGIC_TILE64(NTR)                                                                                  \
    template __global__ void tile_sweep<double, false, NTR, MAGIC_TILE_SHAPE>(SweepArgsT<double>, int);   \
    template __global__ void tile_sweep_list<double, false, NTR, MAGIC_TILE_SHAPE>(                     \
        const SweepArgsT<double> *, const int4 *, int, const int *const *, int *);
MAGIC_TILE64(0)
MAGIC_TILE64(1)
MAGIC_TILE64(2)
MAGIC_TILE64(5)
#undef MAGIC_TILE64

}  // namespace magic
