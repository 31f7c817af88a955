// FP64 instantiations of the tiled sweep, for the tile shape
// MAGIC_TILE_SHAPE: this file is compiled once per shape.  Products and sums
// may contract into fused multiply-adds here, except in the pressures and
// the gravity jump of the Riemann problem (pressure_nc, flux_jump_nc,
// gravity_jump_nc), which are rounded as the reference rounds them.  Which
// pairs fuse is the compiler's choice for each compiled kernel; one grid and
// a batch of grids run the same kernel, so they round alike.
#pragma clang fp contract(fast)
#include "sweep_tile_body.h"

#ifndef MAGIC_TILE_SHAPE
#error "compile once per tile shape with -DMAGIC_TILE_SHAPE=0..TILE_NSHAPE-1"
#endif

namespace magic {

#define MAGIC_TILE64(NTR)                                                                                  \
    template __global__ void tile_sweep_list<double, false, NTR, MAGIC_TILE_SHAPE>(                     \
        const SweepArgsT<double> *, const int4 *, int, const int *const *, int *, int4);
MAGIC_TILE64(0)
MAGIC_TILE64(1)
MAGIC_TILE64(2)
MAGIC_TILE64(3)
MAGIC_TILE64(4)
MAGIC_TILE64(5)
#undef MAGIC_TILE64

}  // namespace magic
