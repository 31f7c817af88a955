// FP64 instantiations of the tiled sweep, for the tile shape
// MAGIC_TILE_SHAPE: this file is compiled once per shape.  Products and sums
// may contract into fused multiply-adds here, except in the pressures and
// the gravity jump of the Riemann problem (pressure_nc, flux_jump_nc,
// gravity_jump_nc), which are rounded as the reference rounds them.
#pragma clang fp contract(fast)
#include "sweep_tile_body.h"

#ifndef MAGIC_TILE_SHAPE
#error "compile once per tile shape with -DMAGIC_TILE_SHAPE=0..TILE_NSHAPE-1"
#endif

namespace magic {

template __global__ void tile_sweep<double, false, 0, MAGIC_TILE_SHAPE>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, MAGIC_TILE_SHAPE>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, MAGIC_TILE_SHAPE>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, MAGIC_TILE_SHAPE>(SweepArgsT<double>, int);

}  // namespace magic
