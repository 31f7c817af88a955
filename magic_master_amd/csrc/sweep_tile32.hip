// FP32 perturbation-form instantiations of the tiled sweep, for the tile
// shape MAGIC_TILE_SHAPE (this file is compiled once per shape).  This form
// has its own error bar (a fraction of the field peak against the FP64
// form), so products and sums may contract into fused multiply-adds.
#pragma clang fp contract(fast)
#include "sweep_tile_body.h"

#ifndef MAGIC_TILE_SHAPE
#error "compile once per tile shape with -DMAGIC_TILE_SHAPE=0..TILE_NSHAPE-1"
#endif

namespace magic {

#define MAGIC_TILE32(NTR)                                                                                  \
    template __global__ void tile_sweep_list<float, true, NTR, MAGIC_TILE_SHAPE>(                        \
        const SweepArgsT<float> *, const int4 *, int, const int *const *, int *, int4);
MAGIC_TILE32(0)
MAGIC_TILE32(1)
MAGIC_TILE32(2)
MAGIC_TILE32(3)
MAGIC_TILE32(4)
MAGIC_TILE32(5)
#undef MAGIC_TILE32

}  // namespace magic
