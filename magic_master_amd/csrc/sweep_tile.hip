// FP64 instantiations of the tiled sweep (built without contraction of
// products and sums into fused multiply-adds, as the reference rounds).
#include "sweep_tile_body.h"

namespace magic {

template __global__ void tile_sweep<double, false, 0, true>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, false>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, true>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, false>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, true>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, false>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, true>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, false>(SweepArgsT<double>, int);

}  // namespace magic
