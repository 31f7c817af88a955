// FP64 instantiations of the tiled sweep (built without contraction of
// products and sums into fused multiply-adds, as the reference rounds).
#include "sweep_tile_body.h"

namespace magic {

template __global__ void tile_sweep<double, false, 0, true, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, true, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, true, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, false, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, false, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 0, false, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, true, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, true, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, true, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, false, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, false, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 1, false, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, true, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, true, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, true, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, false, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, false, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 2, false, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, true, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, true, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, true, 2>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, false, 0>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, false, 1>(SweepArgsT<double>, int);
template __global__ void tile_sweep<double, false, 5, false, 2>(SweepArgsT<double>, int);

}  // namespace magic
