// FP32 perturbation-form instantiations of the tiled sweep.  This form has
// its own error bar (a fraction of the field peak against the FP64 form), so
// products and sums may contract into fused multiply-adds.
#pragma clang fp contract(fast)
#include "sweep_tile_body.h"

namespace magic {

template __global__ void tile_sweep<float, true, 0, true>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 0, false>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 1, true>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 1, false>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 2, true>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 2, false>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 5, true>(SweepArgsT<float>, int);
template __global__ void tile_sweep<float, true, 5, false>(SweepArgsT<float>, int);
template __global__ void tile_sweep_list<float, true, 0, true>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 0, false>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 1, true>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 1, false>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 2, true>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 2, false>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 5, true>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);
template __global__ void tile_sweep_list<float, true, 5, false>(const SweepArgsT<float> *, const int4 *, int,
                                                                   const int *const *, int *);

}  // namespace magic
