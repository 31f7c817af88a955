// Compensated FP64 global sums of a grid's state (sums.hip).
#pragma once
#include "ctx.h"

namespace magic {
// A running sum s and the accumulated rounding error c of its additions:
// the value is s + c.
struct CSum {
    double s, c;
};
}  // namespace magic

// Enqueues the layer sums and their total for context c on its stream:
// partial holds mz * meqn entries, out_dev receives meqn doubles (the sums of
// every component over the interior cells, times the cell volume).
int sums_launch(magic_ctx *c, magic::CSum *partial, double *out_dev);
// Copies the meqn results at out_dev to the host array sums (through the
// context's pinned words) once the context's stream is done.
int sums_fetch(magic_ctx *c, const double *out_dev, double *sums);
