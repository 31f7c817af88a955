// magic_global_sums without any allocation: the same kernels and the same
// order of additions as magic_global_sums, with the partial sums in a
// file-scope device array instead of context memory allocated on the first
// call.  For callers that cannot allocate device memory (code recorded into
// a HIP graph, a serving loop that must not call hipMalloc, harnesses that
// forbid it).
//
// The array holds up to SCRATCH_LAYERS interior layers of up to
// SCRATCH_NC components (160 KiB plus the results); every GPU the library
// runs on holds its own copy until the process ends.  Every call on one GPU
// uses that GPU's copy: calls that overlap on one GPU (two host threads, or
// two streams) race on it.  Each call writes every entry it reads before
// reading it, so nothing needs clearing between calls.
#include <mutex>

#include "ctx.h"
#include "sums.h"

namespace {

constexpr int SCRATCH_LAYERS = 1024;
constexpr int SCRATCH_NC = 10;
constexpr size_t SCRATCH_PARTIAL = (size_t)SCRATCH_LAYERS * SCRATCH_NC;   // CSum entries

// partial sums, then SCRATCH_NC results
__device__ __align__(256) magic::CSum scratch[SCRATCH_PARTIAL + SCRATCH_NC / 2];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_dev[MAX_DEVICES] = {};   // device address of scratch on each GPU

// The device address of this GPU's copy of scratch (looked up once per
// device, with that device current: the array's name in host code is only
// a host-side stand-in).
int scratch_ptr(void **out)
{
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return MAGIC_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "magic_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return MAGIC_EHIP;
        }
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return MAGIC_OK;
}

}  // namespace

extern "C" int magic_global_sums_noalloc(magic_ctx *c, double *sums)
{
    if (!c || !sums || !c->have_aux) return MAGIC_EARG;
    if (c->mz > SCRATCH_LAYERS || c->meqn > SCRATCH_NC) {
        fprintf(stderr,
                "magic_global_sums_noalloc: %d layers x %d components exceed its static scratch of %d x %d "
                "(use magic_global_sums)\n",
                c->mz, c->meqn, SCRATCH_LAYERS, SCRATCH_NC);
        return MAGIC_EARG;
    }
    void *base = nullptr;
    TRY(scratch_ptr(&base));
    CSum *partial = static_cast<CSum *>(base);
    double *out_dev = reinterpret_cast<double *>(partial + SCRATCH_PARTIAL);
    TRY(sums_launch(c, partial, out_dev));
    return sums_fetch(c, out_dev, sums);
}
