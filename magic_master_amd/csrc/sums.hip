// Global sums of the state over the interior cells, in FP64 with
// compensated (Neumaier) summation: the mass sum(rho) dV, the momenta, the
// total energy sum(E) dV and the tracer masses, every component of q.
//
// The reference has no such diagnostic; its conservation check is implicit
// (periodic x and y with the fluid at rest change sum(rho) only through the
// z faces).  In the FP32 perturbation form the background part of every
// layer, n_xy * q_bar(k), enters as an exact FP64 product (its rounding error
// is carried as the compensation term) and the FP32 perturbations are
// summed against it in FP64, so the sum of the full state is as accurate as
// in the FP64 form.
//
// The order of every addition is fixed (one workgroup per interior layer,
// lanes strided over the layer's cells, then a fixed tree; the layers'
// partial sums are added in layer order), so the result is the same on
// every run and for every stream or graph it is launched from.
#include "ctx.h"
#include "sums.h"

namespace magic {

// a += x with the rounding error of the addition kept in a.c
// (Neumaier's variant of Kahan summation)
__device__ __forceinline__ void csum_add(CSum &a, double x)
{
    const double t = a.s + x;
    a.c += fabs(a.s) >= fabs(x) ? (a.s - t) + x : (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void csum_merge(CSum &a, const CSum &b)
{
    csum_add(a, b.s);
    a.c += b.c;
}

// One workgroup per interior layer k = blockIdx.x + 1; partial[(k-1)*nc + m]
// receives that layer's compensated sum of component m.  PERT: q is the
// FP32 perturbation and bg = {rho_bar, E_bar} (FP64 columns, index k+mbc-1).
template <typename T, bool PERT>
__global__ __launch_bounds__(256) void k_layer_sums(Dims d, const T *__restrict__ q, int nc,
                                                    const double *__restrict__ bg_rho,
                                                    const double *__restrict__ bg_E, CSum *partial)
{
    __shared__ CSum sh[256];
    const int k = (int)blockIdx.x + 1;
    const int t = (int)threadIdx.x;
    const int nxy = d.mx * d.my;
    const long long layer = (long long)(k + d.mbc - 1) * d.NY * d.NX;
    for (int m = 0; m < nc; m++) {
        CSum a = {0.0, 0.0};
        if (PERT && t == 0 && (m == 0 || m == 4)) {
            // the background of the layer: n_xy * q_bar(k) as an exact sum of
            // the product and its rounding error
            const double b = m == 0 ? bg_rho[k + d.mbc - 1] : bg_E[k + d.mbc - 1];
            const double p = (double)nxy * b;
            a.s = p;
            a.c = fma((double)nxy, b, -p);
        }
        const T *plane = q + (long long)m * d.P + layer;
        for (int e = t; e < nxy; e += 256) {
            const int j = e / d.mx, i = e - j * d.mx;
            csum_add(a, (double)plane[(long long)(j + d.mbc) * d.NX + (i + d.mbc)]);
        }
        sh[t] = a;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (t < off) csum_merge(sh[t], sh[t + off]);
            __syncthreads();
        }
        if (t == 0) partial[(long long)(k - 1) * nc + m] = sh[0];
        __syncthreads();
    }
}

// The layers' partial sums in layer order (one workgroup of 256 lanes:
// lane t adds layers t, t+256, ... in turn, then a fixed tree), times the
// cell volume: out[m] = (s + c) * vol.
__global__ __launch_bounds__(256) void k_sum_layers(const CSum *__restrict__ partial, int nlayer, int nc, double vol,
                                                    double *out)
{
    __shared__ CSum sh[256];
    const int t = (int)threadIdx.x;
    for (int m = 0; m < nc; m++) {
        CSum a = {0.0, 0.0};
        for (int k = t; k < nlayer; k += 256) csum_merge(a, partial[(long long)k * nc + m]);
        sh[t] = a;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (t < off) csum_merge(sh[t], sh[t + off]);
            __syncthreads();
        }
        if (t == 0) out[m] = (sh[0].s + sh[0].c) * vol;
        __syncthreads();
    }
}

}  // namespace magic

// Enqueues the two reductions of context c on its stream: partial holds
// mz * meqn CSum entries, out_dev meqn doubles (device memory).
int sums_launch(magic_ctx *c, CSum *partial, double *out_dev)
{
    const Dims &d = c->d;
    const double vol = c->dx * c->dy * c->dz;
    const int NZ = d.NZ;
    if (c->form == MAGIC_FORM_FP32P)
        k_layer_sums<float, true><<<d.mz, 256, 0, c->stream>>>(d, c->q32, c->meqn, c->bg64, c->bg64 + NZ, partial);
    else
        k_layer_sums<double, false><<<d.mz, 256, 0, c->stream>>>(d, c->q, c->meqn, nullptr, nullptr, partial);
    TRY(launch_check());
    k_sum_layers<<<1, 256, 0, c->stream>>>(partial, d.mz, c->meqn, vol, out_dev);
    return launch_check();
}

// The meqn results from device memory to the caller through the context's
// pinned words (no staging by the runtime), after the stream is done.
int sums_fetch(magic_ctx *c, const double *out_dev, double *sums)
{
    double *pin = reinterpret_cast<double *>(c->hscal + HSCAL_SUMS);
    static_assert(HSCAL_SUMS + 10 <= HSCAL_WORDS, "room for meqn <= 10 sums");
    HIPCHECK(hipMemcpyAsync(pin, out_dev, (size_t)c->meqn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    memcpy(sums, pin, (size_t)c->meqn * sizeof(double));
    return MAGIC_OK;
}

extern "C" int magic_global_sums(magic_ctx *c, double *sums)
{
    if (!c || !sums || !c->have_aux) return MAGIC_EARG;
    if (!c->sums_d) {
        // mz * meqn partial sums, then meqn results
        const size_t bytes = (size_t)c->mz * c->meqn * sizeof(CSum) + (size_t)c->meqn * sizeof(double);
        if (hipMalloc(&c->sums_d, bytes) != hipSuccess) {
            c->sums_d = nullptr;
            return MAGIC_ENOMEM;
        }
    }
    CSum *partial = (CSum *)c->sums_d;
    double *out_dev = (double *)(partial + (size_t)c->mz * c->meqn);
    TRY(sums_launch(c, partial, out_dev));
    return sums_fetch(c, out_dev, sums);
}
