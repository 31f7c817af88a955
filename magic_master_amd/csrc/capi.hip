// C ABI of the MAGIC step on MI355X (declared in include/magic_hip.h).
//
// A context owns the device-resident state of one grid: q and its backup
// (the reference's copyq3 work array), the background planes the step reads,
// the bc3 z-boundary density ratios, and the source-pass scratch planes.
// Everything runs on one HIP stream of default priority; every entry point
// returns after its work is complete.
#include <mutex>
#include <thread>

#include "ctx.h"

// ------------------------------------------------------ host transfers ---
// q crosses between the caller's host memory (pageable: a Fortran array)
// and the device through two page-locked bounce chunks shared by every
// context (one transfer at a time): a chunk's DMA runs while host threads
// copy the neighbouring chunk between the caller's memory and the other
// bounce chunk.
namespace {
constexpr size_t BOUNCE = size_t(32) << 20;   // bytes per bounce chunk
constexpr int COPY_THREADS = 8;
std::mutex bounce_mu;
char *bounce_buf[2] = {nullptr, nullptr};
hipEvent_t bounce_ev[2] = {nullptr, nullptr};

int bounce_setup()
{
    if (bounce_buf[0]) return MAGIC_OK;
    for (int k = 0; k < 2; k++) {
        if (hipHostMalloc((void **)&bounce_buf[k], BOUNCE, hipHostMallocDefault) != hipSuccess) return MAGIC_ENOMEM;
        HIPCHECK(hipEventCreateWithFlags(&bounce_ev[k], hipEventDisableTiming));
    }
    return MAGIC_OK;
}

// memcpy of n bytes split over COPY_THREADS host threads
void par_memcpy(char *dst, const char *src, size_t n)
{
    if (n < (size_t(1) << 20)) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = (n + COPY_THREADS - 1) / COPY_THREADS;
    std::thread th[COPY_THREADS];
    for (int t = 0; t < COPY_THREADS; t++) {
        const size_t o = std::min(n, t * part), e = std::min(n, o + part);
        th[t] = std::thread([=] { memcpy(dst + o, src + o, e - o); });
    }
    for (auto &t : th) t.join();
}
}  // namespace

// host -> device (dev on the context's stream; returns with the copies queued)
static int upload(magic_ctx *c, void *dev, const void *host, size_t bytes)
{
    std::lock_guard<std::mutex> lock(bounce_mu);
    TRY(bounce_setup());
    for (size_t off = 0, i = 0; off < bytes; off += BOUNCE, i++) {
        const size_t n = std::min(BOUNCE, bytes - off);
        const int k = (int)(i & 1);
        HIPCHECK(hipEventSynchronize(bounce_ev[k]));   // the copy that last read this chunk is done
        par_memcpy(bounce_buf[k], (const char *)host + off, n);
        HIPCHECK(hipMemcpyAsync((char *)dev + off, bounce_buf[k], n, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipEventRecord(bounce_ev[k], c->stream));
    }
    // the bounce chunks are read until these copies run: wait before unlocking
    for (auto e : bounce_ev) HIPCHECK(hipEventSynchronize(e));
    return MAGIC_OK;
}

// device -> host (after the work queued on the context's stream)
static int download(magic_ctx *c, void *host, const void *dev, size_t bytes)
{
    std::lock_guard<std::mutex> lock(bounce_mu);
    TRY(bounce_setup());
    size_t prev_off = 0, prev_n = 0;
    for (size_t off = 0, i = 0; off < bytes; off += BOUNCE, i++) {
        const size_t n = std::min(BOUNCE, bytes - off);
        const int k = (int)(i & 1);
        HIPCHECK(hipMemcpyAsync(bounce_buf[k], (const char *)dev + off, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipEventRecord(bounce_ev[k], c->stream));
        if (i > 0) {   // the previous chunk, while this one's copy runs
            HIPCHECK(hipEventSynchronize(bounce_ev[k ^ 1]));
            par_memcpy((char *)host + prev_off, bounce_buf[k ^ 1], prev_n);
        }
        prev_off = off;
        prev_n = n;
    }
    if (prev_n) {
        const int k = (int)(((bytes + BOUNCE - 1) / BOUNCE - 1) & 1);
        HIPCHECK(hipEventSynchronize(bounce_ev[k]));
        par_memcpy((char *)host + prev_off, bounce_buf[k], prev_n);
    }
    return MAGIC_OK;
}

extern "C" const char *magic_strerror(int s)
{
    switch (s) {
    case MAGIC_OK: return "ok";
    case MAGIC_ERHO: return "rho <= 0 in a Riemann problem";
    case MAGIC_EA2: return "a**2 <= 0 in a Roe average";
    case MAGIC_EBC: return "boundary condition needs MPI or a user routine (mthbc 0 or 4)";
    case MAGIC_EUNSUPPORTED: return "method or option not supported";
    case MAGIC_EARG: return "invalid argument";
    case MAGIC_EHIP: return "HIP runtime error";
    case MAGIC_ENOMEM: return "device allocation failed";
    default: return "unknown status";
    }
}


extern "C" int magic_create(int mx, int my, int mz, int meqn, int mbc, int maux, double xlower,
                            double ylower, double zlower, double dx, double dy, double dz,
                            magic_ctx **out)
{
    return magic_create_form(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz,
                             MAGIC_FORM_FP64, out);
}

extern "C" int magic_create_form(int mx, int my, int mz, int meqn, int mbc, int maux, double xlower,
                                 double ylower, double zlower, double dx, double dy, double dz, int form,
                                 magic_ctx **out)
{
    if (!out) return MAGIC_EARG;
    *out = nullptr;
    if (form != MAGIC_FORM_FP64 && form != MAGIC_FORM_FP32P) return MAGIC_EARG;
    if (form == MAGIC_FORM_FP32P && mbc > 8) return MAGIC_EUNSUPPORTED;
    if (mx < 1 || my < 1 || mz < 1 || mbc < 2 || meqn < 5 || meqn > 10 || maux < 19) return MAGIC_EARG;
    if (!sweep_supported(meqn - 5)) return MAGIC_EUNSUPPORTED;
    if (mz > mx) {
        // conduct.f:86-91 writes S = mz-mx elements past each layer; the
        // device replays those writes after the in-row fills, which is the
        // reference's order only while the rows they land in come before
        // the rows that write them.
        const long long S = mz - mx, NX = mx + 2 * mbc, NY = my + 2 * mbc;
        if (2 * ((S - 1) / NX) + 1 >= NY) return MAGIC_EUNSUPPORTED;
    }
    magic_ctx *c = new magic_ctx();
    c->form = form;
    c->mx = mx; c->my = my; c->mz = mz; c->meqn = meqn; c->mbc = mbc; c->maux = maux;
    c->xlower = xlower; c->ylower = ylower; c->zlower = zlower; c->dx = dx; c->dy = dy; c->dz = dz;
    c->d.mx = mx; c->d.my = my; c->d.mz = mz; c->d.mbc = mbc;
    c->d.NX = mx + 2 * mbc; c->d.NY = my + 2 * mbc; c->d.NZ = mz + 2 * mbc;
    c->P = c->d.P = (long long)c->d.NX * c->d.NY * c->d.NZ;
    const size_t P = (size_t)c->P;
    const size_t NXY = (size_t)c->d.NX * c->d.NY;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&c->stage, P * meqn * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->dscal, 8 * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipMalloc(&c->sink, 4096) == hipSuccess;
    ok = ok && hipHostMalloc(&c->hscal, HSCAL_WORDS * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
    if (ok && form == MAGIC_FORM_FP32P) {
        const size_t NZ = (size_t)c->d.NZ;
        ok = hipMalloc(&c->q32, P * meqn * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->qold32, P * meqn * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->u32, 3 * P * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->qt32, P * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->ut32, P * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->col32, NC32 * NZ * sizeof(float)) == hipSuccess &&
             hipMalloc(&c->bg64, NB64 * NZ * sizeof(double)) == hipSuccess;
        ok = ok && hipMemsetAsync(c->q32, 0, P * meqn * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->qold32, 0, P * meqn * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->u32, 0, 3 * P * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->qt32, 0, P * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->ut32, 0, P * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->col32, 0, NC32 * NZ * sizeof(float), c->stream) == hipSuccess &&
             hipMemsetAsync(c->bg64, 0, NB64 * NZ * sizeof(double), c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
        if (!ok) {
            magic_destroy(c);
            (void)hipGetLastError();   // reported here; not left for the next launch check
            return MAGIC_ENOMEM;
        }
        *out = c;
        return MAGIC_OK;
    }
    ok = ok && hipMalloc(&c->q, P * meqn * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->qold, P * meqn * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->aux, P * NAUXP * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->zratio, 2 * mbc * NXY * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->u, 3 * P * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->qt, P * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->ut, P * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&c->col64, (size_t)NAUXP * c->d.NZ * sizeof(double)) == hipSuccess;
    if (ok) {
        ok = hipMemsetAsync(c->q, 0, P * meqn * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->qold, 0, P * meqn * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->aux, 0, P * NAUXP * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->zratio, 0, 2 * mbc * NXY * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->u, 0, 3 * P * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->qt, 0, P * sizeof(double), c->stream) == hipSuccess &&
             hipMemsetAsync(c->ut, 0, P * sizeof(double), c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
    }
    if (!ok) {
        magic_destroy(c);
        (void)hipGetLastError();   // reported here; not left for the next launch check
        return MAGIC_ENOMEM;
    }
    *out = c;
    return MAGIC_OK;
}

extern "C" int magic_destroy(magic_ctx *c)
{
    if (!c) return MAGIC_OK;
    if (c->cstream) hipStreamSynchronize(c->cstream);   // a snapshot's copy out of stage
    if (c->stream) hipStreamSynchronize(c->stream);
    drain_timing(c);
    for (auto e : c->pool) hipEventDestroy(e);
    hipFree(c->q); hipFree(c->qold); hipFree(c->aux); hipFree(c->zratio); hipFree(c->stage);
    hipFree(c->u); hipFree(c->qt); hipFree(c->ut); hipFree(c->dscal); hipFree(c->sink);
    hipFree(c->q32); hipFree(c->qold32); hipFree(c->u32); hipFree(c->qt32); hipFree(c->ut32);
    hipFree(c->col32); hipFree(c->bg64); hipFree(c->col64); hipFree(c->sums_d);
    hipFree(c->largs); hipFree(c->lrun); hipFree(c->lone); hipFree(c->lcnt);
    for (auto &d : c->litems)
        for (int4 *p : d) hipFree(p);
    if (c->side) hipStreamSynchronize(c->side);
    hipFree(c->dq[0]);
    hipFree(c->dq[1]);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->side) hipStreamDestroy(c->side);
    if (c->hscal) hipHostFree(c->hscal);
    if (c->snap_conv) hipEventDestroy(c->snap_conv);
    if (c->snap_done) hipEventDestroy(c->snap_done);
    if (c->cstream) hipStreamDestroy(c->cstream);
    if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return MAGIC_OK;
}

extern "C" int magic_set_params(magic_ctx *c, const magic_params *p)
{
    if (!c || !p) return MAGIC_EARG;
    c->prm = *p;
    c->have_params = true;
    return MAGIC_OK;
}

// aux slots (1-based) of the ten background planes, in AX_ order
static const int AUX_SLOT[NAUXP] = {1, 2, 3, 4, 5, 11, 12, 17, 18, 19};

// FP32 form: the background must be one column at rest.  Keeps FP64
// columns of rho_bar, E_bar, gamma, R, aux17 and the background's own z jump
// of the well-balanced f-wave (p_bar(k) - p_bar(k-1) - .5 (rho_bar(k) +
// rho_bar(k-1)) aux17(k), p_bar = (gamma-1) E_bar as rpn3euX.f:166-169
// evaluates it for a state at rest; zero to FP64 rounding), which the
// form's FP64 Riemann solve and conversions read; float columns of nu and
// kappa for the diffusion passes; and the bc3 z factors per layer.  get(slot
// index, cell) returns the background value of a padded cell.
template <class F>
static int set_aux32(magic_ctx *c, F get, bool check_columns)
{
    const int NX = c->d.NX, NY = c->d.NY, NZ = c->d.NZ, mbc = c->mbc, mz = c->mz;
    std::vector<double> col((size_t)NAUXP * NZ);
    for (int kk = 0; kk < NZ; kk++) {
        const long long c0 = (long long)kk * NY * NX;
        for (int s = 0; s < NAUXP; s++) {
            const double v0 = get(s, c0);
            if ((s == AX_MX || s == AX_MY || s == AX_MZ) && v0 != 0.0) {
                fprintf(stderr, "magic_hip: FP32 form needs a background at rest (aux slot %d = %g)\n",
                        AUX_SLOT[s], v0);
                return MAGIC_EUNSUPPORTED;
            }
            if (check_columns)
                for (long long e = 0; e < (long long)NY * NX; e++)
                    if (get(s, c0 + e) != v0) {
                        fprintf(stderr, "magic_hip: FP32 form needs a background column (aux slot %d varies "
                                        "in layer %d)\n", AUX_SLOT[s], kk + 1 - mbc);
                        return MAGIC_EUNSUPPORTED;
                    }
            col[(size_t)s * NZ + kk] = v0;
        }
    }
    auto cv = [&](int s, int kk) { return col[(size_t)s * NZ + kk]; };
    std::vector<float> h((size_t)NC32 * NZ, 0.0f);
    std::vector<double> b64((size_t)NB64 * NZ, 0.0);
    for (int kk = 0; kk < NZ; kk++) {
        h[C32_NU * NZ + kk] = (float)cv(AX_NU, kk);
        h[C32_KAP * NZ + kk] = (float)cv(AX_KAP, kk);
        b64[B64_RHO * NZ + kk] = cv(AX_RHO, kk);
        b64[B64_E * NZ + kk] = cv(AX_E, kk);
        b64[B64_GAM * NZ + kk] = cv(AX_GAM, kk);
        b64[B64_R * NZ + kk] = cv(AX_R, kk);
        b64[B64_VG * NZ + kk] = cv(AX_VG, kk);
        if (kk > 0) {
            const double pb = (cv(AX_GAM, kk) - 1.0) * (cv(AX_E, kk) - 0.5 * 0.0 / cv(AX_RHO, kk));
            const double pa = (cv(AX_GAM, kk - 1) - 1.0) * (cv(AX_E, kk - 1) - 0.5 * 0.0 / cv(AX_RHO, kk - 1));
            b64[B64_RES * NZ + kk] = (pb - pa) - 0.5 * (cv(AX_RHO, kk) + cv(AX_RHO, kk - 1)) * cv(AX_VG, kk);
        }
    }
    for (int side = 0; side < 2; side++)
        for (int kbc = 1; kbc <= mbc; kbc++) {
            const int kg = side == 0 ? 1 - kbc : mz + kbc, kb = side == 0 ? 1 : mz;
            c->zr32.s[side][kbc - 1] = std::pow(cv(AX_RHO, kg + mbc - 1) / cv(AX_RHO, kb + mbc - 1), 0.5);
        }
    HIPCHECK(hipMemcpyAsync(c->col32, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->bg64, b64.data(), b64.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->aux11_top = cv(AX_NU, mz + mbc - 1);
    c->aux12_top = cv(AX_KAP, mz + mbc - 1);
    c->have_aux = true;
    return MAGIC_OK;
}

// FP64 form: the ten slots as planes, the bc3 z factors per column.
template <class F>
static int set_aux64(magic_ctx *c, F get)
{
    const long long P = c->P;
    std::vector<double> h((size_t)P * NAUXP);
    for (int s = 0; s < NAUXP; s++) {
        double *dst = h.data() + (size_t)s * P;
        for (long long k = 0; k < P; k++) dst[k] = get(s, k);
    }
    HIPCHECK(hipMemcpyAsync(c->aux, h.data(), (size_t)P * NAUXP * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    {   // layered background: every slot constant on each layer, bit for bit
        const long long NXY = (long long)c->d.NX * c->d.NY;
        const int NZ = c->d.NZ;
        bool layered = true;
        for (int s = 0; s < NAUXP && layered; s++)
            for (int k = 0; k < NZ && layered; k++) {
                const double *row = h.data() + (size_t)s * P + (size_t)k * NXY;
                for (long long e = 1; e < NXY; e++)
                    if (memcmp(row + e, row, sizeof(double)) != 0) {
                        layered = false;
                        break;
                    }
            }
        c->aux_layered = layered;
        if (layered) {
            std::vector<double> cols((size_t)NAUXP * NZ);
            for (int s = 0; s < NAUXP; s++)
                for (int k = 0; k < NZ; k++) cols[(size_t)s * NZ + k] = h[(size_t)s * P + (size_t)k * NXY];
            HIPCHECK(hipMemcpyAsync(c->col64, cols.data(), cols.size() * sizeof(double), hipMemcpyHostToDevice,
                                    c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
    }
    // bc3 z-face scale factors (aux(1,ghost)/aux(1,boundary))**0.5 in FP64
    // (bc3_mpi.f:560-573, 676-689), evaluated once on the host.
    const int NX = c->d.NX, NY = c->d.NY, mbc = c->mbc, mz = c->mz;
    const long long NXY = (long long)NX * NY;
    std::vector<double> zr((size_t)2 * mbc * NXY);
    const double *rho = h.data() + (size_t)AX_RHO * P;
    for (int side = 0; side < 2; side++)
        for (int kbc = 1; kbc <= mbc; kbc++)
            for (long long col = 0; col < NXY; col++) {
                const int kg = side == 0 ? 1 - kbc : mz + kbc;
                const int kb = side == 0 ? 1 : mz;
                const double g = rho[(long long)(kg + mbc - 1) * NXY + col];
                const double b = rho[(long long)(kb + mbc - 1) * NXY + col];
                zr[((size_t)side * mbc + (kbc - 1)) * NXY + col] = std::pow(g / b, 0.5);
            }
    HIPCHECK(hipMemcpyAsync(c->zratio, zr.data(), zr.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    // substep counts read the top cell of column (1,1): src3.f90:315, 319
    const long long c11 = ((long long)(mz + mbc - 1) * NY + (1 + mbc - 1)) * NX + (1 + mbc - 1);
    c->aux11_top = get(AX_NU, c11);
    c->aux12_top = get(AX_KAP, c11);
    c->have_aux = true;
    return MAGIC_OK;
}

extern "C" int magic_set_aux(magic_ctx *c, const double *aux)
{
    if (!c || !aux) return MAGIC_EARG;
    const int maux = c->maux;
    auto get = [&](int s, long long cell) { return aux[cell * maux + AUX_SLOT[s] - 1]; };
    if (c->form == MAGIC_FORM_FP32P) return set_aux32(c, get, true);
    return set_aux64(c, get);
}

extern "C" int magic_set_aux_column(magic_ctx *c, const double *col)
{
    if (!c || !col) return MAGIC_EARG;
    const int maux = c->maux;
    const long long NXY = (long long)c->d.NX * c->d.NY;
    auto get = [&](int s, long long cell) { return col[(cell / NXY) * maux + AUX_SLOT[s] - 1]; };
    if (c->form == MAGIC_FORM_FP32P) return set_aux32(c, get, false);
    return set_aux64(c, get);
}

// stage may be read by a snapshot's copy still in flight: the stream waits
// for it before anything writes stage
static int stage_ready(magic_ctx *c)
{
    if (c->snap_pending) HIPCHECK(hipStreamWaitEvent(c->stream, c->snap_done, 0));
    return MAGIC_OK;
}

// q in the reference layout q(meqn, i, j, k), FP64, into stage (queued)
static int stage_q(magic_ctx *c)
{
    TRY(stage_ready(c));
    if (c->form == MAGIC_FORM_FP32P) {
        const int NZ = c->d.NZ;
        k32_pert_to_aos<<<padded_cells(c->d), cell_block(), 0, c->stream>>>(c->d, c->stage, c->q32, c->meqn,
                                                               c->bg64 + B64_RHO * NZ, c->bg64 + B64_E * NZ);
    } else {
        k_planes_to_aos<<<blocks(c->P, 256), 256, 0, c->stream>>>(c->stage, c->q, c->P, c->meqn);
    }
    return launch_check();
}

extern "C" int magic_set_q(magic_ctx *c, const double *q)
{
    if (!c || !q) return MAGIC_EARG;
    const size_t n = (size_t)c->P * c->meqn;
    TRY(stage_ready(c));
    if (c->form == MAGIC_FORM_FP32P) {
        if (!c->have_aux) return MAGIC_EARG;   // the perturbation needs the background
        const int NZ = c->d.NZ;
        TRY(upload(c, c->stage, q, n * sizeof(double)));
        k32_aos_to_pert<<<padded_cells(c->d), cell_block(), 0, c->stream>>>(c->d, c->q32, c->stage, c->meqn,
                                                               c->bg64 + B64_RHO * NZ, c->bg64 + B64_E * NZ);
        TRY(launch_check());
        HIPCHECK(hipMemcpyAsync(c->qold32, c->q32, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        return MAGIC_OK;
    }
    TRY(upload(c, c->stage, q, n * sizeof(double)));
    k_aos_to_planes<<<blocks(c->P, 256), 256, 0, c->stream>>>(c->q, c->stage, c->P, c->meqn);
    TRY(launch_check());
    HIPCHECK(hipMemcpyAsync(c->qold, c->q, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

// qinit.f:74-82: q = aux(1:5) (fluid at rest on the background), tracers 0,
// on every cell including ghosts.
extern "C" int magic_qinit_rest(magic_ctx *c)
{
    if (!c || !c->have_aux) return MAGIC_EARG;
    const size_t P = (size_t)c->P, n = P * c->meqn;
    if (c->form == MAGIC_FORM_FP32P) {
        HIPCHECK(hipMemsetAsync(c->q32, 0, n * sizeof(float), c->stream));
        HIPCHECK(hipMemsetAsync(c->qold32, 0, n * sizeof(float), c->stream));
    } else {
        HIPCHECK(hipMemsetAsync(c->q, 0, n * sizeof(double), c->stream));
        HIPCHECK(hipMemcpyAsync(c->q, c->aux, 5 * P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(c->qold, c->q, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

extern "C" int magic_get_q(magic_ctx *c, double *q)
{
    if (!c || !q) return MAGIC_EARG;
    const size_t n = (size_t)c->P * c->meqn;
    TRY(stage_q(c));
    return download(c, q, c->stage, n * sizeof(double));
}

extern "C" int magic_snapshot_begin(magic_ctx *c, double *q, void **done)
{
    if (!c || !q || !done) return MAGIC_EARG;
    *done = nullptr;
    if (!c->cstream) {
        HIPCHECK(hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&c->snap_conv, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
    }
    const size_t n = (size_t)c->P * c->meqn;
    TRY(stage_q(c));
    HIPCHECK(hipEventRecord(c->snap_conv, c->stream));
    HIPCHECK(hipStreamWaitEvent(c->cstream, c->snap_conv, 0));
    HIPCHECK(hipMemcpyAsync(q, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->cstream));
    HIPCHECK(hipEventRecord(c->snap_done, c->cstream));
    c->snap_pending = true;
    hipEvent_t ev = nullptr;
    HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (hipEventRecord(ev, c->cstream) != hipSuccess) {
        hipEventDestroy(ev);
        return MAGIC_EHIP;
    }
    *done = ev;
    return MAGIC_OK;
}

extern "C" int magic_snapshot_wait(void *done)
{
    if (!done) return MAGIC_EARG;
    hipEvent_t ev = (hipEvent_t)done;
    const hipError_t e = hipEventSynchronize(ev);
    hipEventDestroy(ev);
    return e == hipSuccess ? MAGIC_OK : MAGIC_EHIP;
}

extern "C" int magic_host_alloc(size_t bytes, void **p)
{
    if (!p || bytes == 0) return MAGIC_EARG;
    *p = nullptr;
    return hipHostMalloc(p, bytes, hipHostMallocDefault) == hipSuccess ? MAGIC_OK : MAGIC_ENOMEM;
}

extern "C" int magic_host_free(void *p)
{
    if (p) HIPCHECK(hipHostFree(p));
    return MAGIC_OK;
}

// ------------------------------------------------------------- bc3 -------
static int bc3_launch(magic_ctx *c, const int *mthbc)
{
    for (int f = 0; f < 6; f++)
        if (mthbc[f] == 0 || mthbc[f] >= 4) return MAGIC_EBC;
    Span sp(c, TC_BC, 3);
    const Dims &d = c->d;
    const double *A = c->aux;
    auto lohi = [](int m) { return m < 0 ? 0 : m; };
    if (c->form == MAGIC_FORM_FP32P) {
        if (mthbc[0] >= 0 || mthbc[1] >= 0) {
            k_bc_x<<<dim3(blocks(d.NY, 64), d.NZ), 64, 0, c->stream>>>(d, c->q32, c->meqn, lohi(mthbc[0]),
                                                                     lohi(mthbc[1]), NoGate{});
            TRY(launch_check());
        }
        if (mthbc[2] >= 0 || mthbc[3] >= 0) {
            k_bc_y<<<dim3(blocks(d.NX, 64), d.NZ), 64, 0, c->stream>>>(d, c->q32, c->meqn, lohi(mthbc[2]),
                                                                     lohi(mthbc[3]), NoGate{});
            TRY(launch_check());
        }
        if (mthbc[4] >= 0 || mthbc[5] >= 0) {
            k32_bc_z<<<dim3(blocks(d.NX, 64), d.NY), 64, 0, c->stream>>>(d, c->q32, c->meqn, lohi(mthbc[4]),
                                                                       lohi(mthbc[5]), c->zr32, NoGate{});
            TRY(launch_check());
        }
        return MAGIC_OK;
    }
    if (mthbc[0] >= 0 || mthbc[1] >= 0) {
        k_bc_x<<<dim3(blocks(d.NY, 64), d.NZ), 64, 0, c->stream>>>(d, c->q, c->meqn, lohi(mthbc[0]),
                                                                 lohi(mthbc[1]), NoGate{});
        TRY(launch_check());
    }
    if (mthbc[2] >= 0 || mthbc[3] >= 0) {
        k_bc_y<<<dim3(blocks(d.NX, 64), d.NZ), 64, 0, c->stream>>>(d, c->q, c->meqn, lohi(mthbc[2]),
                                                                 lohi(mthbc[3]), NoGate{});
        TRY(launch_check());
    }
    if (mthbc[4] >= 0 || mthbc[5] >= 0) {
        k_bc_z<<<dim3(blocks(d.NX, 64), d.NY), 64, 0, c->stream>>>(
            d, c->q, c->meqn, lohi(mthbc[4]), lohi(mthbc[5]), A + AX_RHO * c->P, A + AX_MX * c->P,
            A + AX_MY * c->P, A + AX_E * c->P, c->zratio, NoGate{});
        TRY(launch_check());
    }
    return MAGIC_OK;
}

extern "C" int magic_bc3(magic_ctx *c, const int mthbc[6])
{
    if (!c || !mthbc) return MAGIC_EARG;
    if (!c->have_aux) return MAGIC_EARG;
    TRY(bc3_launch(c, mthbc));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

static int b4step3_launch(magic_ctx *c)
{
    if (c->meqn > 5) {
        Span sp(c, TC_OTHER, 1);
        if (c->form == MAGIC_FORM_FP32P) k_clamp<<<blocks(c->P, 256), 256, 0, c->stream>>>(c->q32, c->P, c->meqn, NoGate{});
        else k_clamp<<<blocks(c->P, 256), 256, 0, c->stream>>>(c->q, c->P, c->meqn, NoGate{});
        TRY(launch_check());
    }
    return MAGIC_OK;
}

extern "C" int magic_b4step3(magic_ctx *c)
{
    if (!c) return MAGIC_EARG;
    TRY(b4step3_launch(c));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

// copyq3: q -> backup (to_backup) or backup -> q
static int copy_state(magic_ctx *c, bool to_backup)
{
    Span sp(c, TC_OTHER, 1);
    const size_t n = (size_t)c->P * c->meqn;
    if (c->form == MAGIC_FORM_FP32P) {
        HIPCHECK(hipMemcpyAsync(to_backup ? c->qold32 : c->q32, to_backup ? c->q32 : c->qold32,
                                n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    } else {
        HIPCHECK(hipMemcpyAsync(to_backup ? c->qold : c->q, to_backup ? c->q : c->qold, n * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
    }
    return MAGIC_OK;
}

extern "C" int magic_backup(magic_ctx *c)
{
    if (!c) return MAGIC_EARG;
    TRY(copy_state(c, true));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

extern "C" int magic_restore(magic_ctx *c)
{
    if (!c) return MAGIC_EARG;
    TRY(copy_state(c, false));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MAGIC_OK;
}

// ------------------------------------------------------------ step3 ------
__global__ void k_zero_scalars(unsigned long long *s)
{
    if (threadIdx.x < 8) s[threadIdx.x] = 0ull;
}

template <int NTR>
static int launch_sweep(magic_ctx *c, SweepArgs a, int *nl)
{
    if (a.n + 3 > 512) return MAGIC_EUNSUPPORTED;   // one workgroup per pencil
    const int NI = a.n + 3;
    const int threads = ((NI + 63) / 64) * 64;
    const size_t smem = (size_t)(7 + NTR) * NI * sizeof(double) + (size_t)NI * sizeof(int);
    for (int ca = 0; ca < 3; ca++)
        for (int cb = 0; cb < 3; cb++) {
            a.cA = ca;
            a.cB = cb;
            dim3 grid((unsigned)((a.nA + 1 - ca) / 3 + 1), (unsigned)((a.nB + 1 - cb) / 3 + 1));
            sweep_kernel<NTR><<<grid, threads, smem, c->stream>>>(a);
            TRY(launch_check());
            ++*nl;
        }
    return MAGIC_OK;
}

// Argument blocks of colours k0 .. k1-1 of direction dir (slot 6 dir + k),
// written on the stream of their launches, so that a later call may change
// them while these launches are still queued
template <typename T>
__global__ void k_put_args(SweepArgsT<T> *slots, SweepArgsT<T> a, int dir, int k0, int k1)
{
    const int k = k0 + (int)threadIdx.x;
    if (k >= k1) return;
    SweepArgsT<T> s = a;
    s.cA = k / 3;
    s.cB = k % 3;
    slots[dir * 6 + k] = s;
}

// The work-list buffers of the context (first tiled sweep)
template <typename T>
static int list_setup(magic_ctx *c)
{
    if (c->largs) return MAGIC_OK;
    if (hipMalloc(&c->largs, 18 * sizeof(SweepArgsT<T>)) != hipSuccess ||
        hipMalloc(&c->lrun, 18 * sizeof(int *)) != hipSuccess || hipMalloc(&c->lone, sizeof(int)) != hipSuccess ||
        hipMalloc(&c->lcnt, 36 * sizeof(int)) != hipSuccess) {
        hipFree(c->largs); hipFree(c->lrun); hipFree(c->lone); hipFree(c->lcnt);
        c->largs = nullptr; c->lrun = nullptr; c->lone = nullptr; c->lcnt = nullptr;
        return MAGIC_ENOMEM;
    }
    const int one = 1;
    const int *run[18];
    for (auto &r : run) r = c->lone;
    const int zero[36] = {};
    HIPCHECK(hipMemcpy(c->lone, &one, sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->lrun, run, sizeof(run), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->lcnt, zero, sizeof(zero), hipMemcpyHostToDevice));
    return MAGIC_OK;
}

// The tiles of direction dir in shape sh, colour by colour (built once)
static int list_items(magic_ctx *c, int dir, int sh, int nA, int nB, int a_is_g)
{
    if (c->litems[dir][sh]) return MAGIC_OK;
    const int TA = tile_ta(sh);
    const int ntiles = (nA + 2 + TA - 1) / TA;
    std::vector<int4> items;
    int *off = c->litem_off[dir][sh];
    for (int ca = 0; ca < 2; ca++)
        for (int cb = 0; cb < 3; cb++) {
            const int k = 3 * ca + cb;
            off[k] = (int)items.size();
            const int gx = (ntiles - ca + 1) / 2, gy = (nB + 1 - cb) / 3 + 1;
            for (int by = 0; by < gy; by++)
                for (int bx = 0; bx < gx; bx++) items.push_back(make_int4(dir * 6 + k, a_is_g, bx, by));
        }
    off[6] = (int)items.size();
    int4 *d = nullptr;
    HIPCHECK(hipMalloc(&d, (items.size() ? items.size() : 1) * sizeof(int4)));
    if (hipMemcpy(d, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return MAGIC_EHIP;
    }
    c->litems[dir][sh] = d;
    return MAGIC_OK;
}

// Colour launches k0 .. k1-1 (k = 3 ca + cb) of one direction: the sweep
// kernel over this grid's tiles of each colour.  The FP64 form runs it as a
// persistent grid over the work list (925 Mcell-updates/s on the full-domain
// workload, against 918 for one tile per workgroup), the FP32 form one tile
// per workgroup (1785 against 1728), which lets the two streams of the sweep
// overlap share the compute units tile by tile.
template <typename T, bool PERT, int NTR>
static int launch_tile(magic_ctx *c, SweepArgsT<T> a, int sh, int a_is_g, int *nl, hipStream_t s, int k0, int k1,
                       bool put)
{
    const void *fn = tile_list_fn<T, PERT, NTR>(sh);
    if (!fn) return MAGIC_EARG;
    const int dir = a.sweep - 1;
    TRY(list_setup<T>(c));
    TRY(list_items(c, dir, sh, a.nA, a.nB, a_is_g));
    if (put) {   // else the caller wrote every slot of this step (k_put_step_args)
        k_put_args<T><<<1, 64, 0, s>>>((SweepArgsT<T> *)c->largs, a, dir, k0, k1);
        TRY(launch_check());
    }
    constexpr bool PER_TILE = sizeof(T) == 4;
    if (!PER_TILE && !c->lgrid[sh]) {
        int dev = 0, ncu = 0, occ = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, 0));
        c->lgrid[sh] = ncu * (occ > 0 ? occ : 1);
    }
    const int *off = c->litem_off[dir][sh];
    for (int k = k0; k < k1; k++) {
        int n = off[k + 1] - off[k];
        if (n <= 0) continue;
        const SweepArgsT<T> *args = (const SweepArgsT<T> *)c->largs;
        const int4 *items = c->litems[dir][sh] + off[k];
        const int *const *run = c->lrun;
        int *cnt = c->lcnt + 2 * (dir * 6 + k);
        // one tile per workgroup: tiles (bx, by) of colour k as the list holds them
        const int TA = tile_ta(sh), ntiles = (a.nA + 2 + TA - 1) / TA, ca = k / 3;
        int4 grid = PER_TILE ? make_int4((ntiles - ca + 1) / 2, dir * 6 + k, a_is_g, 0) : make_int4(0, 0, 0, 0);
        void *kargs[] = {&args, &items, &n, &run, &cnt, &grid};
        HIPCHECK(hipLaunchKernel(fn, dim3(PER_TILE ? n : std::min(c->lgrid[sh], n)), dim3(256), kargs, 0, s));
        ++*nl;
    }
    return MAGIC_OK;
}

template <typename T, bool PERT>
static int tile_dispatch(magic_ctx *c, const SweepArgsT<T> &a, int sh, int a_is_g, int *nl, hipStream_t s, int k0,
                         int k1, bool put)
{
    switch (ntracer(c)) {
    case 0: return launch_tile<T, PERT, 0>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    case 1: return launch_tile<T, PERT, 1>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    case 2: return launch_tile<T, PERT, 2>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    case 3: return launch_tile<T, PERT, 3>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    case 4: return launch_tile<T, PERT, 4>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    case 5: return launch_tile<T, PERT, 5>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    default: return MAGIC_EUNSUPPORTED;
    }
}

static bool use_tile()
{
    const char *e = getenv("MAGIC_SWEEP");
    return !(e && strcmp(e, "pencil") == 0);
}

static int sweep_dispatch(magic_ctx *c, const SweepArgs &a, int *nl)
{
    switch (ntracer(c)) {
    case 0: return launch_sweep<0>(c, a, nl);
    case 1: return launch_sweep<1>(c, a, nl);
    case 2: return launch_sweep<2>(c, a, nl);
    case 3: return launch_sweep<3>(c, a, nl);
    case 4: return launch_sweep<4>(c, a, nl);
    case 5: return launch_sweep<5>(c, a, nl);
    default: return MAGIC_EUNSUPPORTED;
    }
}

// Directional fields of the argument block: dir 0 = x-sweep
// (lib/step3.f:121-280, g = y, h = z), 1 = y-sweep (step3.f:289-450, g = z,
// h = x), 2 = z-sweep (step3.f:459-621, g = x, h = y, gravity on).
template <typename T>
static void dir_args(const magic_ctx *c, SweepArgsT<T> &a, int dir, double dt)
{
    const Dims &d = c->d;
    const long long NXY = (long long)d.NX * d.NY;
    const double dtdx = dt / c->dx, dtdy = dt / c->dy, dtdz = dt / c->dz;
    if (dir == 0) {
        a.comp[0] = 0; a.comp[1] = 1; a.comp[2] = 2; a.comp[3] = 3; a.comp[4] = 4;
        a.n = c->mx; a.nA = c->my; a.nB = c->mz;
        a.sN = 1; a.sA = d.NX; a.sB = NXY;
        a.gravity = 0; a.sweep = 1;
        a.dtdn = dtdx; a.fg = dtdy; a.fh = dtdz;
    } else if (dir == 1) {
        a.comp[0] = 0; a.comp[1] = 2; a.comp[2] = 3; a.comp[3] = 1; a.comp[4] = 4;
        a.n = c->my; a.nA = c->mz; a.nB = c->mx;
        a.sN = d.NX; a.sA = NXY; a.sB = 1;
        a.gravity = 0; a.sweep = 2;
        a.dtdn = dtdy; a.fg = dtdz; a.fh = dtdx;
    } else {
        a.comp[0] = 0; a.comp[1] = 3; a.comp[2] = 1; a.comp[3] = 2; a.comp[4] = 4;
        a.n = c->mz; a.nA = c->mx; a.nB = c->my;
        a.sN = NXY; a.sA = 1; a.sB = d.NX;
        a.gravity = 1; a.sweep = 3;
        a.dtdn = dtdz; a.fg = dtdx; a.fh = dtdy;
    }
}

// One directional sweep with the tiled kernel of shape sh (or the pencil
// kernel of the FP64 form).  The tile axis is g for the x- and z-sweeps and
// x = h for the y-sweep, so that the y- and z-sweeps tile along rows of x.
template <typename T>
static int tile_dir_args(const magic_ctx *c, SweepArgsT<T> &a, int dir, double dt)
{
    dir_args(c, a, dir, dt);
    if (dir == 1) {
        a.nA = c->mx; a.nB = c->mz;
        a.sA = 1; a.sB = (long long)c->d.NX * c->d.NY;
        return 0;
    }
    return 1;
}

// Argument blocks of all 18 launches of one step3 (slot 6 dir + k), from
// each direction's block; dq: the sweep overlap's buffer, which the y-sweep
// and the last three z colours add into (nullptr: no overlap).  Launched on
// the context's stream before any of the step's sweeps, while nothing else
// runs, so it never waits for a compute unit behind a sweep.
template <typename T>
__global__ void k_put_step_args(SweepArgsT<T> *slots, SweepArgsT<T> ax, SweepArgsT<T> ay, SweepArgsT<T> az, T *dq)
{
    const int k = (int)threadIdx.x;
    if (k >= 18) return;
    const int dir = k / 6, kk = k % 6;
    SweepArgsT<T> s = dir == 0 ? ax : dir == 1 ? ay : az;
    s.cA = kk / 3;
    s.cB = kk % 3;
    if (dq && (dir == 1 || (dir == 2 && kk >= 3))) s.qnew = dq;
    slots[k] = s;
}

template <typename T>
static int put_step_args(magic_ctx *c, const SweepArgsT<T> &a, double dt, T *dq)
{
    TRY(list_setup<T>(c));
    SweepArgsT<T> ax = a, ay = a, az = a;
    tile_dir_args(c, ax, 0, dt);
    tile_dir_args(c, ay, 1, dt);
    tile_dir_args(c, az, 2, dt);
    k_put_step_args<T><<<1, 64, 0, c->stream>>>((SweepArgsT<T> *)c->largs, ax, ay, az, dq);
    return launch_check();
}

template <typename T, bool PERT>
static int sweep_dir(magic_ctx *c, SweepArgsT<T> a, int dir, double dt, int sh, bool tile, int *nl,
                     hipStream_t s = nullptr, int k0 = 0, int k1 = 6, bool put = true)
{
    if (!s) s = c->stream;
    if (tile) {
        const int a_is_g = tile_dir_args(c, a, dir, dt);
        return tile_dispatch<T, PERT>(c, a, sh, a_is_g, nl, s, k0, k1, put);
    }
    dir_args(c, a, dir, dt);
    if (s != c->stream || k0 != 0 || k1 != 6) return MAGIC_EUNSUPPORTED;
    if constexpr (!PERT) return sweep_dispatch(c, a, nl);
    return MAGIC_EUNSUPPORTED;
}

// 16-byte vectors of the state's type
template <typename T>
struct Vec16;
template <>
struct Vec16<double> { typedef double2 type; static constexpr int N = 2; };
template <>
struct Vec16<float> { typedef float4 type; static constexpr int N = 4; };

// q += dq over n elements (both 16-byte aligned): 16-byte loads and stores,
// four of each pair in flight per lane
template <typename T>
__device__ __forceinline__ void vadd(typename Vec16<T>::type &a, const typename Vec16<T>::type &b)
{
    a.x = a.x + b.x;
    a.y = a.y + b.y;
    if constexpr (Vec16<T>::N == 4) {
        a.z = a.z + b.z;
        a.w = a.w + b.w;
    }
}

template <typename T>
__global__ void k_merge(T *__restrict__ q, const T *__restrict__ dq, long long n)
{
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::N, U = 4;
    const long long nv = n / N;
    V *qv = reinterpret_cast<V *>(q);
    const V *dv = reinterpret_cast<const V *>(dq);
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long i = t0;
    for (; i + (U - 1) * stride < nv; i += U * stride) {
        V a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            a[u] = qv[i + u * stride];
            b[u] = dv[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            vadd<T>(a[u], b[u]);
            qv[i + u * stride] = a[u];
        }
    }
    for (; i < nv; i += stride) {
        V a = qv[i];
        vadd<T>(a, dv[i]);
        qv[i] = a;
    }
    const long long r = nv * N + t0;   // the last n mod N elements
    if (r < n) q[r] = q[r] + dq[r];
}

template <typename T>
__global__ void k_fill_zero(T *__restrict__ p, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T(0.0);
}

// Whether the y-sweep runs on the second stream (a context that owns its
// stream and has not turned the overlap off; MAGIC_SWEEP_OVERLAP=0 keeps
// the three sweeps on one stream for every context).
static bool overlap_sweeps(const magic_ctx *c)
{
    const char *e = getenv("MAGIC_SWEEP_OVERLAP");
    return c->overlap && c->own_stream && !(e && strcmp(e, "0") == 0);
}

template <typename T>
static int side_setup(magic_ctx *c)
{
    if (!c->side) {
        HIPCHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    if (c->dq_clear) return MAGIC_OK;
    const long long n = (long long)c->meqn * c->P;
    for (int k = 0; k < 2; k++) {
        if (!c->dq[k] && hipMalloc(&c->dq[k], (size_t)n * sizeof(T)) != hipSuccess) {
            c->dq[k] = nullptr;
            return MAGIC_ENOMEM;
        }
        k_fill_zero<T><<<(unsigned)std::min<long long>(blocks(n, 256), 8192), 256, 0, c->stream>>>((T *)c->dq[k], n);
        TRY(launch_check());
    }
    c->dq_clear = true;
    return MAGIC_OK;
}

// The three sweeps of step3 (lib/step3.f:121-621) on argument block a,
// whose state pointers and form-specific fields are already set.
//
// The sweeps only add into q, each from the backup, so they can run on two
// streams into two buffers: the x-sweep and the first three colour launches
// of the z-sweep add into q on the context's stream, the y-sweep and the
// other three z launches into a zeroed buffer dq on the second stream (9
// launches each), and q += dq once both are done.  Each launch's last wave
// of workgroups leaves most compute units idle (4-7 rounds of one workgroup
// per CU per launch); with two streams the other stream's workgroups fill
// them.  Two buffers alternate: the one the previous call added to q is
// cleared on the second stream while this call's sweeps run.  The order of
// the additions into every element is fixed (x and the first z colours,
// then dq's sum of y and the other z colours), so the result is the same
// on every run.
template <typename T, bool PERT>
static int step3_sweeps(magic_ctx *c, const SweepArgsT<T> &a, double dt, bool defer_merge = false)
{
    const bool tile = PERT || use_tile();
    if (tile && overlap_sweeps(c)) {
        TRY(side_setup<T>(c));
        // a launch that fails below leaves the buffers unknown: cleared again next time
        c->dq_clear = false;
        const long long n = (long long)c->meqn * c->P;
        T *dq = (T *)c->dq[c->dq_cur], *dq_next = (T *)c->dq[c->dq_cur ^ 1];
        SweepArgsT<T> ay = a;
        ay.qnew = dq;
        TRY(put_step_args<T>(c, a, dt, dq));
        HIPCHECK(hipEventRecord(c->ev_fork, c->stream));
        HIPCHECK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        // dq_next was added to q by the previous call, before the fork
        k_fill_zero<T><<<(unsigned)std::min<long long>(blocks(n, 256), 8192), 256, 0, c->side>>>(dq_next, n);
        TRY(launch_check());
        {
            Span sp(c, TC_SWY, 0, c->side);
            TRY((sweep_dir<T, PERT>(c, ay, 1, dt, c->shape[1], true, &sp.n, c->side, 0, 6, false)));
        }
        {
            Span sp(c, TC_SWZ, 0, c->side);
            TRY((sweep_dir<T, PERT>(c, ay, 2, dt, c->shape[2], true, &sp.n, c->side, 3, 6, false)));
        }
        {
            Span sp(c, TC_SWX, 0);
            TRY((sweep_dir<T, PERT>(c, a, 0, dt, c->shape[0], true, &sp.n, c->stream, 0, 6, false)));
        }
        {
            Span sp(c, TC_SWZ, 0);
            TRY((sweep_dir<T, PERT>(c, a, 2, dt, c->shape[2], true, &sp.n, c->stream, 0, 3, false)));
        }
        HIPCHECK(hipEventRecord(c->ev_join, c->side));
        HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        c->dq_cur ^= 1;
        c->dq_clear = true;
        if (defer_merge) {
            c->merge_dq = dq;
            return MAGIC_OK;
        }
        Span sp(c, TC_OTHER, 1);
        k_merge<T><<<(unsigned)std::min<long long>(blocks(n / Vec16<T>::N + 1, 256), 8192), 256, 0, c->stream>>>(
            a.qnew, dq, n);
        TRY(launch_check());
        return MAGIC_OK;
    }
    if (tile) TRY(put_step_args<T>(c, a, dt, (T *)nullptr));
    for (int dir = 0; dir < 3; dir++) {
        Span sp(c, TC_SWX + dir, 0);
        TRY((sweep_dir<T, PERT>(c, a, dir, dt, c->shape[dir], tile, &sp.n, c->stream, 0, 6, !tile)));
    }
    return MAGIC_OK;
}

// form-specific part of the argument block: state pointers, columns, scratch
static SweepArgsT<float> args32(magic_ctx *c)
{
    SweepArgsT<float> a;
    memset(&a, 0, sizeof(a));
    a.qold = c->qold32;
    a.qnew = c->q32;
    a.P = c->P;
    a.mbc = c->mbc;
    a.nspecadv = c->have_params ? c->prm.nspecadv : 0;
    a.cfl_bits = c->dscal;
    a.status = (int *)(c->dscal + 1);
    for (int m = 5; m < c->meqn; m++) a.comp[m] = m;
    const int NZ = c->d.NZ;
    a.crho = c->bg64 + B64_RHO * NZ;
    a.cE = c->bg64 + B64_E * NZ;
    a.cgam = c->bg64 + B64_GAM * NZ;
    a.cvg = c->bg64 + B64_VG * NZ;
    a.cres = c->bg64 + B64_RES * NZ;
    a.sink = (float *)c->sink;
    return a;
}

static SweepArgs args64(magic_ctx *c)
{
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    a.qold = c->qold;
    a.qnew = c->q;
    a.gam = c->aux + AX_GAM * c->P;
    a.vgrav = c->aux + AX_VG * c->P;
    a.P = c->P;
    a.mbc = c->mbc;
    a.nspecadv = c->have_params ? c->prm.nspecadv : 0;
    a.cfl_bits = c->dscal;
    a.status = (int *)(c->dscal + 1);
    for (int m = 5; m < c->meqn; m++) a.comp[m] = m;
    a.sink = (double *)c->sink;
    return a;
}

// defer_merge: leave the overlapped sweeps' q += dq to the caller (merge_dq)
static int step3_launch(magic_ctx *c, double dt, bool defer_merge = false)
{
    k_zero_scalars<<<1, 64, 0, c->stream>>>(c->dscal);
    TRY(launch_check());
    if (c->form == MAGIC_FORM_FP32P) TRY((step3_sweeps<float, true>(c, args32(c), dt, defer_merge)));
    else TRY((step3_sweeps<double, false>(c, args64(c), dt, defer_merge)));
    HIPCHECK(hipMemcpyAsync(c->hscal, c->dscal, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            c->stream));
    return MAGIC_OK;
}

static int step3_finish(magic_ctx *c, double *cfl)
{
    HIPCHECK(hipStreamSynchronize(c->stream));
    double v;
    memcpy(&v, &c->hscal[0], sizeof(double));
    *cfl = v;
    const int *st = (const int *)(c->hscal + 1);
    if (st[0] != 0) {
        for (int k = 0; k < 4; k++) c->errloc[k] = st[1 + k];
        return st[0] == 1 ? MAGIC_ERHO : MAGIC_EA2;
    }
    return MAGIC_OK;
}

// the pending q += dq of a deferred merge, as its own pass
template <typename T>
static int merge_launch(magic_ctx *c, T *q)
{
    const long long n = (long long)c->meqn * c->P;
    Span sp(c, TC_OTHER, 1);
    k_merge<T><<<(unsigned)std::min<long long>(blocks(n / Vec16<T>::N + 1, 256), 8192), 256, 0, c->stream>>>(
        q, (const T *)c->merge_dq, n);
    return launch_check();
}
static int merge_flush(magic_ctx *c)
{
    if (!c->merge_dq) return MAGIC_OK;
    const int st = c->form == MAGIC_FORM_FP32P ? merge_launch<float>(c, c->q32) : merge_launch<double>(c, c->q);
    c->merge_dq = nullptr;
    return st;
}

extern "C" int magic_step3(magic_ctx *c, double dt, double *cfl)
{
    if (!c || !cfl || !c->have_aux) return MAGIC_EARG;
    TRY(step3_launch(c, dt));
    int s = step3_finish(c, cfl);
    if (c->timing) drain_timing(c);
    return s;
}

// ------------------------------------------------------- tile shapes -----
extern "C" int magic_set_tile_shapes(magic_ctx *c, const int shapes[3])
{
    if (!c || !shapes || !shapes_valid(shapes)) return MAGIC_EARG;
    for (int k = 0; k < 3; k++) c->shape[k] = shapes[k];
    return MAGIC_OK;
}

extern "C" int magic_get_tile_shapes(magic_ctx *c, int shapes[3])
{
    if (!c || !shapes) return MAGIC_EARG;
    for (int k = 0; k < 3; k++) shapes[k] = c->shape[k];
    return MAGIC_OK;
}

extern "C" int magic_set_sweep_overlap(magic_ctx *c, int on)
{
    if (!c) return MAGIC_EARG;
    c->overlap = on != 0;
    return MAGIC_OK;
}

// Autotuner: every tile shape of every sweep direction is timed on this
// grid's current state (one untimed sweep, then reps timed ones, HIP
// events on the grid's stream) and the fastest shape of each direction is
// kept.  The candidates read the backup copy and add into q; q is restored
// from the backup afterwards, so the state is unchanged bit for bit (the
// backup then holds the current state).
static int tune_rounds(magic_ctx *c, double dt, int reps, TuneClock &clk)
{
    for (int r = 0; r <= reps; r++) {   // repetition 0 warms up and is not timed
        for (int dir = 0; dir < 3; dir++)
            for (int sh = 0; sh < TILE_NSHAPE; sh++) {
                int nl = 0;
                TRY(clk.mark(dir, sh, 0, c->stream));
                TRY((c->form == MAGIC_FORM_FP32P ? sweep_dir<float, true>(c, args32(c), dir, dt, sh, true, &nl)
                                                 : sweep_dir<double, false>(c, args64(c), dir, dt, sh, true, &nl)));
                TRY(clk.mark(dir, sh, 1, c->stream));
            }
        if (r > 0) TRY(clk.add_round());
    }
    return MAGIC_OK;
}

// With the sweeps overlapped on two streams the directions share the
// device, and the fastest assignment need not be each direction's fastest
// shape alone.  Starting from those, each direction's shape is varied in
// turn (at most two passes) and the whole set of step3 sweeps is timed on
// the context's stream (one untimed and reps timed runs); a candidate
// replaces the best one when it is at least 1 % faster.
static int tune_joint(magic_ctx *c, double dt, int reps, int best[3])
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return MAGIC_EHIP;
    }
    auto time_assignment = [&](const int sh[3], float &ms) -> int {
        for (int k = 0; k < 3; k++) c->shape[k] = sh[k];
        ms = 0.f;
        for (int r = 0; r <= reps; r++) {
            if (hipEventRecord(e0, c->stream) != hipSuccess) return MAGIC_EHIP;
            TRY((c->form == MAGIC_FORM_FP32P ? step3_sweeps<float, true>(c, args32(c), dt)
                                             : step3_sweeps<double, false>(c, args64(c), dt)));
            float t = 0.f;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                return MAGIC_EHIP;
            if (r > 0) ms += t / reps;
        }
        return MAGIC_OK;
    };
    float bms = 0.f;
    int st = time_assignment(best, bms);
    for (int pass = 0; pass < 2 && st == MAGIC_OK; pass++) {
        bool changed = false;
        for (int d = 0; d < 3 && st == MAGIC_OK; d++)
            for (int sh = 0; sh < TILE_NSHAPE && st == MAGIC_OK; sh++) {
                if (sh == best[d]) continue;
                int cand[3] = {best[0], best[1], best[2]};
                cand[d] = sh;
                float ms = 0.f;
                st = time_assignment(cand, ms);
                if (st == MAGIC_OK && ms < 0.99f * bms) {
                    bms = ms;
                    best[d] = sh;
                    changed = true;
                }
            }
        if (!changed) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return st;
}

extern "C" int magic_tune(magic_ctx *c, double dt, int reps, int shapes_out[3], float *ms_out)
{
    if (!c || !c->have_aux || !(dt > 0.0) || reps < 1) return MAGIC_EARG;
    if (!(c->form == MAGIC_FORM_FP32P || use_tile()) || !sweep_supported(ntracer(c))) return MAGIC_EUNSUPPORTED;
    TuneClock clk;
    if (!clk.ok) return MAGIC_EHIP;
    const int keep[3] = {c->shape[0], c->shape[1], c->shape[2]};
    TRY(copy_state(c, true));
    int st = tune_rounds(c, dt, reps, clk);
    int best[3];
    if (st == MAGIC_OK) {
        clk.result(reps, best, ms_out);
        if (overlap_sweeps(c)) st = tune_joint(c, dt, reps, best);
    }
    if (st != MAGIC_OK)
        for (int k = 0; k < 3; k++) c->shape[k] = keep[k];
    TRY(copy_state(c, false));   // also after a failed round
    TRY(st);
    k_zero_scalars<<<1, 64, 0, c->stream>>>(c->dscal);   // the candidates' Courant numbers and reports
    TRY(launch_check());
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; k++) c->shape[k] = best[k];
    if (shapes_out)
        for (int k = 0; k < 3; k++) shapes_out[k] = best[k];
    return MAGIC_OK;
}

// ------------------------------------------------------------- src3 ------
// The source pass of the FP32 form: the FP64 form's passes (below) on the
// perturbation planes, with nu and kappa as background columns.
static int src3_passes32(magic_ctx *c, const int *mthbc, int ndtr, double dtr, int ndtrc, double dtrc,
                         const Forcing &f)
{
    const Dims &d = c->d;
    const long long P = c->P;
    const int NZ = d.NZ;
    const dim3 gint = interior_grid(d);
    const dim3 gface(blocks(d.NX, 64), d.NY);
    const float *nu = c->col32 + C32_NU * NZ, *kap = c->col32 + C32_KAP * NZ;
    const Col32 col = col32_of(c);
    {
        Span sp(c, TC_SRC, 1);
        // the pending q += dq of the overlapped sweeps, if any, goes with the conversion
        k32_src_convert<<<padded_cells(c->d), cell_block(), 0, c->stream>>>(d, c->q32, col, NoGate{},
                                                                             (const float *)c->merge_dq, c->meqn);
        c->merge_dq = nullptr;
        TRY(launch_check());
    }
    for (int nn = 1; nn <= ndtr; nn++) {
        TRY(bc3_launch(c, mthbc));
        Span sp(c, TC_SRC, 6 + 6 * (ndtrc > 0 ? ndtrc : 0));
        const double fourthirds = 4.0 / 3.0;
        double dtdx2 = (dtr) / (c->dx * c->dx), dtdy2 = (dtr) / (c->dy * c->dy), dtdz2 = (dtr) / (c->dz * c->dz);
        k_zghost_uvw<<<gface, 64, 0, c->stream>>>(d, c->q32, NoGate{});
        TRY(launch_check());
        const double dtr32 = dtr / 6.0;
        k_diff_cross<float, true, NoGate><<<diff_grid(d), cell_block(), 0, c->stream>>>(d, c->u32, c->q32, nu,
                                                             (float)(dtr32 / (c->dx * c->dy)),
                                                             (float)(dtr32 / (c->dx * c->dz)),
                                                             (float)(dtr32 / (c->dy * c->dz)), diff_kb(d), NoGate{});
        TRY(launch_check());
        for (int cc = 1; cc <= 3; cc++) {
            if (cc == 1) {
                dtdx2 = fourthirds * dtdx2;
            } else if (cc == 2) {
                dtdy2 = fourthirds * dtdy2;
                dtdx2 = dtdx2 / fourthirds;
            } else {
                dtdz2 = fourthirds * dtdz2;
                dtdy2 = dtdy2 / fourthirds;
            }
            k_diff_xyz<float, true, NoGate><<<diff_grid(d), cell_block(), 0, c->stream>>>(
                d, c->q32 + cc * P, c->u32 + (cc - 1) * P, nu, (float)dtdx2, (float)dtdy2, (float)dtdz2, diff_kb(d),
                NoGate{});
            TRY(launch_check());
        }
        const float cx = (float)((dtrc) / (c->dx * c->dx)), cy = (float)((dtrc) / (c->dy * c->dy)),
                    cz = (float)((dtrc) / (c->dz * c->dz));
        float *q5 = c->q32 + 4 * P;
        HIPCHECK(hipMemcpyAsync(c->ut32, q5, P * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        for (int k = 1; k <= ndtrc; k++) {
            k_pass<float, true, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->qt32, c->ut32, c->ut32, kap, cx, 0, NoGate{});
            k_pass<float, true, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->ut32, c->qt32, c->ut32, kap, cy, 1, NoGate{});
            k_conduct_zghost<<<gface, 64, 0, c->stream>>>(d, c->ut32, NoGate{});
            if (c->mz > c->mx) {
                const int S = c->mz - c->mx;
                k_conduct_zspill<<<(S + 63) / 64, 64, 0, c->stream>>>(d, c->ut32, S, NoGate{});
            }
            if (k < ndtrc) {
                k_pass<float, true, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->qt32, c->ut32, c->ut32, kap, cz, 2, NoGate{});
                k_copy_interior<<<gint, cell_block(), 0, c->stream>>>(d, c->ut32, c->qt32, 1, NoGate{});
            } else {
                k_pass<float, true, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, q5, c->ut32, c->ut32, kap, cz, 2, NoGate{});
            }
            TRY(launch_check());
        }
    }
    {
        Span sp(c, TC_SRC, 1);
        k32_src_finish<<<gint, cell_block(), 0, c->stream>>>(d, c->q32, c->meqn, col, f, NoGate{});
        TRY(launch_check());
    }
    return MAGIC_OK;
}

// The source pass of the FP64 form (src3.f90:203-634 with diffuse.f and
// conduct.f).  COL: the background (aux 1-5, 11, 12, 18, 19) is read from
// its columns, which hold the same values when it is constant on each layer.
template <bool COL>
static int src3_passes64(magic_ctx *c, const int *mthbc, int ndtr, double dtr, int ndtrc, double dtrc,
                         const Forcing &f)
{
    const Dims &d = c->d;
    const long long P = c->P;
    const int NZ = d.NZ;
    const double *A = COL ? c->col64 : c->aux;
    const long long S = COL ? NZ : P;   // stride between slots
    const Aux ax = {A + AX_RHO * S, A + AX_MX * S, A + AX_MY * S, A + AX_MZ * S, A + AX_E * S, A + AX_R * S,
                    A + AX_GAM * S};
    const double *nu = A + AX_NU * S, *kap = A + AX_KAP * S;
    const dim3 gint = interior_grid(d);
    const dim3 gface(blocks(d.NX, 64), d.NY);
    {
        Span sp(c, TC_SRC, 1);
        // the pending q += dq of the overlapped sweeps, if any, goes with the conversion
        k_src_convert<COL><<<padded_cells(d), cell_block(), 0, c->stream>>>(d, c->q, ax, NoGate{},
                                                                            (const double *)c->merge_dq, c->meqn);
        c->merge_dq = nullptr;
        TRY(launch_check());
    }
    for (int nn = 1; nn <= ndtr; nn++) {
        TRY(bc3_launch(c, mthbc));
        Span sp(c, TC_SRC, 6 + 6 * (ndtrc > 0 ? ndtrc : 0));
        // diffuse: diffuse.f:37-210.  After the z-ghost fill the cross terms
        // of q(2:4) go to u's interior; each component then takes its three
        // directional passes in one kernel, whose ghost input is q's own
        // ghost layers (the reference's copy, stale as there), and whose
        // result replaces q's interior.
        const double fourthirds = 4.0 / 3.0;
        double dtdx2 = (dtr) / (c->dx * c->dx), dtdy2 = (dtr) / (c->dy * c->dy), dtdz2 = (dtr) / (c->dz * c->dz);
        k_zghost_uvw<<<gface, 64, 0, c->stream>>>(d, c->q, NoGate{});
        TRY(launch_check());
        const double dtr32 = dtr / 6.0;
        k_diff_cross<double, COL, NoGate><<<diff_grid(d), cell_block(), 0, c->stream>>>(d, c->u, c->q, nu,
                                                 dtr32 / (c->dx * c->dy), dtr32 / (c->dx * c->dz),
                                                 dtr32 / (c->dy * c->dz), diff_kb(d), NoGate{});
        TRY(launch_check());
        for (int cc = 1; cc <= 3; cc++) {
            if (cc == 1) {
                dtdx2 = fourthirds * dtdx2;
            } else if (cc == 2) {
                dtdy2 = fourthirds * dtdy2;
                dtdx2 = dtdx2 / fourthirds;
            } else {
                dtdz2 = fourthirds * dtdz2;
                dtdy2 = dtdy2 / fourthirds;
            }
            k_diff_xyz<double, COL, NoGate><<<diff_grid(d), cell_block(), 0, c->stream>>>(
                d, c->q + cc * P, c->u + (cc - 1) * P, nu, dtdx2, dtdy2, dtdz2, diff_kb(d), NoGate{});
            TRY(launch_check());
        }
        // conduct: conduct.f:245-332.  ut = q(5) on every cell once; the
        // interior ping-pongs between ut and qt, the ghost layers stay in ut
        // (x/y ghosts never refreshed, z ghosts refreshed for i <= mz+mbc
        // before each z pass); the last z pass writes q(5) directly.
        const double cx = (dtrc) / (c->dx * c->dx), cy = (dtrc) / (c->dy * c->dy), cz = (dtrc) / (c->dz * c->dz);
        double *q5 = c->q + 4 * P;
        HIPCHECK(hipMemcpyAsync(c->ut, q5, P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        for (int k = 1; k <= ndtrc; k++) {
            k_pass<double, COL, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->qt, c->ut, c->ut, kap, cx, 0, NoGate{});
            k_pass<double, COL, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->ut, c->qt, c->ut, kap, cy, 1, NoGate{});
            k_conduct_zghost<<<gface, 64, 0, c->stream>>>(d, c->ut, NoGate{});
            if (c->mz > c->mx) {
                const int S = c->mz - c->mx;
                k_conduct_zspill<<<(S + 63) / 64, 64, 0, c->stream>>>(d, c->ut, S, NoGate{});
            }
            if (k < ndtrc) {
                k_pass<double, COL, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, c->qt, c->ut, c->ut, kap, cz, 2, NoGate{});
                k_copy_interior<<<gint, cell_block(), 0, c->stream>>>(d, c->ut, c->qt, 1, NoGate{});
            } else {
                k_pass<double, COL, NoGate><<<gint, cell_block(), 0, c->stream>>>(d, q5, c->ut, c->ut, kap, cz, 2, NoGate{});
            }
            TRY(launch_check());
        }
    }
    {
        Span sp(c, TC_SRC, 1);
        k_src_finish<COL><<<gint, cell_block(), 0, c->stream>>>(d, c->q, c->meqn, ax, f, NoGate{});
        TRY(launch_check());
    }
    return MAGIC_OK;
}

static int src3_launch(magic_ctx *c, double t, double dt, const int *mthbc_in, int *ndtr_out,
                       int *ndtrc_out)
{
    const magic_params &p = c->prm;
    if (p.omega1 != 0.0) return MAGIC_EUNSUPPORTED;   // time-dependent wind (src3.f90:389-399)
    const int fm = (int)p.forcemth;
    if (!(p.forcemth == 0.0 || p.forcemth == 1.0 || p.forcemth == 3.0)) return MAGIC_EUNSUPPORTED;
    // substep counts: src3.f90:314-322
    double dmin = c->dx < c->dy ? c->dx : c->dy;
    dmin = dmin < c->dz ? dmin : c->dz;
    const double dxyzmin2 = dmin * dmin;
    const double difmax = (4.0 / 3.0) * c->aux11_top;
    const int ndtr = (int)std::ceil((dt * difmax / (dxyzmin2)) / 0.45);
    const double dtr = dt / ndtr;
    // no viscosity at the top: ndtr = 0 and the substep loop runs no pass
    // (src3.f90:341); the conduction count of an infinite dtr is then 0
    const int ndtrc = ndtr > 0 ? (int)std::ceil((dtr * c->aux12_top / (dxyzmin2)) / 0.45) : 0;
    const double dtrc = dtr / ndtrc;
    *ndtr_out = ndtr;
    *ndtrc_out = ndtrc;
    int mthbc[6];
    memcpy(mthbc, mthbc_in, sizeof(mthbc));
    mthbc[4] = -1;
    mthbc[5] = -1;
    Forcing f;
    f.forcemth = fm;
    f.t = t; f.dt = dt;
    f.xlower = c->xlower; f.ylower = c->ylower; f.zlower = c->zlower;
    f.dx = c->dx; f.dy = c->dy; f.dz = c->dz;
    f.amplitude = p.amplitude; f.xpos = p.xpos; f.ypos = p.ypos; f.zpos = p.zpos;
    f.xwidth = p.xwidth; f.ywidth = p.ywidth; f.zwidth = p.zwidth;
    f.tcenter = p.tcenter; f.twidth = p.twidth;
    f.omega = p.omega; f.propx = p.propx; f.propz = p.propz; f.vsrcx = p.vsrcx;
    f.tgauss = std::exp(-.5 * ((t - p.tcenter) * (t - p.tcenter) / (p.twidth * p.twidth)));
    if (c->form == MAGIC_FORM_FP32P) return src3_passes32(c, mthbc, ndtr, dtr, ndtrc, dtrc, f);
    return c->aux_layered ? src3_passes64<true>(c, mthbc, ndtr, dtr, ndtrc, dtrc, f)
                          : src3_passes64<false>(c, mthbc, ndtr, dtr, ndtrc, dtrc, f);
}

extern "C" int magic_src3(magic_ctx *c, double t, double dt, const int mthbc[6], int *ndtr, int *ndtrc)
{
    if (!c || !mthbc || !ndtr || !ndtrc || !c->have_aux || !c->have_params) return MAGIC_EARG;
    TRY(src3_launch(c, t, dt, mthbc, ndtr, ndtrc));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->timing) drain_timing(c);
    return MAGIC_OK;
}

// ------------------------------------------------------------ claw3 ------
extern "C" int magic_claw3(magic_ctx *c, double tstart, double *tend, double dtv[5], double cflv[4],
                           int nv[2], const int method[7], const int mthbc[6], double *log_dt,
                           double *log_cfl, int *log_acc, int maxlog, int *nlog)
{
    if (!c || !tend || !dtv || !cflv || !nv || !method || !mthbc) return MAGIC_EARG;
    if (!c->have_aux || !c->have_params) return MAGIC_EARG;
    if (method[0] != 1 || method[1] != 2 || method[2] != 22 || method[4] != 1 || method[5] != 0)
        return MAGIC_EUNSUPPORTED;
    if (cflv[1] > cflv[0]) return MAGIC_EARG;          // claw3_mpi.f:671-675
    double t = tstart, dt = dtv[0], cflmax = 0.0, dtmin = dt, dtmax = dt, cfl = 0.0;
    const int maxn = nv[0];
    int st = MAGIC_OK, nl = 0;
    nv[1] = 0;
    for (int n = 1; n <= maxn; n++) {
        const double told = t;
        if (told + dt > *tend && tstart < *tend) dt = *tend - told;   // claw3_mpi.f:741
        for (;;) {
            if ((st = bc3_launch(c, mthbc))) goto out;
            if ((st = b4step3_launch(c))) goto out;
            if ((st = copy_state(c, true))) goto out;
            if ((st = step3_launch(c, dt, true))) goto out;
            if ((st = step3_finish(c, &cfl))) goto out;
            t = told + dt;
            if (log_dt && nl < maxlog) {
                log_dt[nl] = dt;
                if (log_cfl) log_cfl[nl] = cfl;
                if (log_acc) log_acc[nl] = cfl <= cflv[0];
            }
            nl++;
            if (cfl <= cflv[0]) {                                  // claw3_mpi.f:861-886
                cflmax = cflmax > cfl ? cflmax : cfl;
                break;
            }
            t = told;
            c->merge_dq = nullptr;   // the rejected step's increments are dropped with it
            if ((st = copy_state(c, false))) goto out;
            dt = dtv[1] < dt * cflv[1] / cfl ? dtv[1] : dt * cflv[1] / cfl;
        }
        {
            int a, b;
            if ((st = src3_launch(c, told, dt, mthbc, &a, &b))) goto out;   // claw3_mpi.f:900-905
        }
        if (cfl == 0.0) dt = dtv[1];                               // claw3_mpi.f:910-919
        else dt = dtv[1] < dt * cflv[1] / cfl ? dtv[1] : dt * cflv[1] / cfl;
        dtmin = dt < dtmin ? dt : dtmin;
        dtmax = dt > dtmax ? dt : dtmax;
        nv[1] = nv[1] + 1;
        if (t >= *tend) break;
    }
out:
    {
        // a step that stopped before its source pass still ends with q += dq
        const int sm = merge_flush(c);
        if (st == MAGIC_OK) st = sm;
    }
    hipStreamSynchronize(c->stream);
    if (c->timing) drain_timing(c);
    if (nlog) *nlog = nl < maxlog ? nl : maxlog;
    *tend = t;
    cflv[2] = cflmax;
    cflv[3] = cfl;
    dtv[2] = dtmin;
    dtv[3] = dtmax;
    dtv[4] = dt;
    return st;
}

extern "C" int magic_error_location(magic_ctx *c, int *sweep, int *i, int *j, int *k)
{
    if (!c) return MAGIC_EARG;
    if (sweep) *sweep = c->errloc[0];
    if (i) *i = c->errloc[1];
    if (j) *j = c->errloc[2];
    if (k) *k = c->errloc[3];
    return MAGIC_OK;
}

extern "C" int magic_timing_enable(magic_ctx *c, int on)
{
    if (!c) return MAGIC_EARG;
    hipStreamSynchronize(c->stream);
    drain_timing(c);
    c->timing = on != 0;
    return MAGIC_OK;
}

extern "C" int magic_timing_read(magic_ctx *c, double ms[6], long long launches[6])
{
    if (!c) return MAGIC_EARG;
    hipStreamSynchronize(c->stream);
    drain_timing(c);
    for (int k = 0; k < NTC; k++) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
    }
    return MAGIC_OK;
}

extern "C" int magic_timing_reset(magic_ctx *c)
{
    if (!c) return MAGIC_EARG;
    hipStreamSynchronize(c->stream);
    drain_timing(c);
    for (int k = 0; k < NTC; k++) {
        c->ms[k] = 0.0;
        c->launches[k] = 0;
    }
    return MAGIC_OK;
}
