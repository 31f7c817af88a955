// Internal declarations shared by the C ABI (capi.hip) and the batched
// driver (batch.hip): the per-grid context, timing spans and host helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/magic_hip.h"
#include "bc_src.h"
#include "magic_dev.h"

using namespace magic;

#define HIPCHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "magic_hip: %s failed: %s (%s:%d)\n", #x,                 \
                    hipGetErrorString(e_), __FILE__, __LINE__);                       \
            return MAGIC_EHIP;                                                        \
        }                                                                             \
    } while (0)

#define TRY(x)                        \
    do {                              \
        int s_ = (x);                 \
        if (s_ != MAGIC_OK) return s_; \
    } while (0)

// pinned host words of a context and where the global sums land in them
constexpr int HSCAL_WORDS = 32, HSCAL_SUMS = 16;

enum { AX_RHO = 0, AX_MX, AX_MY, AX_MZ, AX_E, AX_NU, AX_KAP, AX_VG, AX_R, AX_GAM, NAUXP };
enum { TC_SWX = 0, TC_SWY, TC_SWZ, TC_BC, TC_SRC, TC_OTHER, NTC };

struct TimedSpan {
    int cls;
    hipEvent_t a, b;
    int nlaunch;
};

struct magic_ctx {
    int form = MAGIC_FORM_FP64;
    int mx, my, mz, meqn, mbc, maux;
    double xlower, ylower, zlower, dx, dy, dz;
    Dims d;
    long long P;
    hipStream_t stream = nullptr;
    double *q = nullptr, *qold = nullptr, *aux = nullptr, *zratio = nullptr;
    double *stage = nullptr, *u = nullptr, *qt = nullptr, *ut = nullptr;
    // FP32 perturbation form: state planes and scratch, background columns
    // (float: nu, kappa; FP64: rho, E, gamma, R, aux17, z residual)
    float *q32 = nullptr, *qold32 = nullptr, *u32 = nullptr, *qt32 = nullptr, *ut32 = nullptr;
    float *col32 = nullptr;
    double *bg64 = nullptr;
    Zr32 zr32;
    // FP64 form: when every aux slot the source pass reads is constant on
    // each layer (bitwise), those slots also live as columns (index k+mbc-1,
    // AX_ order) and the source pass reads the columns instead of the planes
    double *col64 = nullptr;
    bool aux_layered = false;
    bool own_stream = true;   // false when a batch lends its stream
    // The y-sweep of a single grid runs on a second stream next to the x-
    // and z-sweeps and adds into dq[dq_cur] (meqn planes of the form's
    // type), added to q after both; the other buffer is cleared meanwhile
    // for the next step3.  Allocated on first use.
    bool overlap = true;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void *dq[2] = {nullptr, nullptr};
    int dq_cur = 0;
    bool dq_clear = false;   // dq[dq_cur] holds zeros (false until set up, and after a failed overlapped step)
    // Inside magic_claw3 the q += dq pass after an overlapped step3 is left
    // to the source pass's conversion (one pass over q instead of two);
    // merge_dq is the buffer still to be added (null: nothing pending).
    void *merge_dq = nullptr;
    // Asynchronous snapshots (magic_snapshot_begin): the conversion into
    // stage runs on the context's stream, the copy out of stage on cstream
    // (created on first use), so that later work on the stream need not
    // wait for the copy.  snap_done marks the last such copy; every other
    // user of stage waits for it first (stage_ready).
    hipStream_t cstream = nullptr;
    hipEvent_t snap_conv = nullptr, snap_done = nullptr;
    bool snap_pending = false;
    unsigned long long *dscal = nullptr;   // [0] cfl bits, [1..] status ints
    void *sink = nullptr;                  // 4 KiB the sweep's tile flush may write (never read back)
    int shape[3] = {0, 0, 0};              // tile shape of the x-, y- and z-sweeps
    // The sweep kernel's operands (set up on the first tiled sweep): 18
    // argument blocks (direction x colour, written by a kernel at the start
    // of each step3 or of each tuner launch), run flags all pointing at a
    // device 1, 18 pairs of work counters, and per direction and shape the
    // tiles of each colour (items of colour k at [litem_off[dir][sh][k],
    // ..[k + 1])
    void *largs = nullptr;
    const int **lrun = nullptr;
    int *lone = nullptr;
    int *lcnt = nullptr;
    int4 *litems[3][TILE_NSHAPE] = {};
    int litem_off[3][TILE_NSHAPE][7] = {};
    int lgrid[TILE_NSHAPE] = {};           // persistent grid of each shape's kernel (0: not yet known)
    unsigned long long *hscal = nullptr;   // pinned host words: [0, 8) mirror dscal, [HSCAL_SUMS, +10) global sums
    void *sums_d = nullptr;                // global sums: partials and results (first magic_global_sums)
    magic_params prm;
    bool have_params = false, have_aux = false;
    double aux11_top = 0.0, aux12_top = 0.0;
    int errloc[4] = {0, 0, 0, 0};
    bool timing = false;
    double ms[NTC] = {0};
    long long launches[NTC] = {0};
    std::vector<TimedSpan> pending;
    std::vector<hipEvent_t> pool;
};

// ------------------------------------------------------------ timing -----
static inline hipEvent_t ev_get(magic_ctx *c)
{
    hipEvent_t e;
    if (!c->pool.empty()) {
        e = c->pool.back();
        c->pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    return e;
}

struct Span {
    magic_ctx *c;
    int cls, n;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Span(magic_ctx *c_, int cls_, int n_, hipStream_t s_ = nullptr) : c(c_), cls(cls_), n(n_), s(s_ ? s_ : c_->stream)
    {
        if (c->timing && (a = ev_get(c))) hipEventRecord(a, s);
    }
    ~Span()
    {
        if (!c->timing) {
            return;
        }
        hipEvent_t b = ev_get(c);
        if (a && b) {
            hipEventRecord(b, s);
            c->pending.push_back({cls, a, b, n});
        }
    }
};

static inline void drain_timing(magic_ctx *c)
{
    for (auto &s : c->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(s.b) == hipSuccess && hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
            c->ms[s.cls] += ms;
            c->launches[s.cls] += s.nlaunch;
        }
        c->pool.push_back(s.a);
        c->pool.push_back(s.b);
    }
    c->pending.clear();
}

// ---------------------------------------------------------- helpers ------
static inline unsigned blocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

static inline int launch_check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "magic_hip: kernel launch failed: %s\n", hipGetErrorString(e));
        return MAGIC_EHIP;
    }
    return MAGIC_OK;
}

static inline int ntracer(const magic_ctx *c) { return c->meqn - 5; }

// Launch shape of the per-cell kernels (INTERIOR_IJK, PADDED_IJK): 64 cells
// along x by 4 rows along y per workgroup, one layer per grid z index.
static inline dim3 cell_block() { return dim3(64, 4, 1); }
static inline dim3 interior_grid(const Dims &d) { return dim3(blocks(d.mx, 64), blocks(d.my, 4), d.mz); }
static inline dim3 padded_cells(const Dims &d) { return dim3(blocks(d.NX, 64), blocks(d.NY, 4), d.NZ); }
// k_diff_xyz: layers per workgroup so that a launch has about 8192
// workgroups (at least 4 layers each; each workgroup recomputes the x and y
// passes of two extra layers, and its layers run one after the other, so
// many short workgroups hide the latency better), and its grid
static inline int diff_kb(const Dims &d)
{
    const long long cols = (long long)blocks(d.mx, 64) * blocks(d.my, 4);
    const long long chunks = (8192 + cols - 1) / cols;
    const int kb = (int)((d.mz + chunks - 1) / chunks);
    return kb < 4 ? 4 : kb;
}
static inline dim3 diff_grid(const Dims &d) { return dim3(blocks(d.mx, 64), blocks(d.my, 4), blocks(d.mz, diff_kb(d))); }

// every tracer count of meqn 5 .. 10 (rpt3euX.f:134-138 splits tracers 6 .. meqn)
static inline bool sweep_supported(int ntr) { return ntr >= 0 && ntr <= 5; }

// ------------------------------------------------------- tile shapes ----
// The tiled sweep kernel of shape sh (0 .. TILE_NSHAPE-1): the persistent
// work-list kernel that one grid and a batch of grids both launch.
template <typename T, bool PERT, int NTR>
static inline const void *tile_list_fn(int sh)
{
    switch (sh) {
    case 0: return reinterpret_cast<const void *>(&tile_sweep_list<T, PERT, NTR, 0>);
    case 1: return reinterpret_cast<const void *>(&tile_sweep_list<T, PERT, NTR, 1>);
    case 2: return reinterpret_cast<const void *>(&tile_sweep_list<T, PERT, NTR, 2>);
    case 3: return reinterpret_cast<const void *>(&tile_sweep_list<T, PERT, NTR, 3>);
    default: return nullptr;
    }
}
static_assert(TILE_NSHAPE == 4, "tile_list_fn lists every shape");

// Event pairs of the autotuners, one per (direction, shape).  The shapes
// are timed round-robin (every round times each candidate once), so that a
// drift of the clocks over the tuning run favours none of them.
struct TuneClock {
    hipEvent_t ev[3][TILE_NSHAPE][2] = {};
    float ms[3][TILE_NSHAPE] = {};
    bool ok = true;
    TuneClock()
    {
        for (auto &d : ev)
            for (auto &s : d)
                for (auto &e : s) ok = ok && hipEventCreate(&e) == hipSuccess;
    }
    ~TuneClock()
    {
        for (auto &d : ev)
            for (auto &s : d)
                for (auto &e : s)
                    if (e) hipEventDestroy(e);
    }
    int mark(int dir, int sh, int end, hipStream_t st)
    {
        return hipEventRecord(ev[dir][sh][end], st) == hipSuccess ? MAGIC_OK : MAGIC_EHIP;
    }
    // adds every pair of the round just recorded
    int add_round()
    {
        for (int d = 0; d < 3; d++)
            for (int s = 0; s < TILE_NSHAPE; s++) {
                float t = 0.f;
                if (hipEventSynchronize(ev[d][s][1]) != hipSuccess ||
                    hipEventElapsedTime(&t, ev[d][s][0], ev[d][s][1]) != hipSuccess)
                    return MAGIC_EHIP;
                ms[d][s] += t;
            }
        return MAGIC_OK;
    }
    // fastest shape per direction; mean ms per sweep over reps rounds
    void result(int reps, int best[3], float *ms_out) const
    {
        for (int d = 0; d < 3; d++) {
            best[d] = 0;
            for (int s = 1; s < TILE_NSHAPE; s++)
                if (ms[d][s] < ms[d][best[d]]) best[d] = s;
            if (ms_out)
                for (int s = 0; s < TILE_NSHAPE; s++) ms_out[d * TILE_NSHAPE + s] = ms[d][s] / (float)reps;
        }
    }
};

static inline bool shapes_valid(const int *sh)
{
    for (int k = 0; k < 3; k++)
        if (sh[k] < 0 || sh[k] >= TILE_NSHAPE) return false;
    return true;
}

// FP32 form's background columns (index kk = k + mbc - 1): float columns of
// the diffusivities, FP64 columns of everything its arithmetic reads
enum { C32_NU = 0, C32_KAP, NC32 };
enum { B64_RHO = 0, B64_E, B64_GAM, B64_R, B64_VG, B64_RES, NB64 };
static inline Col32 col32_of(const magic_ctx *c)
{
    const int NZ = c->d.NZ;
    return Col32{c->bg64 + B64_RHO * NZ, c->bg64 + B64_E * NZ, c->bg64 + B64_GAM * NZ, c->bg64 + B64_R * NZ};
}
