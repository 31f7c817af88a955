// Batched driver: many grids (scenarios) of one form (the FP32 perturbation
// form or the FP64 form) advanced together on one stream, with claw3's step
// control on the device (batch.h) and one attempted step of every scenario
// captured once as a HIP graph and replayed.
//
// One attempt:
//   prepare  (dt clip, per-grid sweep factors, source-pass substep counts)
//   bc3 (x, y, z faces), b4step3, copyq3 backup
//   3 directions x 6 tile colours of the sweep  (persistent work-list launches)
//   decide   (accept / reject from the Courant number)
//   restore of rejected scenarios, src3 of accepted ones
//   finalize (next dt, step count, end of interval)
// Every launch covers all scenarios: the sweeps through tile work lists,
// the other kernels through block lists that lay the scenarios' own grids
// end to end (batch_ops.h), each scenario gated by its control block.  The
// number of launches per attempt does not depend on the number of
// scenarios; the source pass's substep loops run to the largest substep
// counts any scenario was started with (each scenario's own counts gate its
// part).  No host synchronisation inside an attempt; the host reads the
// control blocks back every few attempts to see whether all scenarios are
// done.
#include <algorithm>

#include "batch.h"
#include "batch_ops.h"
#include "ctx.h"

using namespace magic;

struct magic_batch {
    int n = 0;                           // slots in use (a removed scenario leaves an empty slot)
    int cap = 0;                         // slots the device arrays hold
    int meqn = 0, mbc = 0, maux = 0, form = 0;
    hipStream_t stream = nullptr;
    std::vector<magic_ctx *> sc;         // nullptr: empty slot
    std::vector<ScenCtl> ctl;            // host copy of the control blocks
    std::vector<int> mthbc;              // 6 per scenario
    ScenCtl *ctl_d = nullptr;
    ScenCtl *ctl_pin = nullptr;          // pinned read-back buffer
    void *args_d = nullptr;              // SweepArgsT of the form, 18 per slot: direction x colour
    const int **run_d = nullptr;         // per argument block: &ctl_d[s].run
    int4 *items_d = nullptr;
    int item_off[19] = {0};              // work list of direction d, colour k at [item_off[6d+k], ..)
    int *counters_d = nullptr;           // 18 pairs of work counters (tile_sweep_list)
    int shape[3] = {0, 0, 0};            // tile shape of the x-, y- and z-sweeps
    int grid[3] = {0, 0, 0};             // persistent grid size of each direction's kernel
    int *one_d = nullptr;                // a device 1: the run flag of every tile while tuning
    // boundary, copy and source-pass kernels (batch_ops.h)
    void *ops_d = nullptr;               // BatchOps of the form, one per slot
    int *loff_d = nullptr;               // NLG block lists: cap + 1 offsets each
    int4 *lgdim_d = nullptr;             //                  cap grid shapes each
    long long ltotal[NLG] = {0};         // blocks of each list
    int ndtr_cap = 0, ndtrc_cap = 0;     // substep loops of the captured attempt
    int opgrid[NOP] = {0};               // persistent grid of each kernel class
    // Sweep overlap, as for a single grid (off by default): the y-sweep and
    // the last three colour launches of the z-sweep add into each
    // scenario's dq (ctx dq[0]) on a second stream, the rest into q on the
    // main stream; each scenario then adds dq into q and clears it.  The
    // same order of additions as a single grid with the overlap on.  The
    // batch's persistent launches leave few compute units idle, and on the
    // 64-scenario ensemble the merge's traffic outweighs the gain (3 %
    // slower), hence the default.
    bool overlap = false;
    hipStream_t side = nullptr;
    hipEvent_t sw_fork = nullptr, sw_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;                   // control blocks changed on the host
    bool graph_ok = false;
};

static bool fp32(const magic_batch *b) { return b->form == MAGIC_FORM_FP32P; }
static size_t elt(const magic_batch *b) { return fp32(b) ? sizeof(float) : sizeof(double); }
static void *state_q(const magic_ctx *c) { return c->form == MAGIC_FORM_FP32P ? (void *)c->q32 : (void *)c->q; }
static void *state_qold(const magic_ctx *c)
{
    return c->form == MAGIC_FORM_FP32P ? (void *)c->qold32 : (void *)c->qold;
}

// ------------------------------------------------------- control kernels --
template <typename T>
__global__ void k_batch_prepare(ScenCtl *ctl, int n, SweepArgsT<T> *args, int *counters)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s < 36) counters[s] = 0;   // the sweeps' counter pairs (the kernel also resets them)
    if (s >= n) return;
    ScenCtl &c = ctl[s];
    c.accepted = 0;
    c.rejected = 0;
    c.run = c.active != 0 && c.status == 0;
    c.swept = c.run;
    if (!c.run) return;
    if (c.newstep) {                                   // claw3_mpi.f:741
        c.told = c.t;
        if (c.clip && c.told + c.dt > c.tend) c.dt = c.tend - c.told;
    }
    c.cfl_bits = 0ull;
    for (int e = 0; e < 6; e++) c.err[e] = 0;
    const double dt = c.dt;
    const double dtdx = dt / c.dx, dtdy = dt / c.dy, dtdz = dt / c.dz;
    for (int k = 0; k < 6; k++) {
        SweepArgsT<T> &ax = args[s * 18 + 0 * 6 + k], &ay = args[s * 18 + 1 * 6 + k], &az = args[s * 18 + 2 * 6 + k];
        ax.dtdn = dtdx; ax.fg = dtdy; ax.fh = dtdz;
        ay.dtdn = dtdy; ay.fg = dtdz; ay.fh = dtdx;
        az.dtdn = dtdz; az.fg = dtdx; az.fh = dtdy;
    }
    // src3.f90:314-322
    const int ndtr = (int)ceil((dt * c.difmax / (c.dxyzmin2)) / 0.45);
    const double dtr = dt / ndtr;
    // ndtr = 0 (no viscosity at the top): no substep runs (src3.f90:341)
    const int ndtrc = ndtr > 0 ? (int)ceil((dtr * c.kapmax / (c.dxyzmin2)) / 0.45) : 0;
    const double dtrc = dtr / ndtrc;
    if (ndtr > c.ndtr_cap || ndtrc > c.ndtrc_cap) {   // beyond the captured substeps
        c.status = MAGIC_EUNSUPPORTED;
        c.run = 0;
        c.swept = 0;
        return;
    }
    c.ndtr = ndtr;
    c.ndtrc = ndtrc;
    c.dtr = dtr;
    c.dtrc = dtrc;
    // factors of the diffusion passes in the order the host form evaluates them
    const double fourthirds = 4.0 / 3.0;
    double dtdx2 = (dtr) / (c.dx * c.dx), dtdy2 = (dtr) / (c.dy * c.dy), dtdz2 = (dtr) / (c.dz * c.dz);
    const double dtr32 = dtr / 6.0;
    c.fac[FAC_CROSS + 0] = dtr32 / (c.dx * c.dy);
    c.fac[FAC_CROSS + 1] = dtr32 / (c.dx * c.dz);
    c.fac[FAC_CROSS + 2] = dtr32 / (c.dy * c.dz);
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
        c.fac[FAC_DIFF + 3 * (cc - 1) + 0] = dtdx2;
        c.fac[FAC_DIFF + 3 * (cc - 1) + 1] = dtdy2;
        c.fac[FAC_DIFF + 3 * (cc - 1) + 2] = dtdz2;
    }
    c.fac[FAC_COND + 0] = (dtrc) / (c.dx * c.dx);
    c.fac[FAC_COND + 1] = (dtrc) / (c.dy * c.dy);
    c.fac[FAC_COND + 2] = (dtrc) / (c.dz * c.dz);
    c.tgauss = exp(-.5 * ((c.told - c.tcenter) * (c.told - c.tcenter) / (c.twidth * c.twidth)));
}

// claw3_mpi.f:843-886: accept or reject the attempt
__global__ void k_batch_decide(ScenCtl *ctl, int n)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= n) return;
    ScenCtl &c = ctl[s];
    if (!c.run) return;
    const double cfl = __longlong_as_double((long long)c.cfl_bits);
    c.cfl = cfl;
    c.attempts++;
    if (c.err[0] != 0) {
        c.status = c.err[0] == 1 ? MAGIC_ERHO : MAGIC_EA2;
        c.run = 0;
        c.active = 0;
        return;
    }
    if (cfl <= c.cflv1) {
        c.accepted = 1;
        c.t = c.told + c.dt;
        c.cflmax = c.cflmax > cfl ? c.cflmax : cfl;
    } else {
        c.rejected = 1;
        c.rejects++;
        c.t = c.told;
        c.dt = c.dtmax < c.dt * c.cflv2 / cfl ? c.dtmax : c.dt * c.cflv2 / cfl;
        c.newstep = 0;
    }
}

// claw3_mpi.f:910-927: next dt after an accepted step, end of the interval
__global__ void k_batch_finalize(ScenCtl *ctl, int n)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= n) return;
    ScenCtl &c = ctl[s];
    if (!c.accepted) return;
    if (c.cfl == 0.0) c.dt = c.dtmax;
    else c.dt = c.dtmax < c.dt * c.cflv2 / c.cfl ? c.dtmax : c.dt * c.cflv2 / c.cfl;
    c.nsteps++;
    c.newstep = 1;
    if (c.t >= c.tend || c.nsteps >= c.maxsteps) c.active = 0;
}

// --------------------------------------------------------------- helpers --
// Substep counts of the source pass for any dt up to dtbound
// (src3.f90:314-322 at the largest dt).
static void src_caps(const magic_ctx *c, double dtbound, int *ndtr, int *ndtrc)
{
    double dmin = c->dx < c->dy ? c->dx : c->dy;
    dmin = dmin < c->dz ? dmin : c->dz;
    const double d2 = dmin * dmin;
    const double difmax = (4.0 / 3.0) * c->aux11_top;
    *ndtr = (int)std::ceil((dtbound * difmax / d2) / 0.45);
    if (*ndtr < 1) *ndtr = 1;
    // the conduction count is largest for the largest dtr, i.e. dt / 1
    *ndtrc = (int)std::ceil((dtbound * c->aux12_top / d2) / 0.45);
    if (*ndtrc < 1) *ndtrc = 1;
}

// tracers 0, 1, 2 and 5 (the sweep kernels' instantiations)
template <typename T, bool PERT>
static const void *list_fn(const magic_batch *b, int dir)
{
    switch (b->meqn - 5) {
    case 0: return tile_list_fn<T, PERT, 0>(b->shape[dir]);
    case 1: return tile_list_fn<T, PERT, 1>(b->shape[dir]);
    case 2: return tile_list_fn<T, PERT, 2>(b->shape[dir]);
    case 3: return tile_list_fn<T, PERT, 3>(b->shape[dir]);
    case 4: return tile_list_fn<T, PERT, 4>(b->shape[dir]);
    case 5: return tile_list_fn<T, PERT, 5>(b->shape[dir]);
    default: return nullptr;
    }
}

// colour launches k0 .. k1-1 of direction dir on stream st
template <typename T, bool PERT>
static int launch_dir(magic_batch *b, int dir, hipStream_t st = nullptr, int k0 = 0, int k1 = 6)
{
    if (!st) st = b->stream;
    const void *fn = list_fn<T, PERT>(b, dir);
    if (!fn || b->grid[dir] < 1) return MAGIC_EARG;
    for (int k = k0; k < k1; k++) {
        const int lo = b->item_off[6 * dir + k];
        int n = b->item_off[6 * dir + k + 1] - lo;
        if (n <= 0) continue;
        const SweepArgsT<T> *args = (const SweepArgsT<T> *)b->args_d;
        const int4 *items = b->items_d + lo;
        const int *const *run = b->run_d;
        int *cnt = b->counters_d + 2 * (6 * dir + k);
        int4 grid = make_int4(0, 0, 0, 0);   // the work list
        void *kargs[] = {&args, &items, &n, &run, &cnt, &grid};
        HIPCHECK(hipLaunchKernel(fn, dim3(b->grid[dir]), dim3(256), kargs, 0, st));
        TRY(launch_check());
    }
    return MAGIC_OK;
}

template <typename T, bool PERT>
static int launch_sweeps(magic_batch *b)
{
    if (b->overlap) {
        HIPCHECK(hipEventRecord(b->sw_fork, b->stream));
        HIPCHECK(hipStreamWaitEvent(b->side, b->sw_fork, 0));
        TRY((launch_dir<T, PERT>(b, 1, b->side)));
        TRY((launch_dir<T, PERT>(b, 2, b->side, 3, 6)));
        TRY((launch_dir<T, PERT>(b, 0)));
        TRY((launch_dir<T, PERT>(b, 2, b->stream, 0, 3)));
        HIPCHECK(hipEventRecord(b->sw_join, b->side));
        HIPCHECK(hipStreamWaitEvent(b->stream, b->sw_join, 0));
        return MAGIC_OK;
    }
    for (int dir = 0; dir < 3; dir++) TRY((launch_dir<T, PERT>(b, dir)));
    return MAGIC_OK;
}

// Workgroup shape of the kernels that use block list lg
static dim3 list_block(int lg)
{
    switch (lg) {
    case LG_ALL: case LG_CPY: case LG_CP1: case LG_MRG: return dim3(256);
    case LG_PAD: case LG_DIF: case LG_INT: return cell_block();
    default: return dim3(64);
    }
}

// The block list each kernel class walks
static int op_list(int op)
{
    static const int lg[NOP] = {LG_FX,  LG_FY,  LG_BCZ, LG_ALL, LG_CPY, LG_CPY, LG_PAD, LG_FZ, LG_DIF,
                                LG_DIF, LG_CP1, LG_INT, LG_FZ,  LG_SPL, LG_INT, LG_INT, LG_MRG};
    return lg[op];
}

// One launch of kernel class op over its block list: a persistent grid of
// at most as many workgroups as fit on the chip at once
template <typename T>
static int launch_op(magic_batch *b, int op, GateSpec gs, int arg = 0)
{
    const int lg = op_list(op);
    const int total = (int)b->ltotal[lg];
    if (total <= 0) return MAGIC_OK;
    const void *fn = batch_op_fn<T>(op);
    if (!fn || b->opgrid[op] < 1) return MAGIC_EARG;
    const BatchOps<T> *ops = (const BatchOps<T> *)b->ops_d;
    const int *off = b->loff_d + (size_t)lg * (b->cap + 1);
    const int4 *gd = b->lgdim_d + (size_t)lg * b->cap;
    int n = b->n;
    void *kargs[] = {&ops, &off, &gd, &n, &gs, &arg};
    HIPCHECK(hipLaunchKernel(fn, dim3(std::min(total, b->opgrid[op])), list_block(lg), kargs, 0, b->stream));
    return launch_check();
}

static GateSpec gs(int mode, int nn = 0, int kk = 0, int fslot = -1) { return GateSpec{mode, nn, kk, fslot}; }

// src3 of the accepted scenarios: src3_passes32 / src3_passes64's sequence
// with the substep loops run to the captured caps, each scenario gated by
// its own counts (src3.f90:203-634, diffuse.f:37-210, conduct.f:245-332)
template <typename T>
static int launch_src(magic_batch *b)
{
    TRY(launch_op<T>(b, OP_CONVERT, gs(GATE_ACCEPTED)));
    for (int nn = 1; nn <= b->ndtr_cap; nn++) {
        const GateSpec gd = gs(GATE_DIFF, nn);
        // bc3 with the z faces skipped (src3.f90:333-336)
        TRY(launch_op<T>(b, OP_BCX, gd));
        TRY(launch_op<T>(b, OP_BCY, gd));
        TRY(launch_op<T>(b, OP_ZGHOST_UVW, gd));
        TRY(launch_op<T>(b, OP_DIFF_CROSS, gs(GATE_DIFF, nn, 0, FAC_CROSS)));
        for (int cc = 1; cc <= 3; cc++)
            TRY(launch_op<T>(b, OP_DIFF_XYZ, gs(GATE_DIFF, nn, 0, FAC_DIFF + 3 * (cc - 1)), cc));
        TRY(launch_op<T>(b, OP_COPY_Q5, gd));
        for (int kk = 1; kk <= b->ndtrc_cap; kk++) {
            const GateSpec gk = gs(GATE_COND, nn, kk);
            TRY(launch_op<T>(b, OP_PASS, gs(GATE_COND, nn, kk, FAC_COND + 0), PASS_QT_UT * 4 + 0));
            TRY(launch_op<T>(b, OP_PASS, gs(GATE_COND, nn, kk, FAC_COND + 1), PASS_UT_QT * 4 + 1));
            TRY(launch_op<T>(b, OP_CONDUCT_ZGHOST, gk));
            TRY(launch_op<T>(b, OP_ZSPILL, gk));
            if (kk < b->ndtrc_cap) {   // at kk = ndtrc_cap >= ndtrc these never run
                TRY(launch_op<T>(b, OP_PASS, gs(GATE_COND_MORE, nn, kk, FAC_COND + 2), PASS_QT_UT * 4 + 2));
                TRY(launch_op<T>(b, OP_COPY_INTERIOR, gs(GATE_COND_MORE, nn, kk)));
            }
            TRY(launch_op<T>(b, OP_PASS, gs(GATE_COND_LAST, nn, kk, FAC_COND + 2), PASS_Q5_UT * 4 + 2));
        }
    }
    return launch_op<T>(b, OP_FINISH, gs(GATE_ACCEPTED));
}

// one attempt of every scenario, as launched into the capture
template <typename T, bool PERT>
static int record_attempt(magic_batch *b)
{
    const unsigned nb = blocks(b->n > 18 ? b->n : 18, 256);
    k_batch_prepare<T><<<nb, 256, 0, b->stream>>>(b->ctl_d, b->n, (SweepArgsT<T> *)b->args_d, b->counters_d);
    TRY(launch_check());
    const GateSpec run = gs(GATE_RUN);
    TRY(launch_op<T>(b, OP_BCX, run));
    TRY(launch_op<T>(b, OP_BCY, run));
    TRY(launch_op<T>(b, OP_BCZ, run));
    if (b->meqn > 5) TRY(launch_op<T>(b, OP_CLAMP, run));
    TRY(launch_op<T>(b, OP_BACKUP, run));
    TRY((launch_sweeps<T, PERT>(b)));
    k_batch_decide<<<blocks(b->n, 256), 256, 0, b->stream>>>(b->ctl_d, b->n);
    TRY(launch_check());
    if (b->overlap) TRY(launch_op<T>(b, OP_MERGE, gs(GATE_SWEPT)));
    TRY(launch_op<T>(b, OP_RESTORE, gs(GATE_REJECTED)));
    TRY(launch_src<T>(b));
    k_batch_finalize<<<blocks(b->n, 256), 256, 0, b->stream>>>(b->ctl_d, b->n);
    return launch_check();
}

// Form-specific part of a scenario's sweep argument block
static void sweep_form_args(const magic_ctx *c, SweepArgsT<float> &a)
{
    const int NZ = c->d.NZ;
    a.qold = c->qold32;
    a.qnew = c->q32;
    a.crho = c->bg64 + B64_RHO * NZ;
    a.cE = c->bg64 + B64_E * NZ;
    a.cgam = c->bg64 + B64_GAM * NZ;
    a.cvg = c->bg64 + B64_VG * NZ;
    a.cres = c->bg64 + B64_RES * NZ;
    a.sink = (float *)c->sink;
}

static void sweep_form_args(const magic_ctx *c, SweepArgsT<double> &a)
{
    a.qold = c->qold;
    a.qnew = c->q;
    a.gam = c->aux + AX_GAM * c->P;
    a.vgrav = c->aux + AX_VG * c->P;
    a.sink = (double *)c->sink;
}

// Sweep argument blocks and tile work lists (geometry only; the step's dt
// factors are written by k_batch_prepare).  For the autotuner (tuning) every
// tile runs and the dt factors come from each scenario's next dt.
template <typename T>
static int build_lists(magic_batch *b, bool tuning = false)
{
    std::vector<SweepArgsT<T>> args((size_t)b->n * 18);
    std::vector<const int *> run((size_t)b->n * 18);
    std::vector<std::vector<int4>> lists(18);
    // tiles of longer pencils first (they take longest)
    std::vector<int> order;
    for (int s = 0; s < b->n; s++)
        if (b->sc[s]) order.push_back(s);
    for (int dir = 0; dir < 3; dir++) {
        auto len = [&](int s) { return dir == 0 ? b->sc[s]->mx : dir == 1 ? b->sc[s]->my : b->sc[s]->mz; };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return len(x) > len(y); });
        for (int s : order) {
            magic_ctx *c = b->sc[s];
            const Dims &d = c->d;
            const long long NXY = (long long)d.NX * d.NY;
            SweepArgsT<T> a;
            memset(&a, 0, sizeof(a));
            sweep_form_args(c, a);
            a.P = c->P;
            a.mbc = c->mbc;
            a.nspecadv = c->have_params ? c->prm.nspecadv : 0;
            a.cfl_bits = &b->ctl_d[s].cfl_bits;
            a.status = &b->ctl_d[s].err[0];
            for (int m = 5; m < c->meqn; m++) a.comp[m] = m;
            int a_is_g;
            const double dt = b->ctl[s].dt > 0.0 ? b->ctl[s].dt : 1.0;
            if (dir == 0) {          // step3.f:121-280 (g = y, h = z), tile axis y = g
                a.comp[0] = 0; a.comp[1] = 1; a.comp[2] = 2; a.comp[3] = 3; a.comp[4] = 4;
                a.n = c->mx; a.nA = c->my; a.nB = c->mz;
                a.sN = 1; a.sA = d.NX; a.sB = NXY;
                a.gravity = 0; a.sweep = 1;
                a_is_g = 1;
                if (tuning) { a.dtdn = dt / c->dx; a.fg = dt / c->dy; a.fh = dt / c->dz; }
            } else if (dir == 1) {   // step3.f:289-450 (g = z, h = x), tile axis x = h
                a.comp[0] = 0; a.comp[1] = 2; a.comp[2] = 3; a.comp[3] = 1; a.comp[4] = 4;
                a.n = c->my; a.nA = c->mx; a.nB = c->mz;
                a.sN = d.NX; a.sA = 1; a.sB = NXY;
                a.gravity = 0; a.sweep = 2;
                a_is_g = 0;
                if (tuning) { a.dtdn = dt / c->dy; a.fg = dt / c->dz; a.fh = dt / c->dx; }
            } else {                 // step3.f:459-621 (g = x, h = y), tile axis x = g
                a.comp[0] = 0; a.comp[1] = 3; a.comp[2] = 1; a.comp[3] = 2; a.comp[4] = 4;
                a.n = c->mz; a.nA = c->mx; a.nB = c->my;
                a.sN = NXY; a.sA = 1; a.sB = d.NX;
                a.gravity = 1; a.sweep = 3;
                a_is_g = 1;
                if (tuning) { a.dtdn = dt / c->dz; a.fg = dt / c->dx; a.fh = dt / c->dy; }
            }
            const int TA = tile_ta(b->shape[dir]);
            const int ntiles = (a.nA + 2 + TA - 1) / TA;
            for (int ca = 0; ca < 2; ca++)
                for (int cb = 0; cb < 3; cb++) {
                    const int k = ca * 3 + cb, ai = s * 18 + dir * 6 + k;
                    SweepArgsT<T> t = a;
                    t.cA = ca;
                    t.cB = cb;
                    if (!tuning && b->overlap && (dir == 1 || (dir == 2 && k >= 3))) t.qnew = (T *)c->dq[0];
                    args[ai] = t;
                    run[ai] = tuning ? b->one_d : &b->ctl_d[s].run;
                    const int gx = (ntiles - ca + 1) / 2, gy = (a.nB + 1 - cb) / 3 + 1;
                    for (int by = 0; by < gy; by++)
                        for (int bx = 0; bx < gx; bx++) lists[dir * 6 + k].push_back(make_int4(ai, a_is_g, bx, by));
                }
        }
    }
    size_t total = 0;
    for (int k = 0; k < 18; k++) {
        b->item_off[k] = (int)total;
        total += lists[k].size();
    }
    if (total > 0x7fffffff) return MAGIC_EUNSUPPORTED;
    b->item_off[18] = (int)total;
    std::vector<int4> items;
    items.reserve(total);
    for (auto &l : lists) items.insert(items.end(), l.begin(), l.end());
    hipFree(b->items_d);
    b->items_d = nullptr;
    HIPCHECK(hipMalloc(&b->items_d, (total ? total : 1) * sizeof(int4)));
    HIPCHECK(hipMemcpyAsync(b->items_d, items.data(), total * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(b->args_d, args.data(), args.size() * sizeof(args[0]), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(b->run_d, run.data(), run.size() * sizeof(run[0]), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    return MAGIC_OK;
}

// Form-specific operands of a scenario's boundary and source-pass kernels
static void form_ops(const magic_ctx *c, BatchOps<float> &o)
{
    const int NZ = c->d.NZ;
    o.q = c->q32; o.qold = c->qold32; o.u = c->u32; o.qt = c->qt32; o.ut = c->ut32;
    o.col = col32_of(c);
    o.zr = c->zr32;
    o.nu32 = c->col32 + C32_NU * NZ;
    o.kap32 = c->col32 + C32_KAP * NZ;
}

static void form_ops(const magic_ctx *c, BatchOps<double> &o)
{
    const int NZ = c->d.NZ;
    const long long P = c->P;
    o.q = c->q; o.qold = c->qold; o.u = c->u; o.qt = c->qt; o.ut = c->ut;
    const double *A = c->col64;   // the background's columns (magic_batch_start checks they exist)
    o.ax = {A + AX_RHO * NZ, A + AX_MX * NZ, A + AX_MY * NZ, A + AX_MZ * NZ, A + AX_E * NZ, A + AX_R * NZ,
            A + AX_GAM * NZ};
    o.nu64 = A + AX_NU * NZ;
    o.kap64 = A + AX_KAP * NZ;
    o.aux1 = c->aux + AX_RHO * P;
    o.aux2 = c->aux + AX_MX * P;
    o.aux3 = c->aux + AX_MY * P;
    o.aux5 = c->aux + AX_E * P;
    o.zratio = c->zratio;
}

// Layers per workgroup of a scenario's diffusion passes.  A workgroup
// recomputes the x and y passes of the two layers around its own; one grid
// takes few layers (diff_kb) to have enough workgroups, a batch has enough
// from its scenarios and takes at least 16.  The result does not depend on
// the layer count.
static int batch_kb(const Dims &d) { return std::max(diff_kb(d), std::min(16, d.mz)); }

// Grid of list lg for one scenario (x, y, z blocks; x = 0: no blocks)
static int4 list_grid(const magic_batch *b, const magic_ctx *c, const int *bc, int lg)
{
    const Dims &d = c->d;
    const long long P = c->P;
    auto g1 = [](long long x) { return make_int4((int)x, 1, 1, 0); };
    switch (lg) {
    case LG_FX: return bc[0] >= 0 || bc[1] >= 0 ? make_int4(blocks(d.NY, 64), d.NZ, 1, 0) : g1(0);
    case LG_FY: return bc[2] >= 0 || bc[3] >= 0 ? make_int4(blocks(d.NX, 64), d.NZ, 1, 0) : g1(0);
    case LG_BCZ: return bc[4] >= 0 || bc[5] >= 0 ? make_int4(blocks(d.NX, 64), d.NY, 1, 0) : g1(0);
    case LG_FZ: return make_int4(blocks(d.NX, 64), d.NY, 1, 0);
    case LG_ALL: return g1(blocks(P, 256));
    case LG_CPY: return g1(std::min<long long>(8192, blocks(P * b->meqn, 256)));
    case LG_CP1: return g1(std::min<long long>(8192, blocks(P, 256)));
    case LG_PAD: { const dim3 g = padded_cells(d); return make_int4(g.x, g.y, g.z, 0); }
    case LG_DIF: return make_int4(blocks(d.mx, 64), blocks(d.my, 4), blocks(d.mz, batch_kb(d)), 0);
    case LG_INT: { const dim3 g = interior_grid(d); return make_int4(g.x, g.y, g.z, 0); }
    case LG_SPL: return g1(c->mz > c->mx ? blocks(c->mz - c->mx, 64) : 0);
    case LG_MRG: {
        const long long nv = P * b->meqn / (long long)(16 / elt(b)) + 1;
        return g1(std::min<long long>(8192, blocks(nv, 256)));
    }
    default: return g1(0);
    }
}

// Operand blocks and block lists of the boundary, copy and source-pass
// kernels, and the substep loops the capture records
template <typename T>
static int build_ops(magic_batch *b)
{
    std::vector<BatchOps<T>> ops(b->n);
    std::vector<int> off((size_t)NLG * (b->cap + 1), 0);
    std::vector<int4> gdim((size_t)NLG * b->cap, make_int4(0, 1, 1, 0));
    b->ndtr_cap = 0;
    b->ndtrc_cap = 0;
    for (int lg = 0; lg < NLG; lg++) {
        long long total = 0;
        for (int s = 0; s < b->n; s++) {
            off[(size_t)lg * (b->cap + 1) + s] = (int)total;
            magic_ctx *c = b->sc[s];
            if (!c) continue;
            const int4 g = list_grid(b, c, &b->mthbc[6 * s], lg);
            gdim[(size_t)lg * b->cap + s] = g;
            total += (long long)g.x * g.y * g.z;
            if (total > 0x7fffffff) return MAGIC_EUNSUPPORTED;
        }
        off[(size_t)lg * (b->cap + 1) + b->n] = (int)total;
        b->ltotal[lg] = total;
    }
    for (int s = 0; s < b->n; s++) {
        magic_ctx *c = b->sc[s];
        BatchOps<T> &o = ops[s];
        memset(&o, 0, sizeof(o));
        o.ctl = b->ctl_d + s;   // an empty slot's control block is all zeros: every gate off
        if (!c) continue;
        o.d = c->d;
        form_ops(c, o);
        o.dq = (T *)c->dq[0];
        o.nc = c->meqn;
        o.kb = batch_kb(c->d);
        o.spill = c->mz > c->mx ? c->mz - c->mx : 0;
        memcpy(o.bc, &b->mthbc[6 * s], sizeof(o.bc));
        const magic_params &p = c->prm;
        Forcing &f = o.f;
        f.forcemth = (int)p.forcemth;
        f.t = 0.0; f.dt = 0.0; f.tgauss = 0.0;   // from the control block at run time
        f.xlower = c->xlower; f.ylower = c->ylower; f.zlower = c->zlower;
        f.dx = c->dx; f.dy = c->dy; f.dz = c->dz;
        f.amplitude = p.amplitude; f.xpos = p.xpos; f.ypos = p.ypos; f.zpos = p.zpos;
        f.xwidth = p.xwidth; f.ywidth = p.ywidth; f.zwidth = p.zwidth;
        f.tcenter = p.tcenter; f.twidth = p.twidth;
        f.omega = p.omega; f.propx = p.propx; f.propz = p.propz; f.vsrcx = p.vsrcx;
        b->ndtr_cap = std::max(b->ndtr_cap, b->ctl[s].ndtr_cap);
        b->ndtrc_cap = std::max(b->ndtrc_cap, b->ctl[s].ndtrc_cap);
    }
    HIPCHECK(hipMemcpyAsync(b->ops_d, ops.data(), ops.size() * sizeof(ops[0]), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(b->loff_d, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(b->lgdim_d, gdim.data(), gdim.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    return MAGIC_OK;
}

// persistent grid of each sweep direction and of each kernel class of the
// work-list launches: as many workgroups as fit on the chip at once
template <typename T, bool PERT>
static int list_grids(magic_batch *b)
{
    int dev = 0, ncu = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    for (int dir = 0; dir < 3; dir++) {
        const void *fn = list_fn<T, PERT>(b, dir);
        if (!fn) return MAGIC_EUNSUPPORTED;
        int occ = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, 0));
        b->grid[dir] = ncu * (occ > 0 ? occ : 1);
    }
    for (int op = 0; op < NOP; op++) {
        const void *fn = batch_op_fn<T>(op);
        if (!fn) return MAGIC_EUNSUPPORTED;
        const dim3 blk = list_block(op_list(op));
        int occ = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, (int)(blk.x * blk.y * blk.z), 0));
        b->opgrid[op] = ncu * (occ > 0 ? occ : 1);
    }
    return MAGIC_OK;
}

// Retire the captured attempt: wait until no replay runs (the main stream
// and the sweep overlap's stream), then destroy the executable graph and
// the graph itself.  Called before any device buffer its kernels address
// (control blocks, argument blocks, work and block lists, a scenario's
// state) is freed or reallocated, so that no instantiated graph ever holds
// a dangling pointer, and before every capture.
static int drop_graph(magic_batch *b)
{
    int st = MAGIC_OK;
    if (b->stream && hipStreamSynchronize(b->stream) != hipSuccess) st = MAGIC_EHIP;
    if (b->side && hipStreamSynchronize(b->side) != hipSuccess) st = MAGIC_EHIP;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_ok = false;
    return st;
}

template <typename T, bool PERT>
static int build_graph_t(magic_batch *b)
{
    TRY(drop_graph(b));
    TRY((list_grids<T, PERT>(b)));
    if (b->overlap)
        for (magic_ctx *c : b->sc) {
            if (!c || c->dq[0]) continue;
            const size_t bytes = (size_t)c->P * c->meqn * sizeof(T);
            if (hipMalloc(&c->dq[0], bytes) != hipSuccess) {
                c->dq[0] = nullptr;
                return MAGIC_ENOMEM;
            }
            HIPCHECK(hipMemsetAsync(c->dq[0], 0, bytes, b->stream));   // before the capture
        }
    TRY(build_lists<T>(b));
    TRY(build_ops<T>(b));
    HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = record_attempt<T, PERT>(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != MAGIC_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    HIPCHECK(e);
    b->graph = g;
    HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_ok = true;
    return MAGIC_OK;
}

static int build_graph(magic_batch *b)
{
    return fp32(b) ? build_graph_t<float, true>(b) : build_graph_t<double, false>(b);
}

// ------------------------------------------------------------------ ABI --
// device arrays for at least want slots (contents are rebuilt from the host
// copies before the next attempt)
static int reserve(magic_batch *b, int want)
{
    if (want <= b->cap) return MAGIC_OK;
    int cap = b->cap ? b->cap : 8;
    while (cap < want) cap *= 2;
    TRY(drop_graph(b));
    hipFree(b->ctl_d); hipFree(b->args_d); hipFree(b->run_d);
    hipFree(b->ops_d); hipFree(b->loff_d); hipFree(b->lgdim_d);
    if (b->ctl_pin) hipHostFree(b->ctl_pin);
    b->ctl_d = nullptr; b->args_d = nullptr; b->run_d = nullptr; b->ctl_pin = nullptr;
    b->ops_d = nullptr; b->loff_d = nullptr; b->lgdim_d = nullptr;
    b->cap = 0;
    const size_t args_bytes = fp32(b) ? sizeof(SweepArgsT<float>) : sizeof(SweepArgsT<double>);
    const size_t ops_bytes = fp32(b) ? sizeof(BatchOps<float>) : sizeof(BatchOps<double>);
    if (hipMalloc(&b->ctl_d, cap * sizeof(ScenCtl)) != hipSuccess ||
        hipHostMalloc(&b->ctl_pin, cap * sizeof(ScenCtl), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&b->args_d, (size_t)cap * 18 * args_bytes) != hipSuccess ||
        hipMalloc(&b->run_d, (size_t)cap * 18 * sizeof(int *)) != hipSuccess ||
        hipMalloc(&b->ops_d, (size_t)cap * ops_bytes) != hipSuccess ||
        hipMalloc(&b->loff_d, (size_t)NLG * (cap + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&b->lgdim_d, (size_t)NLG * cap * sizeof(int4)) != hipSuccess)
        return MAGIC_ENOMEM;
    b->cap = cap;
    b->graph_ok = false;
    b->dirty = true;
    return MAGIC_OK;
}

extern "C" int magic_batch_add(magic_batch *b, const int dims[3], const double geom[6], int *slot)
{
    if (!b || !dims || !geom || !slot) return MAGIC_EARG;
    magic_ctx *c = nullptr;
    TRY(magic_create_form(dims[0], dims[1], dims[2], b->meqn, b->mbc, b->maux, geom[0], geom[1], geom[2],
                          geom[3], geom[4], geom[5], b->form, &c));
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
    c->stream = b->stream;
    c->own_stream = false;
    int s = 0;
    while (s < b->n && b->sc[s]) s++;
    if (s == BATCH_MAX_SLOTS) {
        magic_destroy(c);
        return MAGIC_EUNSUPPORTED;
    }
    if (s == b->n) {
        const int st = reserve(b, b->n + 1);
        if (st != MAGIC_OK) {
            magic_destroy(c);
            return st;
        }
        b->n++;
        b->sc.push_back(nullptr);
        b->ctl.push_back(ScenCtl());
        b->mthbc.insert(b->mthbc.end(), 6, 1);
    }
    b->sc[s] = c;
    memset(&b->ctl[s], 0, sizeof(ScenCtl));   // inactive until started
    b->graph_ok = false;
    b->dirty = true;
    *slot = s;
    return MAGIC_OK;
}

extern "C" int magic_batch_remove(magic_batch *b, int s)
{
    if (!b || s < 0 || s >= b->n || !b->sc[s]) return MAGIC_EARG;
    TRY(drop_graph(b));
    magic_destroy(b->sc[s]);
    b->sc[s] = nullptr;
    memset(&b->ctl[s], 0, sizeof(ScenCtl));
    b->graph_ok = false;
    b->dirty = true;
    return MAGIC_OK;
}

extern "C" int magic_batch_create(int nscen, const int *dims, int meqn, int mbc, int maux, const double *geom,
                                  int form, magic_batch **out)
{
    if (!out || nscen < 0 || (nscen > 0 && (!dims || !geom))) return MAGIC_EARG;
    *out = nullptr;
    if (form != MAGIC_FORM_FP32P && form != MAGIC_FORM_FP64) return MAGIC_EUNSUPPORTED;
    if (mbc < 2 || mbc > 8 || meqn < 5 || meqn > 10 || maux < 19) return MAGIC_EARG;
    if (!sweep_supported(meqn - 5)) return MAGIC_EUNSUPPORTED;
    magic_batch *b = new magic_batch();
    b->meqn = meqn;
    b->mbc = mbc;
    b->maux = maux;
    b->form = form;
    int st = MAGIC_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess)
        st = MAGIC_EHIP;
    if (st == MAGIC_OK && (hipEventCreateWithFlags(&b->sw_fork, hipEventDisableTiming) != hipSuccess ||
                           hipEventCreateWithFlags(&b->sw_join, hipEventDisableTiming) != hipSuccess))
        st = MAGIC_EHIP;
    if (st == MAGIC_OK) st = reserve(b, nscen > 0 ? nscen : 1);
    if (st == MAGIC_OK) st = hipMalloc(&b->counters_d, 36 * sizeof(int)) == hipSuccess ? MAGIC_OK : MAGIC_ENOMEM;
    if (st == MAGIC_OK) st = hipMalloc(&b->one_d, sizeof(int)) == hipSuccess ? MAGIC_OK : MAGIC_ENOMEM;
    if (st == MAGIC_OK) {
        const int one = 1;
        st = hipMemcpy(b->one_d, &one, sizeof(int), hipMemcpyHostToDevice) == hipSuccess ? MAGIC_OK : MAGIC_EHIP;
    }
    for (int s = 0; s < nscen && st == MAGIC_OK; s++) {
        int slot;
        st = magic_batch_add(b, dims + 3 * s, geom + 6 * s, &slot);
    }
    if (st != MAGIC_OK) {
        magic_batch_destroy(b);
        return st;
    }
    *out = b;
    return MAGIC_OK;
}

extern "C" int magic_batch_destroy(magic_batch *b)
{
    if (!b) return MAGIC_OK;
    drop_graph(b);
    for (auto c : b->sc)
        if (c) magic_destroy(c);
    hipFree(b->ctl_d); hipFree(b->args_d); hipFree(b->run_d); hipFree(b->items_d); hipFree(b->counters_d);
    hipFree(b->one_d); hipFree(b->ops_d); hipFree(b->loff_d); hipFree(b->lgdim_d);
    if (b->ctl_pin) hipHostFree(b->ctl_pin);
    if (b->sw_fork) hipEventDestroy(b->sw_fork);
    if (b->sw_join) hipEventDestroy(b->sw_join);
    if (b->side) hipStreamDestroy(b->side);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return MAGIC_OK;
}

extern "C" int magic_batch_ctx(magic_batch *b, int s, magic_ctx **ctx)
{
    if (!b || !ctx || s < 0 || s >= b->n || !b->sc[s]) return MAGIC_EARG;
    *ctx = b->sc[s];
    b->graph_ok = false;   // the caller may change parameters, aux or q
    return MAGIC_OK;
}

extern "C" int magic_batch_start(magic_batch *b, int s, double t, double tend, double dt, const double dtv[5],
                                 const double cflv[4], int maxsteps, const int mthbc[6])
{
    if (!b || s < 0 || s >= b->n || !b->sc[s] || !dtv || !cflv || !mthbc || maxsteps < 1) return MAGIC_EARG;
    magic_ctx *c = b->sc[s];
    if (!c->have_aux || !c->have_params) return MAGIC_EARG;
    if (cflv[1] > cflv[0]) return MAGIC_EARG;
    for (int f = 0; f < 6; f++)
        if (mthbc[f] == 0 || mthbc[f] >= 4) return MAGIC_EBC;
    const magic_params &p = c->prm;
    if (p.omega1 != 0.0 || !(p.forcemth == 0.0 || p.forcemth == 1.0 || p.forcemth == 3.0))
        return MAGIC_EUNSUPPORTED;
    // the FP64 form's batched source pass reads the background as columns
    if (!fp32(b) && !c->aux_layered) return MAGIC_EUNSUPPORTED;
    ScenCtl &h = b->ctl[s];
    if (memcmp(&b->mthbc[6 * s], mthbc, 6 * sizeof(int)) != 0) b->graph_ok = false;
    memcpy(&b->mthbc[6 * s], mthbc, 6 * sizeof(int));
    h.t = t;
    h.told = t;
    h.tend = tend;
    h.clip = t < tend;
    h.dt = dt;
    h.dtmax = dtv[1];
    h.cflv1 = cflv[0];
    h.cflv2 = cflv[1];
    h.cfl = 0.0;
    h.cflmax = 0.0;
    h.active = 1;
    h.newstep = 1;
    h.nsteps = 0;
    h.attempts = 0;
    h.rejects = 0;
    h.maxsteps = maxsteps;
    h.status = 0;
    h.dx = c->dx; h.dy = c->dy; h.dz = c->dz;
    double dmin = c->dx < c->dy ? c->dx : c->dy;
    dmin = dmin < c->dz ? dmin : c->dz;
    h.dxyzmin2 = dmin * dmin;
    h.difmax = (4.0 / 3.0) * c->aux11_top;
    h.kapmax = c->aux12_top;
    h.tcenter = p.tcenter;
    h.twidth = p.twidth;
    int a, bb;
    src_caps(c, dt > dtv[1] ? dt : dtv[1], &a, &bb);
    // the captured substep loops cover every count up to the batch's largest
    if (a > b->ndtr_cap || bb > b->ndtrc_cap) b->graph_ok = false;
    h.ndtr_cap = a;
    h.ndtrc_cap = bb;
    b->dirty = true;
    return MAGIC_OK;
}

extern "C" int magic_batch_run(magic_batch *b, int max_attempts, int check_every, int *attempts_done)
{
    if (!b || max_attempts < 1 || check_every < 1) return MAGIC_EARG;
    if (attempts_done) *attempts_done = 0;
    bool started = false;
    for (int s = 0; s < b->n; s++) started = started || (b->sc[s] && b->ctl[s].active && b->ctl[s].status == 0);
    if (!started) return MAGIC_OK;
    if (!b->graph_ok) {
        // the captured kernels read the control blocks; upload before building
        HIPCHECK(hipMemcpyAsync(b->ctl_d, b->ctl.data(), b->n * sizeof(ScenCtl), hipMemcpyHostToDevice, b->stream));
        TRY(build_graph(b));
        b->dirty = false;
    }
    if (b->dirty) {
        HIPCHECK(hipMemcpyAsync(b->ctl_d, b->ctl.data(), b->n * sizeof(ScenCtl), hipMemcpyHostToDevice, b->stream));
        b->dirty = false;
    }
    int done = 0;
    bool any = true;
    while (done < max_attempts && any) {
        const int k = std::min(check_every, max_attempts - done);
        for (int i = 0; i < k; i++) HIPCHECK(hipGraphLaunch(b->exec, b->stream));
        done += k;
        HIPCHECK(hipMemcpyAsync(b->ctl_pin, b->ctl_d, b->n * sizeof(ScenCtl), hipMemcpyDeviceToHost, b->stream));
        HIPCHECK(hipStreamSynchronize(b->stream));
        memcpy(b->ctl.data(), b->ctl_pin, b->n * sizeof(ScenCtl));
        any = false;
        for (int s = 0; s < b->n; s++) any = any || (b->ctl[s].active && b->ctl[s].status == 0);
    }
    if (attempts_done) *attempts_done = done;
    // a scenario that failed (rho <= 0, a**2 <= 0, more substeps than
    // captured) has stopped with its status in its control block, which
    // magic_batch_state reports; the others went on
    return MAGIC_OK;
}

extern "C" int magic_batch_state(magic_batch *b, int s, double real_out[5], int int_out[6])
{
    if (!b || s < 0 || s >= b->n || !b->sc[s]) return MAGIC_EARG;
    const ScenCtl &h = b->ctl[s];
    if (real_out) {
        real_out[0] = h.t;
        real_out[1] = h.dt;
        real_out[2] = h.cfl;
        real_out[3] = h.cflmax;
        real_out[4] = h.tend;
    }
    if (int_out) {
        int_out[0] = h.nsteps;
        int_out[1] = h.attempts;
        int_out[2] = h.rejects;
        int_out[3] = h.active;
        int_out[4] = h.status;
        int_out[5] = h.ndtr;
    }
    return MAGIC_OK;
}

extern "C" int magic_batch_set_tile_shapes(magic_batch *b, const int shapes[3])
{
    if (!b || !shapes || !shapes_valid(shapes)) return MAGIC_EARG;
    for (int k = 0; k < 3; k++) b->shape[k] = shapes[k];
    b->graph_ok = false;
    return MAGIC_OK;
}

extern "C" int magic_batch_set_sweep_overlap(magic_batch *b, int on)
{
    if (!b) return MAGIC_EARG;
    b->overlap = on != 0;
    b->graph_ok = false;
    return MAGIC_OK;
}

extern "C" int magic_batch_get_tile_shapes(magic_batch *b, int shapes[3])
{
    if (!b || !shapes) return MAGIC_EARG;
    for (int k = 0; k < 3; k++) shapes[k] = b->shape[k];
    return MAGIC_OK;
}

// Every candidate reads the backup planes and adds into q (one list build
// per shape and round, all directions); q is restored from the backup
// afterwards.  The candidates' Courant numbers and reports land in the
// control blocks' reduction fields, which the next attempt's prepare clears.
template <typename T, bool PERT>
static int tune_rounds(magic_batch *b, int reps, TuneClock &clk)
{
    for (int r = 0; r <= reps; r++) {   // repetition 0 warms up and is not timed
        for (int sh = 0; sh < TILE_NSHAPE; sh++) {
            for (int k = 0; k < 3; k++) b->shape[k] = sh;
            TRY((list_grids<T, PERT>(b)));
            TRY(build_lists<T>(b, true));
            for (int dir = 0; dir < 3; dir++) {
                TRY(clk.mark(dir, sh, 0, b->stream));
                HIPCHECK(hipMemsetAsync(b->counters_d, 0, 36 * sizeof(int), b->stream));
                TRY((launch_dir<T, PERT>(b, dir)));
                TRY(clk.mark(dir, sh, 1, b->stream));
            }
        }
        if (r > 0) TRY(clk.add_round());
    }
    return MAGIC_OK;
}

extern "C" int magic_batch_tune(magic_batch *b, int reps, int shapes_out[3], float *ms_out)
{
    if (!b || reps < 1) return MAGIC_EARG;
    bool any = false;
    for (int s = 0; s < b->n; s++) any = any || b->sc[s] != nullptr;
    if (!any) return MAGIC_EARG;
    TuneClock clk;
    if (!clk.ok) return MAGIC_EHIP;
    const int saved[3] = {b->shape[0], b->shape[1], b->shape[2]};
    TRY(drop_graph(b));   // the lists and argument blocks are rebuilt before the next attempt
    for (int s = 0; s < b->n; s++) {
        magic_ctx *c = b->sc[s];
        if (c)
            HIPCHECK(hipMemcpyAsync(state_qold(c), state_q(c), (size_t)c->P * c->meqn * elt(b),
                                    hipMemcpyDeviceToDevice, b->stream));
    }
    int st = fp32(b) ? tune_rounds<float, true>(b, reps, clk) : tune_rounds<double, false>(b, reps, clk);
    for (int s = 0; s < b->n; s++) {   // also after a failed round
        magic_ctx *c = b->sc[s];
        if (c && hipMemcpyAsync(state_q(c), state_qold(c), (size_t)c->P * c->meqn * elt(b), hipMemcpyDeviceToDevice,
                                b->stream) != hipSuccess)
            st = MAGIC_EHIP;
    }
    if (st == MAGIC_OK && hipStreamSynchronize(b->stream) != hipSuccess) st = MAGIC_EHIP;
    if (st != MAGIC_OK) {
        for (int k = 0; k < 3; k++) b->shape[k] = saved[k];
        return st;
    }
    int best[3];
    clk.result(reps, best, ms_out);
    for (int k = 0; k < 3; k++) {
        b->shape[k] = best[k];
        if (shapes_out) shapes_out[k] = best[k];
    }
    return MAGIC_OK;
}
