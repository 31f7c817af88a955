/*
 * magic_hip.h -- C ABI of the MI355X-native MAGIC wave-propagation step.
 *
 * The reference drives its hot path through Fortran operator calls from the
 * time integrator claw3 (mpilib/claw3_mpi.f:765-905):
 *     bc3 -> b4step3 -> copyq3 -> step3 -> MPI_ALLREDUCE(cfl) -> src3
 * This library keeps q and aux resident on the GPU and exposes each of those
 * operators, plus the claw3 step-control loop itself, as plain C entry
 * points (no torch or HIP types in the signatures).  A Fortran host binds
 * them with iso_c_binding (magic_master_amd/fortran/magic_hip.f90); Python
 * binds them with ctypes (magic_master_amd/hip.py).
 *
 * Host arrays use the reference's own layouts (column-major Fortran):
 *     q  (meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)   REAL*8
 *     aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)   REAL*8
 * (mpilib/claw3ez_mpi.f:33-36).  Every entry point returns 0 on success or
 * a MAGIC_E* status; the reference stops the process instead.
 * Calls on one context are synchronous with respect to the host and are
 * not thread-safe on the same context.
 */
#ifndef MAGIC_HIP_H
#define MAGIC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MAGIC_OK            0
#define MAGIC_ERHO          1   /* rho <= 0 in a Riemann problem   (rpn3euX.f:140-153) */
#define MAGIC_EA2           2   /* a**2 <= 0 in a Roe average      (rpn3euX.f:176-185) */
#define MAGIC_EBC           3   /* mthbc 0 or 4 on one GPU          (bc3_mpi.f:100-105) */
#define MAGIC_EUNSUPPORTED  4   /* method / option not implemented  */
#define MAGIC_EARG          5   /* invalid argument                 */
#define MAGIC_EHIP          6   /* HIP runtime error                */
#define MAGIC_ENOMEM        7   /* device allocation failed         */

typedef struct magic_ctx magic_ctx;

/* Problem parameters read by setprob (setprob.f:31-61), in that order. */
typedef struct {
    int loadprof, nspecadv;
    double gamma, grav, dens, pres, dmolec0, dtherm0;
    double forcemth, omega, amplitude, propx, propz;
    double xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth, vsrcx;
    double amp0, amp1, omega1, prop1z, phi, z1width, z1pos, z1Gpos, windscale1;
} magic_params;

/* Create a context for one grid on the current HIP device.
 * Replaces the automatic arrays of claw3ez_mpi (mpilib/claw3ez_mpi.f:33-37)
 * and the work partition of claw3 (mpilib/claw3_mpi.f:704-728).
 * Requires mbc >= 2, meqn in 5..10, maux >= 19. */
int magic_create(int mx, int my, int mz, int meqn, int mbc, int maux,
                 double xlower, double ylower, double zlower,
                 double dx, double dy, double dz, magic_ctx **out);

/* State forms.  MAGIC_FORM_FP64 reproduces the reference's FP64 step.
 * MAGIC_FORM_FP32P stores q as FP32 perturbations from a background column
 * at rest (aux slots 1-5, 11, 12, 17-19 independent of i and j, momenta 0)
 * and runs the Riemann solves, limiters, transverse splits, increment
 * accumulation and source pass in FP32; the Courant-number reduction and
 * the background stay FP64.  Both forms take and return the reference's
 * FP64 q and aux arrays.  No counterpart in the reference (a precision mode
 * of the same operators). */
#define MAGIC_FORM_FP64   0
#define MAGIC_FORM_FP32P  1
/* magic_create with a state form; magic_create(...) is form MAGIC_FORM_FP64. */
int magic_create_form(int mx, int my, int mz, int meqn, int mbc, int maux,
                      double xlower, double ylower, double zlower,
                      double dx, double dy, double dz, int form, magic_ctx **out);
int magic_destroy(magic_ctx *ctx);
const char *magic_strerror(int status);

/* setprob commons (setprob.f:17-25); nspecadv selects advected tracers in
 * rpn3 (rpn3euX.f:262-268), the rest drive src3 forcing. */
int magic_set_params(magic_ctx *ctx, const magic_params *p);

/* Upload the background aux (setaux output, maux slots) from host memory.
 * Only the slots the step reads are kept on the device (1-5, 11, 12, 17,
 * 18, 19).  Also precomputes the z-boundary density ratios of bc3
 * ((rho_bar_ghost/rho_bar_1)**0.5, bc3_mpi.f:560-573) in FP64.
 * MAGIC_FORM_FP32P keeps one column and returns MAGIC_EUNSUPPORTED when the
 * background is not a column at rest. */
int magic_set_aux(magic_ctx *ctx, const double *aux);
/* A background that depends on z only: col holds aux(1:maux, k) for
 * k = 1-mbc .. mz+mbc (maux * (mz+2mbc) doubles, k slowest), the value of
 * every column of the grid.  Same effect as magic_set_aux on the aux array
 * that repeats it; no counterpart in the reference (the FP32 form keeps
 * only the column anyway). */
int magic_set_aux_column(magic_ctx *ctx, const double *col);
/* Upload / download q (host, reference layout, ghost cells included). */
int magic_set_q(magic_ctx *ctx, const double *q);
/* qinit (qinit.f:74-82) without noise: q = aux(1:5) on every cell (fluid at
 * rest on the background), tracers 0.  Needs the aux first. */
int magic_qinit_rest(magic_ctx *ctx);
int magic_get_q(magic_ctx *ctx, double *q);
/* Snapshot of q for output while the context's work goes on (the
 * reference's out3 frames, out3_mpi_h5sep.f90:66-195, written by the
 * caller): q in the layout of magic_get_q is converted on the context's
 * stream and copied to q (host memory; page-locked memory from
 * magic_host_alloc, else the copy waits for the device) on a copy stream
 * of the context, so that work queued on the context afterwards (more
 * steps, a batch's attempts) runs meanwhile.  *done receives a handle that
 * magic_snapshot_wait waits on and releases; q holds the snapshot once it
 * returns.  The handle stays valid after the context is destroyed (which
 * first waits for the copy).  No counterpart in the reference. */
int magic_snapshot_begin(magic_ctx *ctx, double *q, void **done);
int magic_snapshot_wait(void *done);
/* Page-locked host memory for snapshots. */
int magic_host_alloc(size_t bytes, void **p);
int magic_host_free(void *p);

/* bc3 (bc3_mpi.f:5-812): fill the ghost layers of q on the device.
 * mthbc(k) = 1 extrapolation, 2 periodic, 3 wall, < 0 skip the face. */
int magic_bc3(magic_ctx *ctx, const int mthbc[6]);
/* b4step3 (b4step3.f:35-45): clamp tracers 6..meqn to [0,1] on all cells. */
int magic_b4step3(magic_ctx *ctx);
/* copyq3 (lib/copyq3.f:4-29) as a device copy q -> backup (claw3_mpi.f:807)
 * and backup -> q (claw3_mpi.f:869). */
int magic_backup(magic_ctx *ctx);
int magic_restore(magic_ctx *ctx);
/* step3 (lib/step3.f:3-626, with flux3fw.f, rpn3euX.f, rpt3euX.f,
 * rptt3euX.f): q += unsplit x/y/z f-wave increments computed from the
 * backup copy (qold); *cfl = max Courant number (step3.f:182/350/520). */
int magic_step3(magic_ctx *ctx, double dt, double *cfl);
/* src3 (src3.f90:203-634) with diffuse.f and conduct.f, forcemth 0/1/3,
 * evaluated at time t over dt; returns the substep counts. */
int magic_src3(magic_ctx *ctx, double t, double dt, const int mthbc[6],
               int *ndtr, int *ndtrc);

/* claw3 (mpilib/claw3_mpi.f:631-951) for method(1)=1, method(2)=2,
 * method(3)=22, method(5)=1, method(6)=0: the step-control loop with CFL
 * rejection/retry, on device-resident q.  Arguments mean what they mean
 * in the reference (tend < tstart takes one accepted step).  If log_dt /
 * log_cfl / log_acc are non-NULL, every attempted step is recorded (up to
 * maxlog entries; *nlog returns the count). */
int magic_claw3(magic_ctx *ctx, double tstart, double *tend, double dtv[5],
                double cflv[4], int nv[2], const int method[7], const int mthbc[6],
                double *log_dt, double *log_cfl, int *log_acc, int maxlog, int *nlog);

/* Global sums over the interior cells in FP64 with compensated (Neumaier)
 * summation: sums[m] = sum over i, j, k of q(m+1, i, j, k) * dx*dy*dz for
 * m = 0 .. meqn-1 -- the mass (m = 0), the momenta, the total energy
 * (m = 4) and the tracer masses.  In the FP32 form the background of each
 * layer enters as an exact FP64 product and the FP32 perturbations are
 * summed against it.  Every addition has a fixed order, so the result is
 * the same on every run.  The first call allocates the context's partial
 * sums (mz * meqn * 16 B).  No counterpart in the reference (it checks
 * conservation only implicitly: periodic x and y with the fluid at rest
 * change sum(rho) only through the z faces). */
int magic_global_sums(magic_ctx *ctx, double *sums);
/* magic_global_sums without allocating anything: the same kernels and the
 * same sums, bit for bit, with the partial sums in a static device array of
 * the library instead of context memory.  No counterpart in the reference;
 * for callers that cannot allocate device memory (HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Limit: mz <= 1024 interior
 * layers (MAGIC_EARG beyond it).  Calls that overlap on one GPU race on the
 * array. */
int magic_global_sums_noalloc(magic_ctx *ctx, double *sums);

/* Where the last MAGIC_ERHO / MAGIC_EA2 happened: sweep (1,2,3), the
 * Fortran indices of the pencil's cell, as the reference prints them
 * (rpn3euX.f:145-150). */
int magic_error_location(magic_ctx *ctx, int *sweep, int *i, int *j, int *k);

/* Per-kernel-class timing with HIP events on the library's streams
 * (0 = off): each class's launches are bracketed on the stream they run on
 * (the second stream for the y-sweep and half of the z-sweep while the
 * sweep overlap is on, so those spans include time shared with the other
 * stream).  Classes: 0 sweep x, 1 sweep y, 2 sweep z, 3 bc3, 4 src3,
 * 5 other.  magic_timing_read returns milliseconds and launch counts
 * accumulated since the last reset. */
int magic_timing_enable(magic_ctx *ctx, int on);
int magic_timing_read(magic_ctx *ctx, double ms[6], long long launches[6]);
int magic_timing_reset(magic_ctx *ctx);

/* ---- Tile shapes and the autotuner (no counterpart in the reference) ----
 * The tiled sweep kernel comes in MAGIC_TILE_NSHAPE shapes: workgroups of
 * 256 lanes covering TA neighbouring pencils x 256/TA interfaces, lanes
 * tile-axis-major or pencil-major:
 *   0: TA 16, tile-axis-major   1: TA 8, tile-axis-major
 *   2: TA 8, pencil-major       3: TA 4, pencil-major
 * shapes[0..2] are the shapes of the x-, y- and z-sweeps (default 0, 0, 0).
 * Every shape computes the same increments; the grouping of their sums, and
 * so the last bits of the result, depends on the shape.  magic_tune times
 * every shape of every direction on the grid's current state (ghost cells
 * set, dt > 0: one untimed and reps timed sweeps each), keeps the fastest
 * per direction and returns them in shapes_out (and, if ms_out is not NULL,
 * the mean ms per sweep in ms_out[3 * MAGIC_TILE_NSHAPE], direction-major).
 * With the sweep overlap on (magic_set_sweep_overlap), it then varies each
 * direction's shape in turn, timing the overlapped sweeps of step3 as a
 * whole, and keeps an assignment that is at least 1 % faster.
 * q is unchanged bit for bit; the backup copy then holds q.
 * MAGIC_EUNSUPPORTED for the FP64 pencil kernel (MAGIC_SWEEP=pencil). */
#define MAGIC_TILE_NSHAPE 4
int magic_set_tile_shapes(magic_ctx *ctx, const int shapes[3]);
int magic_get_tile_shapes(magic_ctx *ctx, int shapes[3]);
int magic_tune(magic_ctx *ctx, double dt, int reps, int shapes_out[3], float *ms_out);

/* ---- Order of the sweeps (no counterpart in the reference) -------------
 * on = 1 (default): the y-sweep of step3 and half of the z-sweep's colour
 * launches run on a second stream next to the x-sweep and the other z
 * launches, adding into a scratch copy of the state's shape that is added
 * to q once both streams are done (the compute units the last workgroups of
 * one stream's launches leave idle take the other stream's workgroups).
 * on = 0: the three sweeps add into q in turn on the context's stream, the
 * order of additions of the batched driver, whose scenarios then equal the
 * same grids stepped alone bit for bit.  Either way the result is the same
 * on every run; the two orders differ in the last bits.  The environment
 * variable MAGIC_SWEEP_OVERLAP=0 turns the overlap off for every context. */
int magic_set_sweep_overlap(magic_ctx *ctx, int on);

/* ---- Batched driver: many grids of one form advanced together ----------
 * Each scenario is a grid with its own state (a magic_ctx reachable through
 * magic_batch_ctx for parameters, aux and q); its step control is claw3's
 * (claw3_mpi.f:736-927) but lives on the device, and one attempted step of
 * every scenario is one replay of a HIP graph captured at the first
 * magic_batch_run.  Each kernel of the attempt is one launch over all
 * scenarios (the sweeps as persistent launches per direction and tile
 * colour over a work list of tiles, the boundary, copy and source-pass
 * kernels over the scenarios' grids laid end to end), so the launches per
 * attempt do not grow with the number of scenarios.  No counterpart in the
 * reference. */
typedef struct magic_batch magic_batch;
/* dims: (mx, my, mz) per scenario; geom: (xlower, ylower, zlower, dx, dy, dz)
 * per scenario (nscen may be 0); form: MAGIC_FORM_FP32P or MAGIC_FORM_FP64
 * (every scenario of a batch has the batch's form).  An FP64 scenario's
 * background must be constant on each layer (magic_set_aux_column, or
 * magic_set_aux with such planes): magic_batch_start returns
 * MAGIC_EUNSUPPORTED otherwise. */
int magic_batch_create(int nscen, const int *dims, int meqn, int mbc, int maux, const double *geom,
                       int form, magic_batch **out);
int magic_batch_destroy(magic_batch *b);
/* Continuous batching: add a scenario (a new grid, inactive until started;
 * may reuse the slot of a removed one) or remove one (its device memory is
 * freed).  The next magic_batch_run captures the graph again. */
int magic_batch_add(magic_batch *b, const int dims[3], const double geom[6], int *slot);
int magic_batch_remove(magic_batch *b, int s);
/* The scenario's context (owned by the batch: do not destroy it). */
int magic_batch_ctx(magic_batch *b, int s, magic_ctx **ctx);
/* Start an interval of scenario s: from t to tend (clipped like claw3 when
 * t < tend) with first dt, dtv(2) = dtv[1] as the largest dt, cflv(1:2),
 * at most maxsteps accepted steps, boundary conditions mthbc.  The source
 * pass's substep counts are bounded from max(dt, dtv[1]); a start whose
 * bound exceeds those of the captured graph captures it again. */
int magic_batch_start(magic_batch *b, int s, double t, double tend, double dt, const double dtv[5],
                      const double cflv[4], int maxsteps, const int mthbc[6]);
/* Replay attempts until no started scenario is active (checked every
 * check_every attempts) or max_attempts were made.  Returns only argument
 * and HIP errors: a scenario that fails (MAGIC_ERHO, MAGIC_EA2, or
 * MAGIC_EUNSUPPORTED when its source pass needs more substeps than were
 * captured) stops with that status, which magic_batch_state reports, and
 * the other scenarios go on. */
int magic_batch_run(magic_batch *b, int max_attempts, int check_every, int *attempts_done);
/* real_out: t, dt (next), cfl (last attempt), cflmax, tend;
 * int_out: accepted steps, attempts, rejected attempts, active, status, ndtr. */
int magic_batch_state(magic_batch *b, int s, double real_out[5], int int_out[6]);
/* Tile shapes of the batch's x-, y- and z-sweeps (one per direction for all
 * scenarios; default 0, 0, 0); setting them captures the graph again.
 * magic_batch_tune is the autotuner of the batch: every shape of every
 * direction is timed as one persistent launch per tile colour over the
 * tiles of all scenarios (their current states and next dt; one untimed
 * and reps timed sweeps), and the fastest per direction is kept.  The
 * scenarios' states are unchanged bit for bit. */
int magic_batch_set_tile_shapes(magic_batch *b, const int shapes[3]);
int magic_batch_get_tile_shapes(magic_batch *b, int shapes[3]);
int magic_batch_tune(magic_batch *b, int reps, int shapes_out[3], float *ms_out);
/* The batch's sweeps in the same two orders as magic_set_sweep_overlap.
 * Off by default: the batch's persistent launches leave few compute units
 * idle, and the per-scenario merge costs more than the overlap gains.
 * While on, each scenario holds one more buffer of its state's size.
 * Setting it captures the graph again.  A scenario equals the same grid
 * stepped alone bit for bit when both use the same order and tile shapes.
 * No counterpart in the reference. */
int magic_batch_set_sweep_overlap(magic_batch *b, int on);

#ifdef __cplusplus
}
#endif
#endif /* MAGIC_HIP_H */
