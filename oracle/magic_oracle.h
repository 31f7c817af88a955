/*
 * magic_oracle.h -- CPU restatement of the MAGIC Clawpack 3-D f-wave step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library in
 * magic_master_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * Arrays use the reference's own layouts (column-major Fortran):
 *   q  (meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)   REAL*8
 *   aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)   REAL*8
 * i.e. the component index is fastest, then i, j, k.
 */
#ifndef MAGIC_ORACLE_H
#define MAGIC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int mx, my, mz, meqn, mbc, maux;
    double xlower, ylower, zlower, dx, dy, dz;
} orc_grid;

/* Problem parameters read by setprob (setprob.f:31-61). */
typedef struct {
    int loadprof, nspecadv;
    double gamma, grav, dens, pres, dmolec0, dtherm0;
    double forcemth, omega, amplitude, propx, propz;
    double xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth, vsrcx;
    double amp0, amp1, omega1, prop1z, phi, z1width, z1pos, z1Gpos, windscale1;
} orc_params;

/* Status codes (the reference stops the process instead). */
enum {
    ORC_OK = 0,
    ORC_RHO_NONPOS = 1,      /* rpn3euX.f:140-153 */
    ORC_A2_NONPOS = 2,       /* rpn3euX.f:176-185 */
    ORC_BAD_BC = 3,          /* bc3_mpi.f:100-105 (mthbc 0) or 4 (needs MPI) */
    ORC_UNSUPPORTED = 4
};

/* Riemann operators on one pencil (Fortran index range 1-mbc..n+mbc).
 * roe: 8 arrays of n+2*mbc doubles, order u2v2w2,u,v,w,enth,a,g1a2,euv
 * (common /comroe/, rpn3euX.f:69-71).  fwave is (meqn,3,pencil), s (3,pencil). */
int orc_rpn3(int ixyz, int meqn, int mbc, int n, int nspecadv,
             const double *q1d, const double *aux19, const double *aux17,
             double dtcom, double dxcom,
             double *fwave, double *s, double *amdq, double *apdq, double *roe);
void orc_rpt3(int ixyz, int icoor, int meqn, int mbc, int n, const double *roe,
              const double *asdq, double *bmasdq, double *bpasdq);
void orc_rptt3(int ixyz, int icoor, int meqn, int mbc, int n, const double *roe,
               const double *bsasdq, double *cmbsasdq, double *cpbsasdq);
double orc_cflsupbee(double theta, double beta, double r, double cfl, double cflup);
double orc_suppow(double r, double cfl, double cflup);

/* step3 (lib/step3.f:3-626) with method (*,2,22,*,*,0,maux), mthlim = 0.
 * qnew must enter equal to qold.  Returns a status code; *cfl is the max
 * Courant number over the three sweeps. */
int orc_step3(const orc_grid *g, int nspecadv, const double *qold, double *qnew,
              const double *aux, double dt, double *cfl);

/* bc3 (bc3_mpi.f) branches 1 (extrapolation), 2 (periodic, one rank),
 * 3 (wall); mthbc(k) < 0 skips a face. */
int orc_bc3(const orc_grid *g, double *q, const double *aux, const int *mthbc);
/* b4step3 (b4step3.f:35-45): clamp tracers 6..meqn to [0,1] on all cells. */
void orc_b4step3(const orc_grid *g, double *q);
/* src3 (src3.f90:203-634) forcemth 0/1/3 with diffuse.f and conduct.f.
 * ndtr / ndtrc return the substep counts (src3.f90:314-322). */
int orc_src3(const orc_grid *g, const orc_params *p, double *q, const double *aux,
             double t, double dt, const int *mthbc, int *ndtr, int *ndtrc);

/* One claw3 call (mpilib/claw3_mpi.f:631-951), variable dt (method(1)=1),
 * Godunov source splitting (method(5)=1).  With tend < tstart it takes exactly
 * one accepted step (outstyle 3).  log_dt/log_cfl/log_acc (length maxlog)
 * record every attempted step; *nlog returns how many were written. */
int orc_claw3(const orc_grid *g, const orc_params *p, double *q, const double *aux,
              double tstart, double *tend, double *dtv, double *cflv, int *nv,
              const int *method, const int *mthbc,
              double *log_dt, double *log_cfl, int *log_acc, int maxlog, int *nlog);

#ifdef __cplusplus
}
#endif
#endif
