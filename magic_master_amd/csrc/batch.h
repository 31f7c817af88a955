// Device-resident step control of the batched driver: one control block per
// scenario, read and written only by kernels, so that one attempted step of
// every scenario runs as a fixed kernel sequence (captured once as a HIP
// graph) with no host round trip.  The control logic is claw3's
// (mpilib/claw3_mpi.f:736-927): clip dt to the end of the interval when a step
// starts, accept when cfl <= cflv(1), otherwise restore q and retry with
// dt = min(dtv(2), dt*cflv(2)/cfl); after an accepted step the source pass
// runs with the step's start time and dt, and the next dt is
// min(dtv(2), dt*cflv(2)/cfl) (dtv(2) when cfl = 0).
#pragma once
#include "magic_dev.h"

namespace magic {

enum { GATE_RUN = 0, GATE_ACCEPTED, GATE_REJECTED, GATE_DIFF, GATE_COND, GATE_COND_MORE, GATE_COND_LAST, GATE_SWEPT };

// factor slots of ScenCtl::fac (source pass, src3.f90:314-322 and the
// passes of diffuse.f / conduct.f)
enum { FAC_CROSS = 0, FAC_DIFF = 3, FAC_COND = 12, NFAC = 15 };

struct ScenCtl {
    // step control (claw3)
    double t, told, dt, tend, dtmax, cflv1, cflv2, cfl, cflmax;
    int clip;        // the interval was started with tstart < tend (claw3_mpi.f:741)
    int active;      // still stepping in the current interval
    int newstep;     // the next attempt starts a step (else it retries one)
    int run;         // this attempt runs
    int swept;       // this attempt's sweeps ran (run before the decision, which may clear run)
    int accepted;    // this attempt was accepted
    int rejected;    // this attempt was rejected
    int nsteps, maxsteps, attempts, rejects, status;
    // source pass of the attempt
    int ndtr, ndtrc, ndtr_cap, ndtrc_cap;
    double dtr, dtrc, tgauss;
    double fac[NFAC];
    // constants of the scenario
    double dx, dy, dz, difmax, kapmax, dxyzmin2, tcenter, twidth;
    // sweep reductions
    unsigned long long cfl_bits;
    int err[6];
};

// Pointer into memory that no kernel writes while the reading kernel runs
// (so that its uniform loads may be scalar loads).
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

// *p copied word by word (a struct's copy constructor takes a generic
// reference, which does not bind to such memory)
template <typename T>
__device__ __forceinline__ T cload(cptr<T> p)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    T v;
    unsigned *d = reinterpret_cast<unsigned *>(&v);
    const cptr<unsigned> w = (cptr<unsigned>)p;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) d[i] = w[i];
    return v;
}

// Gate of the shared kernels for one scenario of a batch.
struct StepGate {
    cptr<ScenCtl> c;
    int mode, nn, kk, fslot;
    __device__ __forceinline__ bool on() const
    {
        switch (mode) {
        case GATE_RUN: return c->run != 0;
        case GATE_ACCEPTED: return c->accepted != 0;
        case GATE_REJECTED: return c->rejected != 0;
        case GATE_DIFF: return c->accepted != 0 && nn <= c->ndtr;
        case GATE_COND: return c->accepted != 0 && nn <= c->ndtr && kk <= c->ndtrc;
        case GATE_COND_MORE: return c->accepted != 0 && nn <= c->ndtr && kk < c->ndtrc;
        case GATE_COND_LAST: return c->accepted != 0 && nn <= c->ndtr && kk == c->ndtrc;
        case GATE_SWEPT: return c->swept != 0;
        default: return false;
        }
    }
    template <typename T>
    __device__ __forceinline__ T f(T x, int i) const { return fslot < 0 ? x : (T)c->fac[fslot + i]; }
    // the forcing of src3 is evaluated at the step's start time and dt
    __device__ __forceinline__ void forcing(double &t, double &dt, double &tgauss) const
    {
        t = c->told;
        dt = c->dt;
        tgauss = c->tgauss;
    }
};

}  // namespace magic
