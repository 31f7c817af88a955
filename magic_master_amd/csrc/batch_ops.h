// Boundary, pre-step, copy and source-pass kernels of the batched driver as
// work-list launches: one launch of a kernel class covers every scenario of
// a batch.  The list's blocks are the scenarios' own grids laid end to end
// (scenario s owns blocks off[s] .. off[s+1]-1, a grid of gdim[s] blocks); a
// persistent grid walks the list, finds each block's scenario by a binary
// search over off and runs the single-grid kernel's body (stencil_body.h,
// bc_src.hip, bc_src32.hip) on that scenario's operands, gated by its
// step-control block (batch.h).  The number of launches per attempt does not
// depend on the number of scenarios, and a scenario whose gate is off (not
// running, or past its own substep counts) costs next to nothing.
//
// The kernel template is defined here and instantiated in bc_src32.hip (FP32
// form) and bc_src.hip (FP64 form), each under its form's floating-point
// contraction setting, so a scenario rounds exactly as the same grid stepped
// alone.
#pragma once
#include "batch.h"
#include "bc_src.h"

namespace magic {

// The operands of one scenario (device array, one entry per batch slot).
template <typename T>
struct BatchOps {
    Dims d;
    T *q, *qold, *u, *qt, *ut, *dq;
    int nc;          // meqn
    int kb;          // layers per block of the diffusion kernels (diff_kb)
    int spill;       // conduct.f:86-91 out-of-row writes (mz - mx when mz > mx, else 0)
    int bc[6];       // mthbc (bc3 faces; the source pass's substeps skip z)
    ScenCtl *ctl;
    // FP32 form: background columns, z-face factors, diffusivity columns
    Col32 col;
    Zr32 zr;
    const float *nu32, *kap32;
    // FP64 form: background columns (source pass), aux planes and z factors (bc3)
    Aux ax;
    const double *nu64, *kap64, *aux1, *aux2, *aux3, *aux5, *zratio;
    Forcing f;       // t, dt and the time factor come from the control block
};

// Kernel classes of the batched attempt
enum {
    OP_BCX = 0, OP_BCY, OP_BCZ, OP_CLAMP, OP_BACKUP, OP_RESTORE, OP_CONVERT, OP_ZGHOST_UVW, OP_DIFF_CROSS,
    OP_DIFF_XYZ, OP_COPY_Q5, OP_PASS, OP_CONDUCT_ZGHOST, OP_ZSPILL, OP_COPY_INTERIOR, OP_FINISH, OP_MERGE, NOP
};
// Block lists, one per launch geometry (a scenario whose faces of a bc3
// kernel are all left to the caller has no blocks in that kernel's list):
//   FX  bc3 x faces          (NY/64, NZ)      FY  bc3 y faces      (NX/64, NZ)
//   BCZ bc3 z faces          (NX/64, NY)      FZ  z-ghost fills    (NX/64, NY)
//   ALL one lane per cell    (P/256)          CPY grid-stride copy of meqn planes
//   CP1 grid-stride copy of one plane         PAD padded cells     (NX/64, NY/4, NZ)
//   DIF diffusion passes     (mx/64, my/4, mz/kb)
//   INT interior cells       (mx/64, my/4, mz)
//   SPL conduct.f's out-of-row writes ((mz-mx)/64)
//   MRG grid-stride q += dq of meqn planes
enum { LG_FX = 0, LG_FY, LG_BCZ, LG_FZ, LG_ALL, LG_CPY, LG_CP1, LG_PAD, LG_DIF, LG_INT, LG_SPL, LG_MRG, NLG };
// OP_PASS variants: (out, interior source) = (qt, ut), (ut, qt), (q(5), ut)
enum { PASS_QT_UT = 0, PASS_UT_QT, PASS_Q5_UT };

struct GateSpec {
    int mode, nn, kk, fslot;
};

// Slots a batch may hold (the work-list kernels keep one offset per slot in
// LDS)
constexpr int BATCH_MAX_SLOTS = 2048;

// The slot of active block a: the largest s with pre[s] <= a (slots with
// no active blocks, whose pre equals the next slot's, are passed over)
__device__ __forceinline__ int active_find(const int *pre, int n, int a)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= a) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int lohi(int m) { return m < 0 ? 0 : m; }

// q += dq, dq = 0 over n elements for a scenario whose sweeps ran (the
// sweep overlap of the batch): float4 or double2 vectors, grid-stride over
// nblk blocks of 256 lanes
template <typename T, class G>
__device__ __forceinline__ void b_merge(const uint3 BI, unsigned nblk, T *__restrict__ q, T *__restrict__ dq,
                                        long long n, const G &gate)
{
    if (!gate.on()) return;
    constexpr int N = 16 / sizeof(T);
    const long long nv = n / N;
    const long long stride = (long long)nblk * blockDim.x;
    const long long t0 = (long long)BI.x * blockDim.x + threadIdx.x;
    for (long long i = t0; i < nv; i += stride)
#pragma unroll
        for (int e = 0; e < N; e++) {
            q[i * N + e] = q[i * N + e] + dq[i * N + e];
            dq[i * N + e] = T(0.0);
        }
    const long long r = nv * N + t0;   // the last n mod N elements
    if (r < n) {
        q[r] = q[r] + dq[r];
        dq[r] = T(0.0);
    }
}

// One launch of kernel class OP over the blocks of the scenarios' list
// (off, gdim) whose gates are on.  Each workgroup of the persistent grid
// first evaluates every slot's gate and lays the active slots' blocks end to
// end in LDS (pre: active blocks before each slot), then takes active
// blocks blockIdx.x, blockIdx.x + gridDim.x, ...; a launch whose gates are
// all off ends after that scan.  arg: the velocity component of
// OP_DIFF_XYZ (1-3), PASS_* * 4 + axis of OP_PASS.  The form's own bodies
// are reached through the FORM functor (FP32: bc_src32.hip, FP64:
// bc_src.hip).  n <= BATCH_MAX_SLOTS (the host checks).
template <typename T, int OP, class FORM>
__global__ __launch_bounds__(256) void k_batch_op(const BatchOps<T> *__restrict__ S, const int *__restrict__ off,
                                                  const int4 *__restrict__ gdim, int n, GateSpec gs, int arg)
{
    // No kernel of an attempt writes the operand blocks or the lists, and
    // only prepare, decide and finalize write the control blocks: read
    // through the constant address space, their values stay in scalar
    // registers although the bodies store to q between the reads.
    const cptr<BatchOps<T>> CS = (cptr<BatchOps<T>>)S;
    const cptr<int> coff = (cptr<int>)off;
    const cptr<int> cgdim = (cptr<int>)gdim;   // 4 ints per scenario
    __shared__ int pre[BATCH_MAX_SLOTS + 1];
    __shared__ int wsum[4];
    {
        // thread t owns slots t*per .. t*per+per-1: their active block
        // counts go to pre, then their exclusive prefix replaces them
        const int t = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        const int nt = (int)(blockDim.x * blockDim.y);   // 64 or 256 lanes
        const int per = (n + nt - 1) / nt;
        int sum = 0;
        for (int i = 0; i < per; i++) {
            const int sl = t * per + i;
            if (sl >= n) break;
            // a slot with no blocks here (an empty slot among them) is not
            // looked at further
            int c = coff[sl + 1] - coff[sl];
            if (c > 0) {
                const StepGate g = {(cptr<ScenCtl>)CS[sl].ctl, gs.mode, gs.nn, gs.kk, gs.fslot};
                if (!g.on()) c = 0;
            }
            pre[sl] = c;
            sum += c;
        }
        const int lane = t & 63, wave = t >> 6;
        int v = sum;   // inclusive scan over the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(v, d, 64);
            if (lane >= d) v += y;
        }
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        int acc = v - sum;
        for (int w = 0; w < wave; w++) acc += wsum[w];
        for (int i = 0; i < per; i++) {
            const int sl = t * per + i;
            if (sl >= n) break;
            const int c = pre[sl];
            pre[sl] = acc;
            acc += c;
        }
        if (t == nt - 1) pre[n] = acc;
        __syncthreads();
    }
    const int total = pre[n];
    const int G = (int)gridDim.x;
    int s = 0, lo = 0, hi = 0;   // the current slot and its active blocks [lo, hi)
    for (int b = (int)blockIdx.x; b < total; b += G) {
        if (b < lo || b >= hi) {
            s = __builtin_amdgcn_readfirstlane(active_find(pre, n, b));
            lo = __builtin_amdgcn_readfirstlane(pre[s]);
            hi = __builtin_amdgcn_readfirstlane(pre[s + 1]);
        }
        const BatchOps<T> o = cload(CS + s);
        const StepGate g = {(cptr<ScenCtl>)o.ctl, gs.mode, gs.nn, gs.kk, gs.fslot};
        const int l = b - lo;
        const int gx = cgdim[4 * s], gy = cgdim[4 * s + 1];
        const uint3 bi = make_uint3((unsigned)(l % gx), (unsigned)((l / gx) % gy), (unsigned)(l / (gx * gy)));
        const long long P = o.d.P;
        if constexpr (OP == OP_BCX) {
            b_bc_x(bi, o.d, o.q, o.nc, lohi(o.bc[0]), lohi(o.bc[1]), g);
        } else if constexpr (OP == OP_BCY) {
            b_bc_y(bi, o.d, o.q, o.nc, lohi(o.bc[2]), lohi(o.bc[3]), g);
        } else if constexpr (OP == OP_BCZ) {
            FORM::bc_z(bi, o, CS + s, g);
        } else if constexpr (OP == OP_CLAMP) {
            b_clamp(bi, o.q, P, o.nc, g);
        } else if constexpr (OP == OP_BACKUP) {
            b_copy(bi, (unsigned)gx, o.qold, o.q, P * o.nc, g);
        } else if constexpr (OP == OP_RESTORE) {
            b_copy(bi, (unsigned)gx, o.q, o.qold, P * o.nc, g);
        } else if constexpr (OP == OP_CONVERT) {
            FORM::convert(bi, o, g);
        } else if constexpr (OP == OP_ZGHOST_UVW) {
            b_zghost_uvw(bi, o.d, o.q, g);
        } else if constexpr (OP == OP_DIFF_CROSS) {
            b_diff_cross<T, true>(bi, o.d, o.u, o.q, FORM::nu(o), T(0.0), T(0.0), T(0.0), o.kb, g);
            __syncthreads();   // the next block's writers of the tile arrays wait for this one's readers
        } else if constexpr (OP == OP_DIFF_XYZ) {
            b_diff_xyz<T, true>(bi, o.d, o.q + arg * P, o.u + (arg - 1) * P, FORM::nu(o), T(0.0), T(0.0), T(0.0), o.kb,
                                g);
            __syncthreads();
        } else if constexpr (OP == OP_COPY_Q5) {
            b_copy(bi, (unsigned)gx, o.ut, o.q + 4 * P, P, g);
        } else if constexpr (OP == OP_PASS) {
            const int v = arg >> 2, axis = arg & 3;
            T *out = v == PASS_QT_UT ? o.qt : v == PASS_UT_QT ? o.ut : o.q + 4 * P;
            const T *src = v == PASS_UT_QT ? o.qt : o.ut;
            b_pass<T, true>(bi, o.d, out, src, o.ut, FORM::kap(o), T(0.0), axis, g);
        } else if constexpr (OP == OP_CONDUCT_ZGHOST) {
            b_conduct_zghost(bi, o.d, o.ut, g);
        } else if constexpr (OP == OP_ZSPILL) {
            b_conduct_zspill(bi, o.d, o.ut, o.spill, g);
        } else if constexpr (OP == OP_COPY_INTERIOR) {
            b_copy_interior(bi, o.d, o.ut, o.qt, 1, g);
        } else if constexpr (OP == OP_FINISH) {
            FORM::finish(bi, o, g);
        } else if constexpr (OP == OP_MERGE) {
            b_merge(bi, (unsigned)gx, o.q, o.dq, P * o.nc, g);
        }
    }
}

// The kernel of class op for form T (FP32: float, FP64: double), defined by
// the form's source file.
template <typename T>
const void *batch_op_fn(int op);

}  // namespace magic
