"""Python binding of the MI355X MAGIC step (libmagic_hip.so, C ABI in
include/magic_hip.h), mirroring the reference's operator interface:

    bc3, b4step3, copyq3 (backup/restore), step3, src3 and the claw3
    step-control loop (mpilib/claw3_mpi.f:4-951), plus the outstyle-3 frame
    loop of claw3ez (mpilib/claw3ez_mpi.f:136-188).

q and aux cross the boundary in the reference's Fortran layouts (numpy
arrays of shape (nz+2mbc, ny+2mbc, nx+2mbc, ncomp)); on the device they stay
resident.  There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .problem import ClawData, Grid, ProbData

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmagic_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

STATUS = {0: "ok", 1: "rho <= 0 in a Riemann problem", 2: "a**2 <= 0 in a Roe average",
          3: "boundary condition needs MPI or a user routine", 4: "unsupported option",
          5: "invalid argument", 6: "HIP runtime error", 7: "device allocation failed"}

MAGIC_OK, MAGIC_ERHO, MAGIC_EA2, MAGIC_EBC, MAGIC_EUNSUPPORTED = 0, 1, 2, 3, 4

FORM_FP64 = 0     # MAGIC_FORM_FP64
FORM_FP32P = 1    # MAGIC_FORM_FP32P: FP32 perturbation from a background column at rest

EXPORTS = ["magic_create", "magic_create_form", "magic_destroy", "magic_strerror", "magic_set_params",
           "magic_set_aux", "magic_set_q", "magic_get_q", "magic_bc3", "magic_b4step3",
           "magic_backup", "magic_restore", "magic_step3", "magic_src3", "magic_claw3",
           "magic_error_location", "magic_timing_enable", "magic_timing_read",
           "magic_timing_reset", "magic_set_aux_column", "magic_qinit_rest", "magic_batch_create", "magic_batch_destroy",
           "magic_batch_ctx", "magic_batch_start", "magic_batch_run", "magic_batch_state",
           "magic_batch_add", "magic_batch_remove", "magic_set_tile_shapes", "magic_get_tile_shapes", "magic_tune",
           "magic_batch_set_tile_shapes", "magic_batch_get_tile_shapes", "magic_batch_tune",
           "magic_set_sweep_overlap", "magic_batch_set_sweep_overlap", "magic_global_sums",
           "magic_global_sums_noalloc", "magic_snapshot_begin", "magic_snapshot_wait", "magic_host_alloc",
           "magic_host_free"]

TILE_NSHAPE = 4   # MAGIC_TILE_NSHAPE
# (pencils per tile, lane order) of each tile shape
TILE_SHAPES = {0: (16, "tile-axis-major"), 1: (8, "tile-axis-major"), 2: (8, "pencil-major"),
               3: (4, "pencil-major")}


class PinnedBuffer:
    """Page-locked host memory (magic_host_alloc) of nbytes, with numpy
    views: the destination of asynchronous snapshots."""

    def __init__(self, nbytes: int):
        self.L = load_library()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        st = self.L.magic_host_alloc(self.nbytes, C.byref(p))
        if st != 0:
            raise MagicError(st, "magic_host_alloc")
        self.ptr = p

    def array(self, shape) -> np.ndarray:
        """A float64 view of the first prod(shape) elements (no copy)."""
        n = int(np.prod(shape))
        if n * 8 > self.nbytes:
            raise ValueError(f"{shape} does not fit {self.nbytes} B")
        buf = (C.c_double * n).from_address(self.ptr.value)
        return np.ctypeslib.as_array(buf).reshape(shape)

    def close(self):
        if getattr(self, "ptr", None):
            self.L.magic_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def snapshot_wait(handle):
    """Wait for an asynchronous snapshot (snapshot_begin) and release its handle."""
    st = load_library().magic_snapshot_wait(handle)
    if st != 0:
        raise MagicError(st, "magic_snapshot_wait")


class MagicError(RuntimeError):
    def __init__(self, code: int, where: str, loc=None):
        msg = f"{where}: {STATUS.get(code, 'status %d' % code)}"
        if loc is not None and code in (1, 2):
            msg += " (sweep %d at i=%d j=%d k=%d)" % tuple(loc)
        super().__init__(msg)
        self.code = code
        self.loc = loc


class CParams(C.Structure):
    _fields_ = [("loadprof", C.c_int), ("nspecadv", C.c_int)] + [
        (n, C.c_double) for n in ("gamma", "grav", "dens", "pres", "dmolec0", "dtherm0",
                                  "forcemth", "omega", "amplitude", "propx", "propz", "xpos",
                                  "ypos", "zpos", "xwidth", "ywidth", "zwidth", "tcenter",
                                  "twidth", "vsrcx", "amp0", "amp1", "omega1", "prop1z", "phi",
                                  "z1width", "z1pos", "z1Gpos", "windscale1")]

    @classmethod
    def of(cls, p: ProbData):
        o = cls()
        for name, _ in cls._fields_:
            setattr(o, name, getattr(p, name))
        return o


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libmagic_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.magic_create.argtypes = [C.c_int] * 6 + [C.c_double] * 6 + [C.POINTER(vp)]
    L.magic_create_form.argtypes = [C.c_int] * 6 + [C.c_double] * 6 + [C.c_int, C.POINTER(vp)]
    L.magic_destroy.argtypes = [vp]
    L.magic_strerror.restype = C.c_char_p
    L.magic_strerror.argtypes = [C.c_int]
    L.magic_set_params.argtypes = [vp, C.POINTER(CParams)]
    L.magic_set_aux.argtypes = [vp, _dp]
    L.magic_set_q.argtypes = [vp, _dp]
    L.magic_get_q.argtypes = [vp, _dp]
    L.magic_snapshot_begin.argtypes = [vp, vp, C.POINTER(vp)]
    L.magic_snapshot_wait.argtypes = [vp]
    L.magic_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.magic_host_free.argtypes = [vp]
    L.magic_bc3.argtypes = [vp, _ip]
    L.magic_b4step3.argtypes = [vp]
    L.magic_backup.argtypes = [vp]
    L.magic_restore.argtypes = [vp]
    L.magic_step3.argtypes = [vp, C.c_double, _dp]
    L.magic_src3.argtypes = [vp, C.c_double, C.c_double, _ip, _ip, _ip]
    L.magic_claw3.argtypes = [vp, C.c_double, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip, C.c_int, _ip]
    L.magic_error_location.argtypes = [vp, _ip, _ip, _ip, _ip]
    L.magic_timing_enable.argtypes = [vp, C.c_int]
    L.magic_timing_read.argtypes = [vp, _dp, C.POINTER(C.c_longlong)]
    L.magic_timing_reset.argtypes = [vp]
    L.magic_set_aux_column.argtypes = [vp, _dp]
    L.magic_qinit_rest.argtypes = [vp]
    L.magic_batch_create.argtypes = [C.c_int, _ip, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.POINTER(vp)]
    L.magic_batch_destroy.argtypes = [vp]
    L.magic_batch_ctx.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.magic_batch_start.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int, _ip]
    L.magic_batch_run.argtypes = [vp, C.c_int, C.c_int, _ip]
    L.magic_batch_state.argtypes = [vp, C.c_int, _dp, _ip]
    L.magic_batch_add.argtypes = [vp, _ip, _dp, _ip]
    L.magic_batch_remove.argtypes = [vp, C.c_int]
    _fp = C.POINTER(C.c_float)
    L.magic_set_tile_shapes.argtypes = [vp, _ip]
    L.magic_get_tile_shapes.argtypes = [vp, _ip]
    L.magic_tune.argtypes = [vp, C.c_double, C.c_int, _ip, _fp]
    L.magic_batch_set_tile_shapes.argtypes = [vp, _ip]
    L.magic_batch_get_tile_shapes.argtypes = [vp, _ip]
    L.magic_batch_tune.argtypes = [vp, C.c_int, _ip, _fp]
    L.magic_set_sweep_overlap.argtypes = [vp, C.c_int]
    L.magic_batch_set_sweep_overlap.argtypes = [vp, C.c_int]
    L.magic_global_sums.argtypes = [vp, _dp]
    L.magic_global_sums_noalloc.argtypes = [vp, _dp]
    _LIB = L
    return L


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("arrays must be C-contiguous float64 in the reference layout")
    return a.ctypes.data_as(_dp)


@dataclass
class StepLog:
    dt: List[float]
    cfl: List[float]
    accepted: List[bool]


class MagicHIP:
    """One grid on the GPU (a magic_ctx)."""

    def __init__(self, g: Grid, p: ProbData, aux: np.ndarray, q: Optional[np.ndarray] = None,
                 form: int = FORM_FP64):
        self.L = load_library()
        self.g = g
        self.p = p
        self.form = form
        h = C.c_void_p()
        self._check(self.L.magic_create_form(g.mx, g.my, g.mz, g.meqn, g.mbc, g.maux, g.xlower,
                                             g.ylower, g.zlower, g.dx, g.dy, g.dz, form,
                                             C.byref(h)),
                    "magic_create_form")
        self.h = h
        self._check(self.L.magic_set_params(self.h, C.byref(CParams.of(p))), "magic_set_params")
        aux = np.ascontiguousarray(aux, dtype=np.float64)
        if aux.shape != g.shape_aux:
            raise ValueError(f"aux shape {aux.shape} != {g.shape_aux}")
        self._check(self.L.magic_set_aux(self.h, _ptr(aux)), "magic_set_aux")
        if q is not None:
            self.set_q(q)

    def close(self):
        if getattr(self, "h", None):
            self.L.magic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, where: str):
        if st != 0:
            loc = None
            if st in (1, 2) and getattr(self, "h", None):
                a = [C.c_int() for _ in range(4)]
                self.L.magic_error_location(self.h, *[C.byref(x) for x in a])
                loc = [x.value for x in a]
            raise MagicError(st, where, loc)

    # -- state ------------------------------------------------------------
    def set_q(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.g.shape_q:
            raise ValueError(f"q shape {q.shape} != {self.g.shape_q}")
        self._check(self.L.magic_set_q(self.h, _ptr(q)), "magic_set_q")

    def get_q(self, out: np.ndarray = None) -> np.ndarray:
        """q in the reference layout; into `out` (C-contiguous float64 of
        shape_q) when given, as the drop-in claw3 copies into its caller's
        array."""
        if out is None:
            out = np.empty(self.g.shape_q)
        elif out.shape != self.g.shape_q or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError(f"out must be C-contiguous float64 of shape {self.g.shape_q}")
        self._check(self.L.magic_get_q(self.h, _ptr(out)), "magic_get_q")
        return out

    def snapshot_begin(self, buf: "PinnedBuffer"):
        """Queue a snapshot of q (get_q's layout) into buf and return its
        handle (snapshot_wait); work queued afterwards runs meanwhile."""
        if buf.nbytes < 8 * int(np.prod(self.g.shape_q)):
            raise ValueError("snapshot buffer too small")
        h = C.c_void_p()
        self._check(self.L.magic_snapshot_begin(self.h, buf.ptr, C.byref(h)), "magic_snapshot_begin")
        return h

    def global_sums(self) -> np.ndarray:
        """Compensated FP64 sums of every component over the interior cells
        times the cell volume (mass, momenta, total energy, tracer masses)."""
        out = np.zeros(self.g.meqn)
        self._check(self.L.magic_global_sums(self.h, _ptr(out)), "magic_global_sums")
        return out

    # -- operators (reference signatures in include/magic_hip.h) ----------
    def bc3(self, mthbc: Sequence[int]):
        self._check(self.L.magic_bc3(self.h, (C.c_int * 6)(*mthbc)), "bc3")

    def b4step3(self):
        self._check(self.L.magic_b4step3(self.h), "b4step3")

    def backup(self):
        self._check(self.L.magic_backup(self.h), "copyq3")

    def restore(self):
        self._check(self.L.magic_restore(self.h), "copyq3")

    def step3(self, dt: float) -> float:
        cfl = C.c_double()
        self._check(self.L.magic_step3(self.h, dt, C.byref(cfl)), "step3")
        return cfl.value

    def src3(self, t: float, dt: float, mthbc: Sequence[int]):
        a, b = C.c_int(), C.c_int()
        self._check(self.L.magic_src3(self.h, t, dt, (C.c_int * 6)(*mthbc), C.byref(a), C.byref(b)),
                    "src3")
        return a.value, b.value

    def claw3(self, tstart: float, tend: float, dtv: Sequence[float], cflv: Sequence[float],
              nv: Sequence[int], method: Sequence[int], mthbc: Sequence[int], maxlog: int = 256):
        """One call of claw3; returns (tend, dtv, cflv, nv, StepLog)."""
        te = C.c_double(tend)
        dtv_a = (C.c_double * 5)(*dtv)
        cflv_a = (C.c_double * 4)(*cflv)
        nv_a = (C.c_int * 2)(*nv)
        ldt = (C.c_double * maxlog)()
        lcfl = (C.c_double * maxlog)()
        lacc = (C.c_int * maxlog)()
        nl = C.c_int()
        st = self.L.magic_claw3(self.h, tstart, C.byref(te), dtv_a, cflv_a, nv_a,
                                (C.c_int * 7)(*method), (C.c_int * 6)(*mthbc), ldt, lcfl, lacc,
                                maxlog, C.byref(nl))
        self._check(st, "claw3")
        n = nl.value
        return te.value, list(dtv_a), list(cflv_a), list(nv_a), StepLog(
            list(ldt[:n]), list(lcfl[:n]), [bool(x) for x in lacc[:n]])

    def run_steps(self, c: ClawData, nsteps: int, t0: float = 0.0, dt0: Optional[float] = None):
        """claw3ez outstyle-3 frame loop (mpilib/claw3ez_mpi.f:136-153): one
        accepted step per claw3 call, dtv(1) = dtv(5) between calls.
        Returns (t, dtv, log of every attempted step)."""
        dtv = list(c.dtv)
        if dt0 is not None:
            dtv[0] = dt0
        cflv = list(c.cflv)
        t = t0
        log = StepLog([], [], [])
        for _ in range(nsteps):
            t, dtv, cflv, nv, lg = self.claw3(t, t - 1.0, dtv, cflv, [c.nv[0], 0], c.method, c.mthbc)
            log.dt += lg.dt
            log.cfl += lg.cfl
            log.accepted += lg.accepted
            dtv[0] = dtv[4]
        return t, dtv, log

    # -- timing -----------------------------------------------------------
    def timing(self, on: bool = True):
        self._check(self.L.magic_timing_enable(self.h, 1 if on else 0), "timing")

    def timing_reset(self):
        self._check(self.L.magic_timing_reset(self.h), "timing")

    def timing_read(self):
        ms = (C.c_double * 6)()
        n = (C.c_longlong * 6)()
        self._check(self.L.magic_timing_read(self.h, ms, n), "timing")
        names = ["sweep_x", "sweep_y", "sweep_z", "bc3", "src3", "other"]
        return {k: (ms[i], n[i]) for i, k in enumerate(names)}

    # -- tile shapes --------------------------------------------------------
    def set_tile_shapes(self, shapes: Sequence[int]):
        self._check(self.L.magic_set_tile_shapes(self.h, (C.c_int * 3)(*shapes)), "magic_set_tile_shapes")

    def tile_shapes(self):
        sh = (C.c_int * 3)()
        self._check(self.L.magic_get_tile_shapes(self.h, sh), "magic_get_tile_shapes")
        return list(sh)

    def set_sweep_overlap(self, on: bool):
        """y-sweep on a second stream (True, the default) or the three sweeps
        in turn (False: the batched driver's order of additions)."""
        self._check(self.L.magic_set_sweep_overlap(self.h, 1 if on else 0), "magic_set_sweep_overlap")

    def tune(self, dt: float, reps: int = 3):
        """Autotuner: times every tile shape of every sweep direction on the
        current state and keeps the fastest; returns (shapes, ms[3][NSHAPE])."""
        sh = (C.c_int * 3)()
        ms = (C.c_float * (3 * TILE_NSHAPE))()
        self._check(self.L.magic_tune(self.h, dt, reps, sh, ms), "magic_tune")
        return list(sh), [list(ms[d * TILE_NSHAPE:(d + 1) * TILE_NSHAPE]) for d in range(3)]


class MagicBatch:
    """Many grids of one form (FORM_FP32P, the default, or FORM_FP64)
    advanced together with the step control on the device (magic_batch_*).
    Each scenario takes its parameters, a background column (aux(1:maux, k)
    for every k) and q (None: the fluid at rest on the background, qinit
    without noise)."""

    def __init__(self, grids, params, aux_columns, qs=None, meqn=None, mbc=None, maux=None, form=None):
        self.L = load_library()
        g0 = grids[0] if grids else None
        meqn = g0.meqn if g0 else meqn
        mbc = g0.mbc if g0 else mbc
        maux = g0.maux if g0 else maux
        self.form = FORM_FP32P if form is None else form
        h = C.c_void_p()
        self._check(self.L.magic_batch_create(0, None, meqn, mbc, maux, None, self.form, C.byref(h)),
                    "magic_batch_create")
        self.h = h
        self.meqn, self.mbc, self.maux = meqn, mbc, maux
        self.grids = {}
        self.ctx = {}
        for s, g in enumerate(grids):
            self.add(g, params[s], aux_columns[s], None if qs is None else qs[s])

    def add(self, g, p, aux_column, q=None) -> int:
        """A new scenario (inactive until start); returns its slot."""
        if (g.meqn, g.mbc, g.maux) != (self.meqn, self.mbc, self.maux):
            raise ValueError("all scenarios need the same meqn, mbc, maux")
        slot = C.c_int()
        self._check(self.L.magic_batch_add(self.h, (C.c_int * 3)(g.mx, g.my, g.mz),
                                           (C.c_double * 6)(g.xlower, g.ylower, g.zlower, g.dx, g.dy, g.dz),
                                           C.byref(slot)), "magic_batch_add")
        s = slot.value
        c = C.c_void_p()
        self._check(self.L.magic_batch_ctx(self.h, s, C.byref(c)), "magic_batch_ctx")
        self.grids[s] = g
        self.ctx[s] = c
        self._check(self.L.magic_set_params(c, C.byref(CParams.of(p))), "magic_set_params")
        col = np.ascontiguousarray(aux_column, dtype=np.float64)
        if col.shape != (g.mz + 2 * g.mbc, g.maux):
            raise ValueError(f"aux column shape {col.shape}")
        self._check(self.L.magic_set_aux_column(c, _ptr(col)), "magic_set_aux_column")
        if q is None:
            self._check(self.L.magic_qinit_rest(c), "magic_qinit_rest")
        else:
            self.set_q(s, q)
        return s

    def remove(self, s):
        self._check(self.L.magic_batch_remove(self.h, s), "magic_batch_remove")
        del self.grids[s]
        del self.ctx[s]

    def close(self):
        if getattr(self, "h", None):
            self.L.magic_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, where):
        if st != 0:
            raise MagicError(st, where)

    def set_q(self, s, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.grids[s].shape_q:
            raise ValueError(f"q shape {q.shape} != {self.grids[s].shape_q}")
        self._check(self.L.magic_set_q(self.ctx[s], _ptr(q)), "magic_set_q")

    def get_q(self, s):
        q = np.empty(self.grids[s].shape_q)
        self._check(self.L.magic_get_q(self.ctx[s], _ptr(q)), "magic_get_q")
        return q

    def snapshot_begin(self, s, buf: "PinnedBuffer"):
        """Queue a snapshot of scenario s's q (get_q's layout) into buf and
        return its handle (snapshot_wait); the batch's attempts run on
        while it is copied."""
        if buf.nbytes < 8 * int(np.prod(self.grids[s].shape_q)):
            raise ValueError("snapshot buffer too small")
        h = C.c_void_p()
        self._check(self.L.magic_snapshot_begin(self.ctx[s], buf.ptr, C.byref(h)), "magic_snapshot_begin")
        return h

    def global_sums(self, s) -> np.ndarray:
        """Compensated FP64 sums of scenario s's state (MagicHIP.global_sums)."""
        out = np.zeros(self.grids[s].meqn)
        self._check(self.L.magic_global_sums(self.ctx[s], _ptr(out)), "magic_global_sums")
        return out

    def start(self, s, t, tend, dt, dtv, cflv, maxsteps, mthbc):
        self._check(self.L.magic_batch_start(self.h, s, t, tend, dt, (C.c_double * 5)(*dtv),
                                             (C.c_double * 4)(*cflv), maxsteps,
                                             (C.c_int * 6)(*mthbc)), "magic_batch_start")

    def run(self, max_attempts=100000, check_every=8):
        """Replays attempts; a failed scenario stops with its status in
        state(s)["status"] while the others go on."""
        done = C.c_int()
        self._check(self.L.magic_batch_run(self.h, max_attempts, check_every, C.byref(done)),
                    "magic_batch_run")
        return done.value

    def state(self, s):
        r = (C.c_double * 5)()
        i = (C.c_int * 6)()
        self._check(self.L.magic_batch_state(self.h, s, r, i), "magic_batch_state")
        return {"t": r[0], "dt": r[1], "cfl": r[2], "cflmax": r[3], "tend": r[4], "nsteps": i[0],
                "attempts": i[1], "rejects": i[2], "active": i[3], "status": i[4], "ndtr": i[5]}

    def set_tile_shapes(self, shapes):
        self._check(self.L.magic_batch_set_tile_shapes(self.h, (C.c_int * 3)(*shapes)),
                    "magic_batch_set_tile_shapes")

    def set_sweep_overlap(self, on: bool):
        """The sweeps' order of additions, as MagicHIP.set_sweep_overlap."""
        self._check(self.L.magic_batch_set_sweep_overlap(self.h, 1 if on else 0), "magic_batch_set_sweep_overlap")

    def tile_shapes(self):
        sh = (C.c_int * 3)()
        self._check(self.L.magic_batch_get_tile_shapes(self.h, sh), "magic_batch_get_tile_shapes")
        return list(sh)

    def tune(self, reps: int = 3):
        """Autotuner of the batch (all scenarios' tiles, current states and
        next dt); returns (shapes, ms[3][NSHAPE])."""
        sh = (C.c_int * 3)()
        ms = (C.c_float * (3 * TILE_NSHAPE))()
        self._check(self.L.magic_batch_tune(self.h, reps, sh, ms), "magic_batch_tune")
        return list(sh), [list(ms[d * TILE_NSHAPE:(d + 1) * TILE_NSHAPE]) for d in range(3)]
