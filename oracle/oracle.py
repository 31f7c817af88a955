"""ctypes front ends for the two CPU checkers of the MAGIC step.

TEST INFRASTRUCTURE ONLY (used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path):

  COracle    -- oracle/libmagic_oracle.so, the C restatement in
                oracle/magic_oracle.c (builds anywhere with gcc).
  Reference  -- oracle/_ref/libmagic_ref.so, the reference Fortran itself
                compiled from its own sources by oracle/ref/Makefile (exists
                only where it was built; tests skip it otherwise).

Arrays are numpy float64 of shape (nz+2mbc, ny+2mbc, nx+2mbc, ncomp), the
memory of the reference's Fortran q(meqn, i, j, k) / aux(maux, i, j, k).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libmagic_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmagic_ref.so")
# the same sources built for timing (-O3, FMA, x86-64-v4; oracle/ref/Makefile `speed`)
REF_SPEED_SO = os.path.join(HERE, "_ref", "libmagic_ref_speed.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class OrcGrid(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("meqn", C.c_int),
                ("mbc", C.c_int), ("maux", C.c_int), ("xlower", C.c_double),
                ("ylower", C.c_double), ("zlower", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double)]

    @classmethod
    def of(cls, g):
        return cls(g.mx, g.my, g.mz, g.meqn, g.mbc, g.maux, g.xlower, g.ylower, g.zlower,
                   g.dx, g.dy, g.dz)


class OrcParams(C.Structure):
    _fields_ = [("loadprof", C.c_int), ("nspecadv", C.c_int)] + [
        (n, C.c_double) for n in ("gamma", "grav", "dens", "pres", "dmolec0", "dtherm0",
                                  "forcemth", "omega", "amplitude", "propx", "propz", "xpos",
                                  "ypos", "zpos", "xwidth", "ywidth", "zwidth", "tcenter",
                                  "twidth", "vsrcx", "amp0", "amp1", "omega1", "prop1z", "phi",
                                  "z1width", "z1pos", "z1Gpos", "windscale1")]

    @classmethod
    def of(cls, p):
        o = cls()
        for name, _ in cls._fields_:
            setattr(o, name, getattr(p, name))
        return o


def build_oracle() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return ORACLE_SO


def _big_stack(fn):
    """The reference keeps its automatic arrays on the stack (flang), several
    MB per call at the workload sizes: run each call on a thread with a 1 GiB
    stack."""
    def wrapped(*a, **k):
        box = {}

        def run():
            try:
                box["r"] = fn(*a, **k)
            except BaseException as e:  # pragma: no cover - re-raised below
                box["e"] = e
        old = threading.stack_size(1 << 30)
        try:
            th = threading.Thread(target=run)
            th.start()
            th.join()
        finally:
            threading.stack_size(old)
        if "e" in box:
            raise box["e"]
        return box.get("r")
    return wrapped


class COracle:
    """The C restatement."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build_oracle()
        self.lib = L = C.CDLL(path)
        L.orc_step3.restype = C.c_int
        L.orc_step3.argtypes = [C.POINTER(OrcGrid), C.c_int, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_bc3.argtypes = [C.POINTER(OrcGrid), _dp, _dp, _ip]
        L.orc_b4step3.argtypes = [C.POINTER(OrcGrid), _dp]
        L.orc_src3.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcParams), _dp, _dp, C.c_double,
                               C.c_double, _ip, _ip, _ip]
        L.orc_claw3.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcParams), _dp, _dp, C.c_double,
                                _dp, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip, C.c_int, _ip]
        L.orc_rpn3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                               C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_rpt3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_rptt3.argtypes = L.orc_rpt3.argtypes
        L.orc_cflsupbee.restype = C.c_double
        L.orc_cflsupbee.argtypes = [C.c_double] * 5
        L.orc_suppow.restype = C.c_double
        L.orc_suppow.argtypes = [C.c_double] * 3

    def step3(self, g, q, aux, dt, nspecadv):
        qold = np.ascontiguousarray(q, dtype=np.float64)
        qnew = qold.copy()
        cfl = C.c_double(0.0)
        st = self.lib.orc_step3(C.byref(OrcGrid.of(g)), nspecadv, _ptr(qold), _ptr(qnew),
                                _ptr(np.ascontiguousarray(aux)), dt, C.byref(cfl))
        return qnew, cfl.value, st

    def bc3(self, g, q, aux, mthbc):
        m = (C.c_int * 6)(*mthbc)
        return self.lib.orc_bc3(C.byref(OrcGrid.of(g)), _ptr(q), _ptr(aux), m)

    def b4step3(self, g, q):
        self.lib.orc_b4step3(C.byref(OrcGrid.of(g)), _ptr(q))

    def src3(self, g, p, q, aux, t, dt, mthbc):
        m = (C.c_int * 6)(*mthbc)
        a, b = C.c_int(0), C.c_int(0)
        st = self.lib.orc_src3(C.byref(OrcGrid.of(g)), C.byref(OrcParams.of(p)), _ptr(q),
                               _ptr(aux), t, dt, m, C.byref(a), C.byref(b))
        return st, a.value, b.value

    def claw3(self, g, p, q, aux, tstart, tend, dtv, cflv, nv, method, mthbc, maxlog=64):
        dtv_a = (C.c_double * 5)(*dtv)
        cflv_a = (C.c_double * 4)(*cflv)
        nv_a = (C.c_int * 2)(*nv)
        te = C.c_double(tend)
        ldt = (C.c_double * maxlog)()
        lcfl = (C.c_double * maxlog)()
        lacc = (C.c_int * maxlog)()
        nlog = C.c_int(0)
        st = self.lib.orc_claw3(C.byref(OrcGrid.of(g)), C.byref(OrcParams.of(p)), _ptr(q),
                                _ptr(aux), tstart, C.byref(te), dtv_a, cflv_a, nv_a,
                                (C.c_int * 7)(*method), (C.c_int * 6)(*mthbc), ldt, lcfl, lacc,
                                maxlog, C.byref(nlog))
        log = [(ldt[i], lcfl[i], bool(lacc[i])) for i in range(nlog.value)]
        return st, te.value, list(dtv_a), list(cflv_a), list(nv_a), log

    def rpn3(self, ixyz, q1d, a19, a17, dt, dx, nspecadv):
        """q1d (N, meqn) with N = n + 2*mbc (mbc = 2)."""
        N, meqn = q1d.shape
        mbc = 2
        n = N - 2 * mbc
        fw = np.zeros((N, 3, meqn)); s = np.zeros((N, 3))
        am = np.zeros((N, meqn)); ap = np.zeros((N, meqn)); roe = np.zeros((8, N))
        st = self.lib.orc_rpn3(ixyz, meqn, mbc, n, nspecadv, _ptr(np.ascontiguousarray(q1d)),
                               _ptr(np.ascontiguousarray(a19)), _ptr(np.ascontiguousarray(a17)),
                               dt, dx, _ptr(fw), _ptr(s), _ptr(am), _ptr(ap), _ptr(roe))
        return st, fw, s, am, ap, roe


class Reference:
    """The reference Fortran, compiled from its own sources (oracle/_ref)."""

    _singleton = None

    @classmethod
    def available(cls, speed: bool = False) -> bool:
        return os.path.exists(REF_SPEED_SO if speed else REF_SO)

    @classmethod
    def get(cls, speed: bool = False):
        """The parity build, or (speed) the timing build; one per process."""
        if cls._singleton is None:
            cls._singleton = cls(REF_SPEED_SO if speed else REF_SO)
        return cls._singleton

    def __init__(self, path: str = REF_SO):
        self.lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        self._mpi = False

    def _common(self, name, ctype):
        return ctype.in_dll(self.lib, name)

    def setup(self, p, mthbc):
        """Bring up MPI (1 rank) and load the setprob commons (setprob.f:17-25,
        63-68) from a ProbData."""
        if not self._mpi:
            self.lib.refh_mpi_setup_((C.c_int * 6)(*mthbc))
            self._mpi = True
        opts = self._common("options_", C.c_int * 2)
        opts[0], opts[1] = p.loadprof, p.nspecadv
        par = self._common("param_", C.c_double * 8)
        par[:] = [p.gamma, p.gamma - 1.0, p.grav, p.dens, p.pres,
                  float(np.sqrt(p.pres * p.gamma / p.dens)), p.dtherm0, p.dmolec0]
        fo = self._common("forcing_", C.c_double * 8)
        speed = par[5]
        w_n = float(np.sqrt(((p.grav ** 2) / (speed ** 2)) * (p.gamma - 1.0)))
        w_o = float(np.sqrt((p.gamma ** 2) * (p.grav ** 2) / (4 * (speed ** 2))))
        fo[:] = [p.omega, p.amplitude, p.propx, p.propz, w_n, w_o, p.vsrcx, p.forcemth]
        ga = self._common("gauss_", C.c_double * 8)
        ga[:] = [p.xpos, p.ypos, p.zpos, p.xwidth, p.ywidth, p.zwidth, p.tcenter, p.twidth]
        wi = self._common("wind_", C.c_double * 9)
        wi[:] = [p.omega1, p.amp0, p.amp1, p.prop1z, p.z1width, p.z1pos, p.z1Gpos, p.phi,
                 p.windscale1]

    @staticmethod
    def _i(v):
        return C.byref(C.c_int(v))

    @staticmethod
    def _d(v):
        return C.byref(C.c_double(v))

    @_big_stack
    def step3(self, g, q, aux, dt):
        qold = np.ascontiguousarray(q, dtype=np.float64)
        qnew = qold.copy()
        cfl = C.c_double(0.0)
        i, d = self._i, self._d
        self.lib.refh_step3_(i(g.meqn), i(g.mbc), i(g.mx), i(g.my), i(g.mz), i(g.maux),
                             _ptr(qold), _ptr(qnew), _ptr(np.ascontiguousarray(aux)),
                             d(g.dx), d(g.dy), d(g.dz), d(dt), C.byref(cfl))
        return qnew, cfl.value

    @_big_stack
    def bc3(self, g, q, aux, mthbc, t=0.0, dt=0.0):
        i, d = self._i, self._d
        self.lib.bc3_(i(g.meqn), i(g.mbc), i(g.mx), i(g.my), i(g.mz), d(g.xlower), d(g.ylower),
                      d(g.zlower), d(g.dx), d(g.dy), d(g.dz), _ptr(q), i(g.maux), _ptr(aux),
                      d(t), d(dt), (C.c_int * 6)(*mthbc))

    @_big_stack
    def b4step3(self, g, q, aux):
        i, d = self._i, self._d
        self.lib.b4step3_(i(g.mbc), i(g.mx), i(g.my), i(g.mz), i(g.meqn), _ptr(q), d(g.xlower),
                          d(g.ylower), d(g.zlower), d(g.dx), d(g.dy), d(g.dz), d(0.0), d(0.0),
                          i(g.maux), _ptr(aux))

    @_big_stack
    def src3(self, g, q, aux, t, dt, mthbc):
        i, d = self._i, self._d
        self.lib.src3_(i(g.meqn), i(g.mbc), i(g.mx), i(g.my), i(g.mz), d(g.xlower), d(g.ylower),
                       d(g.zlower), d(g.dx), d(g.dy), d(g.dz), _ptr(q), i(g.maux), _ptr(aux),
                       d(t), d(dt), (C.c_int * 6)(*mthbc))

    @_big_stack
    def setaux(self, g):
        """The reference setaux (setaux.f:4-212); with loadprof = 1 it reads
        profile.data from the working directory."""
        aux = np.zeros(g.shape_aux)
        i, d = self._i, self._d
        self.lib.setaux_(i(g.mbc), i(g.mx), i(g.my), i(g.mz), d(g.xlower), d(g.ylower),
                         d(g.zlower), d(g.dx), d(g.dy), d(g.dz), i(g.maux), _ptr(aux))
        return aux

    @_big_stack
    def claw3(self, g, q, aux, tstart, tend, dtv, cflv, nv, method, mthbc):
        i, d = self._i, self._d
        dtv_a = (C.c_double * 5)(*dtv)
        cflv_a = (C.c_double * 4)(*cflv)
        nv_a = (C.c_int * 2)(*nv)
        te = C.c_double(tend)
        info = C.c_int(0)
        meth = (C.c_int * 7)(*method)
        mb = (C.c_int * 6)(*mthbc)
        self.lib.refh_claw3_(i(g.meqn), i(g.mbc), i(g.mx), i(g.my), i(g.mz), i(g.maux), _ptr(q),
                             _ptr(aux), d(g.xlower), d(g.ylower), d(g.zlower), d(g.dx), d(g.dy),
                             d(g.dz), d(tstart), C.byref(te), dtv_a, cflv_a, nv_a, meth, mb,
                             C.byref(info))
        return info.value, te.value, list(dtv_a), list(cflv_a), list(nv_a)

    @_big_stack
    def rpn3(self, ixyz, q1d, a_full, dt, dx):
        """Calls rpn3 (rpn3euX.f:3) directly; a_full is (N, maux) aux pencil."""
        N, meqn = q1d.shape
        maux = a_full.shape[1]
        mbc = 2
        n = N - 2 * mbc
        self.lib.refh_set_comxyzt_(self._d(dt), self._d(dx), self._d(dx), self._d(dx), self._d(0.0))
        fw = np.zeros((N, 3, meqn)); s = np.zeros((N, 3))
        am = np.zeros((N, meqn)); ap = np.zeros((N, meqn))
        ql = np.ascontiguousarray(q1d)
        au = np.ascontiguousarray(a_full)
        i = self._i
        self.lib.rpn3_(i(ixyz), i(n), i(meqn), i(3), i(maux), i(mbc), i(n), _ptr(ql), _ptr(ql),
                       _ptr(au), _ptr(au), _ptr(fw), _ptr(s), _ptr(am), _ptr(ap))
        roe = np.frombuffer(self._common("comroe_", C.c_double * (8 * 3002)), dtype=np.float64)
        roe = roe.reshape(8, 3002)[:, :N].copy()   # index -1.. maps to 0..
        return fw, s, am, ap, roe
