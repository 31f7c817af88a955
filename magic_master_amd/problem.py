"""Problem setup on the host: input decks, grid, FP64 background (aux) and
initial state (q), in the reference's array layouts.

Layouts (Fortran column-major, as the reference's claw3ez_mpi.f:33-34):
    q  (meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)  REAL*8
    aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)  REAL*8
Here they are numpy arrays of shape (nz+2mbc, ny+2mbc, nx+2mbc, meqn) in C
order, which is the same memory.

Transcendentals use Python's ``math`` module (the C library), one value per
vertical level, so a background built here is bit-identical to the reference
setaux on the same host.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field, asdict
from typing import List, Optional, Sequence

import numpy as np

# aux slots (1-based, setaux.f:55-187)
AUX_RHO, AUX_MX, AUX_MY, AUX_MZ, AUX_E = 1, 2, 3, 4, 5
AUX_NU, AUX_KAPPA, AUX_VGRAV, AUX_R, AUX_GAMMA = 11, 12, 17, 18, 19


def _num(tok: str) -> str:
    return re.sub(r"[dD]", "e", tok)


def _tokens(path):
    """Fortran list-directed records: yields the value tokens of each
    non-blank line (trailing comments are words that are not numbers)."""
    with open(path) as f:
        for line in f:
            toks = line.split()
            if toks:
                yield toks


@dataclass
class ClawData:
    """claw3ez.data, in the order mpilib/claw3ez_mpi_driver.f:81-145 reads it."""
    mx: int = 64
    my: int = 64
    mz: int = 48
    nout: int = 200
    outstyle: int = 3
    tfinal: float = 0.0
    tout: List[float] = field(default_factory=list)
    nstepout: int = 200
    nstop: int = 200
    dtv: List[float] = field(default_factory=lambda: [0.1, 10.0, 0.0, 0.0, 0.0])
    cflv: List[float] = field(default_factory=lambda: [0.95, 0.9, 0.0, 0.0])
    nv: List[int] = field(default_factory=lambda: [10000, 0])
    method: List[int] = field(default_factory=lambda: [1, 2, 22, 0, 1, 0, 32])
    meqn: int = 6
    mwaves: int = 3
    mthlim: List[int] = field(default_factory=lambda: [0, 0, 0])
    t0: float = 0.0
    xlower: float = 0.0
    xupper: float = 256e3
    ylower: float = 0.0
    yupper: float = 256e3
    zlower: float = 0.0
    zupper: float = 96e3
    mbc: int = 2
    mthbc: List[int] = field(default_factory=lambda: [2, 2, 2, 2, 3, 1])
    lx: int = 1
    ly: int = 1
    lz: int = 1

    @property
    def maux(self) -> int:
        return self.method[6]

    @classmethod
    def read(cls, path) -> "ClawData":
        it = _tokens(path)
        nxt = lambda: next(it)
        d = cls()
        d.mx = int(nxt()[0]); d.my = int(nxt()[0]); d.mz = int(nxt()[0])
        d.nout = int(nxt()[0]); d.outstyle = int(nxt()[0])
        if d.outstyle == 1:
            d.tfinal = float(_num(nxt()[0])); d.nstepout = 1
        elif d.outstyle == 2:
            vals: List[float] = []
            while len(vals) < d.nout:
                vals += [float(_num(t)) for t in nxt() if _isnum(t)]
            d.tout = vals[: d.nout]; d.nstepout = 1
        elif d.outstyle == 3:
            t = nxt(); d.nstepout = int(t[0]); d.nstop = int(t[1]); d.nout = d.nstop
        d.dtv = [float(_num(nxt()[0])), float(_num(nxt()[0])), 0.0, 0.0, 0.0]
        d.cflv = [float(_num(nxt()[0])), float(_num(nxt()[0])), 0.0, 0.0]
        d.nv = [int(nxt()[0]), 0]
        d.method = [int(nxt()[0]) for _ in range(7)]
        d.meqn = int(nxt()[0]); d.mwaves = int(nxt()[0])
        lim: List[int] = []
        while len(lim) < d.mwaves:
            lim += [int(t) for t in nxt() if re.fullmatch(r"[+-]?\d+", t)][: d.mwaves - len(lim)]
        d.mthlim = lim
        d.t0 = float(_num(nxt()[0]))
        d.xlower = float(_num(nxt()[0])); d.xupper = float(_num(nxt()[0]))
        d.ylower = float(_num(nxt()[0])); d.yupper = float(_num(nxt()[0]))
        d.zlower = float(_num(nxt()[0])); d.zupper = float(_num(nxt()[0]))
        d.mbc = int(nxt()[0])
        d.mthbc = [int(nxt()[0]) for _ in range(6)]
        d.lx = int(nxt()[0]); d.ly = int(nxt()[0]); d.lz = int(nxt()[0])
        return d

    def write(self, path) -> None:
        L = [f"{self.mx} mx", f"{self.my} my", f"{self.mz} mz", f"{self.nout} nout",
             f"{self.outstyle} outstyle"]
        if self.outstyle == 1:
            L.append(f"{self.tfinal!r} tfinal")
        elif self.outstyle == 2:
            L.append(" ".join(repr(t) for t in self.tout))
        else:
            L.append(f"{self.nstepout} {self.nstop} nstepout nstop")
        L += [f"{self.dtv[0]!r} dtv1", f"{self.dtv[1]!r} dtv2", f"{self.cflv[0]!r} cflv1",
              f"{self.cflv[1]!r} cflv2", f"{self.nv[0]} nv1"]
        L += [f"{m} method" for m in self.method]
        L += [f"{self.meqn} meqn", f"{self.mwaves} mwaves",
              " ".join(str(m) for m in self.mthlim) + " mthlim",
              f"{self.t0!r} t0", f"{self.xlower!r} xlower", f"{self.xupper!r} xupper",
              f"{self.ylower!r} ylower", f"{self.yupper!r} yupper",
              f"{self.zlower!r} zlower", f"{self.zupper!r} zupper", f"{self.mbc} mbc"]
        L += [f"{m} mthbc" for m in self.mthbc]
        L += [f"{self.lx} lx", f"{self.ly} ly", f"{self.lz} lz"]
        with open(path, "w") as f:
            f.write("\n".join(L) + "\n")


def _isnum(t: str) -> bool:
    try:
        float(_num(t))
        return True
    except ValueError:
        return False


_PROB_FIELDS = ["loadprof", "nspecadv", "gamma", "grav", "dens", "pres", "dmolec0",
                "dtherm0", "forcemth", "omega", "amplitude", "propx", "propz", "xpos",
                "ypos", "zpos", "xwidth", "ywidth", "zwidth", "tcenter", "twidth", "vsrcx",
                "amp0", "amp1", "omega1", "prop1z", "phi", "z1width", "z1pos", "z1Gpos",
                "windscale1"]


@dataclass
class ProbData:
    """setprob.data, in the order setprob.f:31-61 reads it."""
    loadprof: int = 0
    nspecadv: int = 1
    gamma: float = 1.4
    grav: float = 9.81
    dens: float = 1.2
    pres: float = 1.2 * 287.0 * 250.0
    dmolec0: float = 1e-5
    dtherm0: float = 1e-5 / 0.7
    forcemth: float = 3.0
    omega: float = 0.0
    amplitude: float = 20.0
    propx: float = 0.0
    propz: float = 0.0
    xpos: float = 128e3
    ypos: float = 128e3
    zpos: float = 6e3
    xwidth: float = 6e3
    ywidth: float = 6e3
    zwidth: float = 3e3
    tcenter: float = 600.0
    twidth: float = 600.0
    vsrcx: float = 0.0
    amp0: float = 0.0
    amp1: float = 0.0
    omega1: float = 0.0
    prop1z: float = 0.0
    phi: float = 0.0
    z1width: float = 0.0
    z1pos: float = 0.0
    z1Gpos: float = 0.0
    windscale1: float = 0.0

    @property
    def gamma1(self) -> float:
        return self.gamma - 1.0

    @property
    def speed(self) -> float:          # setprob.f:65
        return math.sqrt(self.pres * self.gamma / self.dens)

    @classmethod
    def read(cls, path) -> "ProbData":
        it = _tokens(path)
        d = cls()
        for name in _PROB_FIELDS:
            tok = next(it)[0]
            if name in ("loadprof", "nspecadv"):
                setattr(d, name, int(tok))
            else:
                setattr(d, name, float(_num(tok)))
        return d

    def write(self, path) -> None:
        with open(path, "w") as f:
            for name in _PROB_FIELDS:
                v = getattr(self, name)
                f.write(f"{v!r} {name}\n" if isinstance(v, float) else f"{v} {name}\n")

    def as_list(self) -> List[float]:
        return [float(getattr(self, n)) for n in _PROB_FIELDS]


@dataclass
class Grid:
    mx: int
    my: int
    mz: int
    meqn: int = 6
    mbc: int = 2
    maux: int = 32
    xlower: float = 0.0
    ylower: float = 0.0
    zlower: float = 0.0
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0

    @classmethod
    def from_claw(cls, c: ClawData) -> "Grid":
        # mpilib/claw3ez_mpi.f:59-61: dx = (xupper-xlower)/float(mx), float() is REAL*4
        f32 = lambda n: float(np.float32(n))
        return cls(c.mx, c.my, c.mz, c.meqn, c.mbc, c.maux, c.xlower, c.ylower, c.zlower,
                   (c.xupper - c.xlower) / f32(c.mx), (c.yupper - c.ylower) / f32(c.my),
                   (c.zupper - c.zlower) / f32(c.mz))

    @property
    def shape_q(self):
        return (self.mz + 2 * self.mbc, self.my + 2 * self.mbc, self.mx + 2 * self.mbc, self.meqn)

    @property
    def shape_aux(self):
        return (self.mz + 2 * self.mbc, self.my + 2 * self.mbc, self.mx + 2 * self.mbc, self.maux)

    @property
    def ncell(self) -> int:
        return self.mx * self.my * self.mz

    def zcell(self, k: int) -> float:      # 1-based Fortran index, may be a ghost
        return self.zlower + (k - 0.5) * self.dz

    def xcell(self, i: int) -> float:
        return self.xlower + (i - 0.5) * self.dx

    def ycell(self, j: int) -> float:
        return self.ylower + (j - 0.5) * self.dy


def _vgrav(col):
    """aux(17) from setaux.f:176-187 on a vertical column dict of arrays
    (index 0 = k = 1-mbc).  Entry 0 is never read by the step; it is set to
    the value of entry 1."""
    n = len(col["rho"])
    out = np.zeros(n)
    for k in range(1, n):
        pl = (col["gam"][k - 1] - 1.0) * (col["E"][k - 1] - 0.5 * (col["mx"][k - 1] ** 2 + 0.0 ** 2 + 0.0 ** 2) / col["rho"][k - 1])
        pr = (col["gam"][k] - 1.0) * (col["E"][k] - 0.5 * (col["mx"][k] ** 2 + 0.0 ** 2 + 0.0 ** 2) / col["rho"][k])
        out[k] = (pr - pl) / (0.5 * (col["rho"][k] + col["rho"][k - 1]))
    out[0] = out[1]
    return out


def _fill_aux_from_column(g: Grid, col) -> np.ndarray:
    tmpl = np.zeros((g.shape_aux[0], g.maux))
    for key, slot in (("rho", AUX_RHO), ("mx", AUX_MX), ("E", AUX_E), ("nu", AUX_NU),
                      ("kap", AUX_KAPPA), ("vg", AUX_VGRAV), ("R", AUX_R), ("gam", AUX_GAMMA)):
        tmpl[:, slot - 1] = np.asarray(col[key])
    aux = np.empty(g.shape_aux)
    aux[...] = tmpl[:, None, None, :]          # one contiguous pass
    return aux


def windfnx(p: ProbData, z: float, scale: float) -> float:
    """The analytic background wind in x (fcns.f:100-116) at t = 0, as setaux
    calls it (setaux.f:59-60, 131-132 pass the integer constant 0 for the
    double precision t; this restatement takes t = 0):
    amp1 cos(prop1z (z - z1pos) - phi) exp(-(z - z1Gpos)**2 / (2 z1width**2)),
    times scale when windscale1 = 1, plus amp0."""
    f = np.float64                      # IEEE division: z1width = 0 gives exp(-inf) = 0
    w = f(p.amp1)
    w = w * math.cos(f(p.prop1z) * (f(z) - p.z1pos) - f(p.omega1) * 0.0 - p.phi)
    dz = f(z) - p.z1Gpos
    with np.errstate(divide="ignore", invalid="ignore"):
        arg = f(-0.5) * ((dz * dz) / (f(p.z1width) * p.z1width))
    w = w * math.exp(arg)
    if p.windscale1 == 1.0:
        w = w * scale
    return float(w + p.amp0)


def _wind_scale(rho1: float, rho: float) -> float:
    """setaux.f:58, 131: (aux(1,i,j,1)/aux(1,i,j,k))**0.5 with the C library
    pow, as the reference evaluates it.  In the ghost layers below k = 1 the
    reference reads aux(1,i,j,1) before its loop has set it; rho1 = 0 there
    (a zeroed aux array), so scale = 0."""
    return math.pow(rho1 / rho, 0.5)


def setaux_isothermal(g: Grid, p: ProbData) -> np.ndarray:
    """The formulas of setaux.f:46-87 + 176-187 (its loadprof = 0 branch,
    which the reference never reaches: setaux.f:38-43 sends loadprof <= 0 to
    the aux(1) = 1 special case) without winds: isothermal background
    rho = dens*exp(-z*gamma*grav/speed**2) (fcns.f:22-45), R = 287,
    nu = dmolec0*rho(z_1)/rho, kappa = gamma*dtherm0*rho(z_1)/rho.  The
    analytic winds are restated for the reachable profile branch
    (setaux_profile), not for this one."""
    if p.amp0 != 0.0 or p.amp1 != 0.0:
        raise NotImplementedError("analytic winds on the isothermal background (an unreachable branch of "
                                  "setaux.f) are not restated; use loadprof = 1")
    nzp = g.mz + 2 * g.mbc
    s2 = p.speed ** 2
    dens = lambda z: p.dens * math.exp(-(z * p.gamma * p.grav / s2))
    press = lambda z: p.pres * math.exp(-(z * p.gamma * p.grav / s2))
    rho1 = dens(g.zcell(1))
    col = {k: np.zeros(nzp) for k in ("rho", "mx", "E", "nu", "kap", "R", "gam")}
    for kk in range(nzp):
        k = kk + 1 - g.mbc
        z = g.zcell(k)
        rho = dens(z)
        col["rho"][kk] = rho
        col["E"][kk] = press(z) / p.gamma1 + 0.5 * (0.0 ** 2 + 0.0 ** 2 + 0.0 ** 2) * rho
        col["gam"][kk] = p.gamma
        col["R"][kk] = 287.0
        col["nu"][kk] = p.dmolec0 * (rho1 / rho)
        col["kap"][kk] = (p.dtherm0 * (rho1 / rho)) * p.gamma
    col["vg"] = _vgrav(col)
    return _fill_aux_from_column(g, col)


def read_profile(path: str) -> np.ndarray:
    """profile.data (setaux.f:90-93): one row per height, columns height (km),
    O, N2, O2 number densities, density, temperature (K), H and electron
    number densities, x and y wind (m/s); shape (rows, 10)."""
    rows = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if rows.shape[1] < 10:
        raise ValueError(f"{path}: {rows.shape[1]} columns, profile.data has 10")
    return rows[:, :10]


# Fortran REAL (single precision) literals of setaux.f's diffusivity and
# kinetic-energy formulas, promoted to double as the reference promotes them
_F4 = lambda v: float(np.float32(v))


def setaux_profile(g: Grid, p: ProbData, prof: np.ndarray, cgs: bool = True) -> np.ndarray:
    """setaux.f:38-212 for loadprof = 1 (the MSIS-style vertical profile from
    profile.data): per cell centre height the row with exactly that
    height (in km, rows read in order as the reference reads them; a missing
    height is an error, setaux.f:94-103), then gas constant and gamma from
    the O/N2/O2 composition, density from the number densities (cgs = True:
    the file is in CGS, setaux.f:20-21), energy from the temperature,
    composition ratios in slots 6-10, molecular and thermal diffusivities
    (slots 11, 12; held at the 400th layer above it) and the virtual gravity
    of slot 17 from the pressure jumps (setaux.f:176-187; the lowest ghost
    layer has none, it is 0 here).  The x and y winds come from the profile,
    or with amp0 or amp1 set from the analytic x wind: rho*windfnx + amp0*rho
    (amp0 enters twice, as the reference adds it in windfnx and again,
    setaux.f:128-133) and a y slot of 0 (setaux.f:134).  Same operations in
    the same order as the reference, including its single-precision
    literals."""
    prof = np.asarray(prof, dtype=np.float64)
    winds = not (p.amp0 == 0.0 and p.amp1 == 0.0)
    rho1 = 0.0   # aux(1,i,j,1), set when the loop reaches k = 1
    nzp = g.mz + 2 * g.mbc
    cnv = 1.0e3 if cgs else 1.0e-3
    col = np.zeros((nzp, g.maux))
    pos = 0
    c343, c403, c390, c759, c069 = _F4(3.43), _F4(4.03), _F4(3.90), _F4(75.9), _F4(0.69)
    for kk in range(nzp):
        k = kk + 1 - g.mbc
        target = g.zcell(k) / 1.0e3
        while pos < len(prof) and prof[pos, 0] < target:
            pos += 1
        if pos == len(prof) or prof[pos, 0] > target:
            raise ValueError(f"profile.data has no row at height {target} km (cell centres must match exactly)")
        height, dox, dnit2, dox2, _, temp, dhyd, dne, windx, windy = (float(v) for v in prof[pos])
        pos += 1
        a = col[kk]
        a[17] = 8.31e3 / ((dox * 16.0 + dnit2 * 28.0 + dox2 * 32.0) / (dox + dnit2 + dox2))
        a[18] = (7.0 * (dox2 + dnit2) + 5.0 * dox) / (5.0 * (dox2 + dnit2) + 3.0 * dox)
        roverm, gamma = a[17], a[18]
        gamma1 = gamma - 1.0
        rho = (dox * 16.0 + dox2 * 32.0 + dnit2 * 28.0) * cnv
        rho = rho / 6.022e23
        a[0] = rho
        if k == 1:
            rho1 = rho
        if not winds:
            a[1] = rho * windx
            a[2] = rho * windy
        else:
            a[1] = rho * windfnx(p, g.zcell(k), _wind_scale(rho1, rho))
            a[1] = a[1] + p.amp0 * rho
            a[2] = 0.0
        a[3] = 0.0
        a[4] = rho * temp * roverm / gamma1 + 0.5 * (a[1] ** 2) / rho + 0.5 * (a[2] ** 2) / rho
        a[5] = dox * 16.0 * cnv / (6.022e23 * rho)
        a[6] = dox2 * 32.0 * cnv / (6.022e23 * rho)
        a[7] = dne / (dox + dox2 + dnit2)
        a[8] = dhyd / (dox + dox2 + dnit2)
        a[9] = 0.0
        cp = 1.0 / (1.0 - 1.0 / gamma) * roverm
        dmolec = (c343 * dnit2 + c403 * dox2 + c390 * dox) * 1e-7
        dmolec = dmolec * (temp ** c069) / (dox + dnit2 + dox2)
        dtherm = (56.0 * (dnit2 + dox2) + c759 * dox) * 1e-5
        dtherm = dtherm * (temp ** c069) / (dox + dnit2 + dox2)
        a[10] = dmolec * (1.0 / rho)
        a[11] = (dtherm * (1.0 / (rho * cp))) * gamma
        if k > 400:
            a[10] = col[400 + g.mbc - 1, 10]
            a[11] = col[400 + g.mbc - 1, 11]
        if kk > 0:
            b = col[kk - 1]
            auxpl = (b[18] - 1.0) * (b[4] - 0.5 * (b[1] ** 2 + b[2] ** 2 + b[3] ** 2) / b[0])
            auxpr = (a[18] - 1.0) * (a[4] - 0.5 * (a[1] ** 2 + a[2] ** 2 + a[3] ** 2) / a[0])
            a[16] = (auxpr - auxpl) / (0.5 * (a[0] + b[0]))
    aux = np.empty(g.shape_aux)
    aux[...] = col[:, None, None, :]
    return aux


def temperature_standard(z: float, ts: float = 288.0, lapse: float = 6.5) -> float:
    """Analytic temperature column (K) of the full-domain workload:
    ts (288 K) at the surface, -lapse (6.5 K/km) to 11 km, isothermal to
    20 km, +1 K/km to 47 km, then a tanh rise to 1000 K that is continuous at
    47 km and within 2% of 1000 K above 150 km.  Below the surface (ghost
    cells) the lapse rate continues."""
    zk = z / 1000.0
    if zk <= 11.0:
        return ts - lapse * zk
    t11 = ts - lapse * 11.0
    if zk <= 20.0:
        return t11
    if zk <= 47.0:
        return t11 + (zk - 20.0)
    t47 = t11 + 27.0
    zc, w = 110.0, 20.0
    th0 = math.tanh((47.0 - zc) / w)
    return t47 + (1000.0 - t47) * (math.tanh((zk - zc) / w) - th0) / (1.0 - th0)


def setaux_column(g: Grid, gamma: float = 1.4, grav: float = 9.81, R: float = 287.0,
                  p0: float = 101325.0, mu: float = 1.8e-5, prandtl: float = 0.7,
                  temperature=temperature_standard, as_column: bool = False) -> np.ndarray:
    """Horizontally uniform FP64 background from a temperature column T(z):
    pressure integrated with the trapezoidal hydrostatic balance
        p_k - p_{k-1} = -g dz (rho_k + rho_{k-1})/2,  rho = p/(R T),
    so that aux(17) = (p_k - p_{k-1}) / ((rho_k + rho_{k-1})/2) is -g dz to
    rounding (the well-balanced form rpn3 expects, rpn3euX.f:202-208).
    nu = mu/rho and kappa = gamma*mu/(prandtl*rho) (kinematic, growing with
    height).  p0 is the pressure at z = 0.  as_column returns only the
    column, aux(1:maux, k) for every k, shape (mz+2mbc, maux)."""
    nzp = g.mz + 2 * g.mbc
    zs = [g.zcell(kk + 1 - g.mbc) for kk in range(nzp)]
    T = [temperature(z) for z in zs]
    a = grav * g.dz / (2.0 * R)
    # anchor: first cell above z = 0, from p0 at z = 0 with half a cell
    k1 = g.mbc  # index of k = 1
    p = [0.0] * nzp
    dz0 = zs[k1] - 0.0
    p[k1] = p0 * (1.0 - grav * dz0 / (2.0 * R * temperature(0.0))) / (1.0 + grav * dz0 / (2.0 * R * T[k1]))
    for kk in range(k1 + 1, nzp):
        p[kk] = p[kk - 1] * (1.0 - a / T[kk - 1]) / (1.0 + a / T[kk])
    for kk in range(k1 - 1, -1, -1):
        p[kk] = p[kk + 1] * (1.0 + a / T[kk + 1]) / (1.0 - a / T[kk])
    col = {k: np.zeros(nzp) for k in ("rho", "mx", "E", "nu", "kap", "R", "gam")}
    for kk in range(nzp):
        rho = p[kk] / (R * T[kk])
        col["rho"][kk] = rho
        col["E"][kk] = p[kk] / (gamma - 1.0)
        col["gam"][kk] = gamma
        col["R"][kk] = R
        col["nu"][kk] = mu / rho
        col["kap"][kk] = gamma * mu / (prandtl * rho)
    col["vg"] = _vgrav(col)
    if as_column:
        return _column_array(g, col)
    return _fill_aux_from_column(g, col)


def _column_array(g: Grid, col) -> np.ndarray:
    """The column as aux(1:maux, k), k = 1-mbc..mz+mbc: shape (mz+2mbc, maux)."""
    out = np.zeros((g.mz + 2 * g.mbc, g.maux))
    for key, slot in (("rho", AUX_RHO), ("mx", AUX_MX), ("E", AUX_E), ("nu", AUX_NU),
                      ("kap", AUX_KAPPA), ("vg", AUX_VGRAV), ("R", AUX_R), ("gam", AUX_GAMMA)):
        out[:, slot - 1] = np.asarray(col[key])
    return out


def qinit(g: Grid, aux: np.ndarray, noise: float = 0.0, seed: int = 0,
          tracer_blob: Optional[dict] = None) -> np.ndarray:
    """qinit.f:74-82: q = aux (fluid at rest on the background).  With
    noise > 0 the density gets a relative white-noise perturbation
    noise*U(-1,1) from numpy's PCG64 (seed), applied with the formula of
    qinit.f:56-66 (momenta scaled, kinetic energy replaced).  tracer_blob
    (keys z0, width, amp) sets tracer 6 = amp*exp(-0.5*((z-z0)/width)**2)."""
    q = np.array(aux[..., : g.meqn], copy=True)
    if noise > 0.0:
        rng = np.random.Generator(np.random.PCG64(seed))
        tweak = noise * 2.0 * (rng.random(q.shape[:3]) - 0.5)
        a = aux
        q[..., 0] = a[..., 0] * (1.0 + tweak)
        tm = q[..., 0] / a[..., 0]
        q[..., 1] = tm * a[..., 1]
        q[..., 2] = tm * a[..., 2]
        q[..., 3] = tm * a[..., 3]
        q[..., 4] = (a[..., 4] - 0.5 * (a[..., 1] ** 2 + a[..., 2] ** 2 + a[..., 3] ** 2) / a[..., 0]
                     + 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / q[..., 0])
    if tracer_blob is not None and g.meqn > 5:
        for kk in range(q.shape[0]):
            z = g.zcell(kk + 1 - g.mbc)
            q[kk, :, :, 5] = tracer_blob["amp"] * math.exp(-0.5 * ((z - tracer_blob["z0"]) / tracer_blob["width"]) ** 2)
    return q


# ---------------------------------------------------------------------------
# Named workloads

def config_plain(mx=64, my=64, mz=48, noise=1e-6, seed=1, meqn=6):
    """Plain MAGIC step on an isothermal 250 K atmosphere: 64x64x48 cells,
    dx = dy = 4 km, dz = 2 km, gamma 1.4, R 287, g 9.81, surface rho 1.2,
    mu = 1e-5 constant (Prandtl 0.7), Gaussian plume forcing (forcemth 3,
    6 km / 600 s, centred at the domain middle and t = 600 s, 20 mm/10 min),
    periodic sides, wall bottom, extrapolation top, CFL 0.9, tracer advected
    (meqn - 5 tracers, 6 by default)."""
    c = ClawData(mx=mx, my=my, mz=mz, outstyle=3, nstepout=200, nstop=200, nout=200,
                 xupper=4e3 * mx, yupper=4e3 * my, zupper=2e3 * mz,
                 mthbc=[2, 2, 2, 2, 3, 1], meqn=meqn)
    c.method[3] = 0
    p = ProbData()
    p.pres = p.dens * 287.0 * 250.0
    g = Grid.from_claw(c)
    rho1 = p.dens * math.exp(-(g.zcell(1) * p.gamma * p.grav / p.speed ** 2))
    p.dmolec0 = 1e-5 / rho1
    p.dtherm0 = p.dmolec0 / 0.7
    p.xpos = 0.5 * (c.xlower + c.xupper)
    p.ypos = 0.5 * (c.ylower + c.yupper)
    aux = setaux_isothermal(g, p)
    q = qinit(g, aux, noise=noise, seed=seed)
    return c, p, g, aux, q


def config_full(mx=250, my=200, mz=300, tracer=True):
    """Full-domain FP64 workload: 250x200x300 cells over 500x400x300 km
    (2x2x1 km), standard temperature column (temperature_standard), density
    integrated hydrostatically in FP64, mu = 1.8e-5 (Prandtl 0.7), plume at
    (250, 200) km and t = 2400 s, extrapolation sides and top, wall bottom,
    one advected tracer initialised as a Gaussian blob at 10 km."""
    c = ClawData(mx=mx, my=my, mz=mz, outstyle=3, nstepout=500, nstop=500, nout=500,
                 xupper=2e3 * mx, yupper=2e3 * my, zupper=1e3 * mz,
                 mthbc=[1, 1, 1, 1, 3, 1], meqn=6)
    c.method[3] = 0
    p = ProbData(loadprof=2, forcemth=3.0, amplitude=20.0, xpos=0.5 * 2e3 * mx,
                 ypos=0.5 * 2e3 * my, xwidth=6e3, ywidth=6e3, tcenter=2400.0, twidth=600.0,
                 dmolec0=0.0, dtherm0=0.0, pres=101325.0)
    g = Grid.from_claw(c)
    aux = setaux_column(g)
    q = qinit(g, aux, tracer_blob=dict(z0=10e3, width=2e3, amp=1.0) if tracer else None)
    return c, p, g, aux, q


def ensemble(n: int = 64, seed: int = 2024, nmin=(64, 64, 96), nmax=(200, 200, 300)):
    """Seeded scenarios of the batched ensemble workload: nx, ny uniform in
    [64, 200], nz uniform in [96, 300]; dx = dy = 2 km, dz = 1 km; the
    full-domain background with surface temperature uniform in [280, 305] K
    and lapse rate uniform in [5, 8] K/km; plume centre uniform in the
    interior 60 % of the domain, half-width uniform in [3, 10] km, time
    half-width uniform in [300, 900] s, amplitude log-uniform in [2, 60]
    mm/10 min, peak time uniform in [0, 600] s; extrapolation sides and top,
    wall bottom; one tracer; the fluid starts at rest (q = aux, qinit without
    noise).  Returns a list of (ClawData, ProbData, Grid, aux column
    (mz+2mbc, maux))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        mx, my = (int(rng.integers(nmin[0], nmax[0] + 1)), int(rng.integers(nmin[1], nmax[1] + 1)))
        mz = int(rng.integers(nmin[2], nmax[2] + 1))
        ts = float(rng.uniform(280.0, 305.0))
        lapse = float(rng.uniform(5.0, 8.0))
        c = ClawData(mx=mx, my=my, mz=mz, outstyle=3, nstepout=300, nstop=300, nout=300,
                     xupper=2e3 * mx, yupper=2e3 * my, zupper=1e3 * mz,
                     mthbc=[1, 1, 1, 1, 3, 1], meqn=6)
        c.method[3] = 0
        fx, fy = rng.uniform(0.2, 0.8, size=2)
        width = float(rng.uniform(3e3, 10e3))
        p = ProbData(loadprof=2, forcemth=3.0, amplitude=float(np.exp(rng.uniform(np.log(2.0), np.log(60.0)))),
                     xpos=float(fx * 2e3 * mx), ypos=float(fy * 2e3 * my), xwidth=width, ywidth=width,
                     tcenter=float(rng.uniform(0.0, 600.0)), twidth=float(rng.uniform(300.0, 900.0)),
                     dmolec0=0.0, dtherm0=0.0, pres=101325.0)
        g = Grid.from_claw(c)
        col = setaux_column(g, temperature=lambda z, ts=ts, lapse=lapse: temperature_standard(z, ts, lapse),
                            as_column=True)
        out.append((c, p, g, col))
    return out
