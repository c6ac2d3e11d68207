"""Hydro parameters of evora, restated for the host side of the HIP path.

`HydroParams` holds the namelist `/para/` keys that the hydro step reads
(source/parameter.f90:30-82, defaults as there) and `derive()` restates
`compparameter` (source/parameter.f90:164-265) for the constants the step
uses.  In the Fortran drop-in these values come straight from module `global`
(source/global.f90:45-86); this module exists so that Python callers (tests,
bench) hand the C-ABI exactly what the Fortran host would.

`to_namelist()` writes a parameter file the reference reads with
`readparameter` (source/parameter.f90:84-96).
"""
from __future__ import annotations

import dataclasses
import struct


def _f32(x: float) -> float:
    """A default-kind (single precision) real literal promoted to real(8)."""
    return struct.unpack("f", struct.pack("f", x))[0]


@dataclasses.dataclass
class HydroParams:
    nx: int = 32
    ny: int = 32
    nz: int = 32
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    zmin: float = 0.0
    zmax: float = 1.0
    boundtype: int = 2
    solver: int = 2
    limiter: int = 1
    secondorder: bool = True
    cfl: float = 0.5
    g: float = _f32(1.4)  # parameter.f90:46 assigns the single-precision literal 1.4
    atomic: bool = False
    dualenergy: bool = False
    de_thresh: float = 0.0
    tstart: float = 0.0
    tend: float = 1.0
    nsnap: int = 100  # output count over [tstart, tend] (parameter.f90:55)
    conduction: bool = False  # parameter.f90:78
    c_cond: float = _f32(0.1)
    simplekappa: bool = False
    gravity: bool = False  # parameter.f90:57
    supercomoving: bool = False  # parameter.f90:61-64
    hydro: bool = True  # parameter.f90:44: .false. skips hydrostep and sweep (main.f90:119,155)
    omega0: float = 1.0
    fb: float = 1.0

    # ---- parameter.f90:164-265 ----
    def derive(self) -> "Derived":
        g = _f32(5.0 / 3.0) if self.atomic else self.g  # parameter.f90:169-171
        onedim = self.ny == 1 and self.nz == 1
        twodim = self.nz == 1
        xmin, xmax, ymin, ymax, zmin, zmax = self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax
        dx = (xmax - xmin) / self.nx
        dy = (ymax - ymin) / self.ny
        dz = (zmax - zmin) / self.nz
        if onedim:
            dy = dx
            dz = dx
            ymin, ymax = 0.0, dy
            zmin, zmax = 0.0, dz
            dy = (ymax - ymin) / self.ny
            dz = (zmax - zmin) / self.nz
        if twodim:
            dz = dx
            zmin, zmax = 0.0, dz
            dz = (zmax - zmin) / self.nz
        g1 = g - 1.0
        # units.f90:26-65 without supercomoving expansion: l0 = t0 = d0 = cosmo = 1,
        # m0 = kb (global.f90:103); 37.8 / 40.0 is a single-precision quotient
        kb = 1.3806200e-16
        temp0 = (1.0 ** 2.0) / (1.0 ** 2.0) / (1.0 ** 2.0) * kb / kb
        kappa_coef = 4.6e13 * 1e-20 / _f32(_f32(37.8) / _f32(40.0))
        kappa0 = (1.0 ** 5.0) * (1.0 ** 3.0) / 1.0 / (1.0 ** 4.0) * 0.5 * kappa_coef
        # gravity and the Poisson solver (parameter.f90:235-241,254-263); pi is
        # the single-precision literal of global.f90:98 held in a real(8)
        pi = _f32(3.1415927)
        grav = (1.5 * self.omega0 if self.supercomoving else 8.38553911e-7) if self.gravity else 0.0
        kfx, kfy, kfz = 2.0 * pi / self.nx, 2.0 * pi / self.ny, 2.0 * pi / self.nz
        return Derived(
            nx=self.nx, ny=self.ny, nz=self.nz,
            boundtype=self.boundtype, solver=self.solver, limiter=self.limiter,
            secondorder=int(bool(self.secondorder)), dualenergy=int(bool(self.dualenergy)),
            onedimension=int(onedim), twodimension=int(twodim),
            g=g, g1=g1, g1inv=1.0 / g1, gg1=g / g1, cfl=self.cfl,
            de_thresh1=(1.0 - self.de_thresh) if self.dualenergy else 1.0,
            dx=dx, dy=dy, dz=dz, dxinv=1.0 / dx, dyinv=1.0 / dy, dzinv=1.0 / dz,
            volume=(xmax - xmin) * (ymax - ymin) * (zmax - zmin),
            x0=xmin + dx / 2, y0=ymin + dy / 2, z0=zmin + dz / 2,
            conduction=int(bool(self.conduction)), simplekappa=int(bool(self.simplekappa)),
            c_cond=self.c_cond, temp0=temp0, kappa0=kappa0,
            gravity=int(bool(self.gravity)), supercomoving=int(bool(self.supercomoving)),
            grav=grav, fb=self.fb, h2=dx ** 2.0 / 4.0,
            kfx2=kfx * 0.5, kfy2=kfy * 0.5, kfz2=kfz * 0.5,
            nyqx=max(self.nx // 2, 1), nyqy=max(self.ny // 2, 1), nyqz=max(self.nz // 2, 1),
            dx2inv=0.5 / dx, dy2inv=0.5 / dy, dz2inv=0.5 / dz,
            hydro_off=0 if self.hydro else 1,
        )

    def to_namelist(self) -> str:
        def b(v):
            return ".true." if v else ".false."

        def r(v):
            return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"

        return (
            "$para\n"
            f" hydro = {b(self.hydro)},\n"
            f" secondorder = {b(self.secondorder)}, solver = {self.solver}, limiter = {self.limiter},\n"
            f" cfl = {r(self.cfl)}, g = {r(self.g)}, atomic = {b(self.atomic)},\n"
            f" nx = {self.nx}, ny = {self.ny}, nz = {self.nz},\n"
            f" xmin = {r(self.xmin)}, xmax = {r(self.xmax)},\n"
            f" ymin = {r(self.ymin)}, ymax = {r(self.ymax)},\n"
            f" zmin = {r(self.zmin)}, zmax = {r(self.zmax)},\n"
            f" boundtype = {self.boundtype},\n"
            f" dualenergy = {b(self.dualenergy)}, de_thresh = {r(self.de_thresh)},\n"
            f" tstart = {r(self.tstart)}, tend = {r(self.tend)}, nsnap = {self.nsnap},\n"
            f" conduction = {b(self.conduction)}, c_cond = {r(self.c_cond)}, simplekappa = {b(self.simplekappa)},\n"
            f" gravity = {b(self.gravity)}, supercomoving = {b(self.supercomoving)}, omega0 = {r(self.omega0)},"
            f" fb = {r(self.fb)}\n"
            "/\n"
        )


@dataclasses.dataclass
class Derived:
    """Constants of module `global` that the hydro step reads."""
    nx: int
    ny: int
    nz: int
    boundtype: int
    solver: int
    limiter: int
    secondorder: int
    dualenergy: int
    onedimension: int
    twodimension: int
    g: float
    g1: float
    g1inv: float
    gg1: float
    cfl: float
    de_thresh1: float
    dx: float
    dy: float
    dz: float
    dxinv: float
    dyinv: float
    dzinv: float
    volume: float
    x0: float
    y0: float
    z0: float
    conduction: int = 0
    simplekappa: int = 0
    c_cond: float = 0.1
    temp0: float = 1.0
    kappa0: float = 0.0
    gravity: int = 0
    supercomoving: int = 0
    grav: float = 0.0
    fb: float = 1.0
    h2: float = 0.0
    kfx2: float = 0.0
    kfy2: float = 0.0
    kfz2: float = 0.0
    nyqx: int = 1
    nyqy: int = 1
    nyqz: int = 1
    dx2inv: float = 0.0
    dy2inv: float = 0.0
    dz2inv: float = 0.0
    hydro_off: int = 0

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz


def zslab_ranges(nz: int, nranks: int):
    """The z-slab of each rank, as 0-based [lo, hi) plane ranges: the split of
    decomposition (parallel.f90:53-61), rank r taking planes lz..mz with
    mz = lz + (nz - lz) / (nranks - r) after the ranks below it."""
    if not 1 <= nranks <= nz:
        raise ValueError("every rank needs a plane")  # parallel.f90:42-49
    out, mz = [], 0
    for r in range(nranks):
        lz = mz + 1
        mz = lz + (nz - lz) // (nranks - r)
        out.append((lz - 1, mz))
    return out
