"""Initial conditions for the hydro-step workloads, as conserved q.

Every generator returns q with numpy shape (nz, ny, nx, 6), i.e. the
reference layout q(nq, nx, ny, nz) (source/main.f90:72), float64, plus the
HydroParams the case runs with.  Seeded generators use numpy's PCG64
(`np.random.default_rng(seed)`) so that a dump is reproducible from
(seed, sizes); the same q feeds the CPU oracle and the HIP path.

Cell centres follow `coordinates` (source/comp.f90:150-162):
x = x0 + (ix-1)*dx with x0 = xmin + dx/2.
"""
from __future__ import annotations

import math

import numpy as np

from .params import HydroParams, _f32


def _centres(d):
    x = d.x0 + np.arange(d.nx, dtype=np.float64) * d.dx
    y = d.y0 + np.arange(d.ny, dtype=np.float64) * d.dy
    z = d.z0 + np.arange(d.nz, dtype=np.float64) * d.dz
    return x, y, z


def prim_to_q(d, rho, vx, vy, vz, p):
    """Conserved state from primitives, in the operand order of
    ic/shocktube.f90:88-89 (q5 = 0.5*d*(u**2) + p/g1, q6 = p/(d**g1))."""
    shape = np.broadcast(rho, vx, vy, vz, p).shape
    q = np.empty(shape + (6,))
    q[..., 0] = rho
    q[..., 1] = rho * vx
    q[..., 2] = rho * vy
    q[..., 3] = rho * vz
    u2 = vx * vx + vy * vy + vz * vz
    q[..., 4] = 0.5 * rho * u2 + p / d.g1
    q[..., 5] = p / rho ** d.g1
    return q


# ----------------------------------------------------------------- Sod
def sod_params(nx=1000, **kw) -> HydroParams:
    """1-D Sod tube, gamma = 5/3 (atomic), HLLC, minmod, CFL 0.8, outflow, t_end 0.2."""
    base = dict(nx=nx, ny=1, nz=1, xmin=0.0, xmax=1.0, boundtype=1, solver=2, limiter=1,
                cfl=0.8, atomic=True, tstart=0.0, tend=0.2)
    base.update(kw)
    return HydroParams(**base)


def sod(params: HydroParams, dl=1.0, ul=0.0, pl=1.0, dr=0.125, ur=0.0, pr=0.1, x_diss=0.5, direction=1):
    """ic/shocktube.f90:59-96 for dir = 1, 2 or 3."""
    d = params.derive()
    x, y, z = _centres(d)
    coord = {1: x[None, None, :], 2: y[None, :, None], 3: z[:, None, None]}[direction]
    left = np.broadcast_to(coord < x_diss, (d.nz, d.ny, d.nx))
    dd = np.where(left, dl, dr)
    u = np.where(left, ul, ur)
    p = np.where(left, pl, pr)
    q = np.zeros((d.nz, d.ny, d.nx, 6))
    q[..., 0] = dd
    q[..., direction] = dd * u
    q[..., 4] = 0.5 * dd * (u * u) + p / d.g1
    q[..., 5] = p / dd ** d.g1
    return q


# ----------------------------------------------------------------- Kelvin-Helmholtz
def kh_params(n=256, **kw) -> HydroParams:
    base = dict(nx=n, ny=n, nz=n, boundtype=2, solver=2, limiter=1, cfl=0.8, atomic=True,
                tstart=0.0, tend=1e30)
    base.update(kw)
    return HydroParams(**base)


def kelvin_helmholtz(params: HydroParams, seed=1234, noise=1e-4):
    """Periodic shear layer: p = 2.5; rho = 2, vx = +0.5 for |y-0.5| < 0.25,
    else rho = 1, vx = -0.5; vy = 0.01 sin(4 pi x) (exp(-(y-.25)^2/.005) +
    exp(-(y-.75)^2/.005)); plus U(-noise, noise) on vx, vy, vz."""
    d = params.derive()
    x, y, z = _centres(d)
    X = x[None, None, :]
    Y = y[None, :, None]
    inner = np.abs(Y - 0.5) < 0.25
    shape = (d.nz, d.ny, d.nx)
    rho = np.broadcast_to(np.where(inner, 2.0, 1.0), shape)
    vx = np.broadcast_to(np.where(inner, 0.5, -0.5), shape).copy()
    vy = np.broadcast_to(0.01 * np.sin(4 * math.pi * X) *
                         (np.exp(-(Y - 0.25) ** 2 / 0.005) + np.exp(-(Y - 0.75) ** 2 / 0.005)), shape).copy()
    vz = np.zeros(shape)
    rng = np.random.default_rng(seed)
    vx += rng.uniform(-noise, noise, shape)
    vy += rng.uniform(-noise, noise, shape)
    vz += rng.uniform(-noise, noise, shape)
    return prim_to_q(d, rho, vx, vy, vz, np.full(shape, 2.5))


# ----------------------------------------------------------------- point blast
def blast_params(n=512, solver=2, **kw) -> HydroParams:
    base = dict(nx=n, ny=n, nz=n, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5,
                boundtype=1, solver=solver, limiter=1, cfl=0.6, atomic=True, tstart=0.0, tend=1e30)
    base.update(kw)
    return HydroParams(**base)


def blast(params: HydroParams, seed=42, e0=1.0, p0=1e-5, r_cells=3.5, noise=0.01):
    """ic/blastwave.f90:59-96 with rho = 1 + U(-noise, noise): E = e0/dv in the
    cells with r < r_cells*dx (dv = dx*dy*dz*count), p = p0 elsewhere."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    rng = np.random.default_rng(seed)
    rho = 1.0 * (1.0 + rng.uniform(-noise, noise, shape))
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    inside = r < r_cells * d.dx
    dv = d.dx * d.dy * d.dz * int(inside.sum())
    q = prim_to_q(d, rho, 0.0, 0.0, 0.0, np.full(shape, p0))
    q[..., 4] = np.where(inside, e0 / dv, q[..., 4])
    q[..., 5] = np.where(inside, (e0 * d.g1 / rho ** d.g1) / dv, q[..., 5])
    return q


def blast_reference(params: HydroParams, d0=1.0, p0=1e-5, r0=0.109375, e0=1.0):
    """ic/blastwave.f90:59-96 as it stands: uniform rho = d0, E = e0/dv in
    the cells with r < r0 (dv = dx*dy*dz*count), p = p0 elsewhere; the
    defaults are the blast-wave test's ic_para (bin/test:176-196)."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    inside = r < r0
    dv = d.dx * d.dy * d.dz * int(inside.sum())
    q = np.zeros(shape + (6,))
    q[..., 0] = d0
    q[..., 4] = np.where(inside, e0 / dv, p0 / d.g1)
    q[..., 5] = np.where(inside, (e0 * d.g1 / d0 ** d.g1) / dv, p0 / d0 ** d.g1)
    return q


def kh_reference(params: HydroParams):
    """ic/kevinhelmholtz.f90:40-70 as it stands (the 2-D shear layer of
    para/kevinhelmholtz): rho = 2, u = 0.5 for |y-0.5| <= 0.25, else rho = 1,
    u = -0.5; v = 0.1 sin(4 pi x) (exp(-(y-.25)^2/.05) + exp(-(y-.75)^2/.05))
    with the single-precision pi of global.f90:98; p = 2.5."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    X = np.broadcast_to(x[None, None, :], shape)
    Y = np.broadcast_to(y[None, :, None], shape)
    inner = np.abs(Y - 0.5) <= 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, 0.5, -0.5)
    pi = _f32(3.1415927)
    v = 0.1 * np.sin(4 * pi * X) * (np.exp(-((Y - 0.25) ** 2) / 0.05) + np.exp(-((Y - 0.75) ** 2) / 0.05))
    p = 2.5
    q = np.zeros(shape + (6,))
    q[..., 0] = rho
    q[..., 1] = rho * u
    q[..., 2] = rho * v
    q[..., 4] = 0.5 * rho * ((u ** 2) + (v ** 2)) + p / (d.g - 1)
    q[..., 5] = p / (rho ** (d.g - 1))
    return q


def thermaltest_params(**kw) -> HydroParams:
    """The reference's thermal-conduction test (bin/test:162-173): 100 cells
    on x in [0, 100], hydro off, simplekappa, c_cond 0.9, gamma 5/3, outflow,
    t_end 10."""
    base = dict(nx=100, ny=1, nz=1, xmin=0.0, xmax=100.0, boundtype=1, cfl=0.5, atomic=True, tstart=0.0,
                tend=10.0, hydro=False, conduction=True, c_cond=_f32(0.9), simplekappa=True)
    base.update(kw)
    return HydroParams(**base)


def thermaltest(params: HydroParams, templ=400.0, tempr=800.0, dens=1.0):
    """ic/thermaltest.f90:26-68: rho = 1 at rest, T = 400 for x <= r0 and
    800 beyond, r0 = (xmax - xmin)/2 (e = T/g1/temp0*rho, q6 = g1 e / rho**g1)."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    r0 = 0.5 * (params.xmax - params.xmin)
    el = templ / d.g1 / d.temp0 * dens
    er = tempr / d.g1 / d.temp0 * dens
    e = np.broadcast_to(np.where(x[None, None, :] <= r0, el, er), shape)
    q = np.zeros(shape + (6,))
    q[..., 0] = dens
    q[..., 4] = e
    q[..., 5] = d.g1 * e / (dens ** d.g1)
    return q


# ----------------------------------------------------------------- shock-cloud box
def shock_cloud_params(n=64, **kw) -> HydroParams:
    base = dict(nx=n, ny=n, nz=n, boundtype=1, solver=2, limiter=1, cfl=0.6, atomic=True,
                tstart=0.0, tend=1e30)
    base.update(kw)
    return HydroParams(**base)


def shock_cloud(params: HydroParams, mach=10.0, chi=30.0, radius=0.1, x_shock=0.1,
                centre=(0.3, 0.5, 0.5)):
    """A Mach-`mach` planar shock moving in +x (Rankine-Hugoniot post-shock
    state for x < x_shock) towards a sphere of density `chi` in an ambient
    rho = 1, p = 1 medium at rest."""
    d = params.derive()
    gam = d.g
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    m2 = mach * mach
    c1 = math.sqrt(gam * 1.0 / 1.0)
    rho2 = (gam + 1.0) * m2 / ((gam - 1.0) * m2 + 2.0)
    p2 = (2.0 * gam * m2 - (gam - 1.0)) / (gam + 1.0)
    u2 = mach * c1 * (1.0 - 1.0 / rho2)
    X = np.broadcast_to(x[None, None, :], shape)
    post = X < x_shock
    rr = (X - centre[0]) ** 2 + (y[None, :, None] - centre[1]) ** 2 + (z[:, None, None] - centre[2]) ** 2
    cloud = rr < radius * radius
    rho = np.where(post, rho2, np.where(cloud, chi, 1.0))
    vx = np.where(post, u2, 0.0)
    p = np.where(post, p2, 1.0)
    return prim_to_q(d, rho, vx, np.zeros(shape), np.zeros(shape), p)


def ensemble_sizes(nbox=48, seed=7, edges=(64, 96, 128, 160, 192)):
    """Box edges and (Mach, chi) for the batched shock-cloud ensemble."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(edges), size=nbox)
    mach = rng.uniform(3.0, 30.0, size=nbox)
    chi = rng.uniform(10.0, 100.0, size=nbox)
    return [(int(a), float(b), float(c)) for a, b, c in zip(n, mach, chi)]


# ----------------------------------------------------------------- supersonic random field
def cosmo_params(n=256, **kw) -> HydroParams:
    """Periodic supersonic field, HLLC, CFL 0.5, with the dual-energy switch
    (comp.f90:60-69) on where E_kin/E > 0.99 (de_thresh = 0.01)."""
    base = dict(nx=n, ny=n, nz=n, boundtype=2, solver=2, limiter=1, cfl=0.5, atomic=True,
                tstart=0.0, tend=1e30, dualenergy=True, de_thresh=0.01)
    base.update(kw)
    return HydroParams(**base)


def cosmo_field(params: HydroParams, seed=2024, slope=-2.5, rms=0.5, floor=0.05, mach=10.0, temp=1.0):
    """rho = 1 + delta, delta Gaussian with P(k) ~ k^slope to Nyquist,
    normalised to rms `rms` and clamped at rho >= floor; v = potential flow
    v_k = i k delta_k / k^2 scaled to rms Mach `mach` at uniform temperature
    `temp` (p = rho T, c = sqrt(gamma T))."""
    d = params.derive()
    n = (d.nz, d.ny, d.nx)
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(n)
    wk = np.fft.rfftn(white)
    kz = np.fft.fftfreq(n[0]) * n[0]
    ky = np.fft.fftfreq(n[1]) * n[1]
    kx = np.fft.rfftfreq(n[2]) * n[2]
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)
    amp[0, 0, 0] = 0.0
    dk = wk * amp
    delta = np.fft.irfftn(dk, s=n, axes=(0, 1, 2))
    delta *= rms / delta.std()
    dk = np.fft.rfftn(delta)
    rho = np.maximum(1.0 + delta, floor)
    vel = []
    for K in (KX, KY, KZ):
        vk = 1j * K * dk / k2
        vk[0, 0, 0] = 0.0
        vel.append(np.fft.irfftn(vk, s=n, axes=(0, 1, 2)))
    vx, vy, vz = vel
    c = math.sqrt(d.g * temp)
    vrms = np.sqrt(np.mean(vx ** 2 + vy ** 2 + vz ** 2))
    scale = mach * c / vrms
    return prim_to_q(d, rho, vx * scale, vy * scale, vz * scale, rho * temp)


# ----------------------------------------------------------------- thermal conduction
def conduction_params(n=24, **kw) -> HydroParams:
    """Hot gas with thermal conduction on (simplekappa), HLLC, minmod."""
    base = dict(nx=n, ny=n, nz=n, boundtype=2, solver=2, limiter=1, cfl=0.4, atomic=True,
                tstart=0.0, tend=1e30, conduction=True, simplekappa=True)
    base.update(kw)
    return HydroParams(**base)


def conduction_blob(params: HydroParams, seed=11, t_bg=1000.0, t_blob=3000.0, radius=0.2):
    """rho = 1 with 1% noise, T = t_bg with a Gaussian hot blob of peak
    t_blob at the box centre (p = rho T, units of temp0 = 1), and velocity
    noise of 1% of the background sound speed.  Hot enough that the
    reference's thermal step (timestep.f90:72-104) keeps the explicit
    conduction stable."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.uniform(-1, 1, shape)
    xc = 0.5 * (params.xmin + params.xmax)
    yc = 0.5 * (params.ymin + params.ymax)
    zc = 0.5 * (params.zmin + params.zmax)
    r2 = (x[None, None, :] - xc) ** 2 + (y[None, :, None] - yc) ** 2 + (z[:, None, None] - zc) ** 2
    temp = t_bg + (t_blob - t_bg) * np.exp(-r2 / radius ** 2)
    cs = math.sqrt(d.g * t_bg)
    v = [0.01 * cs * rng.uniform(-1, 1, shape) for _ in range(3)]
    return prim_to_q(d, rho, v[0], v[1], v[2], rho * temp)


# ----------------------------------------------------------------- small parity states
def jeans_params(n=32, **kw) -> HydroParams:
    """Self-gravitating periodic box (gravity on, not comoving: 4 pi G =
    8.38553911e-7 as in parameter.f90:240), HLLC, minmod, gamma 5/3."""
    base = dict(nx=n, ny=n, nz=n, boundtype=2, solver=2, limiter=1, cfl=0.4, atomic=True,
                tstart=0.0, tend=1e30, gravity=True)
    base.update(kw)
    return HydroParams(**base)


def jeans(params: HydroParams, seed=5, rho0=1e6, jeans_length=0.5, amp=0.1, noise=0.01):
    """rho0 (1 + amp cos(2 pi x) cos(2 pi y) + noise) with a uniform pressure
    whose sound speed gives the Jeans length lambda_J = c sqrt(pi / (G rho0))
    = jeans_length (in box units, 4 pi G = grav): the box-scale mode is
    gravitationally unstable and collapses, the short ones oscillate.
    Velocities carry 1% noise of the sound speed."""
    d = params.derive()
    x, y, z = _centres(d)
    shape = (d.nz, d.ny, d.nx)
    rng = np.random.default_rng(seed)
    lx, ly = params.xmax - params.xmin, params.ymax - params.ymin
    mode = np.cos(2 * math.pi * x[None, None, :] / lx) * np.cos(2 * math.pi * y[None, :, None] / ly)
    rho = rho0 * (1.0 + amp * np.broadcast_to(mode, shape) + noise * rng.uniform(-1, 1, shape))
    grav = d.grav if d.grav > 0 else 8.38553911e-7
    cs2 = jeans_length ** 2 * grav * rho0 / (4.0 * math.pi ** 2)
    pres = np.full(shape, rho0 * cs2 / d.g)
    cs = math.sqrt(cs2)
    v = [0.01 * cs * rng.uniform(-1, 1, shape) for _ in range(3)]
    return prim_to_q(d, rho, v[0], v[1], v[2], pres)


def random_smooth(params: HydroParams, seed=0, amp=0.3, jumps=True):
    """A small random state with smooth modes and (optionally) a contact and a
    pressure jump, for one-step parity tests on tiny grids."""
    d = params.derive()
    rng = np.random.default_rng(seed)
    x, y, z = _centres(d)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    X, Y, Z = X.transpose(2, 1, 0), Y.transpose(2, 1, 0), Z.transpose(2, 1, 0)
    L = (d.nx * d.dx, d.ny * d.dy, d.nz * d.dz)

    def mode():
        kx, ky, kz = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, 2 * math.pi, size=3)
        return (np.sin(2 * math.pi * kx * X / L[0] + ph[0]) * np.cos(2 * math.pi * ky * Y / L[1] + ph[1])
                * np.cos(2 * math.pi * kz * Z / L[2] + ph[2]))

    rho = 1.0 + amp * mode() + 0.05 * rng.uniform(-1, 1, X.shape)
    p = 1.0 + amp * mode() + 0.05 * rng.uniform(-1, 1, X.shape)
    vx = amp * mode()
    vy = amp * mode() if d.ny > 1 else np.zeros_like(X)
    vz = amp * mode() if d.nz > 1 else np.zeros_like(X)
    if jumps:
        xm = x.min() + 0.5 * L[0]
        rho = np.where(X < xm, rho, rho * 0.3)
        p = np.where(X < xm * 0.9 + 0.1 * x.min(), p * 3.0, p)
    return prim_to_q(d, rho, vx, vy, vz, p)
