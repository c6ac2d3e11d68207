"""Self-gravity on the GPU (evh_gravpotential, evh_gravintegration) against
the oracle, whose multigreen and gravintegration are bit-identical to the
reference's and whose transforms match numpy's (tests/test_oracle_gravity.py).

Bars: the potential within 1e-12 of the oracle's (relative to max |phi|; the
FFTs differ from the direct DFT sums by rounding only); the fast fused
step as for hydro (relative L-infinity <= 1e-12 after one step, relative
L1 <= 1e-10 after 10); the exact form within 1e-12 after 10 steps (the
FFT rounding rules out bit equality); FP32-mixed relative L1 of rho and p
<= 1e-5.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from conftest import rel_l1, rel_linf
from evora_amd import hydro, ics
from evora_amd.params import HydroParams
from test_gpu_dropin import run_dropin

pytestmark = pytest.mark.gpu


def _potential(d, q, cosmo=1.0, precision=hydro.FP64):
    with hydro.Hydro(d, precision=precision) as h:
        h.upload_q(q)
        h.gravpotential(cosmo)
        return h.download_phi()


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 10, 9), (24, 20, 18), (7, 9, 5)])
def test_potential_matches_oracle(shape):
    nx, ny, nz = shape
    p = ics.jeans_params(nx, ny=ny, nz=nz, xmax=nx / 16, ymax=ny / 16, zmax=nz / 16)
    d = p.derive()
    q = ics.jeans(p)
    phi = _potential(d, q)
    ref = O.Oracle(d).gravpotential(q)
    assert rel_linf(phi, ref) <= 1e-12


@pytest.mark.parametrize("shape", [(15, 32, 16), (34, 8, 64), (40, 64, 32), (7, 2, 128), (16, 12, 32), (9, 24, 8),
                                   (10, 6, 4), (16, 4, 2), (32, 16, 8), (8, 32, 64), (64, 8, 16), (2, 2, 4),
                                   (4, 64, 2), (128, 2, 8), (512, 2, 2), (1024, 2, 4), (4096, 2, 2), (4, 1024, 2),
                                   (2, 4096, 2), (2, 2, 4096)])
def test_potential_pencil_transforms(shape):
    """A power-of-two nz takes the z-pencil kernel (forward z transform,
    Green's function, inverse z transform in LDS).  With other nx or ny,
    hipFFT's 2-D x-y transforms surround it: odd and even nx, power-of-two
    and other ny.  With power-of-two nx and ny as well, the x transforms
    (the source formed on the load, two real rows per complex transform) and
    the y transforms are this library's own: nx, ny and nz from 2 to 4096
    (the longest transform a tile holds: one row pair or column per block),
    one-stage transforms (n 2, 4) and tiles wider than the kx range.  A
    random density puts signal in every bin."""
    nx, ny, nz = shape
    p = ics.jeans_params(nx, ny=ny, nz=nz, xmax=nx / 16, ymax=ny / 16, zmax=nz / 16)
    d = p.derive()
    q = ics.jeans(p)
    rng = np.random.default_rng(nx * ny + nz)
    q[..., 0] = 1e6 * (1.0 + 0.1 * rng.standard_normal(q.shape[:3]))
    phi = _potential(d, q)
    ref = O.Oracle(d).gravpotential(q)
    assert rel_linf(phi, ref) <= 1e-12


def test_potential_unequal_spacing():
    """dx != dy != dz: multigreen's h2 is dx**2 / 4 whatever the other
    spacings (gravitation.f90:80, parameter.f90:254), the force divides by
    each direction's own 2 dx."""
    p = ics.jeans_params(12, ny=10, nz=9, xmax=2.0, ymax=1.5, zmax=0.8)
    d = p.derive()
    q = ics.jeans(p)
    assert rel_linf(_potential(d, q), O.Oracle(d).gravpotential(q)) <= 1e-12
    qo = q.copy()
    O.Oracle(d).run(qo, 2)
    with hydro.Hydro(d) as h:
        h.upload_q(q)
        h.run(2)
        qg = h.download_q()
    for c in range(6):
        assert rel_linf(qg[..., c], qo[..., c]) <= 1e-12, c


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
@pytest.mark.parametrize("shape", [(16, 12, 10), (16, 8, 4)])
def test_potential_supercomoving_source(shape, precision):
    """The comoving source cosmo grav (rho/fb - 1), FP64 and from the FP32
    state of the mixed form (against the oracle on the density rounded to
    FP32); (16, 8, 4) takes this library's own transforms, whose x pass forms
    the source as it loads."""
    nx, ny, nz = shape
    p = HydroParams(nx=nx, ny=ny, nz=nz, gravity=True, supercomoving=True, fb=0.17, atomic=True)
    d = p.derive()
    rng = np.random.default_rng(4)
    q = ics.random_smooth(p, seed=3)
    q[..., 0] = 0.1 + rng.random(q.shape[:3])
    cosmo = 0.37
    qo = q if precision == hydro.FP64 else q.astype(np.float32).astype(np.float64)
    assert rel_linf(_potential(d, q, cosmo, precision), O.Oracle(d).gravpotential(qo, cosmo)) <= 1e-12


@pytest.mark.parametrize("n", [16, 12])
def test_potential_fp32_mixed_source(n):
    """The mixed form takes rho from its FP32 state; the transforms stay FP64
    (16: this library's x transform reads the FP32 density)."""
    p = ics.jeans_params(n)
    d = p.derive()
    q = ics.jeans(p)
    q32 = q.astype(np.float32).astype(np.float64)
    assert rel_linf(_potential(d, q, precision=hydro.FP32_MIXED), O.Oracle(d).gravpotential(q32)) <= 1e-12


@pytest.mark.parametrize("mode", [hydro.MODE_FUSED, hydro.MODE_FUSED_EXACT])
def test_gravity_one_step(mode):
    p = ics.jeans_params(32)
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 1)
    with hydro.Hydro(d, mode=mode) as h:
        h.upload_q(q0)
        h.run(1)
        q = h.download_q()
    for c in range(6):
        assert rel_linf(q[..., c], qo[..., c]) <= 1e-12, c


@pytest.mark.parametrize("mode,tol", [(hydro.MODE_FUSED, 1e-10), (hydro.MODE_FUSED_EXACT, 1e-12)])
def test_gravity_ten_steps(mode, tol):
    p = ics.jeans_params(24, nz=20, zmax=20 / 24)
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    to, dto, _ = O.Oracle(d).run(qo, 10)
    with hydro.Hydro(d, mode=mode) as h:
        h.upload_q(q0)
        t, dt, n = h.run(10)
        q = h.download_q()
    assert n == 10
    assert t == pytest.approx(to, rel=1e-12)
    for c in range(6):
        assert rel_l1(q[..., c], qo[..., c]) <= tol, c


def test_gravity_mixed_precision():
    p = ics.jeans_params(24)
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 5)
    with hydro.Hydro(d, precision=hydro.FP32_MIXED) as h:
        h.upload_q(q0)
        h.run(5)
        q = h.download_q()
    for c in (0, 4):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-5, c


def test_gravity_with_conduction():
    """main.f90:157-175: sweep, then gravity, then conduction (T taken from
    the state after the gravity update)."""
    p = ics.conduction_params(16, boundtype=2, gravity=True)
    d = p.derive()
    d = dataclasses.replace(d, grav=5e4)  # strong enough to matter on this hot blob (Jeans length ~ 1 box)
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 3)
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        h.run(3)
        q = h.download_q()
    qh = q0.copy()
    O.Oracle(dataclasses.replace(d, gravity=0)).run(qh, 3)
    assert rel_l1(qo[..., 1], qh[..., 1]) > 1e-6  # gravity changed the run
    for c in range(6):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-10, c


def test_gravity_fullsize_conservation():
    """256^3 Jeans box, 3 steps: mass conserved to roundoff, and total
    momentum too (periodic hydro and a central-difference force of a
    potential with an even Green's function both conserve it)."""
    p = ics.jeans_params(256)
    d = p.derive()
    q0 = ics.jeans(p)
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        s0 = h.sums()
        h.run(3)
        s1 = h.sums()
        q = h.download_q()
    assert s1[0] == pytest.approx(s0[0], rel=1e-13)
    dm = q[..., 1:4].sum(axis=(0, 1, 2)) - q0[..., 1:4].sum(axis=(0, 1, 2))
    assert np.abs(dm).max() <= 1e-13 * np.abs(q[..., 1:4]).sum()


@pytest.mark.parametrize("exact", [False, True])
def test_gravity_dropin(exact):
    """The Fortran drop-in with gravity on (gravpotential and gravintegration
    dispatched to the GPU) against the oracle's run: the reference's own
    potential cannot be built here (FFTW-2)."""
    p = ics.jeans_params(20, tend=1e30)
    q0 = ics.jeans(p)
    env = {"EVORA_MAXSTEPS": "6"}
    if exact:
        env["EVORA_EXACT"] = "1"
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert len(sh) == 7  # 6 steps and the dt = 0 step that ends the run
    qo = q0.copy()
    O.Oracle(p.derive()).run(qo, 6)
    for c in range(6):
        assert rel_l1(qh[..., c], qo[..., c]) <= 1e-10, c


def test_gravity_refusals():
    """A one-cell x axis is no reduced dimension of the reference (2-D means
    nz = 1, 1-D ny = nz = 1, parameter.f90:173-182) and is refused."""
    with pytest.raises(hydro.EvoraError) as ei:
        hydro.Hydro(HydroParams(nx=1, ny=16, nz=16, gravity=True).derive())
    assert ei.value.code == 8
    p = ics.jeans_params(8)
    with hydro.Hydro(p.derive()) as h:
        h.upload_q(ics.jeans(p))
        with pytest.raises(hydro.EvoraError) as ei:
            h.gravintegration(0.1)
        assert ei.value.code == 10
    with hydro.Hydro(ics.blast_params(8).derive()) as h:
        with pytest.raises(hydro.EvoraError):
            h.gravpotential()


def test_gravity_and_conduction_in_staged_mode():
    """The per-routine kernels (EVH_MODE_STAGED) followed by the gravity and
    conduction passes in the reference's arithmetic: 3 steps against the
    oracle."""
    p = ics.conduction_params(12, boundtype=1, gravity=True)
    d = dataclasses.replace(p.derive(), grav=5e4)
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 3)
    with hydro.Hydro(d, mode=hydro.MODE_STAGED) as h:
        h.upload_q(q0)
        h.run(3)
        q = h.download_q()
    for c in range(6):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-12, c


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 2), (2, 3, 2)])
def test_gravity_smallest_grids(shape):
    """Transforms of length 2 and 3 and a force stencil that wraps onto the
    cell's own neighbours: 2 steps against the oracle."""
    p = ics.jeans_params(2, nx=shape[0], ny=shape[1], nz=shape[2])
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 2)
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        h.run(2)
        q = h.download_q()
    for c in range(6):
        assert rel_linf(q[..., c], qo[..., c]) <= 1e-12, c


@pytest.mark.parametrize("shape", [(16, 16, 1), (12, 10, 1), (9, 12, 1), (64, 32, 1), (2, 2, 1), (16, 1, 1),
                                   (7, 1, 1), (2, 1, 1)])
def test_potential_reduced_dimensions(shape):
    """2-D (nz = 1) and 1-D (ny = nz = 1) grids take hipFFT's 2-D and 1-D
    plans; the reference's 3-D transform is the identity along a one-cell
    axis and the Green's function has that axis's wavenumber 0.  A random
    density puts signal in every bin."""
    nx, ny, nz = shape
    p = ics.jeans_params(nx, ny=ny, nz=nz, xmax=nx / 16, ymax=ny / 16, zmax=nz / 16)
    d = p.derive()
    q = ics.jeans(p)
    rng = np.random.default_rng(nx * ny + nz)
    q[..., 0] = 1e6 * (1.0 + 0.1 * rng.standard_normal(q.shape[:3]))
    phi = _potential(d, q)
    ref = O.Oracle(d).gravpotential(q)
    assert rel_linf(phi, ref) <= 1e-12


@pytest.mark.parametrize("shape,mode,tol", [((24, 20, 1), hydro.MODE_FUSED, 1e-10),
                                            ((24, 20, 1), hydro.MODE_FUSED_EXACT, 1e-12),
                                            ((24, 20, 1), hydro.MODE_STAGED, 1e-12),
                                            ((32, 1, 1), hydro.MODE_STAGED, 1e-12)])
def test_gravity_reduced_dimensions_steps(shape, mode, tol):
    """5 steps of a 2-D (the single-pass 2-D sweep, fast and exact, and the
    per-routine kernels) and a 1-D Jeans box (the per-routine kernels) with
    gravity against the oracle; the force has no z (and in 1-D no y)
    component (gravitation.f90:209,220,232)."""
    nx, ny, nz = shape
    p = ics.jeans_params(nx, ny=ny, nz=nz, xmax=nx / 24, ymax=ny / 24)
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    to, _, _ = O.Oracle(d).run(qo, 5)
    with hydro.Hydro(d, mode=mode) as h:
        h.upload_q(q0)
        t, _, n = h.run(5)
        q = h.download_q()
    assert n == 5
    assert t == pytest.approx(to, rel=1e-12)
    for c in range(6):
        assert rel_l1(q[..., c], qo[..., c]) <= tol, c


def test_gravity_2d_mixed_precision():
    p = ics.jeans_params(32, nz=1)
    d = p.derive()
    q0 = ics.jeans(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 5)
    with hydro.Hydro(d, precision=hydro.FP32_MIXED) as h:
        h.upload_q(q0)
        h.run(5)
        q = h.download_q()
    for c in (0, 4):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-5, c


@pytest.mark.parametrize("exact", [False, True])
def test_gravity_dropin_2d(exact):
    """The Fortran drop-in with gravity on a 2-D grid (twodimension in the
    reference's driver) against the oracle's run."""
    p = ics.jeans_params(24, nz=1, tend=1e30)
    q0 = ics.jeans(p)
    env = {"EVORA_MAXSTEPS": "6"}
    if exact:
        env["EVORA_EXACT"] = "1"
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert len(sh) == 7
    qo = q0.copy()
    O.Oracle(p.derive()).run(qo, 6)
    for c in range(6):
        assert rel_l1(qh[..., c], qo[..., c]) <= 1e-10, c


@pytest.mark.parametrize("shape,mode", [((24, 20, 1), hydro.MODE_FUSED), ((24, 20, 1), hydro.MODE_FUSED_EXACT),
                                        ((48, 1, 1), hydro.MODE_STAGED)])
def test_gravity_with_conduction_reduced_dimensions(shape, mode):
    """main.f90:157-175 on a 2-D (the single-pass 2-D sweep) and a 1-D grid (the staged kernels): sweep, gravity,
    then conduction, 3 steps against the oracle."""
    nx, ny, nz = shape
    p = ics.conduction_params(nx, ny=ny, nz=nz, boundtype=2, gravity=True)
    d = dataclasses.replace(p.derive(), grav=5e4)
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 3)
    with hydro.Hydro(d, mode=mode) as h:
        h.upload_q(q0)
        h.run(3)
        q = h.download_q()
    qh = q0.copy()
    O.Oracle(dataclasses.replace(d, gravity=0)).run(qh, 3)
    assert rel_l1(qo[..., 1], qh[..., 1]) > 1e-6  # gravity changed the run
    for c in range(6):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-10, c
