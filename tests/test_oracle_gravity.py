"""The oracle's self-gravity (gravitation.f90:26-88,172-254; fft.f90:58-80).

The reference's potential needs FFTW-2, which this image lacks, so the
oracle's gravity is pinned in pieces:
* multigreen and gravintegration bit for bit against the reference's own
  compiled routines (fixtures written by tests/golden/make_golden.py);
* the transforms against numpy's FFT (an independent implementation with
  FFTW's sign and normalisation conventions);
* the potential against closed forms: a single Fourier mode is an exact
  eigenvector, so its potential is grav A green(k) times the mode; the
  potential satisfies the 7-point discrete Poisson equation up to the
  single-precision pi of global.f90:98 that the Green's function is built
  with; the supercomoving source is the static one scaled by cosmo / fb;
* conservation: the periodic central-difference force of a potential with an
  even Green's function conserves total momentum.
"""
import numpy as np
import pytest

import oracle as O
from evora_amd import ics
from evora_amd.params import HydroParams


@pytest.fixture(scope="module")
def grav_fixtures():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "grav_fixtures.npz"))


def _params(entry):
    return HydroParams(**entry["params"])


@pytest.mark.parametrize("shape", ["8x6x4", "16x16x16", "5x7x3", "12x10x9", "9x12x1", "16x8x1", "16x1x1", "7x1x1"])
def test_multigreen_matches_reference_bitwise(shape, grav_fixtures, golden_index):
    name = f"green_{shape}"
    d = _params(golden_index[name]).derive()
    out = O.Oracle(d).multigreen(grav_fixtures[name + "/in"])
    ref = grav_fixtures[name + "/out"]
    assert np.array_equal(out.real, ref.real) and np.array_equal(out.imag, ref.imag)


@pytest.mark.parametrize("shape", ["8x6x4", "16x16x16", "5x7x3", "12x10x9", "9x12x1", "16x8x1", "16x1x1", "7x1x1"])
def test_gravintegration_matches_reference_bitwise(shape, grav_fixtures, golden_index):
    name = f"gravint_{shape}"
    e = golden_index[name]
    d = _params(e).derive()
    q = grav_fixtures[name + "/q0"].copy()
    O.Oracle(d).gravintegration(q, grav_fixtures[name + "/phi"], e["dt"])
    assert np.array_equal(q, grav_fixtures[name + "/q"])


@pytest.mark.parametrize("shape", [(8, 6, 4), (5, 7, 3), (16, 1, 1), (9, 12, 1)])
def test_transforms_match_numpy(shape):
    nx, ny, nz = shape
    d = HydroParams(nx=nx, ny=ny, nz=nz, gravity=True).derive()
    o = O.Oracle(d)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
    fwd = np.fft.fftn(a)
    bwd = np.fft.ifftn(a) * a.size
    assert np.abs(o.dft3(a, -1) - fwd).max() <= 1e-14 * np.abs(fwd).max()
    assert np.abs(o.dft3(a, +1) - bwd).max() <= 1e-14 * np.abs(bwd).max()


def test_single_mode_potential_closed_form():
    """rho = rho0 + A cos(2 pi (mx x + my y)) on a cube: the DFT of the mode is
    two spikes, so phi = grav A green(mx, my, 0) cos(...) with
    green = -h2 / (sin^2(kfx2 mx) + sin^2(kfy2 my))."""
    n, mx, my, amp = 16, 3, 2, 0.25
    d = HydroParams(nx=n, ny=n, nz=n, gravity=True).derive()
    i = np.arange(n)
    mode = np.cos(2 * np.pi * (mx * i[None, None, :] + my * i[None, :, None]) / n) * np.ones((n, 1, 1))
    q = np.zeros((n, n, n, 6))
    q[..., 0] = 1.0 + amp * mode
    phi = O.Oracle(d).gravpotential(q)
    green = -d.h2 / (np.sin(d.kfx2 * mx) ** 2 + np.sin(d.kfy2 * my) ** 2)
    expect = d.grav * amp * green * mode
    assert np.abs(phi - expect).max() <= 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("shape", [(12, 10, 8), (12, 10, 1), (16, 1, 1)])
def test_potential_solves_discrete_poisson(shape):
    """The 7-point Laplacian of phi equals grav (rho - mean rho) to the
    relative 1e-7 by which the Green's function's pi (the single-precision
    literal of global.f90:98) differs from pi.  On 2-D and 1-D grids the
    degenerate axes add nothing to the Laplacian (the periodic neighbours of
    a one-cell axis are the cell itself) and nothing to the Green's function
    (kvalue 0 there)."""
    nx, ny, nz = shape
    d = HydroParams(nx=nx, ny=ny, nz=nz, xmax=nx / 10, ymax=ny / 10, zmax=nz / 10, gravity=True).derive()
    assert d.dx == pytest.approx(d.dy) and d.dy == pytest.approx(d.dz)
    rng = np.random.default_rng(3)
    q = np.zeros((nz, ny, nx, 6))
    q[..., 0] = 1.0 + 0.5 * rng.random((nz, ny, nx))
    phi = O.Oracle(d).gravpotential(q)
    lap = sum(np.roll(phi, s, axis=a) for a in range(3) for s in (1, -1)) - 6.0 * phi
    lap /= d.dx ** 2
    src = d.grav * (q[..., 0] - q[..., 0].mean())
    assert np.abs(lap - src).max() <= 1e-6 * np.abs(src).max()
    assert np.abs(lap - src).max() > 0.0  # the pi of the Green's function is not the true pi


def test_supercomoving_source_is_scaled_static_source():
    """cosmo grav (rho/fb - 1) differs from grav (rho - mean) by the factor
    cosmo / fb and a constant, which the Green's function drops (k = 0)."""
    p = HydroParams(nx=10, ny=8, nz=6, gravity=True)
    d = p.derive()
    ds = HydroParams(nx=10, ny=8, nz=6, gravity=True, supercomoving=True, fb=0.17).derive()
    rng = np.random.default_rng(9)
    q = np.zeros((6, 8, 10, 6))
    q[..., 0] = 0.5 + rng.random((6, 8, 10))
    phi_s = O.Oracle(d).gravpotential(q)
    cosmo = 0.37
    phi_c = O.Oracle(ds).gravpotential(q, cosmo)
    expect = phi_s * (cosmo * ds.grav / ds.fb) / d.grav
    assert np.abs(phi_c - expect).max() <= 1e-13 * np.abs(expect).max()


def test_gravity_conserves_momentum():
    p = ics.jeans_params(12, nz=10)
    d = p.derive()
    q = ics.jeans(p)
    o = O.Oracle(d)
    phi = o.gravpotential(q)
    q1 = q.copy()
    dt = 0.05
    o.gravintegration(q1, phi, dt)
    dm = (q1[..., 1:4] - q[..., 1:4]).sum(axis=(0, 1, 2))
    scale = np.abs(q1[..., 1:4] - q[..., 1:4]).sum()
    assert scale > 0
    assert np.abs(dm).max() <= 1e-13 * scale


def test_step_with_gravity_conserves_mass_and_momentum():
    p = ics.jeans_params(16)
    d = p.derive()
    q = ics.jeans(p)
    o = O.Oracle(d)
    s0 = o.sums(q)
    m0 = q[..., 1:4].sum(axis=(0, 1, 2))
    t, dt, n = o.run(q, 10)
    assert n == 10
    assert o.sums(q)[0] == pytest.approx(s0[0], rel=1e-14)
    assert np.abs(q[..., 1:4].sum(axis=(0, 1, 2)) - m0).max() <= 1e-14 * np.abs(q[..., 1:4]).sum()
