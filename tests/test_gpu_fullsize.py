"""The GPU path at the configured problem sizes.

Where one step of the CPU oracle at that size takes seconds (256^3), the
GPU result is compared with the oracle directly, to the same bars as the
small-grid tests: relative L-infinity <= 1e-12 per component after one
step, relative L1 <= 1e-10 after a run.  At 512^3 and for the ensemble,
size-independent properties are checked instead: conservation of mass and
energy while no wave has reached the boundary, x-mirror symmetry of a
symmetric blast, and the FP32-mixed form within its bar (relative L1 of rho
and p <= 1e-5) of the FP64 form.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l1, rel_linf
from evora_amd import hydro, ics
from evora_amd.batch import Batch

pytestmark = pytest.mark.gpu

ONE_STEP_TOL = 1e-12
RUN_L1_TOL = 1e-10
MIXED_TOL = 1e-5


def gpu_run(d, q0, nsteps, precision=hydro.FP64):
    with hydro.Hydro(d, precision=precision) as h:
        h.upload_q(q0)
        s0 = h.sums()
        t, dt, done = h.run(nsteps)
        assert done == nsteps
        return h.download_q(), s0, h.sums(), t


def oracle_run(d, q0, nsteps):
    q = q0.copy()
    t, dt, done = O.Oracle(d).run(q, nsteps)
    assert done == nsteps
    return q, t


def primitive(q, g1):
    rho = q[..., 0]
    kin = 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / rho
    return rho, g1 * (q[..., 4] - kin)


@pytest.mark.parametrize("nsteps", [1, 50])
def test_kh256_matches_oracle(nsteps):
    """3-D Kelvin-Helmholtz 256^3, periodic, HLLC, minmod; the configured
    CFL 0.8 for the first step, and the configured 50 steps at CFL 0.5 (at
    0.8 the reference scheme itself stops after 13-14 steps).

    rho v_z of this configuration is the seeded noise alone (|v_z| < 1e-4)
    while its update is the difference of two face pressures of 2.5: one
    ulp of 2.5 in one face is 7e-13 of max |rho v_z|, so this component
    holds the pressure of every z face to about an ulp of the reference's."""
    p = ics.kh_params(256, cfl=0.8 if nsteps == 1 else 0.5)
    d = p.derive()
    q0 = ics.kelvin_helmholtz(p)
    qg, s0, s1, tg = gpu_run(d, q0, nsteps)
    qo, to = oracle_run(d, q0, nsteps)
    assert tg == pytest.approx(to, rel=1e-12)
    if nsteps == 1:
        for c in range(6):
            assert rel_linf(qg[..., c], qo[..., c]) <= ONE_STEP_TOL, c
    else:
        for c in range(6):
            assert rel_l1(qg[..., c], qo[..., c]) <= RUN_L1_TOL, c
    # periodic box: mass and total energy are conserved to rounding
    assert abs(s1[0] - s0[0]) <= 1e-13 * s0[0]
    assert abs(s1[1] - s0[1]) <= 1e-13 * s0[1]


def test_cosmo256_one_step_matches_oracle():
    """Cosmological-like field 256^3, Mach 10, dual energy on (the entropy
    branch is taken where E_kin/E > 0.99)."""
    p = ics.cosmo_params(256)
    d = p.derive()
    q0 = ics.cosmo_field(p)
    qg, _, _, _ = gpu_run(d, q0, 1)
    qo, _ = oracle_run(d, q0, 1)
    for c in range(6):
        assert rel_linf(qg[..., c], qo[..., c]) <= ONE_STEP_TOL, c


@pytest.mark.parametrize("solver", [2, 1])
def test_blast512_conservation_and_mixed_form(solver):
    """512^3 point blast, HLLC and HLL, CFL 0.4 (at the configured 0.6 the
    reference scheme stops, after 28 steps with HLLC and 2 with HLL), 40
    steps: the blast stays far from the outflow boundary, so FP64 conserves
    mass and energy to rounding; the FP32-mixed form stays within its bar of
    FP64."""
    p = ics.blast_params(512, solver=solver, cfl=0.4)
    d = p.derive()
    q0 = ics.blast(p)
    q64, s0, s1, _ = gpu_run(d, q0, 40)
    assert abs(s1[0] - s0[0]) <= 1e-13 * s0[0]
    assert abs(s1[1] - s0[1]) <= 1e-13 * s0[1]
    q32, m0, m1, _ = gpu_run(d, q0, 40, precision=hydro.FP32_MIXED)
    assert abs(m1[1] - m0[1]) <= 1e-6 * m0[1]
    num = np.zeros(2)
    den = np.zeros(2)
    for z in range(0, 512, 64):  # slabs keep the temporaries small
        a = primitive(q32[z:z + 64], d.g1)
        b = primitive(q64[z:z + 64], d.g1)
        for i in range(2):
            num[i] += np.abs(a[i] - b[i]).sum()
            den[i] += np.abs(b[i]).sum()
    assert (num / den <= MIXED_TOL).all(), num / den


def test_ensemble_boxes_mixed_form():
    """The first 16 boxes of the seeded shock-cloud ensemble (edges from
    {64, ..., 192}, Mach 3-30, contrast 10-100), 50 steps in one batch, FP32-mixed
    against FP64: relative L1 of rho and p <= 1e-5 in every box."""
    spec = ics.ensemble_sizes(48, seed=7)[:16]
    params = [ics.shock_cloud_params(n, cfl=0.6) for n, _, _ in spec]
    q0 = [ics.shock_cloud(p, mach=mach, chi=chi) for (n, mach, chi), p in zip(spec, params)]
    out = {}
    for prec in (hydro.FP64, hydro.FP32_MIXED):
        with Batch(params, precision=prec) as b:
            for i, q in enumerate(q0):
                b.upload_q(i, q)
            b.run(50, graph=True)
            out[prec] = [b.download_q(i) for i in range(len(q0))]
    for i, p in enumerate(params):
        a = primitive(out[hydro.FP32_MIXED][i], p.derive().g1)
        b = primitive(out[hydro.FP64][i], p.derive().g1)
        for c in range(2):
            assert rel_l1(a[c], b[c]) <= MIXED_TOL, (i, c)


@pytest.mark.parametrize("solver", [2, 1])
def test_blast512_mirror_symmetry(solver):
    """The noise-free 512^3 blast (HLLC and HLL, CFL 0.4, 30 steps) is
    symmetric under x -> -x, y -> -y, z -> -z (with the matching momentum
    negated) and under the exchange of x and y; the scheme keeps every
    symmetry to rounding (measured below 1e-14 on MI355X)."""
    p = ics.blast_params(512, solver=solver, cfl=0.4)
    d = p.derive()
    q0 = ics.blast(p, noise=0.0)
    q, _, _, _ = gpu_run(d, q0, 30)
    scale = np.abs(q).max(axis=(0, 1, 2))
    for ax in (2, 1, 0):
        m = np.flip(q, axis=ax).copy()
        m[..., 3 - ax] *= -1
        err = np.abs(m - q).max(axis=(0, 1, 2)) / scale
        assert (err <= 1e-13).all(), (ax, err)
    t = np.transpose(q, (0, 2, 1, 3)).copy()
    t[..., [1, 2]] = t[..., [2, 1]]
    err = np.abs(t - q).max(axis=(0, 1, 2)) / scale
    assert (err <= 1e-13).all(), err
