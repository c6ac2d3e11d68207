"""Pins of the CPU oracle that do not come from the reference's outputs.

* Sod against the exact Riemann solution (Toro ch. 4) at the same gamma.
* Invariants fixed by the equations: a uniform state is unchanged bit for
  bit; periodic mass is conserved to roundoff; x-mirror symmetry with
  u -> -u; axis permutation on a cube; a 3-D run uniform in y, z equals the
  1-D run (within 1 ulp: the 1-D compevo association differs,
  muscl.f90:68 vs :106-107).
* Per-call checks of riemann and comp_slope against closed forms.
Each check is chosen so that a dropped term, a wrong sign or a swapped
index in the oracle fails it.
"""
import math

import numpy as np
import pytest

import exact_riemann
import oracle as O
from conftest import rel_l1
from evora_amd import ics
from evora_amd.params import HydroParams


def test_sod_matches_exact_solution():
    p = ics.sod_params(1000)
    d = p.derive()
    q = ics.sod(p)
    t, dt, n = O.Oracle(d).run(q, 100000, 0.0, p.tend)
    assert t == pytest.approx(0.2, abs=1e-15)
    x = d.x0 + np.arange(d.nx) * d.dx
    rho_e, u_e, p_e = exact_riemann.sample(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, d.g, (x - 0.5) / t)
    rho = q[0, 0, :, 0]
    u = q[0, 0, :, 1] / rho
    pr = d.g1 * (q[0, 0, :, 4] - 0.5 * rho * u * u)
    # first-order convergence at discontinuities: L1 errors of a few 1e-3
    assert np.mean(np.abs(rho - rho_e)) < 4e-3
    assert np.mean(np.abs(u - u_e)) < 6e-3
    assert np.mean(np.abs(pr - p_e)) < 3e-3
    # star state of the exact solution, as seen in the plateau
    ps, us = exact_riemann.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, d.g)
    plateau = (x > 0.5 + 0.2 * t) & (x < 0.5 + 0.8 * us * t)
    assert np.allclose(u[plateau], us, rtol=2e-3)
    assert np.allclose(pr[plateau], ps, rtol=2e-3)
    # total mass exactly conserved with zero-gradient walls before waves arrive
    assert np.sum(rho) * d.dx == pytest.approx(0.5625, rel=1e-14)


def test_exact_solver_known_values():
    # Toro Table 4.2, test 1 (gamma 1.4): p* = 0.30313, u* = 0.92745
    ps, us = exact_riemann.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    assert ps == pytest.approx(0.30313, abs=1e-5)
    assert us == pytest.approx(0.92745, abs=1e-5)
    # Toro test 3: p* = 460.894, u* = 19.5975
    ps, us = exact_riemann.star_state(1.0, 0.0, 1000.0, 1.0, 0.0, 0.01, 1.4)
    assert ps == pytest.approx(460.894, rel=1e-5)
    assert us == pytest.approx(19.5975, rel=1e-5)


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("limiter", [1, 2, 3])
def test_uniform_state_is_unchanged(solver, limiter):
    p = HydroParams(nx=6, ny=5, nz=4, solver=solver, limiter=limiter, boundtype=1, cfl=0.8, atomic=True)
    d = p.derive()
    q = ics.prim_to_q(d, np.full((4, 5, 6), 1.3), np.full((4, 5, 6), 0.2), np.full((4, 5, 6), -0.1),
                      np.full((4, 5, 6), 0.05), np.full((4, 5, 6), 0.7))
    # q + delta*(f - f) == q exactly for the interior; walls see a mirrored state
    p2 = HydroParams(nx=6, ny=5, nz=4, solver=solver, limiter=limiter, boundtype=2, cfl=0.8, atomic=True)
    q0 = q.copy()
    O.Oracle(p2.derive()).run(q, 3)
    assert np.array_equal(q, q0)


@pytest.mark.parametrize("solver", [1, 2])
def test_periodic_mass_conserved(solver):
    p = HydroParams(nx=12, ny=10, nz=8, solver=solver, limiter=2, boundtype=2, cfl=0.3, atomic=True)
    d = p.derive()
    q = ics.random_smooth(p, seed=3, amp=0.1)
    m0 = q[..., 0].sum()
    e0 = q[..., 4].sum()
    O.Oracle(d).run(q, 5)
    assert q[..., 0].sum() == pytest.approx(m0, rel=2e-14)
    assert q[..., 4].sum() == pytest.approx(e0, rel=2e-14)
    # an oracle that dropped a flux term would fail this; so would a wrong index
    assert not np.array_equal(q[..., 0], ics.random_smooth(p, seed=3, amp=0.1)[..., 0])


@pytest.mark.parametrize("solver", [1, 2])
def test_x_mirror_symmetry(solver):
    p = HydroParams(nx=10, ny=6, nz=5, solver=solver, limiter=1, boundtype=1, cfl=0.3, atomic=True)
    d = p.derive()
    q = ics.random_smooth(p, seed=11, amp=0.1, jumps=False)
    qm = q[:, :, ::-1, :].copy()
    qm[..., 1] = -qm[..., 1]
    orc = O.Oracle(d)
    orc.run(q, 3)
    orc.run(qm, 3)
    back = qm[:, :, ::-1, :].copy()
    back[..., 1] = -back[..., 1]
    for c in range(6):
        assert np.max(np.abs(back[..., c] - q[..., c])) <= 1e-13 * max(1.0, np.max(np.abs(q[..., c])))


def test_axis_permutation_on_cube():
    p = HydroParams(nx=8, ny=8, nz=8, solver=2, limiter=3, boundtype=2, cfl=0.3, atomic=True)
    d = p.derive()
    q = ics.random_smooth(p, seed=12, amp=0.1, jumps=False)
    # swap x <-> z: numpy axes (z, y, x) -> (x, y, z), momentum components 1 <-> 3
    qt = q.transpose(2, 1, 0, 3)[..., [0, 3, 2, 1, 4, 5]].copy()
    orc = O.Oracle(d)
    orc.run(q, 2)
    orc.run(qt, 2)
    back = qt.transpose(2, 1, 0, 3)[..., [0, 3, 2, 1, 4, 5]]
    for c in range(6):
        assert np.max(np.abs(back[..., c] - q[..., c])) <= 1e-13 * max(1.0, np.max(np.abs(q[..., c])))


def test_1d_equals_3d_uniform_in_y_z():
    p1 = ics.sod_params(64, tend=0.05)
    q1 = ics.sod(p1)
    p3 = ics.sod_params(64, tend=0.05)
    p3.ny = 4
    p3.nz = 4
    p3.ymax = 4 / 64
    p3.zmax = 4 / 64
    p3.boundtype = 1
    q3 = np.broadcast_to(q1, (4, 4, 64, 6)).copy()
    O.Oracle(p1.derive()).run(q1, 1000, 0.0, p1.tend)
    O.Oracle(p3.derive()).run(q3, 1000, 0.0, p3.tend)
    for c in (0, 1, 4):
        assert rel_l1(q3[1, 2, :, c], q1[0, 0, :, c]) < 1e-14


def test_riemann_closed_forms():
    """HLL/HLLC of identical states is the physical flux; HLLC resolves an
    isolated contact exactly."""
    p = HydroParams(nx=4, ny=4, nz=4, solver=2, atomic=True)
    d = p.derive()
    orc = O.Oracle(d)
    w = np.array([1.3, 0.4, -0.2, 0.1, 0.9])
    for direction in (1, 2, 3):
        f = orc.riemann(w, w, direction)
        ud = w[direction]
        u2 = (w[1:4] ** 2).sum()
        expect = np.array([w[0] * ud, *(w[0] * ud * w[1:4]), ud * (0.5 * w[0] * u2 + d.gg1 * w[4]),
                           ud * w[4] / w[0] ** d.g1])
        expect[direction] += w[4]
        assert np.allclose(f, expect, rtol=1e-14, atol=1e-15)
    # stationary contact: rho jumps, u = 0, p equal -> HLLC mass flux is exactly 0
    f = orc.riemann(np.array([1.0, 0, 0, 0, 1.0]), np.array([0.1, 0, 0, 0, 1.0]), 1)
    assert f[0] == 0.0 and f[1] == pytest.approx(1.0, rel=1e-15)
    p.solver = 1
    orc1 = O.Oracle(p.derive())
    f1 = orc1.riemann(np.array([1.0, 0, 0, 0, 1.0]), np.array([0.1, 0, 0, 0, 1.0]), 1)
    assert f1[0] != 0.0  # HLL diffuses the contact


def test_slope_limiters_closed_forms():
    for limiter, expect in ((1, 0.5), (2, 1.0), (3, 2 * 0.5 * 1.5 / 2.0)):
        p = HydroParams(limiter=limiter)
        orc = O.Oracle(p.derive())
        s = orc.comp_slope(np.full(5, 1.0), np.full(5, 1.5), np.full(5, 3.0))
        assert np.allclose(s, expect, rtol=1e-15)
        # opposite signs -> 0 for every limiter
        assert np.all(orc.comp_slope(np.full(5, 1.0), np.full(5, 1.5), np.full(5, 1.2)) == 0)
    # MC central branch: dwl = 1, dwr = 1.2 -> min(2, 1.1, 2.4) = 1.1
    orc = O.Oracle(HydroParams(limiter=2).derive())
    assert orc.comp_slope(np.zeros(5), np.ones(5), np.full(5, 2.2))[0] == pytest.approx(1.1, rel=1e-15)


def test_bad_limiter_and_solver_codes():
    p = HydroParams(nx=4, ny=4, nz=4, limiter=4)
    with pytest.raises(O.OracleError) as ei:
        O.Oracle(p.derive()).run(ics.random_smooth(p, seed=1), 1)
    assert ei.value.code == 7
    p = HydroParams(nx=4, ny=4, nz=4, solver=3)
    with pytest.raises(O.OracleError) as ei:
        O.Oracle(p.derive()).run(ics.random_smooth(p, seed=1), 1)
    assert ei.value.code == 6


def test_negative_pressure_code():
    p = HydroParams(nx=4, ny=4, nz=4)
    d = p.derive()
    q = ics.random_smooth(p, seed=1)
    q[1, 2, 3, 4] = 0.5 * (q[1, 2, 3, 1:4] ** 2).sum() / q[1, 2, 3, 0] * 0.5  # E below kinetic
    with pytest.raises(O.OracleError) as ei:
        O.Oracle(d).run(q, 1)
    assert ei.value.code == 1 and ei.value.cell == (4, 3, 2)


# ---------------------------------------------------------------- thermal conduction
def test_conduction_linear_profile_closed_form():
    """conduction.f90 with simplekappa: j = -(1/g1) dT/dx at each face.  For
    a temperature linear in x (exactly representable values) every face flux
    is the same number, so interior cells do not change at all; the mirror
    ghosts of `bound` make the two end faces carry no flux, so the hot end
    loses exactly the flux of its inner face and the cold end gains it."""
    p = HydroParams(nx=8, ny=4, nz=4, boundtype=1, atomic=True, conduction=True, simplekappa=True)
    d = p.derive()
    o = O.Oracle(d)
    shape = (d.nz, d.ny, d.nx)
    q = np.zeros(shape + (6,))
    q[..., 0] = 1.0
    temp = 64.0 - 4.0 * np.arange(d.nx)  # 64, 60, ..., 36
    q[..., 4] = np.broadcast_to(temp, shape) / d.g1  # p = rho T (temp0 = 1), u = 0
    q[..., 5] = np.broadcast_to(temp, shape)  # p / rho**g1 with rho = 1
    w = O.new_w(d)
    o.primitive(w, q)
    o.bound(w)
    q1 = q.copy()
    dt = 1e-3
    O._check(o.L.ev_thermalconduction(O.C.byref(o.p), O._ptr(q1), O._ptr(w), dt))
    de = q1[..., 4] - q[..., 4]
    assert np.all(de[..., 1:-1] == 0.0)
    j = -(1.0 / d.g1) * (-4.0) * d.dxinv  # the flux of every inner face
    assert de[..., 0] == pytest.approx(dt * d.dxinv * (0.0 - j), rel=1e-14)
    assert de[..., -1] == pytest.approx(dt * d.dxinv * (j - 0.0), rel=1e-14)
    assert np.all(de[..., 0] < 0) and np.all(de[..., -1] > 0)  # heat flows from hot to cold


def test_conduction_periodic_conserves_energy():
    p = ics.conduction_params(16, boundtype=2)
    q = ics.conduction_blob(p)
    o = O.Oracle(p.derive())
    s0 = o.sums(q)
    o.run(q, 5)
    s1 = o.sums(q)
    assert s1[0] == pytest.approx(s0[0], rel=1e-14)
    assert s1[1] == pytest.approx(s0[1], rel=1e-13)


def test_conduction_uniform_state_unchanged():
    p = ics.conduction_params(12, boundtype=1)
    d = p.derive()
    q = np.zeros((12, 12, 12, 6))
    q[..., 0] = 1.0
    q[..., 4] = 1000.0 / d.g1
    q[..., 5] = 1000.0
    q1 = q.copy()
    O.Oracle(d).run(q1, 3)
    assert np.array_equal(q1, q)


def test_conduction_refuses_full_kappa():
    """The reference's kappa_eff with simplekappa off reads dtempdx before
    assigning it (conduction.f90:134-141): that branch is refused."""
    p = ics.conduction_params(8, simplekappa=False)
    q = ics.conduction_blob(p)
    with pytest.raises(O.OracleError):
        O.Oracle(p.derive()).run(q, 1)


def test_thermal_test_follows_the_error_function(refcase_fixtures, golden_index):
    """The reference's conduction test (bin/test:162-173): T = 400 | 800
    across x = 50 on 100 cells, hydro off, simplekappa (kappa_eff = 1/g1,
    so dT/dt = d2T/dx2 with rho = temp0 = 1).  bin/test:392 overlays the
    solution 600 + 200 erf((x - 50) / (2 sqrt(t))) at t = 10; the oracle's
    run to t = 10 lies within 0.2 K of it everywhere (0.13 K measured, the
    grid's second-order error) and is antisymmetric about x = 50."""
    import math

    from helpers import params_from_index
    p = params_from_index(golden_index["thermal100"])
    d = p.derive()
    q = refcase_fixtures["thermal100/q0"].copy()
    t, dt, n = O.Oracle(d).run(q, 100000, p.tstart, p.tend)
    assert t == p.tend and n == 49
    T = d.g1 * q[0, 0, :, 4] / q[0, 0, :, 0] * d.temp0
    x = d.x0 + np.arange(d.nx) * d.dx
    exact = np.array([600 + 200 * math.erf((xi - 50) / (2 * math.sqrt(t))) for xi in x])
    assert np.abs(T - exact).max() <= 0.2
    assert np.abs(T + T[::-1] - 1200).max() <= 1e-9
    assert np.array_equal(q, refcase_fixtures["thermal100/q"])
