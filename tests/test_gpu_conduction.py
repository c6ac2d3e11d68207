"""Thermal conduction on the GPU (evh_thermalstep, evh_thermalconduction)
against the oracle, itself matched bit for bit to the compiled reference on
the conduction fixtures (timestep.f90:72-104, conduction.f90:26-164,
simplekappa).

Bars: the fast fused form as for hydro (relative L1 <= 1e-10 after the
5-step fixture runs, relative L-infinity <= 1e-12 after one step); the
exact form to 1e-12 (the conduction limit takes T**2.5 and the entropy
update q(1)**g1 from the device power, which may differ from the host
libm's in the last bit); FP32-mixed relative L1 of rho and p <= 1e-5.
"""

import numpy as np
import pytest

import oracle as O
from conftest import rel_l1, rel_linf
from evora_amd import hydro, ics
from test_gpu_dropin import run_dropin

pytestmark = pytest.mark.gpu

FIXTURES = ["cond16_s2_b2", "cond16_s1_b1", "cond16_s2_b3"]


def params_of(index_entry):
    from evora_amd.params import HydroParams
    return HydroParams(**index_entry["params"])


@pytest.mark.parametrize("mode,tol", [(hydro.MODE_FUSED, 1e-10), (hydro.MODE_FUSED_EXACT, 1e-12)])
@pytest.mark.parametrize("name", FIXTURES)
def test_conduction_fixture_runs(name, mode, tol, run_fixtures, golden_index):
    p = params_of(golden_index[name])
    q0 = run_fixtures[name + "/q0"]
    t_dt = run_fixtures[name + "/t_dt"]
    with hydro.Hydro(p.derive(), mode=mode) as h:
        h.upload_q(q0)
        t, dt, done = h.run(len(t_dt))
        q = h.download_q()
    assert done == len(t_dt)
    assert t == pytest.approx(t_dt[-1][0], rel=1e-12)
    assert dt == pytest.approx(t_dt[-1][1], rel=1e-12)
    ref = run_fixtures[name + "/q"]
    for c in range(6):
        assert rel_l1(q[..., c], ref[..., c]) <= tol, c


def test_conduction_one_step_64():
    p = ics.conduction_params(64, boundtype=1)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 1)
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        h.run(1)
        q = h.download_q()
    for c in range(6):
        assert rel_linf(q[..., c], qo[..., c]) <= 1e-12, c


def test_thermalstep_limit():
    p = ics.conduction_params(24, boundtype=2)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    o = O.Oracle(d)
    w = O.new_w(d)
    q = q0.copy()
    o.primitive(w, q)
    o.bound(w)
    dt_o = O.C.c_double(1e99)
    O._check(o.L.ev_thermalstep(O.C.byref(o.p), O._ptr(q), O._ptr(w), O.C.byref(dt_o)))
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        h.primitive()
        dt = h.thermalstep(1e99)
    assert dt == pytest.approx(dt_o.value, rel=1e-14)


def test_conduction_mixed_precision():
    p = ics.conduction_params(32, boundtype=2)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 5)
    with hydro.Hydro(d, precision=hydro.FP32_MIXED) as h:
        h.upload_q(q0)
        h.run(5)
        q = h.download_q()
    for c in (0, 4):
        assert rel_l1(q[..., c], qo[..., c]) <= 1e-5, c


@pytest.mark.parametrize("kw", [dict(simplekappa=False), dict(dualenergy=True, de_thresh=0.01)])
def test_conduction_refused_configs(kw):
    p = ics.conduction_params(16, **kw)
    with pytest.raises(hydro.EvoraError):
        hydro.Hydro(p.derive())


@pytest.mark.parametrize("exact", [False, True])
def test_conduction_dropin(exact):
    """The Fortran drop-in with conduction on: thermalstep and
    thermalconduction dispatched to the GPU, against the compiled
    reference's own routines."""
    p = ics.conduction_params(20, boundtype=3, tend=10.0)
    q0 = ics.conduction_blob(p)
    env = {"EVORA_MAXSTEPS": "6"}
    if exact:
        env["EVORA_EXACT"] = "1"
    qc, sc = run_dropin(p, q0, 1000, "cpu", env)
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert len(sc) == len(sh) == 7
    tol = 1e-12 if exact else 1e-10
    for c in range(6):
        assert rel_l1(qh[..., c], qc[..., c]) <= tol, c


@pytest.mark.parametrize("mode", [hydro.MODE_FUSED, hydro.MODE_FUSED_EXACT])
def test_conduction_linear_profile_closed_form(mode):
    """A temperature linear in x with mirror ghosts (conduction.f90 with
    simplekappa): inner cells do not change, the hot end loses exactly the
    flux of its inner face (-(1/g1) dT/dx) and the cold end gains it.  Called
    on an uploaded state, so T comes from the stand-alone temperature pass."""
    from evora_amd.params import HydroParams
    p = HydroParams(nx=8, ny=4, nz=4, boundtype=1, atomic=True, conduction=True, simplekappa=True)
    d = p.derive()
    shape = (d.nz, d.ny, d.nx)
    q = np.zeros(shape + (6,))
    q[..., 0] = 1.0
    temp = 64.0 - 4.0 * np.arange(d.nx)
    q[..., 4] = np.broadcast_to(temp, shape) / d.g1
    q[..., 5] = np.broadcast_to(temp, shape)
    dt = 1e-3
    with hydro.Hydro(d, mode=mode) as h:
        h.upload_q(q)
        h.thermalconduction(dt)
        q1 = h.download_q()
    de = q1[..., 4] - q[..., 4]
    assert np.all(de[..., 1:-1] == 0.0)
    j = -(1.0 / d.g1) * (-4.0) * d.dxinv
    assert de[..., 0] == pytest.approx(dt * d.dxinv * (0.0 - j), rel=1e-14)
    assert de[..., -1] == pytest.approx(dt * d.dxinv * (j - 0.0), rel=1e-14)
    assert np.array_equal(q1[..., :4], q[..., :4])


def test_conduction_fullsize_conservation_and_forms():
    """256^3 periodic hot blob, 4 steps: mass and energy conserved to
    roundoff, and the fast form within 1e-12 (relative L-infinity per
    component) of the reference-arithmetic form."""
    p = ics.conduction_params(256, boundtype=2)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    out = {}
    for mode in (hydro.MODE_FUSED, hydro.MODE_FUSED_EXACT):
        with hydro.Hydro(d, mode=mode) as h:
            h.upload_q(q0)
            s0 = h.sums()
            h.run(4)
            s1 = h.sums()
            out[mode] = h.download_q()
        assert s1[0] == pytest.approx(s0[0], rel=1e-13)
        assert s1[1] == pytest.approx(s0[1], rel=1e-13)
    for c in range(6):
        assert rel_linf(out[hydro.MODE_FUSED][..., c], out[hydro.MODE_FUSED_EXACT][..., c]) <= 1e-12, c


@pytest.mark.parametrize("shape", [(2, 3, 2), (3, 2, 2)])
def test_conduction_smallest_grids(shape):
    """2 and 3 cells per direction: the ring and the z neighbours of the
    conduction pass come from the ghost rules across the whole grid."""
    p = ics.conduction_params(2, nx=shape[0], ny=shape[1], nz=shape[2], boundtype=1)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    qo = q0.copy()
    O.Oracle(d).run(qo, 2)
    with hydro.Hydro(d) as h:
        h.upload_q(q0)
        h.run(2)
        q = h.download_q()
    for c in range(6):
        assert rel_linf(q[..., c], qo[..., c]) <= 1e-12, c
