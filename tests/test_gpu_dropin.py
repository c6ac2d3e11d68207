"""The Fortran drop-in: the same driver loop (oracle/ref_driver.f90, main.f90's
role) linked with evora_amd/fortran/hydro_dispatch.f90 runs the reference
routines (EVORA_BACKEND=cpu) or the GPU path (EVORA_BACKEND=hip) on the same
input; the GPU result must match the CPU one to the parity bar.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from conftest import rel_l1, rel_linf
from evora_amd import ics
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(O.REF_PATH), "evref_hip")


def run_dropin(params, q0, nsteps, backend, extra_env=None, evo=None):
    """Run the harness; with evo=[] the lines of data/evo are appended to it."""
    env = dict(os.environ, EVORA_BACKEND=backend)
    if extra_env:
        env.update(extra_env)
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "data"))
        para = os.path.join(td, "para")
        with open(para, "w") as fh:
            fh.write(params.to_namelist())
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        np.ascontiguousarray(q0).tofile(fin)
        r = subprocess.run([EXE, para, "run", fin, fout, str(nsteps)], env=env, cwd=td, capture_output=True,
                           text=True, timeout=300)
        assert "ERROR" not in r.stdout, r.stdout[-2000:]
        steps = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("STEP")]
        if evo is not None and os.path.exists(os.path.join(td, "data", "evo")):
            with open(os.path.join(td, "data", "evo")) as fh:
                evo.extend(fh.read().splitlines())
        return np.fromfile(fout).reshape(q0.shape), steps


@pytest.fixture(scope="module", autouse=True)
def need_exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C oracle ref-hip)")


def test_dropin_sod_staged_path():
    p = ics.sod_params(1000)
    q0 = ics.sod(p)
    qc, sc = run_dropin(p, q0, 100000, "cpu")
    qh, sh = run_dropin(p, q0, 100000, "hip")
    assert len(sc) == len(sh) == 575
    assert sc[-1][1:] == sh[-1][1:]  # t and dt printed identically
    for c in range(5):
        assert np.array_equal(qh[..., c], qc[..., c])
    assert rel_linf(qh[..., 5], qc[..., 5]) <= 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dropin_3d_fused_path_with_maxsteps(precision):
    p = HydroParams(nx=30, ny=26, nz=20, solver=2, limiter=1, boundtype=1, cfl=0.4, atomic=True, tend=10.0)
    q0 = ics.random_smooth(p, seed=21, amp=0.1)
    env = {"EVORA_MAXSTEPS": "8", "EVORA_PRECISION": precision}
    qc, sc = run_dropin(p, q0, 1000, "cpu", env)
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert len(sc) == len(sh) == 9  # 8 steps + the dt = 0 step that ends the run
    tol = 1e-10 if precision == "fp64" else 1e-5
    for c in range(6):
        assert rel_l1(qh[..., c], qc[..., c]) <= tol


def test_dropin_exact_matches_reference_bitwise():
    """EVORA_EXACT=1: the GPU runs the fused step with the reference's own
    arithmetic; the run (t and dt of every step, and q) equals the compiled
    reference's bit for bit except the entropy, where a power enters."""
    p = HydroParams(nx=30, ny=26, nz=20, solver=2, limiter=1, boundtype=3, cfl=0.4, atomic=True, tend=10.0)
    q0 = ics.random_smooth(p, seed=22, amp=0.1)
    env = {"EVORA_MAXSTEPS": "8", "EVORA_EXACT": "1"}
    qc, sc = run_dropin(p, q0, 1000, "cpu", env)
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert sc == sh
    for c in range(5):
        assert np.array_equal(qh[..., c], qc[..., c]), c
    assert rel_linf(qh[..., 5], qc[..., 5]) <= 1e-13


def test_dropin_evolution_on_device():
    """With the state on the GPU, evolution (snapshot.f90:159-193) takes the
    totals on the device instead of downloading q.  The lines of data/evo
    follow main.f90's rule (after a step, when (t - tstart)/(tend - tstart)
    * nsnap >= k2), carry the step's t, keep the periodic box's mass, and
    the last one equals the totals of the final state in the reference's
    E16.6 format."""
    p = HydroParams(nx=30, ny=26, nz=20, solver=2, limiter=1, boundtype=2, cfl=0.4, atomic=True,
                    tend=0.1, nsnap=5)
    q0 = ics.random_smooth(p, seed=23, amp=0.1)
    lines = []
    q1, steps = run_dropin(p, q0, 100000, "hip", {"EVREF_EVOLUTION": "1"}, evo=lines)
    rows = [[float(x) for x in ln.split()] for ln in lines if not ln.startswith("#")]
    ts = [float(st[2]) for st in steps]
    expect, k2 = [], 1
    for t in ts:
        if (t - p.tstart) / (p.tend - p.tstart) * p.nsnap >= k2:
            expect.append(t)
            k2 += 1
    assert len(rows) == len(expect) >= p.nsnap
    # E16.6 prints six significant digits: a value is exact to 5e-6 of itself
    for row, t in zip(rows, expect):
        assert row[0] == pytest.approx(t, rel=1e-5)
    vol = p.derive().volume
    totals = [q1[..., 0].sum() * vol, q1[..., 4].sum() * vol, q1[..., 5].sum() * vol]
    for got, want in zip(rows[-1][4:7], totals):
        assert got == pytest.approx(want, rel=1e-5)
    for row in rows:
        assert row[4] == pytest.approx(rows[0][4], rel=1e-5)


def test_dropin_thermal_test_without_hydro():
    """The reference's conduction test (hydro = .false., 100 cells, to t = 10)
    through the Fortran drop-in: the driver calls neither hydrostep nor sweep
    (main.f90:119,155), thermalstep counts the steps; the GPU run takes the
    CPU run's 49 steps and lands within relative L1 1e-12 of it (T**2.5 of the
    conduction limit comes from the device power)."""
    p = ics.thermaltest_params()
    q0 = ics.thermaltest(p)
    qc, sc = run_dropin(p, q0, 100000, "cpu")
    qh, sh = run_dropin(p, q0, 100000, "hip")
    assert len(sc) == len(sh) == 49
    assert float(sh[-1][2]) == float(sc[-1][2]) == p.tend
    for c in range(6):
        assert rel_l1(qh[..., c], qc[..., c]) <= 1e-12, c
    assert np.array_equal(qh[..., :4], q0[..., :4])  # no hydro: mass and momentum untouched


@pytest.mark.parametrize("exact", [False, True])
def test_dropin_2d_single_pass(exact):
    """A 2-D grid (para/kevinhelmholtz's shear layer at 64^2, gamma 1.4)
    through the Fortran drop-in takes the single-pass 2-D form: the exact form
    equals the compiled reference bit for bit in q1-q5 and in every printed
    t and dt, the fast form within relative L1 1e-10 after 12 steps."""
    p = HydroParams(nx=64, ny=64, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.5, atomic=False, tend=12.0)
    q0 = ics.kh_reference(p)
    env = {"EVORA_MAXSTEPS": "12"}
    if exact:
        env["EVORA_EXACT"] = "1"
    qc, sc = run_dropin(p, q0, 1000, "cpu", env)
    qh, sh = run_dropin(p, q0, 1000, "hip", env)
    assert len(sc) == len(sh) == 13
    if exact:
        assert sc == sh
        for c in range(5):
            assert np.array_equal(qh[..., c], qc[..., c]), c
        assert rel_linf(qh[..., 5], qc[..., 5]) <= 1e-13
    else:
        for c in range(6):
            assert rel_l1(qh[..., c], qc[..., c]) <= 1e-10, c
