"""GPU parity of the fused one-pass FP64 step against the CPU oracle.

Bar: relative L-infinity <= 1e-12 per component after one
step, relative L1 <= 1e-10 after a multi-step run.  rel L-inf of a component
is max|gpu - oracle| / max|oracle|.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l1, rel_linf
from evora_amd import hydro, ics
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu

ONE_STEP_TOL = 1e-12
RUN_L1_TOL = 1e-10


def one_step(p, q0):
    d = p.derive()
    _, _, _, q1_o, dt_o = O.Oracle(d).stage(q0)
    with hydro.Hydro(d, mode=hydro.MODE_FUSED) as h:
        h.upload_q(q0)
        h.primitive()
        h.bound()
        dt = h.hydrostep(1e99)
        h.compevo(dt)
        h.sweep(dt)
        q1 = h.download_q()
        dt_next = h.hydrostep(1e99)  # CFL folded into the sweep
    return q1, q1_o, dt, dt_o, dt_next


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("limiter", [1, 2, 3])
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_fused_one_step(solver, limiter, bt):
    p = HydroParams(nx=19, ny=17, nz=11, solver=solver, limiter=limiter, boundtype=bt, cfl=0.4, atomic=True)
    q0 = ics.random_smooth(p, seed=100 * solver + 10 * limiter + bt)
    q1, q1_o, dt, dt_o, dt_next = one_step(p, q0)
    assert dt == dt_o
    for c in range(6):
        assert rel_linf(q1[..., c], q1_o[..., c]) <= ONE_STEP_TOL, c
    # the folded CFL equals a fresh hydrostep on the new state
    d = p.derive()
    w = O.new_w(d)
    qq = q1_o.copy()
    O.Oracle(d).primitive(w, qq)
    assert dt_next == pytest.approx(O.Oracle(d).hydrostep(w, 1e99), rel=1e-13)


@pytest.mark.parametrize("shape", [(30, 28, 20), (16, 45, 9), (64, 16, 70)])
def test_fused_tile_edges(shape):
    """Grids that are not multiples of the 16x12 tile or the z chunk."""
    nx, ny, nz = shape
    p = HydroParams(nx=nx, ny=ny, nz=nz, solver=2, limiter=1, boundtype=1, cfl=0.4, atomic=True)
    q0 = ics.random_smooth(p, seed=nx + ny + nz)
    q1, q1_o, dt, dt_o, _ = one_step(p, q0)
    for c in range(6):
        assert rel_linf(q1[..., c], q1_o[..., c]) <= ONE_STEP_TOL, c


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("limiter", [1, 2, 3])
def test_fused_one_step_dual_energy(solver, limiter):
    """dualenergy on with de_thresh 0.98: cells take both branches of
    comp.f90:60-69, and primitive rewrites q(5) or q(6)."""
    p = HydroParams(nx=21, ny=18, nz=13, solver=solver, limiter=limiter, boundtype=2, cfl=0.4, atomic=True,
                    dualenergy=True, de_thresh=0.98)
    q0 = ics.random_smooth(p, seed=300 + 10 * solver + limiter)
    q1, q1_o, dt, dt_o, _ = one_step(p, q0)
    assert dt == dt_o
    for c in range(6):
        assert rel_linf(q1[..., c], q1_o[..., c]) <= ONE_STEP_TOL, c


@pytest.mark.parametrize("name", ["blast32_hll", "kh24_cfl08", "cosmo16_nodual"])
def test_fused_stops_where_reference_stops(name, run_fixtures, golden_index):
    """At the configured CFL the reference stops on a negative pressure; the
    fused path completes the same steps and stops in the next one."""
    e = golden_index[name]
    p = HydroParams(**e["params"])
    with hydro.Hydro(p, mode=hydro.MODE_FUSED) as h:
        h.upload_q(run_fixtures[name + "/q0"])
        t, dt, n = h.run(e["steps_done"])
        assert n == e["steps_done"]
        with pytest.raises(hydro.EvoraError) as ei:
            h.run(1, t)
    assert ei.value.code == 1


@pytest.mark.parametrize("name", ["run16_s2_l1_b2", "run16_s1_l3_b1", "run16_s2_l2_b3", "cosmo24_de"])
def test_fused_run_matches_reference_fixture(name, run_fixtures, golden_index):
    p = HydroParams(**golden_index[name]["params"])
    q0 = run_fixtures[name + "/q0"]
    t_dt = run_fixtures[name + "/t_dt"]
    with hydro.Hydro(p, mode=hydro.MODE_FUSED) as h:
        h.upload_q(q0)
        t, dt, n = h.run(len(t_dt), p.tstart, p.tend)
        q = h.download_q()
    assert n == len(t_dt)
    assert t == pytest.approx(t_dt[-1][0], rel=1e-13)
    for c in range(6):
        assert rel_l1(q[..., c], run_fixtures[name + "/q"][..., c]) <= RUN_L1_TOL


def test_fused_uniform_state_unchanged():
    p = HydroParams(nx=20, ny=18, nz=16, solver=2, limiter=2, boundtype=2, cfl=0.8, atomic=True)
    d = p.derive()
    sh = (16, 18, 20)
    q0 = ics.prim_to_q(d, np.full(sh, 1.3), np.full(sh, 0.2), np.full(sh, -0.1), np.full(sh, 0.05),
                       np.full(sh, 0.7))
    with hydro.Hydro(d, mode=hydro.MODE_FUSED) as h:
        h.upload_q(q0)
        h.run(3)
        q = h.download_q()
    assert np.array_equal(q, q0)


def test_fused_matches_staged():
    p = HydroParams(nx=24, ny=20, nz=18, solver=2, limiter=3, boundtype=3, cfl=0.4, atomic=True)
    q0 = ics.random_smooth(p, seed=9)
    out = []
    for mode in (hydro.MODE_STAGED, hydro.MODE_FUSED):
        with hydro.Hydro(p, mode=mode) as h:
            h.upload_q(q0)
            h.run(4)
            out.append(h.download_q())
    for c in range(6):
        assert rel_linf(out[1][..., c], out[0][..., c]) <= 1e-12


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 2), (2, 5, 3), (5, 3, 2), (17, 2, 9), (2, 13, 2)])
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_fused_smallest_grids(shape, bt):
    """Grids down to the fused path's minimum of 2 cells per direction, where
    the two ghost layers span the whole grid and the march's prefetch reaches
    past the last plane: 3 steps of the fast form (relative L-infinity
    <= 1e-12), the exact form (density, momentum and energy bit for bit) and
    FP32-mixed (relative L1 of rho and p <= 1e-5) against the oracle."""
    nx, ny, nz = shape
    p = HydroParams(nx=nx, ny=ny, nz=nz, solver=2, limiter=2, boundtype=bt, cfl=0.4, atomic=True)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=nx * 100 + ny * 10 + nz + bt, jumps=False)
    qo = q0.copy()
    O.Oracle(d).run(qo, 3)
    for mode, prec in ((hydro.MODE_FUSED, hydro.FP64), (hydro.MODE_FUSED_EXACT, hydro.FP64),
                       (hydro.MODE_FUSED, hydro.FP32_MIXED)):
        with hydro.Hydro(d, precision=prec, mode=mode) as h:
            h.upload_q(q0)
            h.run(3)
            q = h.download_q()
        if prec == hydro.FP32_MIXED:
            for c in (0, 4):
                assert rel_l1(q[..., c], qo[..., c]) <= 1e-5, (mode, prec, c)
        elif mode == hydro.MODE_FUSED_EXACT:
            assert np.array_equal(q[..., :5], qo[..., :5])
        else:
            for c in range(6):
                assert rel_linf(q[..., c], qo[..., c]) <= 1e-12, (mode, prec, c)


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("limiter", [1, 2, 3])
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_fused_2d(solver, limiter, bt):
    """The single-pass 2-D form (nz = 1, muscl.f90:70-88 and the x/y sweeps):
    one step of the fast form to 1e-12 per component, the exact form equal to
    the oracle bit for bit in q1-q5 over 3 steps."""
    p = HydroParams(nx=37, ny=29, nz=1, solver=solver, limiter=limiter, boundtype=bt, cfl=0.4, atomic=True)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=300 + 100 * solver + 10 * limiter + bt)
    q1, q1_o, dt, dt_o, _ = one_step(p, q0)
    assert dt == dt_o
    for c in range(6):
        assert rel_linf(q1[..., c], q1_o[..., c]) <= ONE_STEP_TOL, c
    qo = q0.copy()
    to, dto, _ = O.Oracle(d).run(qo, 3)
    with hydro.Hydro(d, mode=hydro.MODE_FUSED_EXACT) as h:
        h.upload_q(q0)
        t, dt, n = h.run(3)
        q = h.download_q()
    assert n == 3 and (t, dt) == (to, dto)
    for c in range(5):
        assert np.array_equal(q[..., c], qo[..., c]), c
    assert rel_linf(q[..., 5], qo[..., 5]) <= 1e-13


@pytest.mark.parametrize("shape", [(130, 70), (600, 300)])
@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
def test_fused_2d_run(precision, shape):
    """20 steps of the 2-D form, on one partial row strip and on three strips
    of several row chunks: fast FP64 to relative L1 1e-10, FP32-mixed to
    1e-5 in rho and p."""
    nx, ny = shape
    p = HydroParams(nx=nx, ny=ny, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.4, atomic=True)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=77, amp=0.1)
    qo = q0.copy()
    O.Oracle(d).run(qo, 20)
    with hydro.Hydro(d, precision=precision) as h:
        h.upload_q(q0)
        h.run(20)
        q = h.download_q()
    if precision == hydro.FP64:
        for c in range(6):
            assert rel_l1(q[..., c], qo[..., c]) <= RUN_L1_TOL, c
    else:
        kin = lambda a: 0.5 * (a[..., 1] ** 2 + a[..., 2] ** 2 + a[..., 3] ** 2) / a[..., 0]
        assert rel_l1(q[..., 0], qo[..., 0]) <= 1e-5
        assert rel_l1(q[..., 4] - kin(q), qo[..., 4] - kin(qo)) <= 1e-5


def _run_tuned(d, q0, prec, mode, tuning, steps):
    with hydro.Hydro(d, precision=prec, mode=mode) as h:
        h.set_tuning(*tuning)
        h.upload_q(q0)
        t = 0.0
        for _ in range(steps):
            t, dt, go = h.step(t, 1e30)
        return h.download_q(), t


@pytest.mark.parametrize("form", ["fp64", "fp64_exact", "fp32_mixed"])
@pytest.mark.parametrize("shape", [(24, 24, 24), (40, 36, 20)])
def test_fused_result_independent_of_tiling(form, shape):
    """The tile height, the waves-per-SIMD budget and the z chunk change
    which lanes, call sites and waves form a cell's values (ring lane or
    interior lane, first planes of a chunk or the plane ahead, a y face
    solved by the ring wave as an x face with u and v exchanged): every
    tuning must give the same bits after several steps of a blast."""
    nx, ny, nz = shape
    p = ics.blast_params(nx, solver=2, cfl=0.4, tend=1e30)
    p.ny, p.nz = ny, nz
    d = p.derive()
    q0 = ics.blast(p)
    prec = hydro.FP32_MIXED if form == "fp32_mixed" else hydro.FP64
    mode = hydro.MODE_FUSED_EXACT if form == "fp64_exact" else hydro.MODE_FUSED
    ref, t_ref = _run_tuned(d, q0, prec, mode, (12, 1, 0), 4)
    for ty in (8, 12):
        for hi in (0, 1):
            for zc in (0, 8, 7):
                q, t = _run_tuned(d, q0, prec, mode, (ty, hi, zc), 4)
                assert t == t_ref, (ty, hi, zc)
                diff = np.argwhere(np.any(q != ref, axis=-1))
                assert len(diff) == 0, ((ty, hi, zc), len(diff), diff[:4].tolist())


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
def test_fused_2d_independent_of_row_chunks(precision):
    """2-D form: the row chunk (the z chunk of the tuning) sets where a strip
    starts its march; every chunk gives the same bits."""
    p = HydroParams(nx=300, ny=90, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.5, atomic=False)
    d = p.derive()
    q0 = ics.kh_reference(p)
    ref, t_ref = _run_tuned(d, q0, precision, hydro.MODE_FUSED, (12, 1, 0), 4)
    for zc in (7, 16, 45):
        q, t = _run_tuned(d, q0, precision, hydro.MODE_FUSED, (12, 1, zc), 4)
        assert t == t_ref and np.array_equal(q, ref), zc
