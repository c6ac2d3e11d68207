"""GPU parity of the staged kernels (one kernel per reference routine).

Each routine is compared with the CPU oracle on the same input: w after
primitive+bound, dt from hydrostep, evo and s from compevo, q after sweep.
The staged kernels are built without FP contraction and follow the
reference operation order, so every value is bit-identical except the
entropy component q6, whose pow() comes from a different math library
(tolerance 1e-14 relative, written below).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_linf
from evora_amd import hydro, ics
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu

GRIDS = {"1d": (24, 1, 1), "2d": (10, 9, 1), "3d": (8, 7, 6)}


def assert_q_parity(q_gpu, q_ref, tol6=1e-14):
    for c in range(5):
        assert np.array_equal(q_gpu[..., c], q_ref[..., c]), f"component {c + 1}"
    assert rel_linf(q_gpu[..., 5], q_ref[..., 5]) <= tol6


@pytest.mark.parametrize("dim", list(GRIDS))
@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("limiter", [1, 2, 3])
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_staged_routines_match_oracle(dim, solver, limiter, bt):
    nx, ny, nz = GRIDS[dim]
    p = HydroParams(nx=nx, ny=ny, nz=nz, solver=solver, limiter=limiter, boundtype=bt, cfl=0.4, atomic=True)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=100 * solver + 10 * limiter + bt)
    w_o, evo_o, s_o, q1_o, dt_o = O.Oracle(d).stage(q0)
    with hydro.Hydro(d, mode=hydro.MODE_STAGED) as h:
        h.upload_q(q0)
        h.primitive()
        h.bound()
        dt = h.hydrostep(1e99)
        assert dt == dt_o
        w = h.download_w()
        sl = (slice(None) if nz > 1 else slice(2, 3), slice(None) if ny > 1 else slice(2, 3))
        assert np.array_equal(w[sl], w_o[sl])
        h.compevo(dt)
        evo, s = h.download_evo_s()
        assert np.array_equal(evo, evo_o)
        assert np.array_equal(s, s_o)
        h.sweep(dt)
        assert_q_parity(h.download_q(), q1_o)


STAGES = [n for n in __import__("json").load(open(__import__("os").path.join(
    __import__("os").path.dirname(__file__), "golden", "index.json"))) if n.startswith("stage_")]


@pytest.mark.parametrize("name", STAGES)
def test_staged_matches_reference_fixture(name, stage_fixtures, golden_index):
    """Every routine's output against the reference Fortran's, bit for bit
    (q6 to 1e-14: its pow comes from a different math library)."""
    p = HydroParams(**golden_index[name]["params"])
    d = p.derive()
    with hydro.Hydro(d, mode=hydro.MODE_STAGED) as h:
        h.upload_q(stage_fixtures[name + "/q0"])
        h.primitive()
        h.bound()
        dt = h.hydrostep(1e99)
        assert dt == stage_fixtures[name + "/dt"][0]
        w = h.download_w()
        sl = (slice(None) if d.nz > 1 else slice(2, 3), slice(None) if d.ny > 1 else slice(2, 3))
        ref_w = stage_fixtures[name + "/w"]
        if d.dualenergy:  # w(5) comes from pow in the entropy branch
            assert np.array_equal(w[sl][..., :4], ref_w[sl][..., :4])
            assert rel_linf(w[sl][..., 4], ref_w[sl][..., 4]) <= 1e-14
        else:
            assert np.array_equal(w[sl], ref_w[sl])
        h.compevo(dt)
        evo, s = h.download_evo_s()
        tol = 1e-13 if d.dualenergy else 0.0
        assert rel_linf(evo, stage_fixtures[name + "/evo"]) <= tol
        assert rel_linf(s, stage_fixtures[name + "/s"]) <= tol
        h.sweep(dt)
        q1 = h.download_q()
    if d.dualenergy:
        for c in range(6):
            assert rel_linf(q1[..., c], stage_fixtures[name + "/q1"][..., c]) <= 1e-13
    else:
        assert_q_parity(q1, stage_fixtures[name + "/q1"])


def test_staged_sod_run_matches_reference_fixture(run_fixtures, golden_index):
    p = HydroParams(**golden_index["sod1000"]["params"])
    q0 = run_fixtures["sod1000/q0"]
    t_dt = run_fixtures["sod1000/t_dt"]
    with hydro.Hydro(p) as h:
        h.upload_q(q0)
        t, dt, n = h.run(100000, p.tstart, p.tend)
        q = h.download_q()
    assert n == len(t_dt)
    assert t == t_dt[-1][0] and dt == t_dt[-1][1]
    assert_q_parity(q, run_fixtures["sod1000/q"], tol6=1e-13)


@pytest.mark.parametrize("shape", [(2, 1, 1), (3, 1, 1), (2, 2, 1), (3, 2, 1), (2, 5, 1), (2, 2, 2), (3, 2, 2)])
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_staged_smallest_grids(shape, bt):
    """Grids of 2 and 3 cells per direction, where the two ghost layers of
    `bound` span the whole grid: 3 staged steps against the oracle, bit for
    bit except the entropy."""
    nx, ny, nz = shape
    p = HydroParams(nx=nx, ny=ny, nz=nz, solver=2, limiter=1, boundtype=bt, cfl=0.4, atomic=True)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=nx + ny + nz + bt, jumps=False)
    qo = q0.copy()
    to, dto, _ = O.Oracle(d).run(qo, 3)
    with hydro.Hydro(d, mode=hydro.MODE_STAGED) as h:
        h.upload_q(q0)
        t, dt, n = h.run(3)
        q = h.download_q()
    assert n == 3 and t == to and dt == dto
    assert_q_parity(q, qo)
