"""The CPU oracle against the reference's own outputs (golden fixtures).

The fixtures were written by tests/golden/make_golden.py from the compiled
reference Fortran.  Every comparison here is bitwise (signed zeros compare
equal): the oracle restates the reference in the same operation order.
"""
import numpy as np
import pytest

import oracle as O
from helpers import params_from_index

STAGES = [n for n in __import__("json").load(open(__import__("os").path.join(
    __import__("os").path.dirname(__file__), "golden", "index.json"))) if n.startswith("stage_")]


@pytest.mark.parametrize("name", STAGES)
def test_stage_bitwise(name, stage_fixtures, golden_index):
    p = params_from_index(golden_index[name])
    d = p.derive()
    w, evo, s, q1, dt = O.Oracle(d).stage(stage_fixtures[name + "/q0"])
    ref_w = stage_fixtures[name + "/w"]
    # ghost layers of degenerate dimensions are never written by the reference
    sl = (slice(None) if d.nz > 1 else slice(2, 3), slice(None) if d.ny > 1 else slice(2, 3))
    assert np.array_equal(w[sl[0], sl[1]], ref_w[sl[0], sl[1]])
    assert np.array_equal(evo, stage_fixtures[name + "/evo"])
    assert np.array_equal(s, stage_fixtures[name + "/s"])
    assert np.array_equal(q1, stage_fixtures[name + "/q1"])
    assert dt == stage_fixtures[name + "/dt"][0]


@pytest.mark.parametrize("name", ["sod1000", "run16_s2_l1_b2", "run16_s1_l3_b1", "run16_s2_l2_b3"])
def test_run_bitwise(name, run_fixtures, golden_index):
    p = params_from_index(golden_index[name])
    q = run_fixtures[name + "/q0"].copy()
    t_dt = run_fixtures[name + "/t_dt"]
    t, dt, n = O.Oracle(p.derive()).run(q, len(t_dt), p.tstart, p.tend)
    assert n == len(t_dt)
    assert (t, dt) == tuple(t_dt[-1])
    assert np.array_equal(q, run_fixtures[name + "/q"])


def test_hll_blast_aborts_like_reference(run_fixtures, golden_index):
    """The reference HLL stops on negative pressure in the CFL-0.6 blast."""
    e = golden_index["blast32_hll"]
    p = params_from_index(e)
    q = run_fixtures["blast32_hll/q0"].copy()
    orc = O.Oracle(p.derive())
    with pytest.raises(O.OracleError) as ei:
        orc.run(q, 20)
    assert ei.value.code == 1  # comp.f90:74-83
    assert "negative pressure" in e["error"][0]
    # the same number of completed steps before the stop
    q = run_fixtures["blast32_hll/q0"].copy()
    t, dt, n = orc.run(q, e["steps_done"])
    assert n == e["steps_done"]
