"""Error paths of the HIP path return the reference's stop codes.

Every condition is planted in the input on the host (no kernel is made to
fault): E below the kinetic energy (comp.f90:74-83), the HLL blast that the
reference itself stops on, and invalid solver/limiter codes rejected at
context creation (riemann.f90:37-41, muscl.f90:164-167).
"""
import numpy as np
import pytest

import oracle as O
from evora_amd import hydro, ics
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [hydro.MODE_STAGED, hydro.MODE_FUSED])
def test_negative_pressure_stops(mode):
    p = HydroParams(nx=8, ny=6, nz=5, atomic=True)
    q = ics.random_smooth(p, seed=1)
    q[1, 2, 3, 4] = 0.25 * (q[1, 2, 3, 1:4] ** 2).sum() / q[1, 2, 3, 0]
    with hydro.Hydro(p, mode=mode) as h:
        h.upload_q(q)
        with pytest.raises(hydro.EvoraError) as ei:
            h.run(1)
    assert ei.value.code == 1
    assert ei.value.cell == (4, 3, 2)
    assert "negative pressure" in ei.value.message


@pytest.mark.parametrize("mode", [hydro.MODE_STAGED, hydro.MODE_FUSED])
def test_hll_blast_stops_like_reference(mode, run_fixtures, golden_index):
    e = golden_index["blast32_hll"]
    p = HydroParams(**e["params"])
    q0 = run_fixtures["blast32_hll/q0"]
    with hydro.Hydro(p, mode=mode) as h:
        h.upload_q(q0)
        t, dt, n = h.run(e["steps_done"])
        assert n == e["steps_done"]
        with pytest.raises(hydro.EvoraError) as ei:
            h.run(1, t)
    assert ei.value.code == 1


def test_invalid_codes_rejected():
    for kw, code in ((dict(solver=3), 6), (dict(limiter=4), 7), (dict(boundtype=4), 8)):
        p = HydroParams(nx=4, ny=4, nz=4, **kw)
        with pytest.raises(hydro.EvoraError) as ei:
            hydro.Hydro(p)
        assert ei.value.code == code


def test_calls_out_of_order():
    p = HydroParams(nx=8, ny=8, nz=8, atomic=True)
    with hydro.Hydro(p, mode=hydro.MODE_STAGED) as h:
        h.upload_q(ics.random_smooth(p, seed=2))
        with pytest.raises(hydro.EvoraError) as ei:
            h.bound()
        assert ei.value.code == 10
