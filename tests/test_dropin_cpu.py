"""The Fortran drop-in's CPU routing, without a GPU.

The drop-in harness (oracle/_ref/evref_hip: the driver loop linked with
evora_amd/fortran/hydro_dispatch.f90 and the reference objects whose
hot-path symbols are renamed to evcpu_*) must, with EVORA_BACKEND=cpu,
reproduce the plain reference harness (oracle/_ref/evref) bit for bit: the
dispatch only forwards to the reference routines, and the EVORA_MAXSTEPS
step cap ends both runs the same way.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from evora_amd import ics

HIP_EXE = os.path.join(os.path.dirname(O.REF_PATH), "evref_hip")


def _run(exe, params, q0, nsteps, env):
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "data"))
        para = os.path.join(td, "para")
        with open(para, "w") as fh:
            fh.write(params.to_namelist())
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        np.ascontiguousarray(q0).tofile(fin)
        r = subprocess.run([exe, para, "run", fin, fout, str(nsteps)], env=dict(os.environ, **env), cwd=td,
                           capture_output=True, text=True, timeout=300)
        assert "ERROR" not in r.stdout, r.stdout[-2000:]
        steps = [ln.split()[2:] for ln in r.stdout.splitlines() if ln.startswith("STEP")]
        return np.fromfile(fout).reshape(q0.shape), steps


@pytest.fixture(scope="module", autouse=True)
def need_harnesses():
    if not (O.have_ref() and os.path.exists(HIP_EXE)):
        pytest.skip("the reference harnesses are not built (make -C oracle ref ref-hip)")


@pytest.mark.parametrize("case", ["hydro", "conduction", "conduction_only"])
def test_cpu_backend_is_the_reference(case):
    """conduction_only: the reference's conduction test with hydro = .false.
    (thermalstep, not hydrostep, then counts the steps for EVORA_MAXSTEPS)."""
    if case == "hydro":
        p = ics.blast_params(16, solver=2, cfl=0.4, tend=1e30)
        q0 = ics.blast(p)
    elif case == "conduction":
        p = ics.conduction_params(16, boundtype=3, tend=1e30)
        q0 = ics.conduction_blob(p)
    else:
        p = ics.thermaltest_params(tend=1e30)
        q0 = ics.thermaltest(p)
    env = {"EVORA_BACKEND": "cpu", "EVORA_MAXSTEPS": "4"}
    qd, sd = _run(HIP_EXE, p, q0, 1000, env)
    qr, sr = _run(O.REF_PATH, p, q0, 4, {})
    assert len(sd) == 5  # 4 steps and the dt = 0 step that ends the capped run
    assert sd[:4] == sr  # same t and dt, printed to 16 digits
    assert np.array_equal(qd, qr)
