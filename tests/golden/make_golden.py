"""Write the golden fixtures from the reference's own compiled hot path.

    make -C oracle ref && python tests/golden/make_golden.py [gravity | refcases]

Runs oracle/_ref/evref (the reference Fortran, built in place from the
reference sources) on small seeded states and stores inputs and outputs as
compressed npz files in this folder.  The fixtures are data only: q before,
q after, dt history, and for the one-step cases w, evo and s.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from evora_amd import ics  # noqa: E402
from evora_amd.params import HydroParams  # noqa: E402

# (name, grid, cases) -- one-step stage fixtures over solver x limiter x boundtype
STAGE_GRIDS = {"1d": (24, 1, 1), "2d": (10, 9, 1), "3d": (8, 7, 6)}


def stage_cases():
    for dim, (nx, ny, nz) in STAGE_GRIDS.items():
        for solver in (1, 2):
            for limiter in (1, 2, 3):
                for bt in (1, 2, 3):
                    p = HydroParams(nx=nx, ny=ny, nz=nz, solver=solver, limiter=limiter, boundtype=bt,
                                    cfl=0.4, atomic=True, tend=10.0)
                    seed = 100 * solver + 10 * limiter + bt
                    yield f"stage_{dim}_s{solver}_l{limiter}_b{bt}", p, seed
    # dual energy (comp.f90:60-69): de_thresh 0.98 puts cells in both branches
    for solver in (1, 2):
        for limiter in (1, 2, 3):
            p = HydroParams(nx=8, ny=7, nz=6, solver=solver, limiter=limiter, boundtype=2, cfl=0.4, atomic=True,
                            tend=10.0, dualenergy=True, de_thresh=0.98)
            yield f"stage_3d_s{solver}_l{limiter}_b2_de", p, 500 + 10 * solver + limiter


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/evref is not built: make -C oracle ref")
    index = {}
    # ---- one-step stage fixtures (w, evo, s, q1, dt)
    stage = {}
    for name, p, seed in stage_cases():
        q0 = ics.random_smooth(p, seed=seed)
        w, evo, s, q1, dt = O.ref_stage(p, q0)
        stage[name + "/q0"] = q0
        stage[name + "/w"] = w
        stage[name + "/evo"] = evo
        stage[name + "/s"] = s
        stage[name + "/q1"] = q1
        stage[name + "/dt"] = np.array([dt])
        index[name] = dict(kind="stage", params=p.__dict__, seed=seed)
    np.savez_compressed(os.path.join(HERE, "stage_fixtures.npz"), **stage)

    # ---- multi-step runs
    runs = {}
    sod = ics.sod_params(1000)
    q0 = ics.sod(sod)
    qf, steps, log = O.ref_run(sod, q0, 100000)
    runs["sod1000/q0"] = q0
    runs["sod1000/q"] = qf
    runs["sod1000/t_dt"] = np.array(steps)
    index["sod1000"] = dict(kind="run", params=sod.__dict__, nsteps=len(steps))
    for solver, limiter, bt in ((2, 1, 2), (1, 3, 1), (2, 2, 3)):
        p = HydroParams(nx=16, ny=16, nz=16, solver=solver, limiter=limiter, boundtype=bt, cfl=0.3,
                        atomic=True, tend=10.0)
        q0 = ics.random_smooth(p, seed=7 + bt, amp=0.1)
        qf, steps, log = O.ref_run(p, q0, 10)
        name = f"run16_s{solver}_l{limiter}_b{bt}"
        runs[name + "/q0"] = q0
        runs[name + "/q"] = qf
        runs[name + "/t_dt"] = np.array(steps)
        index[name] = dict(kind="run", params=p.__dict__, nsteps=10)
    # cosmo-like field with the dual-energy switch, 10 steps
    p = ics.cosmo_params(24)
    q0 = ics.cosmo_field(p)
    qf, steps, log = O.ref_run(p, q0, 10)
    assert qf is not None and len(steps) == 10, log[-400:]
    runs["cosmo24_de/q0"] = q0
    runs["cosmo24_de/q"] = qf
    runs["cosmo24_de/t_dt"] = np.array(steps)
    index["cosmo24_de"] = dict(kind="run", params=p.__dict__, nsteps=10)
    # thermal conduction (simplekappa) on a hot blob, 5 steps: thermalstep
    # limits dt, then primitive, bound and thermalconduction follow the sweep
    for solver, bt in ((2, 2), (1, 1), (2, 3)):
        p = ics.conduction_params(16, solver=solver, boundtype=bt)
        q0 = ics.conduction_blob(p)
        qf, steps, log = O.ref_run(p, q0, 5)
        assert qf is not None and len(steps) == 5, log[-400:]
        name = f"cond16_s{solver}_b{bt}"
        runs[name + "/q0"] = q0
        runs[name + "/q"] = qf
        runs[name + "/t_dt"] = np.array(steps)
        index[name] = dict(kind="run", params=p.__dict__, nsteps=5)
    # configured CFL numbers at which the reference scheme itself stops on a
    # negative pressure (3-D unsplit, CFL limit taken per direction)
    aborts = {
        "blast32_hll": (ics.blast_params(32, solver=1), ics.blast, 20),
        "kh24_cfl08": (ics.kh_params(24), ics.kelvin_helmholtz, 40),
        "cosmo16_nodual": (ics.cosmo_params(16, dualenergy=False, de_thresh=0.0), ics.cosmo_field, 20),
    }
    for name, (p, gen, nmax) in aborts.items():
        q0 = gen(p)
        qf, steps, log = O.ref_run(p, q0, nmax)
        err = [ln.strip() for ln in log.splitlines() if "ERROR" in ln][:1]
        assert err, f"{name}: the reference did not stop"
        runs[name + "/q0"] = q0
        runs[name + "/t_dt"] = np.array(steps)
        index[name] = dict(kind="abort", params=p.__dict__, steps_done=len(steps), error=err)
    np.savez_compressed(os.path.join(HERE, "run_fixtures.npz"), **runs)
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)
    print(f"wrote {len(index)} fixtures")


def gravity():
    """multigreen and gravintegration (gravitation.f90:55-89,172-254) of the
    reference on seeded fields, for grids with even, odd and power-of-two
    sizes (grid spacings differ per axis), and 2-D (nz = 1) and 1-D
    (ny = nz = 1) grids.  The reference's full potential needs FFTW-2, which
    this image lacks; these two routines do not."""
    fx = {}
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    for nx, ny, nz in ((8, 6, 4), (16, 16, 16), (5, 7, 3), (12, 10, 9), (9, 12, 1), (16, 8, 1), (16, 1, 1), (7, 1, 1)):
        p = HydroParams(nx=nx, ny=ny, nz=nz, gravity=True, atomic=True, xmax=2.0, ymax=1.5)
        rng = np.random.default_rng(1000 * nx + 10 * ny + nz)
        a = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
        name = f"green_{nx}x{ny}x{nz}"
        fx[name + "/in"] = a
        fx[name + "/out"] = O.ref_green(p, a)
        index[name] = dict(kind="green", params=p.__dict__)
        q = np.zeros((nz, ny, nx, 6))
        q[..., 0] = 1.0 + 0.3 * rng.random((nz, ny, nx))
        q[..., 1:4] = rng.standard_normal((nz, ny, nx, 3))
        q[..., 4] = 5.0 + rng.random((nz, ny, nx))
        q[..., 5] = 1.0
        phi = rng.standard_normal((nz, ny, nx))
        dt = 0.0123
        name = f"gravint_{nx}x{ny}x{nz}"
        fx[name + "/q0"] = q
        fx[name + "/phi"] = phi
        fx[name + "/q"] = O.ref_gravint(p, q, phi, dt)
        index[name] = dict(kind="gravint", params=p.__dict__, dt=dt)
    np.savez_compressed(os.path.join(HERE, "grav_fixtures.npz"), **fx)
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)
    print(f"wrote {len(fx)} gravity arrays")


def refcase_list():
    """The reference's own test cases and parameter files at gamma = 1.4
    (g = 1.4, atomic = .false., the default of parameter.f90:46), the thermal
    conduction test, and dual energy with gamma = 1.4."""
    return {
        # thesis Sod (bin/test:106-119): 1024 cells, boundtype 3, defaults otherwise
        "sod1024_g14_b3": (ics.sod_params(1024, boundtype=3, cfl=0.5, atomic=False),
                           lambda p: ics.sod(p, ur=-0.0), 100000),
        # blast-wave test (bin/test:176-196, ic/blastwave.f90) to its tend
        "blast32_g14": (ics.blast_params(32, solver=2, cfl=0.5, atomic=False, tend=0.0508),
                        ics.blast_reference, 100000),
        # para/kevinhelmholtz (128 x 128 x 1, MC limiter, periodic), 50 steps
        "kh2d128_g14": (HydroParams(nx=128, ny=128, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.5,
                                    atomic=False, tend=12.0), ics.kh_reference, 50),
        # dual energy with gamma = 1.4: the power of g1 reaches p
        "cosmo24_g14_de": (ics.cosmo_params(24, atomic=False), ics.cosmo_field, 10),
        "blast24_g14_de_hll": (ics.blast_params(24, solver=1, cfl=0.4, atomic=False, dualenergy=True,
                                                de_thresh=0.5, limiter=3), ics.blast, 10),
        # thermal conduction test (bin/test:162-173, ic/thermaltest.f90), hydro off, to tend
        "thermal100": (ics.thermaltest_params(), ics.thermaltest, 100000),
        # conduction on 2-D and 1-D grids with hydro on
        "cond2d_g14": (ics.conduction_params(24, nz=1, atomic=False), ics.conduction_blob, 5),
    }


def refcases():
    """Fixtures of refcase_list() (refcase_fixtures.npz), and the point blast
    at its configured CFL 0.6 with HLLC on the seeded 128^3 state: the
    reference completes 28 steps and stops with a negative pressure in the
    29th.  Only its (t, dt) pairs, the error line and a checksum of the
    initial state are stored (the state is regenerated from its seed)."""
    import hashlib
    with open(os.path.join(HERE, "index.json")) as fh:
        index = json.load(fh)
    fx = {}
    for name, (p, gen, nmax) in refcase_list().items():
        q0 = gen(p)
        qf, steps, log = O.ref_run(p, q0, nmax)
        assert qf is not None and "ERROR" not in log, log[-400:]
        fx[name + "/q0"] = q0
        fx[name + "/q"] = qf
        fx[name + "/t_dt"] = np.array(steps)
        index[name] = dict(kind="refcase", params=p.__dict__, nsteps=len(steps))
    p = ics.blast_params(128, solver=2, cfl=0.6)
    q0 = ics.blast(p)
    qf, steps, log = O.ref_run(p, q0, 40)
    err = [ln.strip() for ln in log.splitlines() if "ERROR" in ln][:1]
    assert err, "the reference did not stop"
    fx["blast128_hllc_cfl06/t_dt"] = np.array(steps)
    index["blast128_hllc_cfl06"] = dict(kind="abort_big", params=p.__dict__, steps_done=len(steps), error=err,
                                        q0_sha256=hashlib.sha256(np.ascontiguousarray(q0).tobytes()).hexdigest())
    np.savez_compressed(os.path.join(HERE, "refcase_fixtures.npz"), **fx)
    with open(os.path.join(HERE, "index.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)
    print(f"wrote {len(fx)} arrays")


if __name__ == "__main__":
    if sys.argv[1:] == ["gravity"]:
        gravity()
    elif sys.argv[1:] == ["refcases"]:
        refcases()
    else:
        main()
        gravity()
        refcases()
