#!/usr/bin/env python3
"""Benchmark of the hydro step on one MI355X.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--n 512]

Workload: the 3-D point blast on a 512^3 grid (outflow boundaries,
gamma 5/3, rho = 1 with 1% seeded noise, p = 1e-5, E = 1 in the cells with
r < 3.5 cells), HLLC, minmod.  The configured CFL of 0.6 makes the
reference scheme itself stop on a negative pressure at the blast centre
(step 29 at 128^3, bit-identical in the CPU oracle and on the GPU), so the
timed run uses CFL 0.4, which completes; the work per cell-update does not
depend on the CFL number.  A "step" is one pass of the hydro
step over the grid (primitive, bound, hydrostep, compevo, sweep), so a step
is n^3 cell-updates.  The input is resident in HBM before the timed region.

Printed: one JSON line with the headline number (FP64 fused path) and,
alongside, the roofline of the dominant kernel (HIP-event timing of each
sweep launch on the library's stream), the FP64 fused step in its exact form
(the reference's own arithmetic, bit-identical results), the FP32-mixed and
batched forms, the step with thermal conduction on (whole steps against the
hydro step alone, and the conduction kernel, on a 512^3 hot blob), the step
with self-gravity on (against the hydro step alone, with the potential and
the force kernel timed, on a 512^3 Jeans-unstable box), and the reference's
CPU hot path timed on the host's cores on a bounded sample.
The single-pass 2-D form is timed on the 8192^2 shear layer of
para/kevinhelmholtz (forms.two_d), and one GPU's share of the blast split
into 8 z-slabs, a 512 x 512 x 64 slab with its ghost planes exchanged
through halo buffers every step, on its own (forms.zslab); small grids (the
reference's 128^2 Kelvin-Helmholtz grid, 64^3 and 128^3 blasts) give their
wall time per step (forms.small_grids).  Unless --no-accuracy is given, each FP32-mixed configuration (the blast with
HLLC and with HLL, Kelvin-Helmholtz 256^3 at CFL 0.5 for 50 steps, the
cosmological-like field 256^3 for 100 steps, the Jeans box with gravity and
the hot blob with conduction at 256^3 for 50 steps, and the batched
ensemble) is also advanced in FP64 for the same number of steps, and the
relative L1 errors of rho, p and |v| and the mass and energy drift of both
forms are reported under forms.*.accuracy_vs_fp64 (outside every timed
region).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"fp64": 96.0, "fp32_mixed": 48.0}  # read + write 6 conserved values
COND_BYTES_PER_CELL = {"fp64": 64.0, "fp32_mixed": 32.0}  # read 6 conserved values, write 2
# gravintegration: read rho, momentum, E (5 values) and phi (FP64), write momentum and E (4 values)
GRAV_BYTES_PER_CELL = {"fp64": 80.0, "fp32_mixed": 44.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def primitive_np(q, g1, de_thresh1=None):
    """rho, |v|, p of conserved states q[..., 6] (comp.f90:45-85: p = g1 (E -
    kin), or p = S rho^g1 where the dual-energy switch takes the entropy)."""
    rho = q[..., 0]
    m2 = q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2
    kin = 0.5 * m2 / rho
    p = g1 * (q[..., 4] - kin)
    if de_thresh1 is not None:
        ent = kin / q[..., 4] > de_thresh1
        p = np.where(ent, q[..., 5] * rho ** g1, p)
    return rho, np.sqrt(m2) / rho, p


class L1Error:
    """Relative L1 error sum|a - b| / sum|b| of rho, p and |v|, accumulated
    over z-slabs so that 512^3 states need no full-size temporaries."""

    def __init__(self, d):
        self.g1 = d.g1
        self.de = d.de_thresh1 if d.dualenergy else None
        self.num = np.zeros(3)
        self.den = np.zeros(3)

    def add(self, qa, qb, slab=32):
        for z in range(0, qa.shape[0], slab):
            a = primitive_np(qa[z:z + slab], self.g1, self.de)
            b = primitive_np(qb[z:z + slab], self.g1, self.de)
            for i in range(3):
                self.num[i] += np.abs(a[i] - b[i]).sum()
                self.den[i] += np.abs(b[i]).sum()

    def result(self):
        r = self.num / np.where(self.den > 0, self.den, 1.0)
        return dict(l1_rho=float(r[0]), l1_v=float(r[1]), l1_p=float(r[2]))


def drift(s_end, s_0):
    """Relative change of the mass and energy sums (snapshot.f90:169-172)."""
    return dict(mass=float((s_end[0] - s_0[0]) / s_0[0]), energy=float((s_end[1] - s_0[1]) / s_0[1]))


def run_gpu(d, q0, steps, warmup, precision_name="fp64", tend=1e30, keep_q=False, exact=False, time_xfer=False):
    """Advance q0 warmup + steps steps with the fused path (exact: the FP64
    fused step with the reference's own arithmetic); the `steps` steps are
    timed with the state resident in HBM."""
    from evora_amd import hydro

    prec = hydro.FP64 if precision_name == "fp64" else hydro.FP32_MIXED
    h = hydro.Hydro(d, precision=prec, mode=hydro.MODE_FUSED_EXACT if exact else hydro.MODE_FUSED)
    try:
        log(f"{precision_name}: uploading {q0.nbytes / 1e9:.2f} GB; device bytes {h.device_bytes() / 1e9:.2f} GB")
        t_up = time.perf_counter()
        h.upload_q(q0)
        t_up = time.perf_counter() - t_up
        sums0 = h.sums()
        t = 0.0
        for _ in range(warmup):
            t, dt, go = h.step(t, tend)
        h.synchronize()
        h.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            t, dt, go = h.step(t, tend)
        h.synchronize()
        wall = time.perf_counter() - t0
        kms, launches = h.kernel_stats()
        cms, claunches = h.conduction_stats()
        gps, gpn = h.stage_stats(hydro.STAGE_GRAVPOTENTIAL)
        gis, gin = h.stage_stats(hydro.STAGE_GRAVINTEGRATION)
        sums = h.sums()
        t_dn = time.perf_counter()
        q = h.download_q() if keep_q or time_xfer else None
        t_dn = time.perf_counter() - t_dn
    finally:
        h.close()
    res = dict(
        value=d.ncells * steps / wall,
        ms_per_step=1e3 * wall / steps,
        kernel_ms=kms / max(launches, 1),
        launches=launches,
        cond_ms=cms / max(claunches, 1),
        gravpot_ms=gps / max(gpn, 1),
        gravint_ms=gis / max(gin, 1),
        t_end=t,
        sums0=sums0,
        sums=sums,
        q=q,
    )
    if time_xfer:  # whole-state host transfers (reference layout, pageable host array)
        res["upload_GBs"] = q0.nbytes / t_up / 1e9
        res["download_GBs"] = q0.nbytes / t_dn / 1e9
    log(f"{precision_name}: {res['value'] / 1e9:.2f} G cell-updates/s, {res['ms_per_step']:.3f} ms/step, "
        f"kernel {res['kernel_ms']:.3f} ms")
    return res


def run_conduction(n, steps, warmup, fp64_only=False):
    """The step with thermal conduction on (hydro step, thermalstep and the
    conduction pass of main.f90:121,170-175) on an n^3 hot blob, FP64 fused
    (and FP32-mixed), against the hydro step alone on the same state.  The
    conduction kernel is timed per launch with events on the library's
    stream; its algorithmic bytes are the compulsory ones of the pass (read
    the 6 conserved values, write q(5) and q(6))."""
    from evora_amd import ics

    p = ics.conduction_params(n, boundtype=2)
    d = p.derive()
    dh = ics.conduction_params(n, boundtype=2, conduction=False).derive()
    log(f"building {n}^3 conduction initial condition")
    q0 = ics.conduction_blob(p)
    out = {"workload": f"hot blob {n}^3 (T 1000, peak 3000), periodic, HLLC, minmod, CFL 0.4, simplekappa"}
    for prec in ("fp64",) if fp64_only else ("fp64", "fp32_mixed"):
        r = run_gpu(d, q0, steps, warmup, prec)
        rh = run_gpu(dh, q0, steps, warmup, prec)
        ms = r["cond_ms"]
        nbytes = COND_BYTES_PER_CELL[prec]
        ach = nbytes * d.ncells / (ms * 1e-3) / 1e9
        out[prec] = dict(value=r["value"], ms_per_step=r["ms_per_step"], hydro_only_ms_per_step=rh["ms_per_step"],
                         sweep_kernel_ms=r["kernel_ms"], conduction_kernel_ms=ms, conduction_bytes_per_cell=nbytes,
                         conduction_GBs=ach, conduction_hbm_frac=ach / HBM_PEAK_GBS)
        log(f"conduction {prec}: kernel {ms:.3f} ms ({ach:.0f} GB/s), step {r['ms_per_step']:.3f} ms "
            f"vs hydro alone {rh['ms_per_step']:.3f} ms")
    return out


def run_gravity(n, steps, warmup, fp64_only=False):
    """The step with self-gravity on (hydro step, then gravpotential and
    gravintegration, main.f90:157-164) on an n^3 Jeans-unstable periodic box,
    FP64 fused (and FP32-mixed), against the hydro step alone on the same
    state.  The potential (mean density, source, forward FFT, Green's
    function, inverse FFT; hipFFT in FP64 for both forms) and the force
    kernel are timed per launch with events on the library's stream; the
    force kernel's algorithmic bytes are its compulsory ones (read rho,
    momentum, E and phi, write momentum and E)."""
    from evora_amd import ics

    p = ics.jeans_params(n)
    d = p.derive()
    dh = ics.jeans_params(n, gravity=False).derive()
    log(f"building {n}^3 Jeans initial condition")
    q0 = ics.jeans(p)
    out = {"workload": f"Jeans-unstable periodic box {n}^3 (rho 1e6 with a 10% box mode, Jeans length 0.5), "
                       f"HLLC, minmod, CFL 0.4, 4 pi G = 8.38553911e-7"}
    for prec in ("fp64",) if fp64_only else ("fp64", "fp32_mixed"):
        r = run_gpu(d, q0, steps, warmup, prec)
        rh = run_gpu(dh, q0, steps, warmup, prec)
        nbytes = GRAV_BYTES_PER_CELL[prec]
        ach = nbytes * d.ncells / (r["gravint_ms"] * 1e-3) / 1e9
        out[prec] = dict(value=r["value"], ms_per_step=r["ms_per_step"], hydro_only_ms_per_step=rh["ms_per_step"],
                         sweep_kernel_ms=r["kernel_ms"], gravpotential_ms=r["gravpot_ms"],
                         gravintegration_kernel_ms=r["gravint_ms"], gravintegration_bytes_per_cell=nbytes,
                         gravintegration_GBs=ach, gravintegration_hbm_frac=ach / HBM_PEAK_GBS)
        log(f"gravity {prec}: potential {r['gravpot_ms']:.3f} ms, force {r['gravint_ms']:.3f} ms ({ach:.0f} GB/s), "
            f"step {r['ms_per_step']:.3f} ms vs hydro alone {rh['ms_per_step']:.3f} ms")
    return out


def run_zslab(n, nranks, steps, warmup, whole_kernel_ms):
    """One GPU's share of an n^3 grid split into nranks z-slabs
    (parallel.f90:53-61): an n x n x n/nranks slab whose two ghost planes on
    each side come from halo buffers (evh_slab) filled every step by the
    exchange communicate makes (parallel.f90:317-381), here the slab's own
    boundary planes copied into its halos on the device (a z-periodic slab)
    in place of the transfer from the neighbours.  FP64 fused, HLLC, the
    blast slab through the blast centre.  For each z chunk the sweep kernel
    is timed per launch; the step (exchange copy, hydrostep, sweep) by wall
    clock.  Its efficiency is the slab kernel's time per cell against the
    whole grid's."""
    from evora_amd import hydro, ics
    from evora_amd.params import zslab_ranges

    a, b = zslab_ranges(n, nranks)[nranks // 2]
    nzs = b - a
    p = ics.blast_params(n, nz=nzs, zmin=0.0, zmax=nzs / n, solver=2, cfl=0.4)
    d = p.derive()
    q0 = ics.blast(p)
    cells = d.ncells
    out = {"workload": f"z-slab {n}x{n}x{nzs} (rank {nranks // 2} of {nranks} of the {n}^3 blast, planes {a}-{b - 1}), "
                       "both ghost-plane pairs from halo buffers, HLLC, minmod, CFL 0.4, FP64 fused",
           "cells": cells, "by_zchunk": {}}
    best = None
    for zc in (64, 32, 16):
        with hydro.Hydro(d) as h:
            h.set_tuning(12, True, zc)
            h.upload_q(q0)
            lo, hi = h.slab(True, True)

            def step(t):
                h.synchronize()
                h.slab_faces(hi, lo)  # the planes the neighbours would send, into this slab's halos
                h.primitive()
                dt = h.hydrostep(1e99)
                h.compevo(dt)
                h.sweep(dt)
                return t + dt

            t = 0.0
            for _ in range(warmup):
                t = step(t)
            h.synchronize()
            h.set_timing(True)
            t0 = time.perf_counter()
            for _ in range(steps):
                t = step(t)
            h.synchronize()
            wall = time.perf_counter() - t0
            kms, nl = h.kernel_stats()
            t1 = time.perf_counter()
            for _ in range(20):
                h.slab_faces(hi, lo)
            faces_ms = 1e3 * (time.perf_counter() - t1) / 20
        kernel_ms = kms / max(nl, 1)
        r = dict(ms_per_step=1e3 * wall / steps, value=cells * steps / wall, kernel_ms=kernel_ms,
                 exchange_copy_ms=faces_ms,
                 kernel_efficiency_vs_whole_grid=(whole_kernel_ms / n ** 3) / (kernel_ms / cells))
        out["by_zchunk"][zc] = r
        log(f"z-slab zchunk {zc}: kernel {kernel_ms:.3f} ms, step {r['ms_per_step']:.3f} ms, "
            f"exchange copy {faces_ms:.3f} ms, per-cell efficiency {r['kernel_efficiency_vs_whole_grid']:.3f}")
        if best is None or kernel_ms < out["by_zchunk"][best]["kernel_ms"]:
            best = zc
    out["best_zchunk"] = best
    out.update({k: out["by_zchunk"][best][k] for k in ("value", "ms_per_step", "kernel_ms",
                                                       "kernel_efficiency_vs_whole_grid")})
    return out


def compare_forms(d, r64, r32):
    """FP32-mixed against FP64 after the same number of steps: L1 errors of
    rho, p, |v|; each form's drift of the mass and energy sums from its own
    start; and the FP32-mixed sums at the end relative to the FP64 ones."""
    e = L1Error(d)
    e.add(r32["q"], r64["q"])
    out = e.result()
    out["drift_fp64"] = drift(r64["sums"], r64["sums0"])
    out["drift_fp32_mixed"] = drift(r32["sums"], r32["sums0"])
    out["sums_vs_fp64"] = drift(r32["sums"], r64["sums"])
    out["fp64_value"] = r64["value"]
    out["fp32_mixed_value"] = r32["value"]
    return out


def run_fp64_rounded_each_step(d, q0, steps):
    """FP64 steps with the state rounded to single precision before every
    step (downloaded, rounded, uploaded): what the FP32 storage of the mixed
    form alone does to a run, its arithmetic aside."""
    from evora_amd import hydro

    with hydro.Hydro(d, precision=hydro.FP64) as h:
        q = q0.astype(np.float32).astype(np.float64)
        t = 0.0
        for _ in range(steps):
            h.upload_q(q)
            t, dt, go = h.step(t, 1e30)
            q = h.download_q().astype(np.float32).astype(np.float64)
    return q


def accuracy_case(name, d, q0, steps, control=False, storage_control=False):
    """One configuration advanced `steps` steps in both forms; returns the
    FP32-mixed errors against FP64.  With `control`, FP64 is also run from
    the initial state rounded to single precision: its L1 distance from the
    FP64 run is what the rounding of the input alone grows to in `steps`
    steps (large for unstable flows such as a shear layer).  With
    `storage_control`, FP64 is run with the state rounded to single precision
    before every step: the error of the FP32 storage alone."""
    from evora_amd.hydro import EvoraError

    try:
        r64 = run_gpu(d, q0, steps, 0, "fp64", keep_q=True)
        r32 = run_gpu(d, q0, steps, 0, "fp32_mixed", keep_q=True)
        out = compare_forms(d, r64, r32)
        del r32
        if control:
            rc = run_gpu(d, q0.astype(np.float32).astype(np.float64), steps, 0, "fp64", keep_q=True)
            e = L1Error(d)
            e.add(rc["q"], r64["q"])
            out["fp64_from_rounded_input"] = e.result()
        if storage_control:
            e = L1Error(d)
            e.add(run_fp64_rounded_each_step(d, q0, steps), r64["q"])
            out["fp64_rounded_each_step"] = e.result()
    except EvoraError as e:  # the scheme stopped, as the reference would
        log(f"accuracy {name}: stopped: {e}")
        return dict(stopped=str(e), steps=steps)
    out["steps"] = steps
    log(f"accuracy {name}: {out}")
    return out


def run_batch(nbox, steps, warmup, seed=7, log_path=None, accuracy=True):
    """The ensemble: nbox independent shock-cloud boxes, edges drawn from
    {64, 96, 128, 160, 192}, Mach in [3, 30], density contrast in [10, 100],
    FP32-mixed, CFL 0.6, per-box dt, one persistent launch per step replayed
    from a hipGraph after the autotuner picked the launch shape.  With
    `accuracy`, the same boxes are then advanced the same number of steps in
    FP64 and compared box by box."""
    from evora_amd import hydro, ics
    from evora_amd.batch import Batch

    spec = ics.ensemble_sizes(nbox, seed=seed)
    params = [ics.shock_cloud_params(n, cfl=0.6) for n, _, _ in spec]
    cells = sum(n ** 3 for n, _, _ in spec)
    log(f"batched: {nbox} boxes, {cells / 1e6:.1f} M cells")
    q0 = [ics.shock_cloud(p, mach=mach, chi=chi) for (n, mach, chi), p in zip(spec, params)]
    acc = None
    with Batch(params, precision=hydro.FP32_MIXED) as b:
        for i, q in enumerate(q0):
            b.upload_q(i, q)
        b.run(warmup, graph=False)
        tune_log = b.autotune(trial_steps=3)
        if log_path:
            with open(log_path, "w") as fh:
                fh.write(tune_log)
        b.run(1, graph=True)  # capture
        t0 = time.perf_counter()
        ms = b.run(steps, graph=True)
        wall = time.perf_counter() - t0
        t, dt, nsteps, go = b.times()
        shape = b.get_tuning()
        if accuracy:
            acc = batch_accuracy(params, q0, b, warmup + 1 + steps)
    res = dict(value=cells * steps / wall, ms_per_step=1e3 * wall / steps, device_ms_per_step=ms / steps,
               boxes=nbox, cells=cells, steps=steps, graph=True,
               tuning=dict(tile_ty=shape[0], wpe_hi=shape[1], zchunk=shape[2], blocks_per_cu=shape[3]),
               hbm_frac=BYTES_PER_CELL["fp32_mixed"] * cells / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
               min_steps_done=int(nsteps.min()))
    if acc is not None:
        res["accuracy_vs_fp64"] = acc
    log(f"batched: {res['value'] / 1e9:.2f} G cell-updates/s, {res['device_ms_per_step']:.3f} ms/step, "
        f"shape {res['tuning']}")
    return res


def batch_accuracy(params, q0, b32, nsteps):
    """The FP32-mixed ensemble b32 after nsteps steps against the same boxes
    advanced nsteps steps in FP64: errors summed over all boxes."""
    from evora_amd import hydro
    from evora_amd.batch import Batch

    ds = [p.derive() for p in params]
    e = L1Error(ds[0])
    with Batch(params, precision=hydro.FP64) as b64:
        for i, q in enumerate(q0):
            b64.upload_q(i, q)
        b64.run(nsteps, graph=False)
        s0 = np.zeros(2)
        s64 = np.zeros(2)
        s32 = np.zeros(2)
        for i, d in enumerate(ds):
            qa, qb = b32.download_q(i), b64.download_q(i)
            e.add(qa, qb)
            vol = d.dx * d.dy * d.dz
            for acc, q in ((s0, q0[i]), (s64, qb), (s32, qa)):
                acc += vol * np.array([q[..., 0].sum(), q[..., 4].sum()])
    out = e.result()
    out["drift_fp64"] = drift(s64, s0)
    out["drift_fp32_mixed"] = drift(s32, s0)
    out["sums_vs_fp64"] = drift(s32, s64)
    out["steps"] = nsteps
    log(f"accuracy batched: {out}")
    return out


def run_2d(steps, warmup, fp64_only, n=8192):
    """The single-pass 2-D form (nz = 1) on the shear layer of
    para/kevinhelmholtz (ic/kevinhelmholtz.f90; gamma 1.4, HLLC, MC limiter,
    periodic, CFL 0.5) at n x n: the fast FP64, exact FP64 and FP32-mixed
    forms, with the kernel's HBM roofline at 96 / 48 B per cell."""
    from evora_amd import ics
    from evora_amd.params import HydroParams

    p = HydroParams(nx=n, ny=n, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.5, atomic=False)
    d = p.derive()
    log(f"building {n}^2 Kelvin-Helmholtz initial condition")
    q0 = ics.kh_reference(p)
    out = {"workload": f"2-D Kelvin-Helmholtz {n}x{n}x1 (para/kevinhelmholtz), HLLC, MC, CFL 0.5, periodic",
           "cells": n * n}
    forms = [("fp64", "fp64", False), ("fp64_exact", "fp64", True)]
    if not fp64_only:
        forms.append(("fp32_mixed", "fp32_mixed", False))
    for name, prec, exact in forms:
        r = run_gpu(d, q0, steps, warmup, prec, exact=exact)
        out[name] = dict(value=r["value"], ms_per_step=r["ms_per_step"], kernel_ms=r["kernel_ms"],
                         hbm_frac=BYTES_PER_CELL[prec] * n * n / (r["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
        log(f"2-D {name}: {r['value'] / 1e9:.2f} G cell-updates/s, kernel {r['kernel_ms']:.3f} ms")
    return out


def run_small(steps=200, warmup=5):
    """Wall time per step of the step loop on small grids, where a step is
    bounded by the march's latency and the host's per-step work rather than by
    throughput: the reference's own 128^2 Kelvin-Helmholtz grid
    (para/kevinhelmholtz) and 64^3 and 128^3 blasts, FP64 fused."""
    from evora_amd import hydro, ics
    from evora_amd.params import HydroParams

    cases = [("kh2d_128", HydroParams(nx=128, ny=128, nz=1, solver=2, limiter=2, boundtype=2, cfl=0.5,
                                      atomic=False)),
             ("blast_64", ics.blast_params(64, solver=2, cfl=0.4)), ("blast_128", ics.blast_params(128, solver=2, cfl=0.4))]
    out = {}
    for name, p in cases:
        d = p.derive()
        q0 = ics.kh_reference(p) if name.startswith("kh") else ics.blast(p)
        with hydro.Hydro(d) as h:
            h.upload_q(q0)
            t = 0.0
            for _ in range(warmup):
                t, dt, go = h.step(t, 1e30)
            h.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                t, dt, go = h.step(t, 1e30)
            h.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / steps
        out[name] = dict(ms_per_step=ms, value=d.ncells / (ms * 1e-3), cells=d.ncells, steps=steps)
        log(f"small grid {name}: {ms:.4f} ms/step")
    return out


def cpu_baseline(n, cores, nsteps=2):
    """The reference's own hot path (oracle/_ref/evref, compiled from the
    reference Fortran at -O3) on `cores` processes, each advancing one z-slab
    of the n^3 blast with its own ghost fill (the reference's z-slab split,
    parallel.f90:53-61, without the halo exchange), `nsteps` steps.  Returns
    cell-updates/s over the slowest process's step-loop time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from evora_amd import ics

    if not O.have_ref():
        return None
    p = ics.blast_params(n)
    q = ics.blast(p)
    d = p.derive()
    bounds = []
    mz = 0
    for i in range(cores):  # parallel.f90:55-60
        lz = mz + 1
        mz = lz + (n - lz) // (cores - i)
        bounds.append((lz, mz))
    procs = []
    with tempfile.TemporaryDirectory() as td:
        for r, (lz, mz) in enumerate(bounds):
            sp = ics.blast_params(n)
            sp.nz = mz - lz + 1
            sp.zmin = d.z0 - 0.5 * d.dz + (lz - 1) * d.dz
            sp.zmax = sp.zmin + sp.nz * d.dz
            para = os.path.join(td, f"para{r}")
            with open(para, "w") as fh:
                fh.write(sp.to_namelist())
            fin = os.path.join(td, f"in{r}.bin")
            np.ascontiguousarray(q[lz - 1:mz]).tofile(fin)
            procs.append((para, fin, os.path.join(td, f"out{r}.bin")))
        del q
        log(f"cpu baseline: {cores} reference processes x {nsteps} steps on {n}^3 z-slabs")
        running = [subprocess.Popen([O.REF_PATH, pa, "run", fi, fo, str(nsteps)], stdout=subprocess.PIPE,
                                    stderr=subprocess.STDOUT, text=True, cwd=td) for pa, fi, fo in procs]
        secs = []
        for pr in running:
            out, _ = pr.communicate(timeout=900)
            s = [float(x.split()[1]) for x in out.splitlines() if x.startswith("LOOP_SECONDS")]
            steps_done = sum(1 for x in out.splitlines() if x.startswith("STEP"))
            if not s or steps_done != nsteps:
                log("cpu baseline process failed:\n" + out[-2000:])
                return None
            secs.append(s[0])
    slowest = max(secs)
    try:
        import psutil
        physical = psutil.cpu_count(logical=False)
    except Exception:  # noqa: BLE001 -- the count is informative only
        physical = None
    usable = len(os.sched_getaffinity(0))
    return dict(value=n ** 3 * nsteps / slowest, unit="cell-updates/s", cores=cores, kind="reference",
                sample=f"{nsteps} steps of the {n}^3 HLLC blast, one reference process per z-slab "
                       f"({cores} slabs, own ghost fill, no halo exchange), slowest slab {slowest:.2f} s",
                host_logical_cpus=os.cpu_count(), host_physical_cores=physical, cpus_usable=usable,
                compiler="AMD flang -O3 (gfortran is not in the image)",
                note=(f"{cores} processes on the {usable} CPUs the process may use (the host-CPU share of one GPU)"
                      if cores >= usable else f"{cores} processes; {usable - cores} of the {usable} usable CPUs "
                      "were left idle"))


def copy_probe():
    """A plain device copy of known size (2 GiB read + 2 GiB written, three
    times): the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters of the rocprofv3 traffic runs."""
    import torch

    n = 2 ** 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {a.numel() * 8} bytes three times")


def stream_probe():
    """The HBM rate plain streaming kernels reach on this part: a device copy
    and a triad c = a + 1.5 b over 2 GiB FP64 arrays, timed with HIP events
    (mean of 10 after a warm-up), as the practical ceiling beside the
    8 TB/s specification."""
    import torch

    n = 2 ** 28
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    out = {}
    for name, fn, nbytes in (("copy", lambda: b.copy_(a), 2 * 8 * n),
                             ("triad", lambda: torch.add(a, b, alpha=1.5, out=c), 3 * 8 * n)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        out[name] = dict(ms=ms, GBs=nbytes / (ms * 1e-3) / 1e9)
        log(f"{name}: {ms:.3f} ms, {out[name]['GBs']:.0f} GB/s ({out[name]['GBs'] / HBM_PEAK_GBS:.2f} of the peak)")
    print(json.dumps({"stream_probe": out, "bytes_per_array": 8 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--cpu-cores", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-mixed against FP64 error and drift runs")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2-D form")
    ap.add_argument("--no-conduction", action="store_true", help="skip the thermal-conduction form")
    ap.add_argument("--conduction-only", action="store_true",
                    help="run only the thermal-conduction form and print its JSON (for profiling)")
    ap.add_argument("--no-gravity", action="store_true", help="skip the self-gravity form")
    ap.add_argument("--no-zslab", action="store_true", help="skip one GPU's z-slab of an 8-way split")
    ap.add_argument("--zslab-only", action="store_true", help="run only the z-slab form")
    ap.add_argument("--two-d-only", action="store_true", help="run only the 2-D form")
    ap.add_argument("--no-small", action="store_true", help="skip the small-grid step times")
    ap.add_argument("--gravity-only", action="store_true",
                    help="run only the self-gravity form and print its JSON (for profiling)")
    ap.add_argument("--copy-probe", action="store_true", help="run only the counter-calibration copy")
    ap.add_argument("--stream-probe", action="store_true", help="time a plain copy and a triad (HBM ceiling)")
    ap.add_argument("--batch-boxes", type=int, default=48)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--autotune-log", default=None,
                    help="file to write the batched form's autotuner table to (the chosen shape is in the JSON)")
    args = ap.parse_args()
    if args.stream_probe:
        stream_probe()
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the hydro step runs as one rank on one GPU)", file=sys.stderr)
        sys.exit(2)

    from evora_amd import ics

    if args.conduction_only:
        print(json.dumps(run_conduction(args.n, args.steps, args.warmup, args.fp64_only)), flush=True)
        return
    if args.gravity_only:
        print(json.dumps(run_gravity(args.n, args.steps, args.warmup, args.fp64_only)), flush=True)
        return
    if args.two_d_only:
        print(json.dumps(run_2d(args.steps, args.warmup, args.fp64_only)), flush=True)
        return
    if args.zslab_only:
        pw = ics.blast_params(args.n, solver=2, cfl=args.cfl)
        whole = run_gpu(pw.derive(), ics.blast(pw), args.steps, args.warmup)
        print(json.dumps(run_zslab(args.n, 8, args.steps, args.warmup, whole["kernel_ms"])), flush=True)
        return
    p = ics.blast_params(args.n, solver=2, cfl=args.cfl)
    d = p.derive()
    log(f"building {args.n}^3 blast initial condition")
    q0 = ics.blast(p)
    acc = not args.no_accuracy and not args.fp64_only
    fp64 = run_gpu(d, q0, args.steps, args.warmup, "fp64", keep_q=acc, time_xfer=True)
    achieved = BYTES_PER_CELL["fp64"] * args.n ** 3 / (fp64["kernel_ms"] * 1e-3) / 1e9
    forms = {"fp64": dict(value=fp64["value"], ms_per_step=fp64["ms_per_step"], kernel_ms=fp64["kernel_ms"],
                          hbm_frac=achieved / HBM_PEAK_GBS, upload_GBs=fp64["upload_GBs"],
                          download_GBs=fp64["download_GBs"])}
    ex = run_gpu(d, q0, args.steps, args.warmup, "fp64", exact=True)
    forms["fp64_exact"] = dict(value=ex["value"], ms_per_step=ex["ms_per_step"], kernel_ms=ex["kernel_ms"],
                               hbm_frac=BYTES_PER_CELL["fp64"] * args.n ** 3 / (ex["kernel_ms"] * 1e-3) / 1e9
                               / HBM_PEAK_GBS)
    if not args.fp64_only:
        mx = run_gpu(d, q0, args.steps, args.warmup, "fp32_mixed", keep_q=acc)
        ach = BYTES_PER_CELL["fp32_mixed"] * args.n ** 3 / (mx["kernel_ms"] * 1e-3) / 1e9
        forms["fp32_mixed"] = dict(value=mx["value"], ms_per_step=mx["ms_per_step"], kernel_ms=mx["kernel_ms"],
                                   hbm_frac=ach / HBM_PEAK_GBS, speedup_vs_fp64=mx["value"] / fp64["value"])
        if acc:
            accuracy = {f"blast{args.n}_hllc": compare_forms(d, fp64, mx)}
            accuracy[f"blast{args.n}_hllc"]["steps"] = args.warmup + args.steps
            log(f"accuracy blast hllc: {accuracy[f'blast{args.n}_hllc']}")
        del fp64["q"], mx["q"]
        if acc:
            ph = ics.blast_params(args.n, solver=1, cfl=args.cfl)
            accuracy[f"blast{args.n}_hll"] = accuracy_case("blast hll", ph.derive(), q0, args.warmup + args.steps)
            del q0
            kp = ics.kh_params(256, cfl=0.5)
            accuracy["kh256"] = accuracy_case("kh", kp.derive(), ics.kelvin_helmholtz(kp), 50, control=True)
            cp = ics.cosmo_params(256)
            accuracy["cosmo256"] = accuracy_case("cosmo", cp.derive(), ics.cosmo_field(cp), 100, control=True)
            jp = ics.jeans_params(256)
            accuracy["jeans256_gravity"] = accuracy_case("jeans gravity", jp.derive(), ics.jeans(jp), 50,
                                                         control=True, storage_control=True)
            bp = ics.conduction_params(256, boundtype=2)
            accuracy["blob256_conduction"] = accuracy_case("blob conduction", bp.derive(), ics.conduction_blob(bp),
                                                           50, control=True)
            forms["fp32_mixed"]["accuracy_vs_fp64"] = accuracy
    if not args.fp64_only and not args.no_batch:
        if args.autotune_log:
            os.makedirs(os.path.dirname(os.path.abspath(args.autotune_log)), exist_ok=True)
        forms["batched"] = run_batch(args.batch_boxes, args.batch_steps, 2, log_path=args.autotune_log,
                                     accuracy=acc)
    if not args.no_2d:
        forms["two_d"] = run_2d(args.steps, args.warmup, args.fp64_only)
    if not args.no_conduction:
        forms["conduction"] = run_conduction(args.n, args.steps, args.warmup, args.fp64_only)
    if not args.no_gravity:
        forms["gravity"] = run_gravity(args.n, args.steps, args.warmup, args.fp64_only)
    if not args.no_zslab:
        forms["zslab"] = run_zslab(args.n, 8, args.steps, args.warmup, fp64["kernel_ms"])
    if not args.no_small:
        forms["small_grids"] = run_small()
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_blast{args.n}_fp64.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.n, min(args.cpu_cores, os.cpu_count() or 1), args.cpu_steps)
    line = {
        "metric": "cell-updates/s, 512^3 HLLC, FP32-mixed vs FP64 on 1 GPU; % of HBM roofline",
        "value": fp64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded point-blast initial condition)",
        "config": {"workload": f"3-D point blast {args.n}^3, HLLC, minmod, CFL {args.cfl}, outflow, FP64 fused sweep",
                   "cells": args.n ** 3, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_fused (FP64)", "kernel_ms": fp64["kernel_ms"],
                     "bytes_per_cell_update": BYTES_PER_CELL["fp64"]},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
