"""One GPU's share of a grid decomposed into z-slabs (evh_slab,
evh_slab_faces).

The reference splits a grid into z-slabs, one per rank (parallel.f90:53-61),
gives each slab the two planes beyond it from its neighbours every step
(communicate, parallel.f90:317-381; bound keeps its z rules on the grid's
outer faces, boundary.f90:90-110) and takes the minimum of the ranks' dt
(reducestep, parallel.f90:155-165).  Here the slabs are contexts on one GPU
and the exchange copies each slab's boundary planes into its neighbours'
halo buffers.  A cell's update depends on its two neighbours in each
direction and on dt only, so the slabs together must reproduce the
whole-grid run bit for bit, in every fused form, for every boundary type
and uneven slab thickness.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from evora_amd import hydro, ics
from evora_amd.params import zslab_ranges

# device buffers for the faces test; imported before the library first touches the GPU
try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

pytestmark = pytest.mark.gpu

FORMS = [(hydro.FP64, hydro.MODE_FUSED), (hydro.FP64, hydro.MODE_FUSED_EXACT), (hydro.FP32_MIXED, hydro.MODE_FUSED)]


def run_slabs(d, q0, nranks, nsteps, precision, mode, tend=1e30):
    """nsteps of main.f90:109-175 on nranks z-slabs, each a context on this GPU:
    with conduction on, thermalstep joins the dt reduction and T of the planes next
    to each slab is exchanged before thermalconduction."""
    ranges = zslab_ranges(d.nz, nranks)
    periodic = d.boundtype == 2 and nranks > 1
    hs = []
    try:
        for a, b in ranges:
            h = hydro.Hydro(dataclasses.replace(d, nz=b - a), precision=precision, mode=mode)
            hs.append(h)
            h.upload_q(q0[a:b])
        halos, thalos = [], []
        for r, h in enumerate(hs):
            lo = r > 0 or periodic
            hi = r < nranks - 1 or periodic
            halos.append(h.slab(lo, hi))
            if d.conduction:
                thalos.append(h.slab_temp())
        t, go, n = 0.0, True, 0
        while go and n < nsteps:
            # communicate: the lowest planes to the neighbour below, the highest to the one above,
            # once every slab's last sweep (which reads its halos) has finished
            for h in hs:
                h.synchronize()
            for r, h in enumerate(hs):
                below, above = (r - 1) % nranks, (r + 1) % nranks
                h.slab_faces(halos[below][1] if halos[r][0] else None, halos[above][0] if halos[r][1] else None)
            dt = 1e99
            for h in hs:
                h.primitive()
                h.bound()
                dt = h.hydrostep(dt)
                if d.conduction:
                    dt = h.thermalstep(dt)
            if tend - t <= dt:  # updatetime, timestep.f90:163-169
                dt = tend - t
                go = False
            t += dt
            for h in hs:
                h.compevo(dt)
                h.sweep(dt)
            if d.conduction:  # main.f90:170-175
                for h in hs:
                    h.synchronize()
                for r, h in enumerate(hs):
                    below, above = (r - 1) % nranks, (r + 1) % nranks
                    h.slab_temp_faces(thalos[below][1] if thalos[r][0] else None,
                                      thalos[above][0] if thalos[r][1] else None)
                for h in hs:
                    h.thermalconduction(dt)
            n += 1
        return np.concatenate([h.download_q() for h in hs]), t, n
    finally:
        for h in hs:
            h.close()


def run_whole(d, q0, nsteps, precision, mode, tend=1e30):
    with hydro.Hydro(d, precision=precision, mode=mode) as h:
        h.upload_q(q0)
        t, _, n = h.run(nsteps, tend=tend)
        return h.download_q(), t, n


@pytest.mark.parametrize("precision,mode", FORMS)
@pytest.mark.parametrize("bt", [1, 2, 3])
def test_slabs_equal_the_whole_grid(bt, precision, mode):
    """Three slabs of 14, 13 and 13 planes of a 40-plane blast, 6 steps."""
    p = ics.blast_params(24, nz=40, boundtype=bt, solver=2, limiter=1, cfl=0.4, tend=1e30)
    d = p.derive()
    q0 = ics.blast(p)
    qs, ts, ns = run_slabs(d, q0, 3, 6, precision, mode)
    qw, tw, nw = run_whole(d, q0, 6, precision, mode)
    assert ns == nw == 6 and ts == tw
    assert np.array_equal(qs, qw)


@pytest.mark.parametrize("nranks", [2, 5])
@pytest.mark.parametrize("solver,limiter", [(1, 2), (2, 3)])
def test_slabs_equal_the_whole_grid_other_schemes(nranks, solver, limiter):
    """HLL and MC, HLLC and van Leer; slabs down to the two planes the exchange sends; a smooth periodic
    state with dual energy (the primitive rewrites q of the neighbours' planes it reads as it does its own)."""
    p = ics.blast_params(20, nz=12, boundtype=2, solver=solver, limiter=limiter, cfl=0.4, tend=1e30)
    p = dataclasses.replace(p, dualenergy=True, de_thresh=0.1)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=11)
    qs, ts, _ = run_slabs(d, q0, nranks, 5, hydro.FP64, hydro.MODE_FUSED)
    qw, tw, _ = run_whole(d, q0, 5, hydro.FP64, hydro.MODE_FUSED)
    assert ts == tw
    assert np.array_equal(qs, qw)


@pytest.mark.parametrize("precision,mode", FORMS)
@pytest.mark.parametrize("bt", [1, 2])
def test_slabs_with_conduction_equal_the_whole_grid(bt, precision, mode):
    """The hot blob with thermal conduction on, 3 slabs of 7, 6 and 6 planes, 4 steps: the temperature of the
    neighbours' planes crosses before each conduction pass, and thermalstep's limit joins the dt minimum."""
    p = ics.conduction_params(16, nz=19, boundtype=bt)
    d = p.derive()
    q0 = ics.conduction_blob(p)
    qs, ts, ns = run_slabs(d, q0, 3, 4, precision, mode)
    qw, tw, nw = run_whole(d, q0, 4, precision, mode)
    assert ns == nw == 4 and ts == tw
    assert np.array_equal(qs, qw)


def test_slabs_fp32_mixed_dual_energy():
    """The FP32-mixed form with dual energy (its FP64 entropy branch on the neighbours' planes too), 4 slabs,
    periodic: the whole-grid run bit for bit."""
    p = ics.blast_params(20, nz=16, boundtype=2, solver=2, limiter=1, cfl=0.4, tend=1e30)
    p = dataclasses.replace(p, dualenergy=True, de_thresh=0.1)
    d = p.derive()
    q0 = ics.blast(p)
    qs, ts, _ = run_slabs(d, q0, 4, 5, hydro.FP32_MIXED, hydro.MODE_FUSED)
    qw, tw, _ = run_whole(d, q0, 5, hydro.FP32_MIXED, hydro.MODE_FUSED)
    assert ts == tw
    assert np.array_equal(qs, qw)


def test_slabs_match_the_oracle():
    """Two slabs against the oracle's whole-grid run: the exact form bit for bit in q1-q5 and t."""
    p = ics.blast_params(16, nz=24, boundtype=1, solver=2, cfl=0.4, tend=1e30)
    d = p.derive()
    q0 = ics.blast(p)
    qo = q0.copy()
    to, _, _ = O.Oracle(d).run(qo, 4)
    qs, ts, ns = run_slabs(d, q0, 2, 4, hydro.FP64, hydro.MODE_FUSED_EXACT)
    assert ns == 4 and ts == to
    assert np.array_equal(qs[..., :5], qo[..., :5])


@pytest.mark.skipif(torch is None, reason="device buffers come from torch")
def test_slab_faces_are_the_boundary_planes():
    p = ics.blast_params(12, nz=9, tend=1e30)
    d = p.derive()
    q0 = ics.random_smooth(p, seed=2)
    with hydro.Hydro(d, precision=hydro.FP32_MIXED) as h:
        h.upload_q(q0)
        lo = torch.full((6, 2, d.ny, d.nx), -1.0, dtype=torch.float32, device="cuda")
        hi = torch.full_like(lo, -1.0)
        h.slab_faces(lo.data_ptr(), hi.data_ptr())
        q32 = q0.astype(np.float32)
        assert np.array_equal(lo.cpu().numpy(), np.moveaxis(q32[:2], -1, 0))
        assert np.array_equal(hi.cpu().numpy(), np.moveaxis(q32[-2:], -1, 0))


def test_slab_refusals():
    p = ics.blast_params(12, nz=8, tend=1e30)
    d = p.derive()
    with hydro.Hydro(d) as h:
        h.slab(True, False)
        with pytest.raises(hydro.EvoraError) as ei:
            h.step(0.0, 1.0)
        assert ei.value.code == 10
        with pytest.raises(hydro.EvoraError) as ei:
            h.set_mode(hydro.MODE_STAGED)
        assert ei.value.code == 8
    with hydro.Hydro(d, mode=hydro.MODE_STAGED) as h:
        with pytest.raises(hydro.EvoraError) as ei:
            h.slab(True, True)
        assert ei.value.code == 8
    pg = ics.jeans_params(12, nz=8)
    with hydro.Hydro(pg.derive()) as h:
        with pytest.raises(hydro.EvoraError) as ei:
            h.slab(True, True)
        assert ei.value.code == 8
    with hydro.Hydro(d) as h:
        with pytest.raises(hydro.EvoraError) as ei:
            h.slab_temp()
        assert ei.value.code == 10
