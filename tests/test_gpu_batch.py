"""The batched form: every box of a batch must evolve exactly as the same box
run alone through the single-grid fused path (same per-cell arithmetic, same
per-box dt), whether the steps run eagerly or replay from a hipGraph."""
import numpy as np
import pytest

from evora_amd import hydro, ics
from evora_amd.batch import Batch
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu


def boxes():
    ps = [HydroParams(nx=24, ny=24, nz=24, solver=2, limiter=1, boundtype=1, cfl=0.4, atomic=True),
          HydroParams(nx=33, ny=20, nz=17, solver=2, limiter=1, boundtype=2, cfl=0.3, atomic=True),
          HydroParams(nx=40, ny=36, nz=30, solver=2, limiter=1, boundtype=3, cfl=0.4, atomic=True)]
    qs = [ics.random_smooth(p, seed=40 + i, amp=0.1) for i, p in enumerate(ps)]
    return ps, qs


def alone(p, q0, nsteps, precision):
    with hydro.Hydro(p, precision=precision) as h:
        h.upload_q(q0)
        t, dt, n = h.run(nsteps)
        return h.download_q(), t, dt


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_runs(precision, graph):
    ps, qs = boxes()
    with Batch(ps, precision=precision) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        b.run(5, graph=graph)
        t, dt, steps, go = b.times()
        out = [b.download_q(i) for i in range(len(ps))]
    assert list(steps) == [5, 5, 5]
    for i, (p, q0) in enumerate(zip(ps, qs)):
        qa, ta, dta = alone(p, q0, 5, precision)
        assert t[i] == ta and dt[i] == dta, i
        assert np.array_equal(out[i], qa), i


def test_batch_graph_equals_eager_over_many_steps():
    ps, qs = boxes()
    res = []
    for graph in (False, True):
        with Batch(ps, precision=hydro.FP32_MIXED) as b:
            for i, q in enumerate(qs):
                b.upload_q(i, q)
            b.run(7, graph=graph)
            b.run(6, graph=graph)
            res.append([b.download_q(i) for i in range(len(ps))] + [b.times()[0]])
    for a, c in zip(res[0], res[1]):
        assert np.array_equal(a, c)


def alone_until(p, q0, tend, precision=hydro.FP64):
    with hydro.Hydro(p, precision=precision) as h:
        h.upload_q(q0)
        t, dt, n = h.run(10000, tend=tend)
        return h.download_q(), t, n


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
def test_batch_tend_and_error_isolation(precision):
    """Box 1 stops on an error in its first primitive, box 2 at its tend
    (updatetime's clamp) after a few steps, box 0 runs on.  Every further
    step of the batch must leave the stopped boxes' states as they were
    (each box ping-pongs between the buffers on its own), after odd and even
    numbers of extra steps alike."""
    ps, qs = boxes()
    qs[1] = qs[1].copy()
    qs[1][3, 4, 5, 4] = 0.25 * (qs[1][3, 4, 5, 1:4] ** 2).sum() / qs[1][3, 4, 5, 0]  # E below kinetic
    tend2 = 0.004
    q2a, t2a, n2a = alone_until(ps[2], qs[2], tend2, precision)
    assert 1 <= n2a < 4
    q1_in = None
    with Batch(ps, precision=precision) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        q1_in = b.download_q(1)  # box 1 as stored (rounded to FP32 in the mixed form)
        b.set_time(tend=np.array([1e30, 1e30, tend2]))
        total = 0
        for extra in (4, 1, 2, 3):
            with pytest.raises(hydro.EvoraError) as ei:
                b.run(extra, graph=extra % 2 == 1)
            total += extra
            assert ei.value.code == 1 and ei.value.box == 1
            t, dt, steps, go = b.times()
            assert steps[1] == 0 and go[1] == 0
            assert steps[0] == total
            assert t[2] == t2a and steps[2] == n2a and go[2] == 0  # updatetime's clamp ended box 2
            assert np.array_equal(b.download_q(2), q2a), total
            assert np.array_equal(b.download_q(1), q1_in), total
            if total == 4:
                q0 = b.download_q(0)
                qa, ta, _ = alone(ps[0], qs[0], 4, precision)
                assert np.array_equal(q0, qa)


def test_batch_autotune_restores_state():
    ps, qs = boxes()
    with Batch(ps, precision=hydro.FP32_MIXED) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        b.run(1)
        log = b.autotune(trial_steps=1)
        assert "best:" in log and log.count("ms_per_step") >= 48
        b.run(3)
        tuned = [b.download_q(i) for i in range(len(ps))]
    with Batch(ps, precision=hydro.FP32_MIXED) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        b.run(4)
        plain = [b.download_q(i) for i in range(len(ps))]
    for a, c in zip(tuned, plain):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("kw", [dict(gravity=True), dict(conduction=True, simplekappa=True)])
def test_batch_rejects_physics_beyond_the_hydro_step(kw):
    params = [ics.blast_params(16), ics.blast_params(16, **kw)]
    with pytest.raises(hydro.EvoraError) as ei:
        Batch(params)
    assert ei.value.code == 8


def test_batch_boxes_match_the_oracle():
    """FP64 boxes of uneven sizes and boundary types, 6 steps in one batch:
    each box against the CPU oracle run alone, relative L-infinity 1e-12 per
    component, the same t."""
    import oracle as O
    from conftest import rel_linf
    ps, qs = boxes()
    ps.append(HydroParams(nx=50, ny=13, nz=21, solver=2, limiter=1, boundtype=1, cfl=0.45, atomic=True))
    qs.append(ics.random_smooth(ps[-1], seed=49, amp=0.1))
    with Batch(ps) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        b.run(6, graph=True)
        t, dt, steps, go = b.times()
        out = [b.download_q(i) for i in range(len(ps))]
    for i, (p, q0) in enumerate(zip(ps, qs)):
        qo = q0.copy()
        to, dto, n = O.Oracle(p.derive()).run(qo, 6)
        assert n == 6 and steps[i] == 6
        assert t[i] == pytest.approx(to, rel=1e-12)
        for c in range(6):
            assert rel_linf(out[i][..., c], qo[..., c]) <= 1e-12, (i, c)


def test_batch_configured_ensemble():
    """The configured ensemble: 48 shock-cloud boxes with edges drawn from
    {64, 96, 128, 160, 192} (seed 7), Mach 3-30, density contrast 10-100,
    CFL 0.6, 200 steps replayed from the hipGraph, FP32-mixed against the
    same batch in FP64.  Every box takes all 200 steps in both forms, with t
    within 1e-5; rho and p within relative L1 1e-5 of FP64 in every box."""
    spec = ics.ensemble_sizes(48, seed=7)
    params = [ics.shock_cloud_params(n, cfl=0.6) for n, _, _ in spec]
    q0 = [ics.shock_cloud(p, mach=mach, chi=chi) for (n, mach, chi), p in zip(spec, params)]
    res = {}
    for prec in (hydro.FP64, hydro.FP32_MIXED):
        with Batch(params, precision=prec) as b:
            for i, q in enumerate(q0):
                b.upload_q(i, q)
            b.run(200, graph=True)
            t, dt, steps, go = b.times()
            assert (steps == 200).all() and (go == 1).all()
            res[prec] = (t, [b.download_q(i) for i in range(len(q0))])
    t64, q64 = res[hydro.FP64]
    t32, q32 = res[hydro.FP32_MIXED]
    assert np.abs(t32 / t64 - 1).max() <= 1e-5
    for i, p in enumerate(params):
        g1 = p.derive().g1
        for c, (a, r) in enumerate(zip(_prim(q32[i], g1), _prim(q64[i], g1))):
            err = np.abs(a - r).sum() / np.abs(r).sum()
            assert err <= 1e-5, (i, c, err)


def _prim(q, g1):
    rho = q[..., 0]
    return rho, g1 * (q[..., 4] - 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / rho)
