"""The allocation-free batched run (evh_batch_run_noalloc) against the batched
form with its own buffers (evh_batch_create + evh_batch_run): on the
caller's device buffers, with every descriptor, queue and per-box status in
the library's fixed scratch array, the same boxes must come out bit for bit
the same, with the same t and step counts -- on two calls in a row, the
second with fewer and smaller boxes, so that the scratch still holds the
first call's contents when the second begins."""
import numpy as np
import pytest

from evora_amd import hydro, ics
from evora_amd.batch import Batch, run_noalloc
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def first_boxes():
    ps = [HydroParams(nx=40, ny=36, nz=30, solver=2, limiter=1, boundtype=3, cfl=0.4, atomic=True),
          HydroParams(nx=33, ny=20, nz=17, solver=2, limiter=1, boundtype=2, cfl=0.3, atomic=True),
          HydroParams(nx=24, ny=24, nz=24, solver=2, limiter=1, boundtype=1, cfl=0.4, atomic=True)]
    return ps, [ics.random_smooth(p, seed=60 + i, amp=0.1) for i, p in enumerate(ps)]


def second_boxes():
    ps = [HydroParams(nx=20, ny=18, nz=16, solver=2, limiter=1, boundtype=1, cfl=0.4, atomic=True),
          HydroParams(nx=17, ny=19, nz=9, solver=2, limiter=1, boundtype=2, cfl=0.35, atomic=True)]
    return ps, [ics.random_smooth(p, seed=70 + i, amp=0.1) for i, p in enumerate(ps)]


def with_buffers(ps, qs, nsteps, precision, tend):
    """Run through the allocation-free entry point on torch-owned buffers."""
    dtype = torch.float64 if precision == hydro.FP64 else torch.float32
    soa = [np.ascontiguousarray(np.moveaxis(q, -1, 0)).ravel() for q in qs]  # (6, nz, ny, nx) per box
    total = sum(a.size for a in soa)
    buf = torch.zeros((2, total), dtype=dtype, device="cuda")
    side = np.array([i % 2 for i in range(len(ps))], dtype=np.int32)  # start boxes in both buffers
    off = 0
    for i, a in enumerate(soa):
        buf[side[i], off:off + a.size] = torch.from_numpy(a).to(device="cuda", dtype=dtype)
        off += a.size
    torch.cuda.synchronize()
    t = np.zeros(len(ps))
    steps = np.zeros(len(ps), dtype=np.int32)
    go = np.ones(len(ps), dtype=np.int32)
    code = run_noalloc(ps, buf[0].data_ptr(), buf[1].data_ptr(), side, t, np.asarray(tend, dtype=np.float64),
                       steps, go, nsteps, precision=precision)
    assert code == 0
    out, off = [], 0
    for i, (p, a) in enumerate(zip(ps, soa)):
        d = p.derive()
        flat = buf[side[i], off:off + a.size].double().cpu().numpy()
        out.append(np.moveaxis(flat.reshape(6, d.nz, d.ny, d.nx), 0, -1))
        off += a.size
    return out, t, steps, go


def with_batch(ps, qs, nsteps, precision, tend):
    with Batch(ps, precision=precision) as b:
        for i, q in enumerate(qs):
            b.upload_q(i, q)
        b.set_time(tend=np.asarray(tend, dtype=np.float64))
        b.run(nsteps, graph=False)
        t, dt, steps, go = b.times()
        return [b.download_q(i) for i in range(len(ps))], t, steps, go


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
def test_noalloc_matches_batched_form_twice(precision):
    for (ps, qs), nsteps, tend in ((first_boxes(), 6, [1e30, 1e30, 0.004]), (second_boxes(), 4, [1e30, 0.003])):
        a, ta, sa, ga = with_buffers(ps, qs, nsteps, precision, tend)
        b, tb, sb, gb = with_batch(ps, qs, nsteps, precision, tend)
        assert list(sa) == list(sb) and list(ga) == list(gb)
        assert np.array_equal(ta, tb)
        for i in range(len(ps)):
            assert np.array_equal(a[i], b[i]), (len(ps), i)


def test_noalloc_rejects_more_boxes_than_the_scratch_holds():
    ps = [ics.blast_params(8)] * 65
    z = np.zeros(65)
    i32 = np.zeros(65, dtype=np.int32)
    with pytest.raises(hydro.EvoraError) as ei:
        run_noalloc(ps, 1, 1, i32.copy(), z.copy(), z.copy(), i32.copy(), i32.copy(), 1)
    assert ei.value.code == 8 and "64" in str(ei.value)
