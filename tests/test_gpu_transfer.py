"""Host <-> device transfers of q (reference layout) through the pinned,
double-buffered chunks: round trips over grids of several chunks with a
ragged last one, in both precisions and for batched boxes."""
import numpy as np
import pytest

from evora_amd import hydro
from evora_amd.batch import Batch
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu

CHUNK_CELLS = (64 << 20) // 48  # cells per pinned chunk


def random_q(shape, seed):
    rng = np.random.default_rng(seed)
    q = rng.uniform(0.5, 2.0, shape + (6,))
    q[..., 1:4] -= 1.25
    return q


@pytest.mark.parametrize("dims", [(7, 5, 3), (100, 110, 150), (160, 160, 160)])
def test_roundtrip_fp64_bitwise(dims):
    nx, ny, nz = dims
    p = HydroParams(nx=nx, ny=ny, nz=nz, atomic=True)
    q = random_q((nz, ny, nx), seed=nx)
    with hydro.Hydro(p.derive()) as h:
        h.upload_q(q)
        back = h.download_q()
    assert np.array_equal(back, q)
    if nx * ny * nz > CHUNK_CELLS:
        assert (nx * ny * nz) % CHUNK_CELLS != 0  # a ragged last chunk


def test_roundtrip_fp32_rounds_once():
    p = HydroParams(nx=100, ny=110, nz=150, atomic=True)
    q = random_q((150, 110, 100), seed=3)
    with hydro.Hydro(p.derive(), precision=hydro.FP32_MIXED) as h:
        h.upload_q(q)
        back = h.download_q()
    assert np.array_equal(back, q.astype(np.float32).astype(np.float64))


def test_batch_roundtrip_uneven_boxes():
    dims = [(40, 30, 20), (120, 120, 110), (9, 8, 7)]
    params = [HydroParams(nx=a, ny=b, nz=c, atomic=True) for a, b, c in dims]
    qs = [random_q((c, b, a), seed=i) for i, (a, b, c) in enumerate(dims)]
    with Batch(params) as bt:
        for i, q in enumerate(qs):
            bt.upload_q(i, q)
        for i, q in enumerate(qs):
            assert np.array_equal(bt.download_q(i), q), i
