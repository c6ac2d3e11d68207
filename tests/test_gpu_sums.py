"""The on-device conservation totals (evh_sums, snapshot.f90:169-172) against
an exactly rounded host sum (math.fsum) of the state the device holds.

The device sum is compensated at every addition (per thread, in the block
tree and over the block partials on the host), so it equals the exactly
rounded sum to about one ulp: the bar is 1e-15 relative, for FP64 and for
the FP32-mixed storage, on grids whose cell count is not a multiple of the
kernel's grid stride, with values spread over six decades."""
import math

import numpy as np
import pytest

from evora_amd import hydro
from evora_amd.params import HydroParams

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [hydro.FP64, hydro.FP32_MIXED])
@pytest.mark.parametrize("shape", [(37, 29, 23), (257, 255, 129)])
def test_sums_equal_exactly_rounded_sum(precision, shape):
    nx, ny, nz = shape
    p = HydroParams(nx=nx, ny=ny, nz=nz, atomic=True)
    rng = np.random.default_rng(nx * ny + nz + precision)
    q = np.zeros((nz, ny, nx, 6))
    q[..., 0] = 10.0 ** rng.uniform(-3, 3, (nz, ny, nx))
    q[..., 1:4] = rng.standard_normal((nz, ny, nx, 3)) * q[..., :1]
    q[..., 4] = 0.5 * (q[..., 1:4] ** 2).sum(-1) / q[..., 0] + 10.0 ** rng.uniform(-3, 3, (nz, ny, nx))
    q[..., 5] = 10.0 ** rng.uniform(-3, 3, (nz, ny, nx))
    with hydro.Hydro(p.derive(), precision=precision) as h:
        h.upload_q(q)
        s = h.sums()
        stored = h.download_q()  # the values the device holds (rounded to FP32 in the mixed form)
    for j, c in enumerate((0, 4, 5)):
        want = math.fsum(stored[..., c].ravel())
        assert abs(s[j] - want) <= 1e-15 * abs(want), (c, s[j], want)
