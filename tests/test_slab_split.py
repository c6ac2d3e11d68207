"""The z-slab split of decomposition (parallel.f90:53-61) on the CPU."""
import pytest

from evora_amd.params import zslab_ranges


@pytest.mark.parametrize("nz,nranks,sizes", [(40, 3, [14, 13, 13]), (32, 2, [16, 16]), (7, 3, [3, 2, 2]),
                                             (512, 8, [64] * 8), (5, 5, [1] * 5), (10, 4, [3, 3, 2, 2])])
def test_split_by_hand(nz, nranks, sizes):
    """Sizes worked out by hand from the loop: rank r takes (nz - lz) / (nranks - r)
    planes after lz (integer division), so the remainder goes to the lowest ranks."""
    assert [b - a for a, b in zslab_ranges(nz, nranks)] == sizes


@pytest.mark.parametrize("nz", range(1, 70))
def test_split_covers_the_grid_evenly(nz):
    for nranks in range(1, nz + 1):
        zs = zslab_ranges(nz, nranks)
        assert zs[0][0] == 0 and zs[-1][1] == nz
        assert all(zs[i][1] == zs[i + 1][0] for i in range(nranks - 1))
        sizes = [b - a for a, b in zs]
        assert set(sizes) <= {nz // nranks, -(-nz // nranks)}
        assert sizes == sorted(sizes, reverse=True)


def test_split_needs_a_plane_per_rank():
    """parallel.f90:42-49 stops when nz < mpisize."""
    with pytest.raises(ValueError):
        zslab_ranges(3, 4)
