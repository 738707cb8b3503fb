"""Host side of the ragged-batch runtime (cg_batch_tiles; CPU only, no GPU call).

The tile list must cover every point of every patch exactly once, never straddle a patch
boundary, keep every tile within the kernel's points per workgroup, and align tiles to
multiples of the tile size in the packed arrays (a short head tile per patch), so that full
tiles read whole cache lines as the contiguous launch does.  Empty and one-point patches
produce no / one tile; bad offsets are refused with an error code.
"""
import numpy as np
import pytest

from chemgen_amd.runtime import ChemgenError, batch_tiles, lib


def _check(offsets, tile):
    st, ln = batch_tiles(offsets, tile)
    off = np.asarray(offsets)
    covered = np.zeros(off[-1], dtype=np.int64)
    for s, n in zip(st, ln):
        assert 1 <= n <= tile
        b = np.searchsorted(off, s, side="right") - 1        # patch of the tile's first point
        assert off[b] <= s and s + n <= off[b + 1]            # inside one patch
        e = s + n
        assert e == off[b + 1] or e % tile == 0                # ends on a tile boundary or the patch end
        assert s == off[b] or s % tile == 0
        covered[s:e] += 1
    assert np.all(covered == 1)
    return st, ln


@pytest.mark.parametrize("tile", [1, 4, 16, 64])
def test_tiles_cover_patches(tile):
    rng = np.random.default_rng(3)
    sizes = np.concatenate([[0, 1, 0, 63, 64, 65, 1000], rng.integers(0, 3000, 40)])
    off = np.concatenate([[0], np.cumsum(sizes)])
    _check(off, tile)


def test_tile_counts():
    st, ln = _check([0, 0, 1, 1, 130], 64)
    assert list(st) == [0, 1, 64, 128] and list(ln) == [1, 63, 64, 2]
    st, ln = _check([0], 64)
    assert len(st) == 0


def test_bad_offsets_refused():
    with pytest.raises(ChemgenError):
        batch_tiles([1, 5], 64)
    with pytest.raises(ChemgenError):
        batch_tiles([0, 10, 5], 64)
    assert lib().cg_batch_tiles(None, 1, 64, None, None, 0) < 0
