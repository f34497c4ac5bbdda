"""Properties of the halo-blocking decomposition (culbm_amd.blocking, the
library's culbm_blocking_axis), checked exhaustively on small axes against a
naive sweep: the valid ranges tile [0, N) disjointly, every tile loads at most
B cells, halos are at most I-1 cells, and after I steps of clamped pull
streaming a tile's valid cells depend only on cells it loaded."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import blocking, cases


def test_axis_partition_properties():
    checked = 0
    for dom in range(1, 90):
        for blk in range(1, 40):
            for it in range(1, 8):
                try:
                    vb = blocking.axis(dom, blk, it)
                except ValueError:
                    continue
                checked += 1
                n = len(vb) - 1
                assert vb[0] == 0 and vb[-1] == dom
                assert all(vb[k] < vb[k + 1] for k in range(n)), (dom, blk, it, vb)
                for k in range(n):
                    lo = max(vb[k] - (it - 1), 0)
                    hi = min(vb[k + 1] + (it - 1), dom)
                    assert hi - lo <= blk, (dom, blk, it, k)
                    # dependency cone: I steps of +-1 streaming from the loaded
                    # range reach the valid range without touching an unloaded cell,
                    # except across the domain faces where the clamp keeps it inside
                    assert vb[k] - lo >= min(it - 1, vb[k])
                    assert hi - vb[k + 1] >= min(it - 1, dom - vb[k + 1])
    assert checked > 5000


def test_reference_readme_configuration():
    """288x272x280 with tile 64x96x76 and I=5 (the reference's tuned run):
    5 x 3 x 4 = 60 tiles, 1.277x redundant work."""
    dims = (288, 272, 280)
    tile = (64, 96, 76)
    assert [len(blocking.axis(d, b, 5)) - 1 for d, b in zip(dims, tile)] == [5, 3, 4]
    assert len(blocking.plan(dims, tile, 5)) == 60
    assert blocking.redundancy(dims, tile, 5) == pytest.approx(1.277, abs=1e-3)


def test_tile_flags_mark_valid_box():
    st = cases.channel(20, 14, 12)
    tile = (9, 8, 7)
    it = 2
    covered = np.zeros(st.shape, np.int32)
    for t in blocking.plan(st.domain, tile, it):
        tf = blocking.tile_flags(st.flag, t, tile)
        (x0, y0, z0), (x1, y1, z1) = t.origin, t.load_end
        sub = tf[: z1 - z0, : y1 - y0, : x1 - x0]
        assert np.array_equal(sub & ~np.uint32(blocking.CORRECT), st.flag[z0:z1, y0:y1, x0:x1])
        assert np.all(tf[z1 - z0:] == 0) and np.all(tf[:, y1 - y0:] == 0) and np.all(tf[:, :, x1 - x0:] == 0)
        mark = (sub & np.uint32(blocking.CORRECT)) != 0
        covered[z0:z1, y0:y1, x0:x1] += mark
    assert np.all(covered == 1)


def test_invalid_axis_is_rejected():
    with pytest.raises(ValueError):
        blocking.axis(100, 4, 4)   # B+1-I = 1 but B+2-2I = -2 with more than two tiles
    with pytest.raises(ValueError):
        blocking.axis(10, 3, 5)    # a tile shorter than its halo
