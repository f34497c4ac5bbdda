"""Host side of the depth selector in culbm_amd.calibrate (no GPU)."""
from __future__ import annotations

import math

from culbm_amd.calibrate import SELECT_CELLS, depth_candidates


def test_depth_candidates_cover_the_launch_sizes_within_the_domain():
    for dom in ((288, 272, 280), (128, 128, 128), (512, 512, 512), (64, 64, 8), (1000, 1000, 3)):
        nx, ny, nz = dom
        c = depth_candidates(dom)
        assert c[0] == 0 and c == sorted(set(c))                # the automatic choice, then ascending
        assert all(2 <= d <= nz for d in c[1:]) or nz < 2
        for cells in SELECT_CELLS:                              # every target size is represented
            d = max(2, min(nz, math.ceil(cells / (nx * ny))))
            assert d in c


def test_depth_candidates_at_the_optimised_workload():
    # 288x272 planes of 78,336 cells: 1.4M cells is 18 planes, 2.6M is 34
    c = depth_candidates((288, 272, 280))
    assert 18 in c and 34 in c and 16 in c


def test_select_rejects_an_empty_candidate_list():
    import pytest

    from culbm_amd.calibrate import select

    with pytest.raises(ValueError):
        select((16, 16, 16), depths=())
    with pytest.raises(ValueError):
        select((16, 16, 16), inner_loops=())
