"""The bandwidth probes behind culbm_amd.calibrate: every mode runs and reports
a rate between a slow PCIe-class floor and twice MI355X's HBM peak, and the
population-pattern probes (modes 5, 6) move the bytes they claim (their
rate at a cache-resident size is at least their rate at an HBM-sized one)."""
from __future__ import annotations

import pytest

from culbm_amd.lattice import BW_MODES, probe_bandwidth

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.mark.parametrize("mode", sorted(BW_MODES))
def test_every_probe_reports_a_plausible_rate(mode):
    g = probe_bandwidth(mode, 64 * MB, iters=3)
    assert 50.0 < g < 16000.0, (mode, g)


def test_population_pattern_is_faster_from_the_cache_than_from_hbm():
    small = probe_bandwidth("pop27", 64 * MB, iters=5)
    big = probe_bandwidth("pop27", 1024 * MB, iters=3)
    assert small > 0.9 * big, (small, big)


def test_selector_times_every_candidate_and_returns_the_fastest():
    from culbm_amd.calibrate import select

    res = select((64, 48, 40), stream_policy=1, storage="fp32", inner_loops=(3,), depths=(0, 4, 8), steps=6)
    rows = res["candidates"]
    assert [r["requested_depth"] for r in rows] == [0, 4, 8]
    assert all(r["mlups"] > 0 for r in rows)
    assert [r["wave_depth"] for r in rows[1:]] == [4, 8]  # explicit depths are used as given
    assert rows[0]["wave_depth"] >= 2                      # the automatic one is reported
    assert res["best"]["mlups"] == max(r["mlups"] for r in rows)
