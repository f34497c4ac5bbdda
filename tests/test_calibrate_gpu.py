"""The bandwidth probes behind culbm_amd.calibrate.

Every mode runs, reports a rate between a slow PCIe-class floor and twice
MI355X's HBM peak, and passes the device-side data check: after the timed
launches every word the probe claims to have moved (copied, shifted, gathered
with the pull's x offsets, written, or read into its per-workgroup XOR) holds
what it should.  The check itself is shown to catch a probe that moved the
wrong data: run against a deliberately wrong expectation (one word off) it
must fail.  The population-pattern probe is at least as fast from the cache
as from HBM."""
from __future__ import annotations

import pytest

from culbm_amd._lib import ERR_STATE, CulbmError
from culbm_amd.lattice import BW_MODES, probe_bandwidth

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.mark.parametrize("mode", sorted(BW_MODES))
def test_every_probe_moves_its_data_at_a_plausible_rate(mode):
    g = probe_bandwidth(mode, 64 * MB, iters=3)  # raises if the data check fails
    assert 50.0 < g < 16000.0, (mode, g)


@pytest.mark.parametrize("mode", sorted(BW_MODES))
@pytest.mark.parametrize("nbytes", [64 * 1024 + 16, 3 * MB + 48])
def test_data_check_on_ragged_sizes(mode, nbytes):
    probe_bandwidth(mode, nbytes, iters=1)


@pytest.mark.parametrize("mode", sorted(BW_MODES))
def test_data_check_catches_wrong_data(mode):
    with pytest.raises(CulbmError) as e:
        probe_bandwidth(mode, 4 * MB, iters=1, check_wrong=True)
    assert e.value.code == ERR_STATE and "moved wrong data" in e.value.message


def test_population_pattern_is_faster_from_the_cache_than_from_hbm():
    small = probe_bandwidth("pop27", 64 * MB, iters=5)
    big = probe_bandwidth("pop27", 1024 * MB, iters=3)
    assert small > 0.9 * big, (small, big)


def test_selector_times_every_candidate_and_returns_the_fastest():
    from culbm_amd.calibrate import select

    res = select((64, 48, 40), stream_policy=1, storage="fp32", inner_loops=(3,), depths=(0, 4, 8), steps=6, develop=10)
    rows = res["candidates"]
    assert [r["requested_depth"] for r in rows] == [0, 4, 8]
    assert all(r["mlups"] > 0 for r in rows)
    assert [r["wave_depth"] for r in rows[1:]] == [4, 8]  # explicit depths are used as given
    assert rows[0]["wave_depth"] >= 2                      # the automatic one is reported
    assert res["best"]["mlups"] == max(r["mlups"] for r in rows)
