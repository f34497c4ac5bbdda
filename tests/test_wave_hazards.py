"""Host-side hazard check of the temporal-blocking schedule (no GPU).

The library's own wavefront pass (culbm_wave_schedule, the launch order the
dynamic and static policies issue) is simulated slot level by slot level for
the twisted in-place layout (both starting parities) and the double-buffered
pull layout: every read must see the level its step expects, no write may land
on a slot another plane of the same launch reads, and after the pass every
slot the next step reads holds the pass's last level.  Exhaustive over
nz <= 40, depth <= 12, steps <= 9, plus the benchmark and validation shapes.
Schedules with a planted mistake (no skew, the skew turned the wrong way,
chunk ends one plane off, chunk order reversed) must be flagged, so the
checker is shown to catch them."""
from __future__ import annotations

import pytest

from culbm_amd import hazards
from culbm_amd.blocking import wave_schedule


def test_schedule_covers_every_plane_step_once():
    for nz, d, t in [(280, 18, 5), (7, 3, 4), (1, 1, 9), (40, 12, 1)]:
        seen = {}
        for z0, z1, s in wave_schedule(nz, d, t):
            for z in range(z0, z1):
                seen[(z, s)] = seen.get((z, s), 0) + 1
        assert seen == {(z, s): 1 for z in range(nz) for s in range(1, t + 1)}


@pytest.mark.parametrize("nz_block", [range(1, 11), range(11, 21), range(21, 31), range(31, 41)])
def test_library_schedule_is_hazard_free_exhaustive(nz_block):
    for nz in nz_block:
        for d in range(1, 13):
            for t in range(1, 10):
                launches = wave_schedule(nz, d, t)
                for layout, par in (("inplace", 1), ("inplace", 0), ("pull", 1)):
                    bad = hazards.check_wave_pass(nz, d, t, layout, par, launches)
                    assert not bad, f"nz={nz} D={d} T={t} {layout} parity {par}: {bad[:3]}"


@pytest.mark.parametrize("nz,d,t", [(280, 18, 5), (280, 35, 5), (280, 16, 5), (512, 6, 5), (512, 7, 5),
                                    (384, 10, 9), (384, 16, 9), (128, 86, 5), (256, 22, 5), (256, 43, 5)])
def test_library_schedule_is_hazard_free_at_the_workload_shapes(nz, d, t):
    launches = wave_schedule(nz, d, t)
    for layout, par in (("inplace", 1), ("inplace", 0), ("pull", 1)):
        assert not hazards.check_wave_pass(nz, d, t, layout, par, launches), (layout, par)


def _unskewed(nz, d, t):
    out = []
    for k in range((nz + d - 1) // d):
        for s in range(1, t + 1):
            out.append((k * d, min(nz, (k + 1) * d), s))
    return out


def _skewed_forward(nz, d, t):
    """The skew turned the wrong way: step t of chunk k starts t - 1 planes later."""
    out = []
    for k in range((nz + d - 1) // d):
        for s in range(1, t + 1):
            lo, hi = min(nz, k * d + s - 1), min(nz, (k + 1) * d + s - 1)
            if lo < hi:
                out.append((lo, hi, s))
    return out


def _chunk_end_off_by_one(nz, d, t):
    """Chunk ends shifted one plane further than the starts: planes left out."""
    out = []
    for k in range((nz + t - 1 + d - 1) // d):
        for s in range(1, t + 1):
            lo, hi = max(0, k * d - (s - 1)), min(nz, (k + 1) * d - s)
            if lo < hi:
                out.append((lo, hi, s))
    return out


def _reversed_chunks(nz, d, t):
    return list(reversed(wave_schedule(nz, d, t)))


@pytest.mark.parametrize("make", [_unskewed, _skewed_forward, _chunk_end_off_by_one, _reversed_chunks])
@pytest.mark.parametrize("layout", ["inplace", "pull"])
def test_planted_schedule_mistakes_are_flagged(make, layout):
    for nz, d, t in [(20, 4, 3), (33, 7, 5), (12, 5, 2)]:
        bad = hazards.check_wave_pass(nz, d, t, layout, 1, make(nz, d, t))
        assert bad, f"{make.__name__} {layout} nz={nz} D={d} T={t} not flagged"


def test_planted_layout_mistake_is_flagged():
    """A twisted step whose z-face cells skip their ghost stores reads stale
    data at the face in the following step."""
    orig = hazards._inplace_stores

    def no_ghost(z, parity, nz):
        for c in (-1, 0, 1):
            yield (-c if parity == 1 else c), z + hazards._p(c)

    hazards._inplace_stores = no_ghost
    try:
        assert hazards.check_wave_pass(10, 3, 3, "inplace", 1)
    finally:
        hazards._inplace_stores = orig
    assert not hazards.check_wave_pass(10, 3, 3, "inplace", 1)
