"""FP64 reductions of the stored populations (culbm_moment_sums): total mass
and momentum over flag-selected cells, for every layout and storage format,
against FP64 sums of the same populations read back with culbm_get_ddf, and
the GPU mass against the FP64 oracle's after a run."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import Lattice, LatticeConfig, cases

pytestmark = pytest.mark.gpu

CX = np.array([i % 3 - 1 for i in range(27)], np.float64)
CY = np.array([(i // 3) % 3 - 1 for i in range(27)], np.float64)
CZ = np.array([i // 9 - 1 for i in range(27)], np.float64)


def host_sums(ddf, sel):
    f = ddf.astype(np.float64)[:, sel]
    return np.array([(1.0 + f.sum(axis=0)).sum(), (CX @ f).sum(), (CY @ f).sum(), (CZ @ f).sum()])


def make(st, **kw):
    lat = Lattice(LatticeConfig(domain=st.domain, **kw))
    lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
    lat.init_ddf("feq")
    return lat


@pytest.mark.parametrize("sp,op,storage,steps", [(0, 0, "fp32", 7), (1, 0, "fp32", 7), (1, 0, "fp32", 8),
                                                 (1, 3, "fp32", 10), (0, 0, "fp16", 5), (1, 0, "bf16", 3)])
def test_sums_match_host_sums_of_the_populations(sp, op, storage, steps):
    st = cases.config_small(seed=9, n=24)
    st.vy[...] = 0.01
    with make(st, inv_tau=0.7, stream_policy=sp, opt_policy=op, inner_loop=5, storage=storage) as lat:
        lat.step(steps)
        ddf = lat.get_ddf()
        for mask in (0, cases.DUMP_VX, cases.BOUNCE_BACK):
            sel = np.ones(st.shape, bool) if mask == 0 else (st.flag & mask) != 0
            got = np.array(lat.moment_sums(mask))
            want = host_sums(ddf, sel)
            scale = np.array([want[0]] + [max(abs(w), 1e-3) for w in want[1:]])
            assert np.all(np.abs(got - want) <= 1e-12 * scale), (mask, got, want)


def test_mass_follows_the_fp64_oracle():
    from oracle import pyoracle as po
    st = cases.config_small(seed=2, n=32)
    with make(st, inv_tau=0.6) as lat:
        m0 = lat.moment_sums()[0]
        lat.step(50)
        m1 = lat.moment_sums()[0]
    sim = po.Sim(st, 0.6, np.float64)
    sim.init_feq()
    w0 = host_sums(sim.ddf(), np.ones(st.shape, bool))[0]
    sim.step(50)
    w1 = host_sums(sim.ddf(), np.ones(st.shape, bool))[0]
    # the inlet and the open outlet change the mass; FP32 storage follows the FP64 run closely
    assert abs(m0 - w0) <= 1e-9 * w0
    assert abs(m1 - w1) <= 1e-6 * w1
    assert abs(w1 - w0) > 1e-6 * w0
