"""The dump fused into the next step (culbm_dump_deferred) against the separate
dump pass (culbm_dump, the reference's StaticD3Q27PullDumpKernel,
kernel.cu:2634-2735), bit for bit, for every policy and storage format.

Context A steps n1, dumps, reads rho/u, steps n2 and dumps again.  Context B
steps n1, arms a deferred dump, steps n2 (whose first step writes the dump),
reads rho/u, then dumps.  Both must hold identical fields at both points and
identical populations at the end.  The flag field includes EQU cells that are
also dumped (their dumped rho/u become their EQU inputs) and cells without the
DUMP bits (their fields keep the initial values)."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import Lattice, LatticeConfig, cases

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def state():
    st = cases.channel(45, 23, 19, u_in=0.06, outlet="zero_gradient")
    cases.add_spheres(st, 4, 2.0, 4.0, seed=3)
    rng = np.random.default_rng(3)
    st.rho[...] = (1.0 + rng.uniform(-2e-3, 2e-3, st.shape)).astype(np.float32)
    # some inlet cells are dumped too; some fluid cells dump only part of the fields
    st.flag[2:9, 3:12, 0] |= np.uint32(cases.DUMP_RHO | cases.DUMP_VY)
    st.flag[5:10, 5:15, 10:20] &= ~np.uint32(cases.DUMP_VX | cases.DUMP_VZ)
    never = (st.flag & cases.DUMP_VY) == 0
    st.vy[...] = np.where(never, np.float32(7.0), np.float32(0.0))  # sentinel where vy is never dumped
    st.vy[:, :, 0] = 0.0  # the inlet's EQU input
    return st


POLICIES = [(0, 0, 5, 0), (1, 0, 5, 0), (0, 3, 5, 4), (1, 3, 5, 4), (1, 3, 3, 2), (1, 1, 5, 3), (0, 1, 4, 0),
            (0, 2, 4, 0), (0, 2, 3, 5)]


@pytest.mark.parametrize("storage,rounding", [("fp32", "nearest"), ("fp16", "nearest"), ("bf16", "stochastic")])
@pytest.mark.parametrize("sp,op,il,wd", POLICIES)
def test_fused_dump_equals_separate_dump(sp, op, il, wd, storage, rounding):
    st = state()

    def make():
        lat = Lattice(LatticeConfig(domain=st.domain, inv_tau=0.7, stream_policy=sp, opt_policy=op, inner_loop=il,
                                    wave_depth=wd, storage=storage, rounding=rounding))
        lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
        lat.init_ddf("feq")
        return lat

    n1, n2 = 7, 9
    with make() as a:
        a.step(n1)
        a.dump()
        m1 = a.macros()
        a.step(n2)
        a.dump()
        m2 = a.macros()
        fa = a.get_ddf()
    with make() as b:
        b.step(n1)
        b.dump_deferred()
        b.step(n2)
        got1 = b.macros()  # written by the first of the n2 steps
        b.dump()
        got2 = b.macros()
        fb = b.get_ddf()
    what = f"({sp},{op}) I={il} D={wd} {storage} {rounding}"
    for x, y, name in zip(got1, m1, ("rho", "vx", "vy", "vz")):
        assert np.array_equal(bits(x), bits(y)), f"{what}: fused {name} differs from the dump pass"
    for x, y in zip(got2, m2):
        assert np.array_equal(bits(x), bits(y)), f"{what}: fields after the run"
    assert np.array_equal(bits(fa), bits(fb)), f"{what}: populations after the run"
    # the fields were really written where dumped and kept elsewhere
    never = (st.flag & cases.DUMP_VY) == 0
    assert np.all(got1[2][never] == st.vy[never]) and np.all(got1[2][~never] != 7.0)


def test_deferred_dump_is_flushed_by_a_read_without_steps():
    st = state()
    cfg = LatticeConfig(domain=st.domain, inv_tau=0.7, stream_policy=1, opt_policy=3, inner_loop=5)
    with Lattice(cfg) as a, Lattice(cfg) as b:
        for lat in (a, b):
            lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
            lat.init_ddf("feq")
            lat.step(6)
        a.dump()
        b.dump_deferred()
        for x, y in zip(a.macros(), b.macros()):
            assert np.array_equal(bits(x), bits(y))
        # and the pending dump was consumed: the next steps do not dump again
        a.step(3)
        b.step(3)
        for x, y in zip(a.macros(), b.macros()):
            assert np.array_equal(bits(x), bits(y))
