"""The allocation-free moment sums (culbm_moment_sums_noalloc) against the
allocating entry point (culbm_moment_sums), through the C ABI.

Both run the same kernels (a bounded grid of partial sums, then their totals
in a fixed order), the first on the library's file-scope scratch array, the
second on memory it allocates, so the four FP64 sums must agree bit for bit.
Two calls in a row, the second on a smaller domain (fewer partial sums: the
scratch still holds the first call's larger set, which must not leak into the
totals), each with and without a flag mask, in place and pull, FP32 and FP16.
Checked against the FP64 host sums of the same populations to 1e-12 as well."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from culbm_amd import Lattice, LatticeConfig, _lib, cases

pytestmark = pytest.mark.gpu


def noalloc(lat, mask=0):
    out = (ctypes.c_double * 4)()
    fn = _lib.lib().culbm_moment_sums_noalloc
    _lib.check(fn(lat._ctx, ctypes.c_uint32(mask), out), lat._ctx)
    return tuple(out)


def host_sums(ddf, flag, mask):
    cx, cy, cz = (np.array([(i % 3) - 1 for i in range(27)]), np.array([(i // 3) % 3 - 1 for i in range(27)]),
                  np.array([i // 9 - 1 for i in range(27)]))
    f = ddf.astype(np.float64).reshape(27, -1)
    sel = np.ones(f.shape[1], bool) if mask == 0 else (flag.reshape(-1) & mask) != 0
    f = f[:, sel]
    return (float(sel.sum() + f.sum()), float((cx[:, None] * f).sum()), float((cy[:, None] * f).sum()),
            float((cz[:, None] * f).sum()))


def run_case(domain, sp, op, storage, seed):
    st = cases.channel(*domain, u_in=0.05, outlet="zero_gradient")
    rng = np.random.default_rng(seed)
    st.rho[...] = (1.0 + rng.uniform(-1e-3, 1e-3, st.shape)).astype(np.float32)
    cfg = LatticeConfig(domain=domain, inv_tau=0.6, stream_policy=sp, opt_policy=op, inner_loop=4, storage=storage)
    with Lattice(cfg) as lat:
        lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
        lat.init_ddf("feq")
        lat.step(9)
        ddf = lat.get_ddf()
        for mask in (0, cases.DUMP_RHO):
            want = lat.moment_sums(mask)
            got = noalloc(lat, mask)
            assert got == want, (domain, sp, op, storage, mask, got, want)
            ref = host_sums(ddf, st.flag, mask)
            for g, r in zip(got, ref):
                assert abs(g - r) <= 1e-12 * max(abs(r), 1.0), (g, r)


@pytest.mark.parametrize("sp,op,storage", [(1, 3, "fp32"), (0, 0, "fp32"), (1, 0, "fp16"), (0, 3, "fp16")])
def test_noalloc_sums_equal_the_allocating_entry_point(sp, op, storage):
    run_case((160, 96, 80), sp, op, storage, 1)   # 1.2M cells: the full 1024 partial sums
    run_case((37, 11, 9), sp, op, storage, 2)     # 3663 cells: 15 partial sums, stale ones above


def test_noalloc_rejects_null_arguments():
    fn = _lib.lib().culbm_moment_sums_noalloc
    out = (ctypes.c_double * 4)()
    assert fn(None, 0, out) == _lib.ERR_INVALID
    with Lattice(LatticeConfig(domain=(8, 8, 8))) as lat:
        assert fn(lat._ctx, 0, None) == _lib.ERR_INVALID
