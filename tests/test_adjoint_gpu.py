"""GPU discrete adjoint (culbm_adjoint) against the FP64 host adjoint of the
oracle, across checkpoint schedules, storage formats and policies.

The adjoint has no reference counterpart; the oracle's FP64 adjoint is pinned
by central finite differences of its own forward run (test_adjoint_oracle.py),
and its forward run is the reference update.  The GPU adjoint runs in FP32
with FP64 accumulators, so it is compared with a tolerance relative to the
largest gradient entry, stated per test."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import CulbmError, Lattice, LatticeConfig, cases
from culbm_amd.lattice import adjoint_plan

pytestmark = pytest.mark.gpu

NX, NY, NZ = 16, 12, 10
INV_TAU = 0.6
N, T0, T1 = 24, 14, 24


def make_case(seed: int = 5, frozen: bool = True):
    st = cases.channel(NX, NY, NZ, u_in=0.05, outlet="zero_gradient")
    st.flag[4:7, 5:8, 6:9] = cases.WALL
    if frozen:
        st.flag[3, 3, 3] = cases.LOAD | cases.COLLIDE | cases.DUMP  # no STORE: never changes
    prof, _ = cases.inlet_profile(NY, NZ, seed)
    inlet = st.flag[:, :, 0] == cases.EQUILIBRIUM
    st.vx[:, :, 0] = np.where(inlet, prof, 0.0)
    rng = np.random.default_rng(seed)
    st.rho[...] = (1.0 + rng.uniform(-1e-3, 1e-3, st.shape)).astype(np.float32)
    w = np.zeros(st.shape)
    w[1:-1, 1:-1, -1] = 1.0 / ((NY - 2) * (NZ - 2) * (T1 - T0 + 1))  # mean outlet u_x
    w[3, 3, 3] = 0.01     # a cell without STORE
    w[0, 2, 5] = -0.02    # a wall cell on a face
    return st, w


def lattice(st, storage="fp32", sp=0, op=0, inner=5, rounding="nearest"):
    lat = Lattice(LatticeConfig(domain=st.domain, inv_tau=INV_TAU, stream_policy=sp, opt_policy=op,
                                inner_loop=inner, storage=storage, rounding=rounding))
    lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
    lat.init_ddf("feq")
    return lat


def rel_max(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def reference():
    from oracle import pyoracle as po
    st, w = make_case()
    with lattice(st) as lat:
        f0 = lat.get_ddf().astype(np.float64)
    J, g = po.adjoint_f64(st, INV_TAU, N, T0, T1, w, f0)
    return st, w, J, g


@pytest.fixture(scope="module")
def gpu_fp32():
    st, w = make_case()
    with lattice(st) as lat:
        return lat.adjoint(w, N, (T0, T1), want_f0=True, timing=True)


def test_objective_matches_fp64(reference, gpu_fp32):
    _, _, J, _ = reference
    print(f"J gpu {gpu_fp32.objective:.10e} fp64 {J:.10e}", flush=True)
    # J differs from FP64 by the FP32 forward run's rounding (about 1e-6 relative)
    assert abs(gpu_fp32.objective - J) <= 2e-5 * abs(J)


@pytest.mark.parametrize("key", ["rho", "ux", "uy", "uz"])
def test_inlet_gradients_match_fp64(reference, gpu_fp32, key):
    st, _, _, g = reference
    got = getattr(gpu_fp32, "grad_" + key)
    inlet = (st.flag & cases.EQU) != 0
    assert np.all(got[~inlet] == 0.0)
    err = rel_max(got[inlet], g[key][inlet])
    print(f"grad_{key}: max rel err {err:.3e}", flush=True)
    assert err <= 2e-4


def test_inv_tau_and_f0_gradients_match_fp64(reference, gpu_fp32):
    _, _, _, g = reference
    e_tau = abs(gpu_fp32.grad_inv_tau - g["inv_tau"]) / abs(g["inv_tau"])
    e_f0 = rel_max(gpu_fp32.grad_f0, g["f0"])
    print(f"dJ/dinvTau err {e_tau:.3e}, dJ/df0 err {e_f0:.3e}", flush=True)
    assert e_tau <= 2e-4
    assert e_f0 <= 2e-4
    # the frozen cell's carry reaches dJ/df0
    assert np.abs(gpu_fp32.grad_f0[:, 3, 3, 3]).max() > 0.0


def test_schedules_give_identical_bits(gpu_fp32):
    st, w = make_case()
    runs = [dict(checkpoint_every=k) for k in (1, 3, 5, 7, N)] + \
           [dict(max_slots=s) for s in (1, 2, 3, 6)]
    for kw in runs:
        with lattice(st) as lat:
            r = lat.adjoint(w, N, (T0, T1), want_f0=True, **kw)
        k = kw.get("checkpoint_every", 0)
        assert r.stats["forward_steps"] == adjoint_plan(N, k, kw.get("max_slots", 1))[0], kw
        assert r.objective == gpu_fp32.objective, kw
        assert r.grad_inv_tau == gpu_fp32.grad_inv_tau, kw
        for key in ("grad_rho", "grad_ux", "grad_uy", "grad_uz", "grad_f0"):
            assert np.array_equal(getattr(r, key), getattr(gpu_fp32, key)), (kw, key)


@pytest.mark.parametrize("sp,op,storage,rounding", [(0, 0, "fp32", "nearest"), (1, 0, "fp32", "nearest"),
                                                   (1, 3, "fp32", "nearest"), (0, 2, "fp32", "nearest"),
                                                   (1, 3, "fp16", "stochastic"), (0, 2, "bf16", "stochastic")])
def test_context_ends_nsteps_on(sp, op, storage, rounding):
    """The adjoint's forward run reproduces the context's own trajectory bit for
    bit (with stochastic rounding it continues the context's step numbers)."""
    st, w = make_case(frozen=False)
    kw = dict(sp=sp, op=op, inner=4, storage=storage, rounding=rounding)
    with lattice(st, **kw) as a, lattice(st, **kw) as b:
        a.step(3)
        b.step(3)
        a.adjoint(w, N, (T0, T1), checkpoint_every=5)
        b.step(N)
        assert a.info()["step"] == b.info()["step"] == N + 3
        assert np.array_equal(a.get_ddf(), b.get_ddf())
        a.step(3)
        b.step(3)
        assert np.array_equal(a.get_ddf(), b.get_ddf())


def test_objective_equals_dumped_velocities():
    st, w = make_case()
    with lattice(st) as lat:
        r = lat.adjoint(w, N, (T0, T1))
    want = 0.0
    with lattice(st) as lat:
        for t in range(N + 1):
            if t >= T0:
                lat.dump()
                vx = lat.macros()[1].astype(np.float64)
                m = (st.flag & cases.DUMP_VX) != 0
                want += float((w[m] * vx[m]).sum())
                # cells without the DUMP bit are not dumped: add them from the populations
                if (~m & (w != 0)).any():
                    f = lat.get_ddf().astype(np.float64)
                    for z, y, x in zip(*np.nonzero(~m & (w != 0))):
                        want += w[z, y, x] * _pulled_ux(f, x, y, z)
            if t < N:
                lat.step(1)
    # the same FP32 velocities, summed in another order (one wall cell's u_x in FP64)
    assert abs(r.objective - want) <= 1e-9 * abs(want)


def _pulled_ux(f, x, y, z):
    """x-velocity of the populations cell (x, y, z) pulls (clamped at faces)."""
    from oracle import pyoracle as po
    nz, ny, nx = f.shape[1:]
    g = np.empty(27)
    for i in range(27):
        cx, cy, cz = i % 3 - 1, (i // 3) % 3 - 1, i // 9 - 1
        sx = min(max(x - cx, 0), nx - 1) if 0 <= x - cx < nx else x
        sy = min(max(y - cy, 0), ny - 1) if 0 <= y - cy < ny else y
        sz = min(max(z - cz, 0), nz - 1) if 0 <= z - cz < nz else z
        g[i] = f[i, sz, sy, sx]
    return po.moments(g)[1]


@pytest.mark.parametrize("storage,tol,tol_j", [("fp16", 1e-4, 1e-3), ("bf16", 1e-3, 5e-3)])
def test_16bit_forward_adjoint_close_to_fp32(gpu_fp32, storage, tol, tol_j):
    st, w = make_case()
    with lattice(st, storage=storage) as lat:
        r = lat.adjoint(w, N, (T0, T1), max_slots=4)
    inlet = (st.flag & cases.EQU) != 0
    err = rel_max(r.grad_ux[inlet], gpu_fp32.grad_ux[inlet])
    ej = abs(r.objective - gpu_fp32.objective) / abs(gpu_fp32.objective)
    print(f"{storage}: grad_ux err vs fp32 {err:.3e}, J err {ej:.3e}", flush=True)
    assert err <= tol and ej <= tol_j


def test_invalid_specs_raise():
    st, w = make_case()
    with lattice(st) as lat:
        with pytest.raises(CulbmError, match="window"):
            lat.adjoint(w, N, (T0, N + 1))
        with pytest.raises(CulbmError, match="memory cap"):
            lat.adjoint(w, N, (T0, T1), mem_cap=1 << 16)
        with pytest.raises(CulbmError, match="memory cap"):
            lat.adjoint(w, N, (T0, T1), checkpoint_every=2, mem_cap=1 << 20)
        with pytest.raises(CulbmError, match="grad_f0|want_f0"):
            from culbm_amd import _lib
            import ctypes
            spec = _lib.AdjointSpec()
            spec.nsteps, spec.t_begin, spec.t_end, spec.want_f0 = N, T0, T1, 1
            st_ = _lib.AdjointStats()
            wd = np.ascontiguousarray(w, np.float64)
            _lib.check(_lib.lib().culbm_adjoint(lat._ctx, ctypes.byref(spec), _lib.ptr(wd),
                                                ctypes.byref(st_), None, None, None, None, None), lat._ctx)
        assert lat.info()["step"] == 0  # failed calls leave the state alone


def test_autograd_function_returns_adjoint_gradients():
    import torch
    from culbm_amd.autograd import InletProblem, inlet_objective
    st, w = make_case(frozen=False)
    inlet = (st.flag & cases.EQU) != 0
    inlet[:, :, -1] = False
    prob = InletProblem(st, inlet, w, N, (T0, T1), init="rest")
    p = torch.tensor(prob.default_profile(), requires_grad=True)
    tau = torch.tensor(INV_TAU, dtype=torch.float64, requires_grad=True)
    J = inlet_objective(p, tau, prob)
    (2.0 * J).backward()
    with Lattice(LatticeConfig(domain=st.domain, inv_tau=INV_TAU)) as lat:
        lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
        lat.init_ddf("rest")
        r = lat.adjoint(w, N, (T0, T1))
    assert float(J.detach()) == r.objective
    assert np.array_equal(p.grad.numpy(), 2.0 * r.grad_ux[inlet])
    assert float(tau.grad) == 2.0 * r.grad_inv_tau
    prob.close()
