"""The FP64 discrete adjoint of the oracle against central finite differences
of its own FP64 forward run (the adjoint has no reference counterpart, so the
forward model pins it).  The case covers an EQU inlet, a zero-gradient
(FLUID, clamped) outlet, bounce-back walls on the faces, a solid block, a
cell without STORE, and an objective window ending at the last step."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import cases
from oracle import pyoracle as po

NX, NY, NZ = 9, 6, 7
INV_TAU = 0.7
NSTEPS, T0, T1 = 7, 3, 7


def setup():
    st = cases.channel(NX, NY, NZ, u_in=0.05, outlet="zero_gradient")
    st.flag[3, 2:4, 4:6] = cases.WALL
    st.flag[2, 3, 2] = cases.LOAD | cases.COLLIDE  # no STORE: frozen populations
    rng = np.random.default_rng(21)
    st.vx[:, :, 0] = np.where(st.flag[:, :, 0] == cases.EQUILIBRIUM,
                              0.05 + 0.01 * rng.uniform(-1, 1, (NZ, NY)), 0.0).astype(np.float32)
    st.rho[...] = 1.0
    w = np.zeros(st.shape)
    w[1:-1, 1:-1, -1] = 1.0 / ((NY - 2) * (NZ - 2) * (T1 - T0 + 1))
    f0 = np.zeros((27,) + st.shape)
    for z, y, x in np.ndindex(st.shape):
        f0[:, z, y, x] = po.equ(1.0 + 1e-3 * rng.uniform(-1, 1), *(0.02 * rng.uniform(-1, 1, 3)))
    return st, w, f0


def J_of(st, w, f0, inv_tau=INV_TAU, **fields):
    return po.adjoint_f64(st, inv_tau, NSTEPS, T0, T1, w, f0, **fields)[0]


@pytest.fixture(scope="module")
def case():
    st, w, f0 = setup()
    J, g = po.adjoint_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0)
    return st, w, f0, J, g


def fd(fun, x0, h):
    return (fun(x0 + h) - fun(x0 - h)) / (2 * h)


def test_objective_is_mean_outlet_velocity(case):
    st, w, f0, J, g = case
    assert 0.0 < J < 0.1


def test_inlet_velocity_gradient(case):
    st, w, f0, J, g = case
    base = st.vx.astype(np.float64)
    for (z, y) in [(1, 1), (2, 3), (5, 4), (3, 2)]:
        def fun(v):
            ux = base.copy()
            ux[z, y, 0] = v
            return J_of(st, w, f0, ux=ux)
        want = fd(fun, base[z, y, 0], 1e-6)
        assert abs(g["ux"][z, y, 0] - want) <= 1e-6 * max(abs(want), 1e-8) + 1e-12, (z, y)


def test_inlet_density_and_transverse_velocity_gradient(case):
    st, w, f0, J, g = case
    z, y = 2, 2
    for key, field in (("rho", st.rho), ("uy", st.vy), ("uz", st.vz)):
        base = field.astype(np.float64)

        def fun(v):
            a = base.copy()
            a[z, y, 0] = v
            return J_of(st, w, f0, **{key: a})
        want = fd(fun, base[z, y, 0], 1e-6)
        assert abs(g[key][z, y, 0] - want) <= 1e-6 * max(abs(want), 1e-8) + 1e-12, key


def test_inv_tau_gradient(case):
    st, w, f0, J, g = case
    want = fd(lambda v: J_of(st, w, f0, inv_tau=v), INV_TAU, 1e-6)
    assert abs(g["inv_tau"] - want) <= 1e-6 * abs(want) + 1e-12


def test_initial_state_gradient(case):
    st, w, f0, J, g = case
    rng = np.random.default_rng(3)
    picks = [(int(rng.integers(27)), int(rng.integers(NZ)), int(rng.integers(NY)), int(rng.integers(NX)))
             for _ in range(12)] + [(5, 2, 3, 2), (14, 3, 3, NX - 1), (22, 0, 2, 4)]
    for (i, z, y, x) in picks:
        def fun(v):
            f = f0.copy()
            f[i, z, y, x] = v
            return J_of(st, w, f)
        want = fd(fun, f0[i, z, y, x], 1e-6)
        assert abs(g["f0"][i, z, y, x] - want) <= 1e-6 * max(abs(want), 1e-8) + 1e-12, (i, z, y, x)


@pytest.mark.parametrize("every,nth", [(1, 1), (2, 3), (3, 8), (NSTEPS, 4), (100, 2)])
def test_checkpointed_parallel_adjoint_is_bit_identical(case, every, nth):
    st, w, f0, J, g = case
    J2, g2 = po.adjoint_ckpt_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0, every, nthreads=nth)
    assert J2 == J
    assert g2["inv_tau"] == g["inv_tau"]
    for k in ("rho", "ux", "uy", "uz", "f0"):
        assert np.array_equal(g2[k], g[k]), k


def test_forward_objective_matches_the_adjoint_run(case):
    st, w, f0, J, g = case
    for nth in (1, 5):
        assert po.objective_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0, nthreads=nth) == J


def test_forward_objective_steps_like_the_fp64_simulator(case):
    # the objective's forward run is the oracle's FP64 step
    st, w, f0, J, g = case
    _, f_end = po.objective_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0, f_end=True)
    sim = po.Sim(st, INV_TAU, np.float64)
    sim.set_ddf(f0)
    sim.step(NSTEPS)
    assert np.array_equal(f_end, sim.ddf())
