"""GPU adjoint gradients against central finite differences of the FP64 forward
run (h = 1e-4) and against the FP64 host adjoint, on the differentiation
workload: channel with seeded spheres, smooth inlet profile on the EQU cells of
x = 0, zero-gradient outlet, J = mean outlet u_x over the last steps.

The GPU computes in FP32, so the finite differences use the FP64 forward run of
the oracle (an FP32 difference quotient at h = 1e-4 is dominated by rounding).
The reduced case (40^3, 80 steps) is compared live against the FP64 oracle.
The full 128^3, 300-step workload (checkpoints every 20 steps, J over steps
251-300, 32 random inlet cells) is compared against FP64 values committed as a
fixture, since its host adjoint and 64 finite-difference runs take about half an
hour of host time; tests/golden/make_adjoint_fixture.py regenerates them."""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np
import pytest

from culbm_amd import Lattice, LatticeConfig, cases

pytestmark = pytest.mark.gpu

INV_TAU = 0.5
H = 1e-4


def outlet_weights(st, t0, t1):
    w = np.zeros(st.shape)
    fluid = cases.LOAD | cases.COLLIDE | cases.STORE
    out = (st.flag[:, :, -1] & 0x1f) == fluid
    w[:, :, -1][out] = 1.0 / (out.sum() * (t1 - t0 + 1))
    return w


def gpu_adjoint(st, nsteps, t0, t1, every, sp=1, op=3):
    w = outlet_weights(st, t0, t1)
    cfg = LatticeConfig(domain=st.domain, inv_tau=INV_TAU, stream_policy=sp, opt_policy=op, inner_loop=5)
    with Lattice(cfg) as lat:
        lat.set_state(st.flag, st.rho, st.vx, st.vy, st.vz)
        lat.init_ddf("rest")
        t = time.time()
        r = lat.adjoint(w, nsteps, (t0, t1), checkpoint_every=every, timing=True)
        print(f"GPU adjoint {time.time() - t:.2f} s, stats {r.stats}", flush=True)
    return w, r


def inlet_cells(st):
    inlet = (st.flag & cases.EQU) != 0
    inlet[:, :, 1:] = False
    return inlet


def compare_host_adjoint(st, nsteps, t0, t1, every):
    """GPU adjoint against the FP64 host adjoint: J, dJ/dinvTau, dJ/du_x on every inlet cell."""
    from oracle import pyoracle as po

    w, r = gpu_adjoint(st, nsteps, t0, t1, every)
    f0 = np.zeros((27,) + st.shape)
    t = time.time()
    J64, g64 = po.adjoint_ckpt_f64(st, INV_TAU, nsteps, t0, t1, w, f0, every)
    print(f"FP64 host adjoint {time.time() - t:.1f} s", flush=True)
    inlet = inlet_cells(st)
    gi, hi = r.grad_ux[inlet], g64["ux"][inlet]
    res = {"J_gpu": r.objective, "J_fp64": J64, "J_rel": abs(r.objective - J64) / abs(J64),
           "tau_gpu": r.grad_inv_tau, "tau_fp64": g64["inv_tau"],
           "tau_rel": abs(r.grad_inv_tau - g64["inv_tau"]) / abs(g64["inv_tau"]),
           "grad_ux_max_rel": float(np.abs(gi - hi).max() / np.abs(hi).max()),
           "inlet_cells": int(inlet.sum())}
    print(res, flush=True)
    return res


def compare_fd(st, nsteps, t0, t1, every, ncheck, seed, part=slice(None)):
    """GPU adjoint dJ/du_x against central differences (h = 1e-4) of the FP64
    forward run on `ncheck` random inlet cells (optionally a slice of them)."""
    from oracle import pyoracle as po

    w, r = gpu_adjoint(st, nsteps, t0, t1, every)
    f0 = np.zeros((27,) + st.shape)
    rng = np.random.default_rng(seed)
    cells = np.argwhere(inlet_cells(st))
    pick = cells[rng.choice(len(cells), size=min(ncheck, len(cells)), replace=False)][part]
    base = st.vx.astype(np.float64)
    fd, adj = [], []
    for k, (z, y, x) in enumerate(pick):
        t = time.time()
        vals = []
        for s in (+1, -1):
            ux = base.copy()
            ux[z, y, x] += s * H
            vals.append(po.objective_f64(st, INV_TAU, nsteps, t0, t1, w, f0, ux=ux))
        fd.append((vals[0] - vals[1]) / (2 * H))
        adj.append(r.grad_ux[z, y, x])
        print(f"FD cell {k + 1}/{len(pick)} (z={z}, y={y}): fd {fd[-1]:.9e} gpu {adj[-1]:.9e} "
              f"rel {abs(adj[-1] - fd[-1]) / abs(fd[-1]):.2e} ({time.time() - t:.1f} s)", flush=True)
    fd, adj = np.asarray(fd), np.asarray(adj)
    res = {"cells": len(fd), "max_rel": float(np.abs(adj - fd).max() / np.abs(fd).max()),
           "worst_cell_rel": float(np.max(np.abs(adj - fd) / np.abs(fd)))}
    print(res, flush=True)
    return res


REDUCED = dict(nsteps=80, t0=71, t1=80, every=10)


def reduced_state():
    return cases.config_adjoint(40, seed=4, count=4, r_min=3.0, r_max=5.0)[0]


def test_reduced_case_vs_host_adjoint():
    res = compare_host_adjoint(reduced_state(), **REDUCED)
    assert res["J_rel"] <= 1e-5 and res["tau_rel"] <= 1e-4 and res["grad_ux_max_rel"] <= 1e-4


def test_reduced_case_vs_finite_differences():
    res = compare_fd(reduced_state(), **REDUCED, ncheck=8, seed=11)
    assert res["worst_cell_rel"] <= 1e-4


FIXTURE = Path(__file__).resolve().parent / "golden" / "adjoint_128.npz"


def test_full_case_vs_fp64_fixture():
    """The differentiation workload itself: 128^3, 20 spheres, smooth inlet
    profile, 300 steps, checkpoints every 20, J over the states 251..300.  The
    GPU adjoint against the FP64 values committed in tests/golden/adjoint_128.npz
    (written by tests/golden/make_adjoint_fixture.py from the FP64 oracle): J,
    dJ/dinvTau and dJ/du_x on all 15,876 inlet cells against the FP64 host
    adjoint, and dJ/du_x on 32 random inlet cells against central differences
    (h = 1e-4) of the FP64 forward run.  Bars: 1e-4 relative."""
    g = np.load(FIXTURE)
    st = cases.config_adjoint(128)[0]
    nsteps, t0, t1, every = (int(v) for v in g["steps"])
    w, r = gpu_adjoint(st, nsteps, t0, t1, every)
    z, y, x = g["inlet_cells"].T
    gi, hi = r.grad_ux[z, y, x], g["grad_ux_inlet"]
    fz, fy, fx = g["fd_cells"].T
    fd, adj = g["fd_grad_ux"], r.grad_ux[fz, fy, fx]
    res = {"J_rel": abs(r.objective - float(g["objective"])) / abs(float(g["objective"])),
           "tau_rel": abs(r.grad_inv_tau - float(g["grad_inv_tau"])) / abs(float(g["grad_inv_tau"])),
           "grad_ux_max_rel": float(np.abs(gi - hi).max() / np.abs(hi).max()),
           "fd_worst_cell_rel": float(np.max(np.abs(adj - fd) / np.abs(fd))),
           "inlet_cells": int(gi.size), "fd_cells": int(fd.size)}
    print(res, flush=True)
    assert res["inlet_cells"] == 15876 and res["fd_cells"] == 32
    # J carries the FP32 forward run's rounding over 300 steps
    assert res["J_rel"] <= 1e-4 and res["tau_rel"] <= 1e-4
    assert res["grad_ux_max_rel"] <= 1e-4
    assert res["fd_worst_cell_rel"] <= 1e-4
