"""Write the FP64 reference values of the differentiation workload.

Workload: 128^3 channel, 20 seeded spheres (radius 4-8), smooth inlet profile
u_x(y, z) on the EQU cells of x = 0, zero-gradient outlet, invTau 0.5, 300
steps from rest, J = mean outlet u_x over the states 251..300.

Values written (tests/golden/adjoint_128.npz), all from the FP64 oracle:
    J, dJ/dinvTau and dJ/du_x on every inlet cell, from the checkpointed FP64
        host adjoint (oracle/oracle.cpp orc_adjoint_ckpt_f64, checkpoints every
        20 steps; the oracle's adjoint is itself pinned against central
        differences of its own forward run by tests/test_adjoint_oracle.py);
    central differences (J(u + h) - J(u - h)) / 2h, h = 1e-4, of the FP64
        forward run on 32 random inlet cells (orc_objective_f64);
    SHA-256 digests of the generated flag and vx fields, so a CPU test can
        tell when the generator no longer produces this workload.

The GPU test then runs only the GPU adjoint and compares against these.
Run from the repository root (about 15-30 minutes on 8 cores):
    python tests/golden/make_adjoint_fixture.py [--threads N]
"""
from __future__ import annotations

import argparse
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from culbm_amd import cases  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

OUT = Path(__file__).resolve().parent / "adjoint_128.npz"

N = 128
INV_TAU = 0.5
NSTEPS, T0, T1, EVERY = 300, 251, 300, 20
H = 1e-4
NCHECK, SEED = 32, 12


def outlet_weights(st, t0, t1):
    """w of J: the FLUID cells of the x = nx-1 face, each 1 / (cells * states)."""
    w = np.zeros(st.shape)
    fluid = cases.LOAD | cases.COLLIDE | cases.STORE
    out = (st.flag[:, :, -1] & 0x1f) == fluid
    w[:, :, -1][out] = 1.0 / (out.sum() * (t1 - t0 + 1))
    return w


def inlet_cells(st):
    inlet = (st.flag & cases.EQU) != 0
    inlet[:, :, 1:] = False
    return inlet


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fd_cells(st):
    """The 32 inlet cells checked by central differences, (z, y, x) rows."""
    rng = np.random.default_rng(SEED)
    cells = np.argwhere(inlet_cells(st))
    return cells[rng.choice(len(cells), size=min(NCHECK, len(cells)), replace=False)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ns = ap.parse_args(argv)
    nth = ns.threads or None
    st, _ = cases.config_adjoint(N)
    w = outlet_weights(st, T0, T1)
    f0 = np.zeros((27,) + st.shape)
    inlet = inlet_cells(st)

    t = time.time()
    J, g = po.adjoint_ckpt_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0, EVERY, nthreads=nth)
    print(f"host adjoint: J {J:.12e}, dJ/dinvTau {g['inv_tau']:.12e} ({time.time() - t:.0f} s)", flush=True)

    pick = fd_cells(st)
    base = st.vx.astype(np.float64)
    fd = []
    for k, (z, y, x) in enumerate(pick):
        t = time.time()
        vals = []
        for s in (+1, -1):
            ux = base.copy()
            ux[z, y, x] += s * H
            vals.append(po.objective_f64(st, INV_TAU, NSTEPS, T0, T1, w, f0, ux=ux, nthreads=nth))
        fd.append((vals[0] - vals[1]) / (2 * H))
        print(f"FD cell {k + 1}/{len(pick)} (z={z}, y={y}): {fd[-1]:.12e}, adjoint {g['ux'][z, y, x]:.12e} "
              f"({time.time() - t:.0f} s)", flush=True)

    np.savez_compressed(
        OUT,
        objective=np.float64(J),
        grad_inv_tau=np.float64(g["inv_tau"]),
        inlet_cells=np.argwhere(inlet).astype(np.int32),
        grad_ux_inlet=g["ux"][inlet],
        fd_cells=pick.astype(np.int32),
        fd_grad_ux=np.asarray(fd, np.float64),
        h=np.float64(H),
        steps=np.array([NSTEPS, T0, T1, EVERY], np.int32),
        inv_tau=np.float64(INV_TAU),
        flag_sha256=np.array(digest(st.flag)),
        vx_sha256=np.array(digest(st.vx)),
    )
    print(f"wrote {OUT}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
