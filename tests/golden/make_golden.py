"""Write the golden vectors of the per-cell physics and the blocking math.

The expected outputs come from the reference's own functions, compiled on the
host from the headers in the reference checkout (oracle/_ref/libculbm_ref.so,
built by `make -C oracle ref`):
    calcEqu<27>, calcRhoU<27>, collision2<27>, collision<27>
        src/lbm_core/device_function.hpp:305-463, 592-754
    validBlkAxisConfig / calcBlkNum / calcValidPrev
        src/blocking_core/blocking_alogrithm.hpp:8-41
    VelSet3D<27> direction and weight tables
        src/basic/velocity_set.hpp:189-255

Run from the repository root:  python tests/golden/make_golden.py
The fixtures are data only (inputs and expected outputs), so tests on hosts
without the reference checkout still compare against the reference.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def cells(rng, n):
    """Population vectors spanning rest, near-equilibrium and strongly
    non-equilibrium cells, in the shifted (f - w) form."""
    f = np.zeros((n, 27), np.float32)
    k = n // 4
    f[:k] = 0.0
    f[k:2 * k] = rng.normal(0, 1e-3, (k, 27))
    f[2 * k:3 * k] = rng.normal(0, 2e-2, (k, 27))
    f[3 * k:] = rng.uniform(-0.05, 0.05, (n - 3 * k, 27))
    return f.astype(np.float32)


def main() -> int:
    r = pyoracle.ref()
    rng = np.random.default_rng(20240601)
    n = 768

    cx = np.zeros(27, np.int32)
    cy = np.zeros(27, np.int32)
    cz = np.zeros(27, np.int32)
    w = np.zeros(27, np.float32)
    r.ref_velset(cx, cy, cz, w)

    f_in = cells(rng, n)
    inv_taus = np.array([0.5, 0.2857142857142857, 1.0, 1.9, 0.8], np.float32)
    collide_out = np.zeros((len(inv_taus), n, 27), np.float32)
    for k, it in enumerate(inv_taus):
        g = f_in.copy()
        r.ref_collide_batch(float(it), g, n)
        collide_out[k] = g

    q = np.zeros((n, 4), np.float32)
    q[:, 0] = 1.0 + rng.uniform(-0.02, 0.02, n)
    q[:, 1:] = rng.uniform(-0.12, 0.12, (n, 3))
    q[0] = (1.0, 0.0, 0.0, 0.0)
    q[1] = (1.0, 0.05, 0.0, 0.0)
    equ_out = np.zeros((n, 27), np.float32)
    r.ref_equ_batch(np.ascontiguousarray(q[:, 0]), np.ascontiguousarray(q[:, 1]),
                    np.ascontiguousarray(q[:, 2]), np.ascontiguousarray(q[:, 3]), equ_out, n)

    mom_out = np.zeros((n, 4), np.float32)
    r.ref_moments_batch(f_in, mom_out, n)

    np.savez_compressed(
        OUT / "cell_physics.npz",
        cx=cx, cy=cy, cz=cz, w=w,
        f_in=f_in, inv_taus=inv_taus, collide_out=collide_out,
        q_in=q, equ_out=equ_out, moments_out=mom_out,
    )

    rows = []
    for dom in range(1, 97):
        for blk in range(1, 40):
            for it in range(1, 10):
                valid = r.ref_blk_valid(dom, blk, it)
                # calcBlkNum divides by blk+2-2*it once dom > 2*(blk+1-it): only
                # query it where that stays a positive integer division
                safe = dom <= blk or (blk + 1 > it and dom <= 2 * (blk + 1 - it)) or blk + 2 > 2 * it
                num = r.ref_blk_num(dom, blk, it) if valid and safe else 0
                prev = [r.ref_valid_prev(k, blk, num, it, dom) for k in range(num + 1)] if num else []
                rows.append((dom, blk, it, valid, num, prev))
    # flatten the variable-length valid-start lists
    offs = np.zeros(len(rows) + 1, np.int64)
    for k, row in enumerate(rows):
        offs[k + 1] = offs[k] + len(row[5])
    starts = np.concatenate([np.array(row[5], np.int32) for row in rows if row[5]])
    np.savez_compressed(
        OUT / "blocking.npz",
        dom=np.array([x[0] for x in rows], np.int32),
        blk=np.array([x[1] for x in rows], np.int32),
        it=np.array([x[2] for x in rows], np.int32),
        valid=np.array([x[3] for x in rows], np.int32),
        num=np.array([x[4] for x in rows], np.int32),
        starts=starts, offs=offs,
    )
    print("wrote", OUT / "cell_physics.npz", OUT / "blocking.npz")
    return 0


if __name__ == "__main__":
    sys.exit(main())
