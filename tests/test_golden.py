"""The oracle and the library's blocking math against golden vectors written
from the reference's own functions (tests/golden/make_golden.py).  Runs on any
host: the vectors are committed data."""
from __future__ import annotations

import numpy as np

from culbm_amd import blocking
from oracle import pyoracle as po


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_velocity_set(golden_dir):
    g = np.load(golden_dir / "cell_physics.npz")
    cx, cy, cz, w = po.velset()
    assert np.array_equal(g["cx"], cx) and np.array_equal(g["cy"], cy) and np.array_equal(g["cz"], cz)
    assert np.array_equal(g["w"], w.astype(np.float32))


def test_collision_equilibrium_moments(golden_dir):
    g = np.load(golden_dir / "cell_physics.npz")
    f_in = g["f_in"]
    n = f_in.shape[0]
    for k, it in enumerate(g["inv_taus"]):
        a = f_in.copy()
        po.lib().orc_collide_batch_f32(float(it), a, n)
        assert np.array_equal(bits(a), bits(g["collide_out"][k])), f"invTau {it}"
    q = g["q_in"]
    e = np.zeros((n, 27), np.float32)
    po.lib().orc_equ_batch_f32(*(np.ascontiguousarray(q[:, j]) for j in range(4)), e, n)
    assert np.array_equal(bits(e), bits(g["equ_out"]))
    m = np.zeros((n, 4), np.float32)
    po.lib().orc_moments_batch_f32(f_in, m, n)
    assert np.array_equal(bits(m), bits(g["moments_out"]))


def test_blocking_tables(golden_dir):
    with np.load(golden_dir / "blocking.npz") as z:
        g = {k: z[k] for k in z.files}  # NpzFile re-reads an array on every access
    for r in range(g["dom"].size):
        dom, blk, it = int(g["dom"][r]), int(g["blk"][r]), int(g["it"][r])
        assert bool(g["valid"][r]) == blocking.valid_ref(dom, blk, it)
        assert int(g["valid"][r]) == po.lib().orc_blk_valid(dom, blk, it)
        num = int(g["num"][r])
        if num == 0:
            continue
        starts = list(g["starts"][g["offs"][r]:g["offs"][r + 1]])
        # the library's own axis math (stricter validity) agrees wherever it accepts
        try:
            mine = blocking.axis(dom, blk, it)
        except ValueError:
            assert blk + 1 - it <= 0 or (dom > 2 * (blk + 1 - it) and blk + 2 - 2 * it <= 0)
            continue
        assert mine == [int(v) for v in starts], (dom, blk, it)
