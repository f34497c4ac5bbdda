"""The oracle against the reference's own code, compiled on the host from the
headers in the reference checkout (oracle/_ref/libculbm_ref.so, built by
`make -C oracle`): velocity set, calcEqu/calcRhoU/collision2/collision
(src/lbm_core/device_function.hpp:305-754) and the blocking core
(src/blocking_core/blocking_alogrithm.hpp:8-41).  Skipped where the checkout,
hence the binding, is absent; tests/test_golden.py covers those hosts.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(), reason="reference-header binding not built")

F32 = np.float32


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_velocity_set_identical():
    cx = np.zeros(27, np.int32)
    cy = np.zeros(27, np.int32)
    cz = np.zeros(27, np.int32)
    w = np.zeros(27, np.float32)
    po.ref().ref_velset(cx, cy, cz, w)
    ocx, ocy, ocz, ow = po.velset()
    assert np.array_equal(cx, ocx) and np.array_equal(cy, ocy) and np.array_equal(cz, ocz)
    assert np.array_equal(w, ow.astype(np.float32))


@pytest.mark.parametrize("inv_tau", [0.5, 0.2857142857142857, 1.0, 1.85, 0.01])
def test_collision_bit_identical(inv_tau):
    rng = np.random.default_rng(11)
    n = 50000
    f = np.concatenate([rng.normal(0, 1e-3, (n // 2, 27)), rng.uniform(-0.05, 0.05, (n - n // 2, 27))]).astype(F32)
    a = f.copy()
    b = f.copy()
    po.lib().orc_collide_batch_f32(inv_tau, a, n)
    po.ref().ref_collide_batch(inv_tau, b, n)
    assert np.array_equal(bits(a), bits(b))


def test_collision_reports_moments_it_used():
    rng = np.random.default_rng(12)
    for _ in range(200):
        f = rng.normal(0, 0.01, 27).astype(F32)
        g = f.copy()
        m = np.zeros(4, F32)
        po.ref().ref_collide_moments(0.6, g, m)
        assert np.array_equal(bits(m), bits(po.moments(f)))
        assert np.array_equal(bits(g), bits(po.collide(f, 0.6)))


def test_equilibrium_and_moments_bit_identical():
    rng = np.random.default_rng(13)
    n = 50000
    rho = (1 + rng.uniform(-0.05, 0.05, n)).astype(F32)
    u = rng.uniform(-0.2, 0.2, (3, n)).astype(F32)
    a = np.zeros((n, 27), F32)
    b = np.zeros((n, 27), F32)
    po.lib().orc_equ_batch_f32(rho, u[0].copy(), u[1].copy(), u[2].copy(), a, n)
    po.ref().ref_equ_batch(rho, u[0].copy(), u[1].copy(), u[2].copy(), b, n)
    assert np.array_equal(bits(a), bits(b))
    f = rng.normal(0, 0.02, (n, 27)).astype(F32)
    ma = np.zeros((n, 4), F32)
    mb = np.zeros((n, 4), F32)
    po.lib().orc_moments_batch_f32(f, ma, n)
    po.ref().ref_moments_batch(f, mb, n)
    assert np.array_equal(bits(ma), bits(mb))


def test_fp64_oracle_tracks_reference_fp32():
    rng = np.random.default_rng(14)
    for _ in range(300):
        f = rng.normal(0, 0.02, 27)
        g = f.astype(F32)
        po.ref().ref_collide(0.7, g)
        assert np.abs(g - po.collide(f, 0.7)).max() < 1e-7


def test_blocking_math_identical():
    for dom in range(1, 130, 3):
        for blk in range(1, 70, 2):
            for it in range(1, 10):
                assert po.ref().ref_blk_valid(dom, blk, it) == po.lib().orc_blk_valid(dom, blk, it)
                if blk + 2 - 2 * it <= 0 and dom > blk:
                    continue  # the reference divides by blk+2-2*it here
                num = po.ref().ref_blk_num(dom, blk, it)
                assert num == po.lib().orc_blk_num(dom, blk, it)
                for k in range(num + 1):
                    assert po.ref().ref_valid_prev(k, blk, num, it, dom) == \
                        po.lib().orc_valid_prev(k, blk, num, it, dom)
