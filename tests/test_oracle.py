"""Pins of the CPU oracle that need neither the reference checkout nor a GPU.

Each check is chosen so a plausible restatement mistake fails it: the
equilibrium is compared with the textbook second-order form built from the
direction table and weights (a wrong sign, index or weight fails), moments
and conservation laws are exact identities, the lattice step is compared with
an independent brute-force pure-Python implementation of the pull semantics
(kernel.cu:2489-2632: clamp, flag order, STORE) on tiny grids, and the 16-bit
codecs are compared with numpy's own round-to-nearest-even conversions.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from culbm_amd import cases
from oracle import pyoracle as po

CX, CY, CZ, W = po.velset()
C = np.stack([CX, CY, CZ], axis=1).astype(np.float64)


def feq_textbook(rho, u):
    cu = C @ u
    return W * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u)) - W


# ---- velocity set --------------------------------------------------------------

def test_direction_numbering_and_opposites():
    for i in range(27):
        assert i == (CX[i] + 1) + 3 * (CY[i] + 1) + 9 * (CZ[i] + 1)
        assert (CX[26 - i], CY[26 - i], CZ[26 - i]) == (-CX[i], -CY[i], -CZ[i])
    assert (CX[13], CY[13], CZ[13]) == (0, 0, 0)


def test_weights_isotropy():
    assert abs(W.sum() - 1) < 1e-15
    assert np.abs(W @ C).max() < 1e-15
    assert np.abs((C.T * W) @ C - np.eye(3) / 3).max() < 1e-15
    # fourth moment: (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9
    d = np.eye(3)
    m4 = np.einsum("i,ia,ib,ic,id->abcd", W, C, C, C, C)
    want = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
            + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.abs(m4 - want).max() < 1e-15


# ---- per-cell physics (FP64) -------------------------------------------------------

def test_equilibrium_matches_textbook_form():
    rng = np.random.default_rng(0)
    for _ in range(200):
        rho = 1 + rng.uniform(-0.05, 0.05)
        u = rng.uniform(-0.15, 0.15, 3)
        assert np.abs(po.equ(rho, *u) - feq_textbook(rho, u)).max() < 2e-16


def test_equilibrium_moments_exact():
    rng = np.random.default_rng(1)
    for _ in range(100):
        rho = 1 + rng.uniform(-0.05, 0.05)
        u = rng.uniform(-0.15, 0.15, 3)
        f = po.equ(rho, *u)
        assert abs(f.sum() - (rho - 1)) < 1e-15
        assert np.abs(C.T @ f - rho * u).max() < 1e-15
        # second moment of feq (unshifted): rho (u u + I/3)
        pi = np.einsum("i,ia,ib->ab", f + W, C, C)
        assert np.abs(pi - rho * (np.outer(u, u) + np.eye(3) / 3)).max() < 1e-14


def test_rest_equilibrium_is_zero():
    assert np.all(po.equ(1.0, 0.0, 0.0, 0.0) == 0.0)
    assert np.all(po.equ(1.0, 0.0, 0.0, 0.0, dtype=np.float32) == 0.0)


def test_moments_definition():
    rng = np.random.default_rng(2)
    f = rng.normal(0, 0.02, 27)
    rho, ux, uy, uz = po.moments(f)
    assert abs(rho - (1 + f.sum())) < 1e-15
    assert np.abs(np.array([ux, uy, uz]) - C.T @ f / rho).max() < 1e-15


@pytest.mark.parametrize("inv_tau", [0.0, 0.2857142857142857, 0.5, 1.0, 1.7])
def test_collision_conserves_and_relaxes(inv_tau):
    rng = np.random.default_rng(3)
    for _ in range(50):
        f = rng.normal(0, 0.02, 27)
        g = po.collide(f, inv_tau)
        assert abs(g.sum() - f.sum()) < 1e-15
        assert np.abs(C.T @ g - C.T @ f).max() < 1e-15
        rho, ux, uy, uz = po.moments(f)
        feq = po.equ(rho, ux, uy, uz)
        assert np.abs(g - (inv_tau * feq + (1 - inv_tau) * f)).max() < 1e-15
        if inv_tau == 0.0:
            assert np.array_equal(g, f)
    # equilibrium is a fixed point
    f = po.equ(1.01, 0.05, -0.02, 0.03)
    assert np.abs(po.collide(f, inv_tau) - f).max() < 1e-15


def test_fp32_restatement_close_to_fp64():
    rng = np.random.default_rng(4)
    for _ in range(100):
        f = rng.normal(0, 0.02, 27)
        g64 = po.collide(f, 0.6)
        g32 = po.collide(f.astype(np.float32), 0.6)
        assert np.abs(g32 - g64).max() < 1e-7  # rho = 1 + sum f carries FP32 rounding of ~6e-8


# ---- lattice step vs a brute-force pure-Python pull --------------------------------

def brute_step(flag, rho, vx, vy, vz, inv_tau, src, dst):
    """Independent pure-Python restatement of one pull step (FP64)."""
    _, nz, ny, nx = src.shape
    out = dst.copy()
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        fl = int(flag[z, y, x])
        f = np.zeros(27)
        if fl & cases.LOAD:
            for i in range(27):
                sx, sy, sz = x - CX[i], y - CY[i], z - CZ[i]
                sx = x if not 0 <= sx < nx else sx
                sy = y if not 0 <= sy < ny else sy
                sz = z if not 0 <= sz < nz else sz
                f[i] = src[i, sz, sy, sx]
        if fl & cases.EQU:
            f = po.equ(float(rho[z, y, x]), float(vx[z, y, x]), float(vy[z, y, x]), float(vz[z, y, x]))
        if fl & cases.BOUNCE_BACK:
            f = f[::-1].copy()
        if fl & cases.COLLIDE:
            f = po.collide(f, inv_tau)
        if fl & cases.STORE:
            out[:, z, y, x] = f
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 5, 6), (6, 2, 3), (1, 4, 5)])
def test_step_matches_brute_force(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(sum(shape))
    bits = [cases.LOAD, cases.EQU, cases.COLLIDE, cases.BOUNCE_BACK, cases.STORE]
    flag = np.zeros(shape, np.uint32)
    for b in bits:
        flag |= np.where(rng.random(shape) < 0.6, np.uint32(b), np.uint32(0))
    flag |= np.where(rng.random(shape) < 0.5, np.uint32(cases.DUMP), np.uint32(0))
    st = cases.State(flag, (1 + rng.uniform(-0.01, 0.01, shape)).astype(np.float32),
                     rng.uniform(-0.05, 0.05, shape).astype(np.float32),
                     rng.uniform(-0.05, 0.05, shape).astype(np.float32),
                     rng.uniform(-0.05, 0.05, shape).astype(np.float32))
    sim = po.Sim(st, 0.7, np.float64, nthreads=1)
    f0 = rng.normal(0, 0.01, (27,) + shape)
    sim.src[...] = f0
    sim.dst[...] = rng.normal(0, 0.01, (27,) + shape)  # stale values for cells without STORE
    want = brute_step(flag, st.rho, st.vx, st.vy, st.vz, 0.7, sim.src, sim.dst)
    sim.step(1)
    assert np.array_equal(sim.src, want)


def test_dump_writes_only_flagged_fields():
    shape = (3, 4, 5)
    rng = np.random.default_rng(7)
    flag = np.full(shape, cases.FLUID, np.uint32)
    flag[0] &= ~np.uint32(cases.DUMP_VX)
    st = cases.State(flag, np.full(shape, 2.0, np.float32), np.full(shape, 3.0, np.float32),
                     np.zeros(shape, np.float32), np.zeros(shape, np.float32))
    sim = po.Sim(st, 0.5, np.float32, nthreads=1)
    sim.src[...] = rng.normal(0, 0.01, sim.src.shape).astype(np.float32)
    rho, vx, vy, vz = sim.dump()
    assert np.all(vx[0] == 3.0) and np.all(vx[1:] != 3.0)
    assert np.all(rho != 2.0)


def test_uniform_equilibrium_is_steady_with_face_clamp():
    """An all-FLUID box at a uniform equilibrium stays there: the face clamp
    pulls a face cell's own (identical) populations, so nothing enters or leaves."""
    n = 7
    st = cases.channel(n, n, n)
    st.flag[...] = cases.FLUID
    st.rho[...] = 1.004
    st.vx[...] = 0.03
    st.vy[...] = -0.01
    sim = po.Sim(st, 0.8, np.float64, nthreads=1)
    sim.init_feq()
    f0 = sim.src.copy()
    sim.step(15)
    assert np.abs(sim.src - f0).max() < 1e-15


def test_channel_flow_develops_from_inlet():
    """The inlet (EQU, u_in) drives flow into a resting channel."""
    st = cases.channel(16, 10, 10, u_in=0.05)
    sim = po.Sim(st, 0.8, np.float64, nthreads=1)
    sim.step(30)
    rho, vx, vy, vz = sim.dump()
    fluid = st.flag == cases.FLUID
    assert vx[fluid].max() > 0.01
    assert np.abs(vy[fluid]).max() < vx[fluid].max()


# ---- 16-bit codecs -----------------------------------------------------------------

def test_fp16_codec_matches_numpy_round_to_nearest_even():
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.normal(0, 1, 200000) * 10.0 ** rng.integers(-9, 5, 200000),
                        (np.arange(1, 4000) + 0.5) * 2.0 ** -24]).astype(np.float32)
    enc = np.zeros(x.size, np.uint16)
    po.lib().orc_encode16(1, x, enc, x.size)
    with np.errstate(over="ignore"):
        want = (x + np.float32(0.0)).astype(np.float16).view(np.uint16)
    assert np.array_equal(enc, want)
    # every finite binary16 value decodes exactly
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    finite = np.isfinite(codes.view(np.float16))
    dec = np.zeros(codes.size, np.float32)
    po.lib().orc_decode16(1, codes, dec, codes.size)
    assert np.array_equal(dec[finite], codes.view(np.float16)[finite].astype(np.float32))


def test_negative_zero_is_stored_as_positive_zero_in_fp16():
    x = np.array([-0.0], np.float32)
    enc = np.zeros(1, np.uint16)
    po.lib().orc_encode16(1, x, enc, 1)
    assert enc[0] == 0


def test_bf16_codec_matches_torch_round_to_nearest_even():
    """BF16 encode against torch's float32 -> bfloat16 conversion (round to
    nearest, ties to even), an implementation independent of the oracle's.  The
    inputs include exact ties (the 16 dropped bits = 0x8000) with both an even
    and an odd kept bit, values one ulp either side of a tie, and the carry
    into the exponent."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    x = (rng.normal(0, 1, 100000) * 10.0 ** rng.integers(-20, 20, 100000)).astype(np.float32)
    hi = rng.integers(0, 1 << 16, 20000).astype(np.uint32) << 16
    hi &= np.uint32(0xBF7FFFFF)  # exponent bits 30 and 23 cleared: finite values
    ties = np.concatenate([hi | 0x8000, hi | 0x7FFF, hi | 0x8001, np.array([0x3F7FFFFF, 0x3F7F8000, 0x3F808000,
                                                                              0x3F818000, 0x7F7F7FFF], np.uint32)])
    x = np.concatenate([x, ties.view(np.float32)])
    x = x[np.isfinite(x)]
    enc = np.zeros(x.size, np.uint16)
    po.lib().orc_encode16(2, x, enc, x.size)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(enc, want), np.flatnonzero(enc != want)[:5]
    # ties really occurred, on both sides
    tie = (x.view(np.uint32) & 0xFFFF) == 0x8000
    assert tie.sum() > 1000 and len(set((x.view(np.uint32)[tie] >> 16) & 1)) == 2
    # decode is exact: the code's 16 bits are the upper half of the FP32 value
    dec = np.zeros(x.size, np.float32)
    po.lib().orc_decode16(2, enc, dec, x.size)
    assert np.array_equal(dec, torch.from_numpy(enc.view(np.int16)).view(torch.bfloat16).float().numpy())


# ---- stochastic rounding of the 16-bit formats ---------------------------------

def _sr_encode(fmt, x, cells, state):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.size, np.uint16)
    po.lib().orc_encode16_sr(1 if fmt == "fp16" else 2, x.reshape(-1), out, x.size, cells, state)
    return out


def _decode(fmt, h):
    out = np.zeros(h.size, np.float32)
    po.lib().orc_decode16(1 if fmt == "fp16" else 2, h, out, h.size)
    return out


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_stochastic_rounding_picks_a_neighbour_and_is_unbiased(fmt):
    rng = np.random.default_rng(5)
    cells = 4096
    x = rng.uniform(-40.0, 40.0, cells).astype(np.float32)
    near = np.zeros(cells, np.uint16)
    po.lib().orc_encode16(1 if fmt == "fp16" else 2, x, near, cells)
    acc = np.zeros(cells)
    states = 400
    # the format's quantum at x: 2^(floor(log2|x|) - mantissa bits)
    quantum = 2.0 ** (np.floor(np.log2(np.abs(x.astype(np.float64)))) - (10 if fmt == "fp16" else 7))
    for state in range(1, states + 1):
        v = _decode(fmt, _sr_encode(fmt, x, cells, state)).astype(np.float64)
        # always one of the two representable neighbours of x
        assert np.all(np.abs(v - x) < quantum)
        acc += v
    # the mean over independent draws approaches x (unbiased to the 1/256 resolution),
    # which round to nearest, always on the same side, does not
    err = np.abs(acc / states - x)
    rn_err = np.abs(_decode(fmt, near).astype(np.float64) - x)
    assert np.median(err / quantum) < 0.05
    assert err.mean() < 0.25 * rn_err.mean()


def test_stochastic_rounding_is_a_function_of_cell_direction_and_state():
    x = np.linspace(-3.0, 3.0, 27 * 64).astype(np.float32)
    a = _sr_encode("fp16", x, 64, 7)
    b = _sr_encode("fp16", x, 64, 7)
    c = _sr_encode("fp16", x, 64, 8)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
