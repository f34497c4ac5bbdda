"""GPU parity tests: the HIP path through the C ABI against the oracle.

Bars:
  * FP32 plain pull == FP32 CPU restatement, bit for bit (same operation order);
  * FP32 rho/u vs the FP64 oracle: max |drho| <= 1e-5 (relative to rho ~ 1),
    max |du| <= 1e-5 * u_in;
  * every in-place / blocked variant == plain pull of the same storage, bit for bit;
  * 16-bit storage == CPU 16-bit restatement, bit for bit;
  * per-cell physics == the reference's own functions (golden vectors), bit for bit.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from culbm_amd import Lattice, LatticeConfig, cases
from culbm_amd import lattice as L
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint16)


def assert_bitexact(a, b, what=""):
    assert a.shape == b.shape, what
    ba, bb = bits(a), bits(b)
    bad = np.flatnonzero(ba.reshape(-1) != bb.reshape(-1))
    assert bad.size == 0, f"{what}: {bad.size} values differ, first at flat index {bad[:5]}"


def make(state, init="rest", **cfg):
    lat = Lattice(LatticeConfig(domain=state.domain, **cfg))
    lat.set_state(state.flag, state.rho, state.vx, state.vy, state.vz)
    lat.init_ddf(init)
    return lat


# ---- per-cell physics vs the reference's functions --------------------------

def test_cell_physics_matches_reference_golden(golden_dir):
    g = np.load(golden_dir / "cell_physics.npz")
    for k, it in enumerate(g["inv_taus"]):
        out = L.probe_collide(g["f_in"], float(it))
        assert_bitexact(out, g["collide_out"][k], f"collision2<27> invTau={it}")
    assert_bitexact(L.probe_equilibrium(g["q_in"]), g["equ_out"], "calcEqu<27>")
    assert_bitexact(L.probe_moments(g["f_in"]), g["moments_out"], "calcRhoU<27>")


def test_codec_matches_host_restatement():
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.normal(0, 1e-2, 20000), rng.normal(0, 1e-6, 5000), rng.uniform(-20, 20, 5000),
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 1e-8, 6.0e-8, 2.0 ** -24 * 1.5,
                  2.0 ** -14, 2.0 ** -14 * (1 - 2.0 ** -11)])
    ]).astype(np.float32)
    for storage, fmt, scale in (("fp16", 1, 1.0), ("fp16", 1, 4096.0), ("bf16", 2, 1.0), ("bf16", 2, 256.0)):
        enc, dec = L.probe_codec(x, storage, scale)
        want = np.zeros(x.shape, np.uint16)
        scaled = x * np.float32(scale)
        pyoracle.lib().orc_encode16(fmt, scaled, want, x.size)
        assert np.array_equal(enc, want), f"{storage} encode (scale {scale})"
        back = np.zeros(x.shape, np.float32)
        pyoracle.lib().orc_decode16(fmt, want, back, x.size)
        assert_bitexact(dec, back * np.float32(1.0 / scale), f"{storage} decode")


# ---- plain pull vs the oracle -------------------------------------------------

def test_plain_pull_small_config_vs_oracle():
    """64^3 channel, zero-gradient outlet, rho = 1 + U(-1e-3, 1e-3), 100 steps."""
    st = cases.config_small(seed=1)
    inv_tau = 0.5
    lat = make(st, init="feq", inv_tau=inv_tau)
    lat.step(100)
    lat.dump()
    g_rho, g_vx, g_vy, g_vz = lat.macros()
    g_ddf = lat.get_ddf()

    o32 = pyoracle.Sim(st, inv_tau, np.float32)
    o32.init_feq()
    o32.step(100)
    r32 = o32.dump()
    assert_bitexact(g_ddf, o32.ddf(), "FP32 populations after 100 steps")
    for a, b, name in zip((g_rho, g_vx, g_vy, g_vz), r32, ("rho", "vx", "vy", "vz")):
        assert_bitexact(a, b, f"FP32 {name}")

    o64 = pyoracle.Sim(st, inv_tau, np.float64)
    o64.init_feq()
    o64.step(100)
    rho64, vx64, vy64, vz64 = o64.dump()
    fluid = (st.flag & cases.DUMP) != 0
    drho = np.abs(g_rho.astype(np.float64) - rho64)[fluid].max()
    du = max(np.abs(a.astype(np.float64) - b)[fluid].max() for a, b in ((g_vx, vx64), (g_vy, vy64), (g_vz, vz64)))
    assert drho <= 1e-5, drho
    assert du <= 1e-5 * 0.05, du
    # the flow really developed (not a trivially resting state)
    assert np.abs(vx64[fluid]).max() > 1e-3


@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("storage,fmt", [("fp16", "fp16"), ("bf16", "bf16")])
def test_plain_pull_16bit_vs_oracle(storage, fmt, rounding):
    st = cases.config_small(seed=7, n=40)
    lat = make(st, init="feq", inv_tau=0.6, storage=storage, rounding=rounding)
    scale = lat.info()["storage_scale"]
    lat.step(30)
    lat.dump()
    g = lat.macros()
    o = pyoracle.Sim(st, 0.6, fmt, scale=scale, rounding=rounding)
    o.init_feq()
    o.step(30)
    r = o.dump()
    assert_bitexact(lat.get_ddf(), o.ddf(), f"{storage} {rounding} populations")
    for a, b in zip(g, r):
        assert_bitexact(a, b, f"{storage} {rounding} macros")


# ---- streaming KAT: index tags through pure streaming ------------------------

def stream_expect(tags, steps):
    """Pure streaming with the reference's face clamp, repeated `steps` times."""
    cx, cy, cz, _ = pyoracle.velset()
    _, nz, ny, nx = tags.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cur = tags.copy()
    for _ in range(steps):
        nxt = np.empty_like(cur)
        for i in range(27):
            sx, sy, sz = x - cx[i], y - cy[i], z - cz[i]
            sx = np.where((sx < 0) | (sx >= nx), x, sx)
            sy = np.where((sy < 0) | (sy >= ny), y, sy)
            sz = np.where((sz < 0) | (sz >= nz), z, sz)
            nxt[i] = cur[i][sz, sy, sx]
        cur = nxt
    return cur


POLICIES = [(0, 0), (1, 0), (0, 3), (1, 3), (0, 1), (1, 1), (0, 2)]


@pytest.mark.parametrize("sp,op", POLICIES)
@pytest.mark.parametrize("nx", [13, 14])
def test_stream_index_tags(sp, op, nx):
    """Every population carries its own index; after 1 and 4 steps each cell
    holds exactly the indices the clamped pull brings it.  Odd nx runs the
    in-place step with x ghost stores, even nx the one whose x-face cells read
    their clamped value in place."""
    ny, nz = 9, 11
    st = cases.channel(nx, ny, nz)
    st.flag[...] = cases.LOAD | cases.STORE
    lat = Lattice(LatticeConfig(domain=(nx, ny, nz), stream_policy=sp, opt_policy=op, inner_loop=3,
                                wave_depth=2))
    lat.set_state(st.flag)
    n = nx * ny * nz
    tags = (np.arange(27)[:, None] * n + np.arange(n)[None, :]).astype(np.float32).reshape(27, nz, ny, nx)
    lat.set_ddf(tags)
    for steps in (1, 4):
        lat.set_ddf(tags)
        lat.step(steps)
        got = lat.get_ddf()
        want = stream_expect(tags, steps)
        assert np.array_equal(got, want), f"policy ({sp},{op}) after {steps} steps"


# ---- every variant equals plain pull bit for bit -----------------------------

def obstacle_case(nx=37, ny=29, nz=31, outlet="zero_gradient", seed=11):
    st = cases.channel(nx, ny, nz, u_in=0.06, outlet=outlet)
    cases.add_spheres(st, 6, 2.5, 5.0, seed)
    rng = np.random.default_rng(seed)
    st.rho[...] = (1.0 + rng.uniform(-2e-3, 2e-3, st.shape)).astype(np.float32)
    return st


@pytest.mark.parametrize("storage,rounding", [("fp32", "nearest"), ("fp16", "nearest"), ("bf16", "nearest"),
                                              ("fp16", "stochastic"), ("bf16", "stochastic")])
@pytest.mark.parametrize("outlet", ["zero_gradient", "equilibrium"])
def test_variants_bitexact_vs_plain(storage, rounding, outlet):
    st = obstacle_case(outlet=outlet)
    steps = 17
    ref = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding)
    ref.step(steps)
    ref.dump()
    want_m = ref.macros()
    want_f = ref.get_ddf()
    for sp, op, il, wd in [(1, 0, 5, 0), (0, 3, 5, 3), (1, 3, 5, 3), (1, 3, 1, 2), (1, 3, 4, 0),
                           (0, 1, 3, 4), (1, 1, 5, 2), (1, 1, 2, 0), (0, 2, 5, 0), (0, 2, 2, 3),
                           (0, 2, 4, 1), (0, 2, 3, 7)]:
        lat = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding, stream_policy=sp,
                   opt_policy=op, inner_loop=il, wave_depth=wd)
        lat.step(steps)
        lat.dump()
        what = f"({sp},{op}) I={il} D={wd} {storage} {rounding} {outlet}"
        assert_bitexact(lat.get_ddf(), want_f, what + " populations")
        for a, b in zip(lat.macros(), want_m):
            assert_bitexact(a, b, what + " macros")
        lat.close()


def test_window_semantics():
    st = obstacle_case()
    lat = make(st, stream_policy=1, opt_policy=3, inner_loop=4)
    steps, ms = lat.run_window(10)
    assert steps == 12 and ms > 0
    assert lat.info()["step"] == 12
    plain = make(st)
    steps, _ = plain.run_window(10)
    assert steps == 10


# ---- dump semantics ------------------------------------------------------------

def test_dump_writes_only_flagged_cells():
    st = obstacle_case()
    st.flag[:, :, : st.shape[2] // 2] &= ~np.uint32(cases.DUMP_VY)
    st.vy[...] = 7.0  # sentinel in the EQU/dump buffer
    st.vy[:, :, 0] = 0.0
    lat = make(st, inv_tau=0.6)
    lat.step(5)
    lat.dump()
    rho, vx, vy, vz = lat.macros()
    no_dump = (st.flag & cases.DUMP_VY) == 0
    assert np.all(vy[no_dump] == st.vy[no_dump])
    assert np.all(vy[~no_dump] != 7.0)
    o = pyoracle.Sim(st, 0.6, np.float32)
    o.step(5)
    r = o.dump()
    for a, b in zip((rho, vx, vy, vz), r):
        assert_bitexact(a, b, "dump")


@pytest.mark.parametrize("sp,op", [(0, 0), (1, 0), (1, 3)])
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
@pytest.mark.parametrize("nx", [19, 20])
def test_set_get_ddf_roundtrip(sp, op, storage, nx):
    """set_ddf then get_ddf returns the populations (16-bit: as the codec
    rounds them); even nx with FP32 in place goes through the packed rows'
    ghost-column region for the last cell of every row."""
    st = obstacle_case(nx=nx, ny=17, nz=15)
    lat = make(st, stream_policy=sp, opt_policy=op, storage=storage)
    rng = np.random.default_rng(3)
    f = rng.normal(0, 1e-2, (27,) + st.shape).astype(np.float32)
    lat.set_ddf(f)
    if storage == "fp16":  # the scale set_ddf derived from f and the state
        scale = lat.info()["storage_scale"]
        enc = np.zeros(f.size, np.uint16)
        pyoracle.lib().orc_encode16(1, (f * np.float32(scale)).reshape(-1), enc, f.size)
        dec = np.zeros(f.size, np.float32)
        pyoracle.lib().orc_decode16(1, enc, dec, f.size)
        f = (dec * np.float32(1 / scale)).reshape(f.shape)
    assert_bitexact(lat.get_ddf(), f, "roundtrip")


@pytest.mark.slow
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_fp16_scale_keeps_headroom_over_a_long_run(rounding):
    """The benchmark channel (288x272x280, 60 spheres, u_in 0.05) in FP16 with the
    per-run scale, 6000 steps in place + wavefront from rest: every population
    stays finite and within a quarter of the scale's range (|f - w| * s below
    16376 of 65504), and the velocity field tracks FP32's."""
    st = cases.config_obstacles()
    steps = 6000
    lat = make(st, storage="fp16", rounding=rounding, stream_policy=1, opt_policy=3, inner_loop=5)
    scale = lat.info()["storage_scale"]
    assert scale == 2.0 ** 16
    lat.step(steps)
    f = lat.get_ddf()
    assert np.isfinite(f).all()
    assert np.abs(f).max() * scale < 65504.0 / 4
    lat.dump()
    u16 = lat.macros()[1]
    lat.close()
    ref = make(st, stream_policy=1, opt_policy=3, inner_loop=5)
    ref.step(steps)
    ref.dump()
    u32 = ref.macros()[1]
    ref.close()
    fluid = (st.flag & cases.DUMP) != 0
    assert np.abs(u32[fluid]).max() > 0.04  # the flow has developed
    err = np.abs(u16[fluid].astype(np.float64) - u32[fluid]).max() / 0.05
    print(f"fp16 {rounding} {steps} steps: max |f - w| * s = {np.abs(f).max() * scale:.1f} of 65504, "
          f"max |u16 - u32| / u_in = {err:.2e}", flush=True)
    assert err < 2e-2, err


def test_16bit_scale_is_derived_per_run():
    """FP16's scale is the largest power of two with 16 B s <= 65504, B the
    bound on |f_i - w_i| of the run's state (at least 1/32) or of the
    populations given to set_ddf; BF16 takes 1; an explicit scale is kept."""
    st = obstacle_case(nx=19, ny=17, nz=15)  # rho 1 +- 2e-3, u_in 0.06: B = 0.0604

    def scale_of(state, storage="fp16", **kw):
        lat = make(state, storage=storage, **kw)
        s = lat.info()["storage_scale"]
        lat.close()
        return s

    assert scale_of(st) == 2.0 ** 16  # 65504 / (16 B) = 67.7k
    assert scale_of(st, storage="bf16") == 1.0
    assert scale_of(st, storage_scale=512.0) == 512.0
    fast = cases.State(st.flag, st.rho, st.vx * 8, st.vy, st.vz)  # u_in 0.48: B = 0.84 -> 4.9k
    assert scale_of(fast) == 2.0 ** 12
    lat = make(st, storage="fp16")
    f = np.zeros((27,) + st.shape, np.float32)
    f[3, 2, 2, 2] = 2.0  # 65504 / 32 = 2047
    lat.set_ddf(f)
    assert lat.info()["storage_scale"] == 2.0 ** 10
    assert lat.get_ddf()[3, 2, 2, 2] == 2.0
    lat.close()


# ---- the optimised workload at full size --------------------------------------

FULL_SIZE_FORMS = {
    "fp32": [(1, 3), (0, 3), (0, 2), (1, 1), (1, 0)],
    "fp16": [(1, 3), (0, 3), (0, 2)],
    "bf16": [(1, 3)],
}


@pytest.mark.slow
@pytest.mark.parametrize("storage,rounding,steps", [("fp32", "nearest", 1000), ("fp16", "nearest", 200),
                                                    ("fp16", "stochastic", 200), ("bf16", "nearest", 200)])
def test_full_size_blocked_equals_plain(storage, rounding, steps):
    """288x272x280, 60 spheres, I = 5, automatic depths and workgroup orders (the
    benchmark's configurations): every blocked and in-place form vs plain pull.
    The populations start at feq(rho, u) with rho = 1 + U(-1e-3, 1e-3) from a
    seed, so every cell, sphere and chunk seam carries non-zero, distinct values
    from the first step; FP32 runs the workload's 1000 steps."""
    st = cases.config_obstacles()
    rng = np.random.default_rng(21)
    st.rho[...] = (1.0 + rng.uniform(-1e-3, 1e-3, st.shape)).astype(np.float32)

    def run(sp, op):
        lat = make(st, init="feq", inv_tau=0.5, stream_policy=sp, opt_policy=op, inner_loop=5, storage=storage,
                   rounding=rounding)
        lat.step(steps)
        lat.dump()
        r = (lat.macros(), lat.get_ddf())
        lat.close()
        return r

    want_m, want_f = run(0, 0)
    # the flow is developed everywhere: hardly a fluid cell has u_x exactly 0
    # (16-bit storage can cancel to an exact zero in a few cells of 20M)
    fluid = (st.flag & cases.DUMP) != 0
    assert np.count_nonzero(want_m[1][fluid] == 0.0) <= 1e-5 * np.count_nonzero(fluid)
    for sp, op in FULL_SIZE_FORMS[storage]:
        got_m, got_f = run(sp, op)
        what = f"full-size ({sp},{op}) {storage} {rounding} {steps} steps"
        for a, b in zip(got_m, want_m):
            assert_bitexact(a, b, what + " macros")
        assert_bitexact(got_f, want_f, what + " populations")
        del got_m, got_f


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_precision_workload_512_blocked_equals_plain(storage):
    """512^3 (134M cells, the largest configuration) in each storage format, the
    flow started from feq: in-place + wavefront blocking against plain pull over
    five full passes and a partial one.  Every cell's dumped rho/u must agree bit
    for bit and the FP64 sums of the stored populations to 1e-12 (reading back
    up to 14.5 GB of populations is left out)."""
    st = cases.config_precision(512)
    steps = 27
    out = []
    for sp, op in ((0, 0), (1, 3)):
        lat = make(st, init="feq", inv_tau=0.5, stream_policy=sp, opt_policy=op, inner_loop=5,
                   storage=storage)
        lat.step(steps)
        lat.dump()
        out.append((lat.macros(), lat.moment_sums()))
        lat.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert_bitexact(a, b, f"512^3 {storage} macros")
    for a, b in zip(out[0][1], out[1][1]):
        assert abs(a - b) <= 1e-12 * max(abs(a), 1.0), (out[0][1], out[1][1])


def test_mixed_with_inactive_cells_and_midrun_flag_change():
    """Cells without STORE keep their populations in both buffers; the fused
    two-step kernel reads them from src, plain pull alternates buffers: both
    must agree, also after the flag field changes between runs."""
    st = obstacle_case(nx=70, ny=20, nz=17)
    rng = np.random.default_rng(9)
    st.flag[rng.random(st.shape) < 0.03] = 0
    flag2 = st.flag.copy()
    flag2[(rng.random(st.shape) < 0.02) & (flag2 != 0)] = cases.LOAD | cases.COLLIDE  # no STORE
    res = []
    for op in (0, 2):
        lat = make(st, init="feq", inv_tau=0.8, opt_policy=op, inner_loop=4)
        lat.step(6)
        lat.set_state(flag2, st.rho, st.vx, st.vy, st.vz)
        lat.step(7)
        lat.dump()
        res.append((lat.get_ddf(), lat.macros()))
        lat.close()
    assert_bitexact(res[0][0], res[1][0], "populations")
    for a, b in zip(res[0][1], res[1][1]):
        assert_bitexact(a, b, "macros")


@pytest.mark.parametrize("storage,rounding", [("fp32", "nearest"), ("fp16", "stochastic")])
def test_mixed_pairs_do_not_depend_on_call_boundaries(storage, rounding):
    """The mixed policy takes its steps in pairs across passes and calls (one
    plain step for an odd count): any split of 19 steps into calls, windows
    included, gives the bits of 19 plain pull steps."""
    st = obstacle_case(nx=45, ny=26, nz=23)
    ref = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding)
    ref.step(19)
    want = ref.get_ddf()
    ref.close()
    for split in ([19], [1, 18], [5, 3, 9, 2], [2, 2, 2, 13], [7, 7, 5]):
        lat = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding, opt_policy=2, inner_loop=5)
        for n in split:
            lat.step(n)
        assert lat.info()["step"] == 19
        assert_bitexact(lat.get_ddf(), want, f"{storage} {rounding} split {split}")
        lat.close()
    # windows advance by innerLoop: three windows of 7 steps run 10 each
    lat = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding, opt_policy=2, inner_loop=5)
    done = [lat.run_window(7)[0] for _ in range(3)]
    assert done == [10, 10, 10]
    ref = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding)
    ref.step(30)
    assert_bitexact(lat.get_ddf(), ref.get_ddf(), f"{storage} {rounding} windows")
    lat.close()
    ref.close()


# ---- ragged and tiny domains -------------------------------------------------

RAGGED = [(1, 1, 1), (1, 1, 6), (2, 2, 2), (3, 1, 4), (64, 1, 1), (65, 3, 2), (7, 5, 3), (129, 2, 3),
          (33, 9, 17), (34, 9, 17), (66, 3, 2), (130, 2, 3)]


@pytest.mark.parametrize("dom", RAGGED)
def test_ragged_domains_all_policies_match_the_oracle(dom):
    """Every cell FLUID (so every face clamps), random near-equilibrium start:
    each policy equals the FP32 oracle bit for bit on domains smaller than a
    wave, a tile or a wavefront chunk."""
    nx, ny, nz = dom
    rng = np.random.default_rng(nx * 10007 + ny * 101 + nz)
    flag = np.full((nz, ny, nx), cases.FLUID, np.uint32)
    rho = (1.0 + rng.uniform(-1e-3, 1e-3, (nz, ny, nx))).astype(np.float32)
    u = [rng.uniform(-0.02, 0.02, (nz, ny, nx)).astype(np.float32) for _ in range(3)]
    state = cases.State(flag, rho, *u)
    steps = 7
    o = pyoracle.Sim(state, 0.8, np.float32)
    o.init_feq()
    o.step(steps)
    want = o.ddf()
    for sp, op, il, wd in [(0, 0, 5, 0), (1, 0, 5, 0), (0, 3, 3, 0), (1, 3, 3, 1), (1, 3, 5, 0),
                           (0, 1, 2, 0), (1, 1, 3, 2), (0, 2, 3, 0), (0, 2, 2, 1)]:
        lat = Lattice(LatticeConfig(domain=dom, inv_tau=0.8, stream_policy=sp, opt_policy=op,
                                    inner_loop=il, wave_depth=wd))
        lat.set_state(flag, rho, *u)
        lat.init_ddf("feq")
        lat.step(steps)
        assert_bitexact(lat.get_ddf(), want, f"{dom} ({sp},{op}) I={il} D={wd}")
        lat.close()


# ---- 16-bit storage: every policy against the host restatement ---------------------

@pytest.mark.parametrize("storage,rounding", [("fp16", "nearest"), ("bf16", "nearest"), ("fp16", "stochastic")])
@pytest.mark.parametrize("outlet", ["zero_gradient", "equilibrium"])
def test_16bit_variants_bitexact_vs_oracle(storage, rounding, outlet):
    """Every policy (pull, in place, wavefront, static, mixed) in FP16 / BF16,
    nearest and stochastic, against the 16-bit host restatement, bit for bit."""
    st = obstacle_case(nx=38, ny=29, nz=31, outlet=outlet)
    steps = 17
    fmt = storage
    lat0 = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding)
    scale = lat0.info()["storage_scale"]
    lat0.close()
    o = pyoracle.Sim(st, 0.7, fmt, scale=scale, rounding=rounding)
    o.init_feq()
    o.step(steps)
    want_f = o.ddf()
    want_m = o.dump()
    for sp, op, il, wd in [(0, 0, 5, 0), (1, 0, 5, 0), (0, 3, 5, 3), (1, 3, 5, 3), (1, 3, 1, 2), (1, 3, 4, 0),
                           (0, 1, 3, 4), (1, 1, 5, 2), (0, 2, 5, 0), (0, 2, 3, 7)]:
        lat = make(st, init="feq", inv_tau=0.7, storage=storage, rounding=rounding, stream_policy=sp,
                   opt_policy=op, inner_loop=il, wave_depth=wd)
        lat.step(steps)
        lat.dump()
        what = f"({sp},{op}) I={il} D={wd} {storage} {rounding} {outlet}"
        assert_bitexact(lat.get_ddf(), want_f, what + " populations")
        for a, b in zip(lat.macros(), want_m):
            assert_bitexact(a, b, what + " macros")
        lat.close()


RAGGED16 = [(2, 1, 1), (2, 3, 2), (4, 1, 5), (64, 2, 3), (66, 3, 2), (130, 2, 2), (8, 9, 17)]


@pytest.mark.parametrize("dom", RAGGED16)
def test_16bit_ragged_domains_match_the_oracle(dom):
    """Every cell FLUID (every face clamps), FP16 from a random near-equilibrium
    start, on domains down to nx = 2: each policy equals the FP16 host
    restatement bit for bit."""
    nx, ny, nz = dom
    rng = np.random.default_rng(nx * 10007 + ny * 101 + nz)
    flag = np.full((nz, ny, nx), cases.FLUID, np.uint32)
    rho = (1.0 + rng.uniform(-1e-3, 1e-3, (nz, ny, nx))).astype(np.float32)
    u = [rng.uniform(-0.02, 0.02, (nz, ny, nx)).astype(np.float32) for _ in range(3)]
    state = cases.State(flag, rho, *u)
    steps = 7
    want = None
    for sp, op, il, wd in [(0, 0, 5, 0), (1, 0, 5, 0), (0, 3, 3, 0), (1, 3, 3, 1), (1, 3, 5, 0),
                           (0, 1, 2, 0), (1, 1, 3, 2), (0, 2, 3, 0)]:
        lat = Lattice(LatticeConfig(domain=dom, inv_tau=0.8, stream_policy=sp, opt_policy=op,
                                    inner_loop=il, wave_depth=wd, storage="fp16"))
        lat.set_state(flag, rho, *u)
        lat.init_ddf("feq")
        if want is None:  # the scale derived from this state
            o = pyoracle.Sim(state, 0.8, "fp16", scale=lat.info()["storage_scale"])
            o.init_feq()
            o.step(steps)
            want = o.ddf()
        lat.step(steps)
        assert_bitexact(lat.get_ddf(), want, f"{dom} ({sp},{op}) I={il} D={wd} fp16")
        lat.close()


@pytest.mark.parametrize("sp,op", POLICIES)
@pytest.mark.parametrize("nx", [5, 6])
def test_16bit_stream_index_tags(sp, op, nx):
    """Cell-id tags through pure streaming in FP16 with scale 1 (every tag an
    exact binary16 integer): the 16-bit addressing, face clamp, ghost copies
    (odd nx) and in-place x-face reads (even nx) against the clamped-streaming
    expectation."""
    ny, nz = 3, 4
    st = cases.channel(nx, ny, nz)
    st.flag[...] = cases.LOAD | cases.STORE
    lat = Lattice(LatticeConfig(domain=(nx, ny, nz), stream_policy=sp, opt_policy=op, inner_loop=3,
                                wave_depth=2, storage="fp16", storage_scale=1.0))
    lat.set_state(st.flag)
    n = nx * ny * nz
    tags = (np.arange(27)[:, None] * n + np.arange(n)[None, :]).astype(np.float32).reshape(27, nz, ny, nx)
    assert tags.max() < 2048
    for steps in (1, 4):
        lat.set_ddf(tags)
        lat.step(steps)
        got = lat.get_ddf()
        want = stream_expect(tags, steps)
        assert np.array_equal(got, want), f"policy ({sp},{op}) after {steps} steps"
