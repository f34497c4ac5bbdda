"""Generators and the on-disk state format (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

from culbm_amd import cases


def reference_channel_flags(nx, ny, nz):
    """The assignment sequence of the reference's leftInletRightOutlet
    generator (scripts/state_initialization.ipynb), restated."""
    f = np.zeros((nz, ny, nx), np.uint32)
    f[1:-1, 1:-1, 1:-1] = cases.FLUID
    f[:, 0, :] = cases.WALL
    f[:, -1, :] = cases.WALL
    f[0, 1:-1, :] = cases.WALL
    f[-1, 1:-1, :] = cases.WALL
    f[1:-1, 1:-1, 0] = cases.EQUILIBRIUM
    f[1:-1, 1:-1, -1] = cases.EQUILIBRIUM
    return f


def test_flag_bits_match_reference_values():
    assert (cases.LOAD, cases.EQU, cases.COLLIDE, cases.BOUNCE_BACK, cases.STORE) == (1, 2, 4, 8, 16)
    assert (cases.DUMP_RHO, cases.DUMP_VX, cases.DUMP_VY, cases.DUMP_VZ) == (1 << 27, 1 << 28, 1 << 29, 1 << 30)


def test_channel_matches_reference_convention():
    st = cases.channel(12, 9, 7, u_in=0.07)
    assert np.array_equal(st.flag, reference_channel_flags(12, 9, 7))
    assert np.all(st.vx[1:-1, 1:-1, 0] == np.float32(0.07))
    assert np.all(st.vx[:, :, 1:] == 0) and np.all(st.rho == 1)


def test_zero_gradient_outlet_is_fluid_face():
    st = cases.channel(12, 9, 7, outlet="zero_gradient")
    assert np.all(st.flag[1:-1, 1:-1, -1] == cases.FLUID)
    assert np.all(st.flag[1:-1, 1:-1, 0] == cases.EQUILIBRIUM)


def test_spheres_are_walls_and_keep_inlet():
    st = cases.channel(40, 30, 30)
    c = cases.add_spheres(st, 5, 3.0, 6.0, seed=3)
    z, y, x = np.indices(st.shape)
    for cx, cy, cz, r in c:
        inside = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r) & (x > 0) & (x < 39)
        assert np.all(st.flag[inside] == cases.WALL)
    assert np.all(st.flag[1:-1, 1:-1, 0] == cases.EQUILIBRIUM)


def test_inlet_profile_range_and_determinism():
    p1, _ = cases.inlet_profile(32, 24, seed=5)
    p2, _ = cases.inlet_profile(32, 24, seed=5)
    assert np.array_equal(p1, p2)
    assert p1.shape == (24, 32)
    assert np.isclose(np.abs(p1 - 0.05).max(), 0.01, rtol=1e-5)


def test_config_small_density_field():
    st = cases.config_small(seed=1, n=16)
    rng = np.random.default_rng(1)
    want = (1.0 + rng.uniform(-1e-3, 1e-3, (16, 16, 16))).astype(np.float32)
    assert np.array_equal(st.rho, want)


def test_state_roundtrip_and_defaults(tmp_path):
    st = cases.config_obstacles(20, 16, 12, count=3, r_min=2, r_max=4, seed=1)
    st.vy[...] = np.random.default_rng(0).normal(size=st.shape).astype(np.float32)
    cases.write_state(tmp_path / "a", st)
    back = cases.read_state(tmp_path / "a", st.domain)
    for k in ("flag", "rho", "vx", "vy", "vz"):
        assert np.array_equal(getattr(back, k), getattr(st, k))
    # raw little-endian x-fastest layout, exactly nx*ny*nz values
    raw = np.fromfile(tmp_path / "a" / "flag.dat", dtype="<u4")
    assert raw.size == 20 * 16 * 12 and np.array_equal(raw.reshape(12, 16, 20), st.flag)
    (tmp_path / "a" / "vz.dat").unlink()
    assert np.all(cases.read_state(tmp_path / "a", st.domain).vz == 0)
    (tmp_path / "a" / "flag.dat").unlink()
    with pytest.raises(FileNotFoundError, match="flag.dat doesn't exist"):
        cases.read_state(tmp_path / "a", st.domain)


def test_correctness_case_follows_the_reference_generator_order():
    """cases.config_correctness against a per-cell restatement of the
    reference generator's assignment order (analysis/correctness/
    generate.py:30-42): FLUID everywhere, then EQU|STORE on both x faces, then
    the WALL cube, then WALL on both y faces, then WALL on both z faces; vx =
    0.05 on x = 0, then 0 on the y and z faces.  Small shape, same rules."""
    nx, ny, nz, (a, b) = 24, 12, 10, (4, 8)
    st = cases.config_correctness(nx, ny, nz, cube=(a, b))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                f = cases.FLUID
                if x in (0, nx - 1):
                    f = cases.EQUILIBRIUM
                if a <= x < b and a <= y < b and a <= z < b:
                    f = cases.WALL
                if y in (0, ny - 1) or z in (0, nz - 1):
                    f = cases.WALL
                v = 0.05 if x == 0 and y not in (0, ny - 1) and z not in (0, nz - 1) else 0.0
                assert st.flag[z, y, x] == f, (x, y, z)
                assert st.vx[z, y, x] == np.float32(v), (x, y, z)
    assert np.all(st.rho == 1.0) and not st.vy.any() and not st.vz.any()


def test_write_state_subset_of_fields(tmp_path):
    st = cases.config_correctness(16, 8, 8, cube=(2, 4))
    cases.write_state(tmp_path, st, fields=("vx",))
    assert sorted(p.name for p in tmp_path.iterdir()) == ["flag.dat", "vx.dat"]
    back = cases.read_state(tmp_path, st.domain)
    assert np.array_equal(back.flag, st.flag) and np.array_equal(back.vx, st.vx)
    assert np.all(back.rho == 1.0)
