"""CPU model of the in-place (esoteric twist) layout's location rules.

Restates, per cell and direction, where the in-place step loads and stores
(culbm_amd/csrc/lbm_kernels.hip: TwistLoad, TwistStore, TwistGhostYZ,
twist_stored_index; layout in culbm_amd/csrc/capi.hip: layout_of) and checks
them against what they must achieve, on tiny domains and pure streaming:

* the populations each cell loads equal the clamped pull of the step before
  (the reference's face clamp, src/simulator/single_dev_expt/kernel.cu:2489-2632),
  step after step and through the layout conversions;
* no location is written twice in one step;
* a location one cell loads and another cell stores in the same step is only
  ever an x-face pair of adjacent lanes in one 64-lane wave (even nx: x = 0 / 1
  and nx-2 / nx-1), whose loads all return before its stores issue.

Both x-face schemes are covered (odd nx: ghost copies; even nx: x faces read in
place) and both row layouts (padded pitch; FP32 rows packed at pitch nx with
the ghost column in its own region).  Host-only: no GPU.
"""
from itertools import product

import numpy as np
import pytest

C = [(cx, cy, cz) for cz in (-1, 0, 1) for cy in (-1, 0, 1) for cx in (-1, 0, 1)]  # i = (cx+1)+3(cy+1)+9(cz+1)
WAVE = 64


def p(c):
    return 1 if c == 1 else 0


class Layout:
    def __init__(self, nx, ny, nz, packed):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.xr = nx % 2 == 0  # x faces read in place, no x ghost copies
        self.packed = packed and self.xr
        self.px = nx if self.packed else nx + 1
        self.py, self.pz = ny + 1, nz + 1
        self.gcol = self.px * self.py * self.pz if self.packed else None

    def idx(self, x, y, z):
        assert 0 <= x < self.px and 0 <= y < self.py and 0 <= z < self.pz, (x, y, z)
        return x + self.px * (y + self.py * z)

    def gidx(self, y, z):
        assert 0 <= y < self.py and 0 <= z < self.pz
        return self.gcol + y + self.py * z

    def size(self):
        return self.px * self.py * self.pz + (self.py * self.pz if self.packed else 0)


def load_loc(L, P, i, x, y, z):
    """(slot, index) a step of parity P loads direction i of cell (x, y, z) from (TwistLoad)."""
    cx, cy, cz = C[i]
    slot = i if P == 1 else 26 - i
    xo = 1 if cx == -1 else 0
    if L.xr and cx == 1:
        xo = 1 if x == 0 else 0
    if L.xr and cx == -1:
        xo = 0 if x == L.nx - 1 else 1
    return slot, L.idx(x + xo, y + p(-cy), z + p(-cz))


def store_locs(L, P, i, x, y, z):
    """(slot, index) locations a step of parity P stores direction i of cell (x, y, z)
    to: the regular store, the x ghost copy and the y/z ghost copies (TwistStore,
    TwistGhostYZ)."""
    cx, cy, cz = C[i]
    slot = 26 - i if P == 1 else i
    nx, ny, nz = L.nx, L.ny, L.nz
    out = []
    if L.packed and cx == 1 and x == nx - 1:
        out.append((slot, L.gidx(y + p(cy), z + p(cz))))
    else:
        out.append((slot, L.idx(x + p(cx), y + p(cy), z + p(cz))))
    fx = (x == 0, x == nx - 1)
    if not L.xr and cx != 0 and (fx[0] if cx == 1 else fx[1]):
        out.append((slot, L.idx(x + (1 if cx == -1 else 0), y + p(cy), z + p(cz))))
    fy, fz = (y == 0, y == ny - 1), (z == 0, z == nz - 1)
    if (fy[0] or fy[1] or fz[0] or fz[1]) and (cy != 0 or cz != 0):
        g = lambda c, f: (f[0] if c == 1 else f[1] if c == -1 else False)
        gm = int(g(cx, fx)) | (int(g(cy, fy)) << 1) | (int(g(cz, fz)) << 2)
        out_x = cx == 1 and x == nx - 1
        for m in range(2, 8):
            if (m & ~gm) != 0 or (L.xr and (m & 1)) or (out_x and not (m & 1)):
                continue
            xx = x + ((1 if cx == -1 else 0) if m & 1 else p(cx))
            yy = y + ((1 if cy == -1 else 0) if m & 2 else p(cy))
            zz = z + ((1 if cz == -1 else 0) if m & 4 else p(cz))
            assert not (L.packed and xx == nx), "a y/z ghost copy would reach the next row"
            out.append((slot, L.idx(xx, yy, zz)))
    return out


def stored_loc(L, P, i, x, y, z):
    """Where a step of parity P left the post-collision value i of cell (x, y, z)
    (twist_stored_index: the regular store slot, or the ghost column)."""
    return store_locs(L, P, i, x, y, z)[0]


def cells(L):
    return product(range(L.nz), range(L.ny), range(L.nx))


def step(L, buf, P, z0=0):
    """One in-place step of parity P over the whole domain (pure streaming):
    every load first, then every store, as a wave does.  Checks the step's
    hazards; returns the loaded (pre-collision) populations [27][nz][ny][nx]."""
    nx, ny = L.nx, L.ny
    lane = lambda x, y, z: x + nx * (y + ny * (z - z0))  # flattened index in the launch (locate)
    f = np.zeros((27, L.nz, L.ny, L.nx), np.float64)
    readers = {}
    for z, y, x in cells(L):
        for i in range(27):
            loc = load_loc(L, P, i, x, y, z)
            f[i, z, y, x] = buf[loc]
            readers.setdefault(loc, set()).add((x, y, z))
    writers = {}
    for z, y, x in cells(L):
        for i in range(27):
            for loc in store_locs(L, P, i, x, y, z):
                w = writers.setdefault(loc, set())
                w.add((x, y, z))
                assert len(w) == 1, f"location {loc} written by {w} in one step"
                buf[loc] = f[i, z, y, x]
    for loc, ws in writers.items():
        (b,) = ws
        for a in readers.get(loc, ()):
            if a == b:
                continue
            same_wave = lane(*a) // WAVE == lane(*b) // WAVE
            x_pair = a[1:] == b[1:] and {a[0], b[0]} in ({0, 1}, {nx - 2, nx - 1})
            assert same_wave and x_pair and L.xr, f"cell {b} overwrites {loc}, which cell {a} loads"
    return f


def clamped_pull(f):
    """One pull step of pure streaming with the face clamp."""
    _, nz, ny, nx = f.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.empty_like(f)
    for i, (cx, cy, cz) in enumerate(C):
        sx, sy, sz = x - cx, y - cy, z - cz
        sx = np.where((sx < 0) | (sx >= nx), x, sx)
        sy = np.where((sy < 0) | (sy >= ny), y, sy)
        sz = np.where((sz < 0) | (sz >= nz), z, sz)
        out[i] = f[i][sz, sy, sx]
    return out


def upload(L, plain):
    """plain -> layout as left by an E step (k_plain_to_twist<S, 0, XG = odd nx>)."""
    buf = np.full((27, L.size()), np.nan)
    for z, y, x in cells(L):
        for i in range(27):
            for loc in store_locs(L, 0, i, x, y, z):
                buf[loc] = plain[i, z, y, x]
    return buf


def download(L, buf, P):
    out = np.empty((27, L.nz, L.ny, L.nx))
    for z, y, x in cells(L):
        for i in range(27):
            out[i, z, y, x] = buf[stored_loc(L, P, i, x, y, z)]
    return out


DOMAINS = [(1, 1, 1), (2, 1, 1), (3, 2, 2), (4, 3, 2), (5, 3, 3), (6, 2, 3), (7, 1, 2), (14, 3, 2), (66, 2, 2),
           (130, 1, 2)]


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("dom", DOMAINS)
def test_twisted_steps_equal_the_clamped_pull(dom, packed):
    nx, ny, nz = dom
    L = Layout(nx, ny, nz, packed)
    n = nx * ny * nz
    tags = (np.arange(27)[:, None] * n + np.arange(n)[None, :]).astype(np.float64).reshape(27, nz, ny, nx)
    buf = upload(L, tags)
    assert np.array_equal(download(L, buf, 0), tags)  # set_ddf -> get_ddf
    want = tags
    P = 1  # the state was written with the E rules: the next step is an O step
    for s in range(4):
        want = clamped_pull(want)
        got = step(L, buf, P, z0=s % 2)
        assert np.array_equal(got, want), f"{dom} packed={packed} step {s + 1}"
        assert np.array_equal(download(L, buf, P), want), f"{dom} packed={packed} get_ddf after step {s + 1}"
        P ^= 1


def _tags(nx, ny, nz):
    n = nx * ny * nz
    return np.arange(27 * n, dtype=np.float64).reshape(27, nz, ny, nx)


def test_hazard_check_flags_an_x_face_pair_across_waves():
    """Planted mistake: the even-nx rule (x faces read in place) on odd nx.  At
    nx = 65 the cells x = 63 and 64 of the first row sit on lanes 63 and 64,
    two waves, and cell 63 overwrites the slot cell 64 reads."""
    L = Layout(65, 2, 1, packed=False)
    L.xr = True
    with pytest.raises(AssertionError, match="overwrites"):
        step(L, upload(L, _tags(65, 2, 1)), 1)


def test_hazard_check_flags_packed_rows_without_the_ghost_column():
    """Planted mistake: packed rows whose last cell still stores its c_x = +1
    values at x = nx, i.e. into the next row."""
    L = Layout(4, 2, 2, packed=True)
    L.gidx = lambda y, z: L.idx(L.nx, y, z)  # no separate region
    with pytest.raises(AssertionError):
        step(L, upload(L, _tags(4, 2, 2)), 1)
