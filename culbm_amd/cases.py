"""Initial-state generators and the on-disk state format.

Format (the reference's `--initStateFolder`): `flag.dat` (u32), `rho.dat`,
`vx.dat`, `vy.dat`, `vz.dat` (f32), each nx*ny*nz values, little-endian,
x fastest (simulator_expt_platform.cu:95-118, 200-228).  Dumps are
`<folder>/<field>_<step>.dat`, same layout (simulator_expt_platform.cu:572-606).

Boundary convention of the reference generators
(scripts/state_initialization.ipynb, leftInletRightOutlet; and
analysis/correctness/generate.py:19-42):
  interior FLUID = LOAD|COLLIDE|STORE|DUMP*
  y = 0, ny-1 planes and z = 0, nz-1 planes (y interior) = LOAD|BOUNCE_BACK|STORE
  x = 0 and x = nx-1 (y, z interior)                    = EQU|STORE
  vx = u_in on the x = 0 face; obstacles = LOAD|BOUNCE_BACK|STORE
The outlet face of that convention is an equilibrium (rho 1, u 0) face.  The
"zero_gradient" outlet makes the x = nx-1 face FLUID instead: its pulls from
x = nx are clamped onto the face itself, so every population entering through
the outlet copies the value the x = nx-2 cell receives in the same step
(f_i(nx-1) = f_i(nx-2) for c_ix = -1), a zero-gradient extrapolation.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

LOAD = 1 << 0
EQU = 1 << 1
COLLIDE = 1 << 2
BOUNCE_BACK = 1 << 3
STORE = 1 << 4
DUMP_RHO = 1 << 27
DUMP_VX = 1 << 28
DUMP_VY = 1 << 29
DUMP_VZ = 1 << 30
DUMP = DUMP_RHO | DUMP_VX | DUMP_VY | DUMP_VZ

FLUID = LOAD | COLLIDE | STORE | DUMP
WALL = LOAD | BOUNCE_BACK | STORE
EQUILIBRIUM = EQU | STORE

FIELDS = ("rho", "vx", "vy", "vz")


@dataclass
class State:
    """Flag and macroscopic fields, each shaped (nz, ny, nx)."""

    flag: np.ndarray
    rho: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray

    @property
    def shape(self) -> Tuple[int, int, int]:
        return self.flag.shape

    @property
    def domain(self) -> Tuple[int, int, int]:
        nz, ny, nx = self.flag.shape
        return (nx, ny, nz)


def channel(nx: int, ny: int, nz: int, u_in: float = 0.05, outlet: str = "equilibrium") -> State:
    """Left-inlet / right-outlet channel in the reference convention."""
    if outlet not in ("equilibrium", "zero_gradient"):
        raise ValueError("outlet must be 'equilibrium' or 'zero_gradient'")
    shape = (nz, ny, nx)
    flag = np.zeros(shape, dtype=np.uint32)
    flag[1:-1, 1:-1, 1:-1] = FLUID
    flag[:, 0, :] = WALL
    flag[:, -1, :] = WALL
    flag[0, 1:-1, :] = WALL
    flag[-1, 1:-1, :] = WALL
    flag[1:-1, 1:-1, 0] = EQUILIBRIUM
    flag[1:-1, 1:-1, -1] = EQUILIBRIUM if outlet == "equilibrium" else FLUID
    rho = np.ones(shape, dtype=np.float32)
    vx = np.zeros(shape, dtype=np.float32)
    vx[1:-1, 1:-1, 0] = u_in
    vy = np.zeros(shape, dtype=np.float32)
    vz = np.zeros(shape, dtype=np.float32)
    return State(flag, rho, vx, vy, vz)


def add_spheres(state: State, count: int, r_min: float, r_max: float, seed: int,
                rng: Optional[np.random.Generator] = None) -> np.ndarray:
    """Solid spheres (LOAD|BOUNCE_BACK|STORE), centres uniform in the interior,
    radius uniform in [r_min, r_max], overlaps allowed.  Returns (count, 4)
    rows (cx, cy, cz, r) in cell units.  The inlet/outlet faces keep their
    flags (spheres never cover x = 0 or x = nx-1)."""
    nz, ny, nx = state.shape
    rng = rng if rng is not None else np.random.default_rng(seed)
    c = np.empty((count, 4), dtype=np.float64)
    c[:, 0] = rng.uniform(1, nx - 1, count)
    c[:, 1] = rng.uniform(1, ny - 1, count)
    c[:, 2] = rng.uniform(1, nz - 1, count)
    c[:, 3] = rng.uniform(r_min, r_max, count)
    z, y, x = np.ogrid[:nz, :ny, :nx]
    solid = np.zeros((nz, ny, nx), dtype=bool)
    for cx, cy, cz, r in c:
        z0, z1 = max(0, int(cz - r) - 1), min(nz, int(cz + r) + 2)
        y0, y1 = max(0, int(cy - r) - 1), min(ny, int(cy + r) + 2)
        x0, x1 = max(1, int(cx - r) - 1), min(nx - 1, int(cx + r) + 2)
        if x0 >= x1:
            continue
        d2 = (x[:, :, x0:x1] - cx) ** 2 + (y[:, y0:y1, :] - cy) ** 2 + (z[z0:z1] - cz) ** 2
        solid[z0:z1, y0:y1, x0:x1] |= d2 <= r * r
    state.flag[solid] = WALL
    state.vx[solid] = 0.0
    state.vy[solid] = 0.0
    state.vz[solid] = 0.0
    return c


def inlet_profile(ny: int, nz: int, seed: int, base: float = 0.05, amp: float = 0.01,
                  modes: int = 8) -> Tuple[np.ndarray, np.ndarray]:
    """u_x(y, z) = base + amp * s(y, z), s a sum of `modes` random-phase
    sinusoids sin(2 pi (a_k y/ny + b_k z/nz) + phi_k) normalised to [-1, 1].
    Returns (profile (nz, ny) float32, parameters (modes, 3) = a, b, phi)."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 4, modes).astype(np.float64)
    b = rng.integers(1, 4, modes).astype(np.float64)
    phi = rng.uniform(0, 2 * np.pi, modes)
    y = np.arange(ny)[None, :] / ny
    z = np.arange(nz)[:, None] / nz
    s = np.zeros((nz, ny))
    for k in range(modes):
        s += np.sin(2 * np.pi * (a[k] * y + b[k] * z) + phi[k])
    s = s / max(np.abs(s).max(), 1e-30)
    return (base + amp * s).astype(np.float32), np.stack([a, b, phi], axis=1)


# ---- the measurement configurations ----------------------------------------

def config_small(seed: int = 1, n: int = 64) -> State:
    """64^3 channel, u_in 0.05, zero-gradient outlet, rho = 1 + U(-1e-3, 1e-3)
    from default_rng(seed) over (z, y, x), u = 0 (populations start at
    feq(rho, u), init mode "feq")."""
    st = channel(n, n, n, u_in=0.05, outlet="zero_gradient")
    rng = np.random.default_rng(seed)
    st.rho[...] = (1.0 + rng.uniform(-1e-3, 1e-3, st.shape)).astype(np.float32)
    return st


def config_obstacles(nx: int = 288, ny: int = 272, nz: int = 280, count: int = 60,
                     r_min: float = 6.0, r_max: float = 14.0, seed: int = 2,
                     outlet: str = "equilibrium", u_in: float = 0.05) -> State:
    """Channel with `count` seeded solid spheres (the optimised FP32 workload:
    288x272x280, 60 spheres of radius 6-14)."""
    st = channel(nx, ny, nz, u_in=u_in, outlet=outlet)
    add_spheres(st, count, r_min, r_max, seed)
    return st


def config_precision(n: int = 512, seed: int = 3) -> State:
    """Precision workload generator: the small configuration's generator at
    n^3 (zero-gradient outlet, u_in 0.05, rho = 1 + U(-1e-3, 1e-3))."""
    return config_small(seed=seed, n=n)


def config_adjoint(n: int = 128, seed: int = 4, count: int = 20, r_min: float = 4.0,
                   r_max: float = 8.0) -> Tuple[State, np.ndarray]:
    """Differentiation workload: n^3 channel, seeded spheres, smooth inlet
    profile on x = 0.  Returns (state, inlet profile (nz, ny))."""
    st = channel(n, n, n, u_in=0.05, outlet="zero_gradient")
    add_spheres(st, count, r_min, r_max, seed)
    prof, _ = inlet_profile(n, n, seed)
    inlet = st.flag[:, :, 0] == EQUILIBRIUM
    st.vx[:, :, 0] = np.where(inlet, prof, 0.0)
    return st, prof


def config_correctness(nx: int = 768, ny: int = 384, nz: int = 384, cube: Tuple[int, int] = (128, 256),
                       u_in: float = 0.05) -> State:
    """The reference's own bitwise correctness case (analysis/correctness/
    generate.py:6-42): the channel convention (equilibrium outlet) at
    768x384x384 with a solid cube over [128, 256) on every axis.  Its case
    folder holds flag.dat and vx.dat only (rho, vy, vz take the defaults)."""
    st = channel(nx, ny, nz, u_in=u_in, outlet="equilibrium")
    a, b = cube
    st.flag[a:b, a:b, a:b] = WALL
    st.vx[a:b, a:b, a:b] = 0.0
    return st


# ---- files -----------------------------------------------------------------

def write_state(folder, state: State, fields=FIELDS) -> None:
    """flag.dat plus the named fields (a field left out takes its default when
    read back, as a missing file does for the reference)."""
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    state.flag.astype("<u4").tofile(folder / "flag.dat")
    for name in fields:
        getattr(state, name).astype("<f4").tofile(folder / f"{name}.dat")


def read_state(folder, domain: Tuple[int, int, int]) -> State:
    """Read an init folder; a missing field falls back to its default
    (rho 1, u 0), a missing flag.dat raises, as initFlag/initRhoU do."""
    folder = Path(folder)
    nx, ny, nz = domain
    shape = (nz, ny, nx)
    fpath = folder / "flag.dat"
    if not fpath.exists():
        raise FileNotFoundError(f"File: {folder}/flag.dat doesn't exist!")
    flag = np.fromfile(fpath, dtype="<u4", count=nx * ny * nz).reshape(shape)
    out = {}
    for name, dflt in zip(FIELDS, (1.0, 0.0, 0.0, 0.0)):
        p = folder / f"{name}.dat"
        if p.exists():
            out[name] = np.fromfile(p, dtype="<f4", count=nx * ny * nz).reshape(shape)
        else:
            out[name] = np.full(shape, dflt, dtype=np.float32)
    return State(flag.astype(np.uint32), **{k: v.astype(np.float32) for k, v in out.items()})


def read_dump(folder, field: str, step: int, domain: Tuple[int, int, int]) -> np.ndarray:
    nx, ny, nz = domain
    return np.fromfile(Path(folder) / f"{field}_{step}.dat", dtype="<f4").reshape(nz, ny, nx)
