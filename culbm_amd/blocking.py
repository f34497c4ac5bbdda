"""Halo (temporal) blocking decomposition: the reference's blocking core.

Axis math restated from src/blocking_core/blocking_alogrithm.hpp:8-41 and the
tile placement of src/simulator/single_dev_expt/simulator_expt_platform.cu:120-198,
523-540, 792-795 (tile origin = max(valid start - (I-1), 0); tile-local flags =
global flags plus CORRECT on the valid box, 0 outside the domain).  The same
axis math is compiled into the library (culbm_blocking_axis); `axis()` calls it
so the Python and native paths cannot drift apart.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import _lib

CORRECT = 1 << 31


def valid_ref(dom: int, blk: int, it: int) -> bool:
    """validBlkAxisConfig of the reference, with its unsigned arithmetic."""
    m = 0xFFFFFFFF
    if blk < dom and (blk + 1) & m <= it:
        return False
    if (2 * ((blk + 1 - it) & m)) & m < dom and (blk + 2) & m <= it:
        return False
    return True


def axis(dom: int, blk: int, it: int) -> List[int]:
    """Valid-range boundaries [vb(0)=0, vb(1), ..., vb(n)=dom] of the tiles of
    one axis (n = len - 1 tiles).  Raises ValueError for an ill-posed axis."""
    num = ctypes.c_int32()
    rc = _lib.lib().culbm_blocking_axis(int(dom), int(blk), int(it), ctypes.byref(num), None, 0)
    if rc != _lib.OK:
        raise ValueError(_lib.lib().culbm_last_error(None).decode())
    starts = np.zeros(num.value + 1, np.int32)
    _lib.check(_lib.lib().culbm_blocking_axis(int(dom), int(blk), int(it), ctypes.byref(num),
                                              _lib.ptr(starts), starts.size))
    return [int(v) for v in starts]


@dataclass
class Tile:
    index: Tuple[int, int, int]        # (kx, ky, kz)
    origin: Tuple[int, int, int]       # first loaded cell (x, y, z)
    load_end: Tuple[int, int, int]     # one past the last loaded cell
    valid: Tuple[Tuple[int, int], Tuple[int, int], Tuple[int, int]]  # [begin, end) per axis


def plan(domain: Tuple[int, int, int], tile: Tuple[int, int, int], it: int) -> List[Tile]:
    """All tiles of a domain in the reference's sweep order (x fastest)."""
    starts = [axis(d, b, it) for d, b in zip(domain, tile)]
    tiles = []
    for kz in range(len(starts[2]) - 1):
        for ky in range(len(starts[1]) - 1):
            for kx in range(len(starts[0]) - 1):
                ks = (kx, ky, kz)
                valid = tuple((starts[a][ks[a]], starts[a][ks[a] + 1]) for a in range(3))
                origin = tuple(max(valid[a][0] - (it - 1), 0) for a in range(3))
                end = tuple(min(valid[a][1] + (it - 1), domain[a]) for a in range(3))
                tiles.append(Tile(ks, origin, end, valid))
    return tiles


def tile_flags(flag: np.ndarray, t: Tile, tile: Tuple[int, int, int]) -> np.ndarray:
    """Tile-local flag block (initBlockingFlag): global flags over the loaded
    box, CORRECT on the valid box, 0 where the tile extends past the domain."""
    bx, by, bz = tile
    out = np.zeros((bz, by, bx), dtype=np.uint32)
    (x0, y0, z0), (x1, y1, z1) = t.origin, t.load_end
    out[: z1 - z0, : y1 - y0, : x1 - x0] = flag[z0:z1, y0:y1, x0:x1]
    (vx0, vx1), (vy0, vy1), (vz0, vz1) = t.valid
    out[vz0 - z0: vz1 - z0, vy0 - y0: vy1 - y0, vx0 - x0: vx1 - x0] |= np.uint32(CORRECT)
    return out


def redundancy(domain: Tuple[int, int, int], tile: Tuple[int, int, int], it: int) -> float:
    """Loaded cells over domain cells: the recomputation factor of a blocking."""
    total = sum(int(np.prod([e - o for o, e in zip(t.origin, t.load_end)]))
                for t in plan(domain, tile, it))
    return total / float(np.prod(domain))


def wave_schedule(nz: int, depth: int, steps: int) -> List[Tuple[int, int, int]]:
    """Launches (z0, z1, t) of one pass of the dynamic / static policies in issue
    order, as the library generates them (culbm_wave_schedule): step t
    (1-based) over planes [z0, z1)."""
    cnt = ctypes.c_int32()
    _lib.check(_lib.lib().culbm_wave_schedule(int(nz), int(depth), int(steps), None, 0, ctypes.byref(cnt)))
    out = np.zeros(3 * cnt.value, np.int32)
    _lib.check(_lib.lib().culbm_wave_schedule(int(nz), int(depth), int(steps), _lib.ptr(out), cnt.value,
                                              ctypes.byref(cnt)))
    return [tuple(int(v) for v in out[3 * k: 3 * k + 3]) for k in range(cnt.value)]
