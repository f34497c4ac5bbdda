"""Host-side hazard checker of the temporal-blocking launch schedule.

The reference has no race detection: the safety of its in-place blocking rests
on Esoteric Twist's "read set = write set" property and on launch boundaries
(src/simulator/single_dev_expt/kernel.cu:954-1019, 1195-1348).  This module
simulates, launch by launch, which state level every population slot holds
while a wavefront pass runs, and reports every read that would see another
level than the step expects and every write that lands on a slot another
plane of the same launch still reads.

Model.  A launch updates every cell of its z-planes, so hazards between
launches live along z; within a launch the twisted layout's per-cell slot sets
are disjoint (each slot is read and written by one cell, which the streaming
index-tag tests pin on the GPU).  One exception stays inside a wave: for even
nx the cells on an x face read their clamped x value at the neighbour one cell
inward (no x ghost copies), a slot that neighbour then overwrites; the two are
adjacent lanes of one wave, whose loads all return before any of its stores
issue, so no ordering between launches or workgroups is involved.  Along z only
c_z of a direction matters, so
the 27 slots of a cell fall into three classes c = c_z in {-1, 0, +1}:

  in place, O step (parity 1): load class c at plane z + p(-c), store class -c
      at z + p(c);  E step (parity 0): load class -c at z + p(-c), store class c
      at z + p(c)  (p(c) = max(c, 0); the layout has one padding plane, nz);
      a z-face cell whose consumer is clamped onto it (c = +1 at z = 0,
      c = -1 at z = nz-1) also stores the same class at z + p(-c) ("ghost");
  pull: step t loads class c at plane clamp(z - c) of buffer (cur + t - 1) mod 2
      and stores every class at plane z of buffer (cur + t) mod 2.

Every slot read at step t must hold level t - 1 (levels counted from the start
of the pass), and after the pass every slot the next step reads holds level T.
"""
from __future__ import annotations

from typing import Iterable, List, Optional, Tuple

Launch = Tuple[int, int, int]  # (z0, z1, t): step t over planes [z0, z1)


def _p(c: int) -> int:
    return c if c > 0 else 0


def _inplace_loads(z: int, parity: int):
    for c in (-1, 0, 1):
        yield (c if parity == 1 else -c), z + _p(-c)


def _inplace_stores(z: int, parity: int, nz: int):
    for c in (-1, 0, 1):
        cls = -c if parity == 1 else c
        yield cls, z + _p(c)
        if (c == 1 and z == 0) or (c == -1 and z == nz - 1):
            yield cls, z + _p(-c)


def check_inplace(nz: int, launches: Iterable[Launch], steps: int, parity0: int = 1,
                  limit: int = 20) -> List[str]:
    """Hazards of an in-place (twisted) pass starting with a step of parity
    `parity0`; empty when the schedule is safe."""
    level = {c: [0] * (nz + 1) for c in (-1, 0, 1)}
    done = [0] * nz  # steps each plane has completed
    out: List[str] = []
    for z0, z1, t in launches:
        par = (parity0 + t - 1) & 1
        reads = {}
        for z in range(z0, z1):
            if done[z] != t - 1:
                out.append(f"launch {(z0, z1, t)}: plane {z} has done {done[z]} steps, not {t - 1}")
            for cls, q in _inplace_loads(z, par):
                if level[cls][q] != t - 1:
                    out.append(f"launch {(z0, z1, t)}: plane {z} reads class {cls} at plane {q} "
                               f"holding level {level[cls][q]}, expected {t - 1}")
                reads[(cls, q)] = z
        for z in range(z0, z1):
            for cls, q in _inplace_stores(z, par, nz):
                r = reads.get((cls, q))
                if r is not None and r != z:
                    out.append(f"launch {(z0, z1, t)}: plane {z} writes class {cls} at plane {q}, "
                               f"read by plane {r} in the same launch")
                level[cls][q] = t
            done[z] = t
        if len(out) >= limit:
            return out
    out += _final_inplace(nz, level, done, steps, parity0)
    return out[:limit]


def _final_inplace(nz, level, done, steps, parity0) -> List[str]:
    out = []
    par = (parity0 + steps) & 1
    for z in range(nz):
        if done[z] != steps:
            out.append(f"after the pass plane {z} has done {done[z]} of {steps} steps")
        for cls, q in _inplace_loads(z, par):
            if level[cls][q] != steps:
                out.append(f"after the pass plane {z}'s next load (class {cls}, plane {q}) "
                           f"holds level {level[cls][q]}, not {steps}")
    return out


def check_pull(nz: int, launches: Iterable[Launch], steps: int, limit: int = 20) -> List[str]:
    """Hazards of a double-buffered pull pass (the current state in buffer 0)."""
    level = [{c: [0] * nz for c in (-1, 0, 1)}, {c: [-1] * nz for c in (-1, 0, 1)}]
    done = [0] * nz
    out: List[str] = []
    for z0, z1, t in launches:
        src, dst = level[(t - 1) & 1], level[t & 1]
        for z in range(z0, z1):
            if done[z] != t - 1:
                out.append(f"launch {(z0, z1, t)}: plane {z} has done {done[z]} steps, not {t - 1}")
            for c in (-1, 0, 1):
                q = min(max(z - c, 0), nz - 1)
                if src[c][q] != t - 1:
                    out.append(f"launch {(z0, z1, t)}: plane {z} pulls class {c} from plane {q} "
                               f"holding level {src[c][q]}, expected {t - 1}")
        for z in range(z0, z1):
            for c in (-1, 0, 1):
                dst[c][z] = t
            done[z] = t
        if len(out) >= limit:
            return out
    fin = level[steps & 1]
    for z in range(nz):
        if done[z] != steps:
            out.append(f"after the pass plane {z} has done {done[z]} of {steps} steps")
        for c in (-1, 0, 1):
            if fin[c][z] != steps:
                out.append(f"after the pass buffer {steps & 1} plane {z} class {c} holds level "
                           f"{fin[c][z]}, not {steps}")
    return out[:limit]


def check_wave_pass(nz: int, depth: int, steps: int, layout: str = "inplace", parity0: int = 1,
                    launches: Optional[List[Launch]] = None) -> List[str]:
    """Hazards of the library's own wavefront pass (or of `launches`, when
    given) for a domain of nz planes, chunks of `depth` planes, `steps` steps."""
    if launches is None:
        from .blocking import wave_schedule

        launches = wave_schedule(nz, depth, steps)
    if layout == "inplace":
        return check_inplace(nz, launches, steps, parity0)
    if layout == "pull":
        return check_pull(nz, launches, steps)
    raise ValueError("layout must be 'inplace' or 'pull'")
