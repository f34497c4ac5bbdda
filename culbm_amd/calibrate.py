"""Memory-hierarchy calibration of the GPU for tile and depth selection.

MI355X counterparts of the reference's microbenchmarks:
  * lambda: DRAM vs L2 bandwidth ratio (analysis/lambda/measure.cu:14-63); here
    the HBM vs Infinity-Cache (MALL) ratio, from streaming probes whose
    working set sits below / above the 256 MB cache.
  * gamma: partial-sector store penalty (analysis/gamma/measure.cu:14-69); here
    a copy whose every store is shifted by one element (each wave's store
    straddles one extra 128-byte line) against the aligned copy.
  * the ceiling of the population access pattern itself: 27 arrays read and
    written once per cell (pop27), with the pull gather's x offsets
    (pop27_shifted).  No step kernel can stream faster than this without
    cache reuse.

    python -m culbm_amd.calibrate [--quick] [--device N]

Prints one JSON object.  With --quick only two sizes run (used as the
known-size copy when reading hardware byte counters).

The tile/I selector (the counterpart of analysis/selector.py:15-69, which
scores tile shapes and I with a model built on its lambda/gamma probes):
the z-wavefront's one free parameter is the chunk depth D, and its optimum
is set by where a chunk's working set meets the Infinity Cache, by launch
ramp-up and tail and by the workgroup order, which no closed form captured
across sizes.  So the selector measures: it times the dynamic policy over
candidate depths (launches of 0.6M to 3.2M cells, plus the automatic choice)
and innerLoop values on this GPU and domain, and reports the fastest
(--waveDepth / --innerLoop of the CLI):

    python -m culbm_amd.calibrate --select --domain 288 272 280 [--storage fp16]
        [--stream 1] [--inner-loop 5 6] [--init-state DIR]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from typing import Iterable, Optional, Sequence, Tuple

from .lattice import probe_bandwidth

MB = 1 << 20


def measure(device: int = 0, quick: bool = False) -> dict:
    sizes = (64, 2048) if quick else (16, 64, 128, 256, 512, 1024, 2048, 4096)
    res = {"sizes_mb_per_buffer": list(sizes)}
    for mode in ("copy", "read", "write", "copy_scalar", "copy_shifted", "pop27", "pop27_shifted"):
        res[mode] = [round(probe_bandwidth(mode, s * MB, iters=10 if s < 1024 else 4, device=device), 1)
                     for s in sizes]
    small = sizes.index(64)
    big = sizes.index(2048)
    res["lambda_hbm_over_mall"] = round(res["copy"][big] / res["copy"][small], 4)
    res["gamma_shifted_over_aligned_hbm"] = round(res["copy_shifted"][big] / res["copy_scalar"][big], 4)
    res["gamma_shifted_over_aligned_mall"] = round(res["copy_shifted"][small] / res["copy_scalar"][small], 4)
    # FP32 MLUPS that the pattern's HBM-regime rate allows at 216 B per update
    res["pop27_mlups_ceiling_fp32"] = round(res["pop27_shifted"][big] * 1e3 / 216, 0)
    return res


# launch sizes (cells per chunk) the depth candidates are drawn from
SELECT_CELLS = (0.6e6, 0.9e6, 1.25e6, 1.4e6, 1.75e6, 2.1e6, 2.6e6, 3.2e6)


def depth_candidates(domain: Tuple[int, int, int]) -> list:
    """Chunk depths (planes) for launches of SELECT_CELLS cells, plus 0 (the
    automatic depth), within [2, nz]."""
    nx, ny, nz = domain
    plane = nx * ny
    return sorted({0} | {max(2, min(nz, math.ceil(c / plane))) for c in SELECT_CELLS})


def select(domain: Tuple[int, int, int], stream_policy: int = 1, storage: str = "fp32",
           inner_loops: Iterable[int] = (5,), depths: Optional[Sequence[int]] = None,
           state=None, rounding: str = "nearest", device: int = 0, steps: Optional[int] = None,
           develop: int = 1000) -> dict:
    """Time the dynamic policy (optPolicy 3) over candidate depths and innerLoop
    values; return every candidate's MLUPS (best of three timed runs) and the
    fastest.  `state` defaults to the seeded obstacle channel of `domain`.
    Each candidate first runs `develop` untimed steps from rest: the step rate
    depends on the population values (a flow at rest, whose shifted
    populations are zero, runs up to 7 % faster than the developed flow)."""
    from . import cases
    from .lattice import Lattice, LatticeConfig

    nx, ny, nz = domain
    n = nx * ny * nz
    if state is None:
        state = cases.config_obstacles(nx, ny, nz) if min(domain) >= 64 else cases.channel(nx, ny, nz)
    cands = list(depths) if depths is not None else depth_candidates(domain)
    inner_loops = list(inner_loops)
    if not cands or not inner_loops:
        raise ValueError("select needs at least one depth and one innerLoop value")

    def timed(il, d):
        lat = Lattice(LatticeConfig(domain=domain, stream_policy=stream_policy, opt_policy=3, inner_loop=il,
                                    storage=storage, wave_depth=d, rounding=rounding, device=device))
        try:
            lat.set_state(state.flag, state.rho, state.vx, state.vy, state.vz)
            lat.init_ddf("rest")
            lat.step(develop + 2 * il)
            k = steps or il * max(2, math.ceil(4e8 / (n * il)))
            ms = min(lat.time_steps(k) for _ in range(3))
            return n * k / (ms * 1e-3) / 1e6, lat.info()["wave_depth"]
        finally:
            lat.close()

    timed(inner_loops[0], cands[0])  # untimed: the first context of a process runs cold
    rows = []
    for il in inner_loops:
        for d in cands:
            mlups, used = timed(il, d)
            rows.append({"inner_loop": il, "requested_depth": d, "wave_depth": used, "mlups": round(mlups, 1)})
    best = max(rows, key=lambda r: r["mlups"])
    return {"domain": list(domain), "stream_policy": stream_policy, "storage": storage, "rounding": rounding,
            "best": best, "candidates": rows}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--select", action="store_true", help="time wavefront depths / innerLoop values instead")
    ap.add_argument("--domain", type=int, nargs=3, default=(288, 272, 280))
    ap.add_argument("--stream", type=int, default=1, help="streamPolicy: 0 pull, 1 in place")
    ap.add_argument("--storage", default="fp32", choices=("fp32", "fp16", "bf16"))
    ap.add_argument("--rounding", default="nearest", choices=("nearest", "stochastic"))
    ap.add_argument("--inner-loop", type=int, nargs="+", default=[5])
    ap.add_argument("--init-state", default=None, help="folder of flag/rho/vx/vy/vz.dat to time on")
    ap.add_argument("--develop", type=int, default=1000, help="untimed steps from rest before each timing")
    ns = ap.parse_args(argv)
    if ns.select:
        from . import cases

        dom = tuple(ns.domain)
        st = cases.read_state(ns.init_state, dom) if ns.init_state else None
        res = select(dom, ns.stream, ns.storage, ns.inner_loop, state=st, rounding=ns.rounding, device=ns.device,
                     develop=ns.develop)
    else:
        res = measure(ns.device, ns.quick)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
