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
"""
from __future__ import annotations

import argparse
import json
import sys

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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--device", type=int, default=0)
    ns = ap.parse_args(argv)
    print(json.dumps(measure(ns.device, ns.quick)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
