"""Build the native parts in-tree.

    python -m culbm_amd.build            # library + CLI + oracle
    python -m culbm_amd.build --lib      # only libculbm.so

Products (git-ignored, they travel with the source tree):
    culbm_amd/libculbm.so                   HIP kernels + C ABI (gfx950)
    culbm_amd/bin/single_dev_expt_main      drop-in CLI over the C ABI
    oracle/liboracle.so                     CPU oracle (test infrastructure)
    oracle/_ref/libculbm_ref.so             reference headers bound on the
                                            host (only where the reference
                                            checkout exists)
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libculbm.so"
CLI = PKG / "bin" / "single_dev_expt_main"

# -ffp-contract=off: only explicit fmaf() calls fuse, so every kernel variant
# (and the host FP32 restatement) performs the same FP32 operations.
# The packed-FP32 feature is disabled: hipcc would otherwise emit v_pk_*_f32.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, **kw):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, **kw)


def build_lib(verbose_resources: bool = False) -> Path:
    cmd = [_hipcc(), *HIP_FLAGS, "-shared", CSRC / "capi.hip", "-o", str(LIB) + ".tmp"]
    if verbose_resources:
        cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
    _run(cmd)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


def build_cli() -> Path:
    CLI.parent.mkdir(exist_ok=True)
    cmd = [
        "g++", "-O2", "-std=c++17", "-Wall",
        CSRC / "single_dev_expt_main.cpp",
        f"-I{ROOT / 'include'}",
        f"-L{PKG}", "-lculbm",
        "-Wl,-rpath,$ORIGIN/..",
        "-o", CLI,
    ]
    _run(cmd)
    return CLI


def build_oracle() -> None:
    _run(["make", "-s", "-C", ROOT / "oracle"])


def build_all() -> None:
    build_lib()
    build_cli()
    build_oracle()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only the HIP library")
    ap.add_argument("--resources", action="store_true", help="print kernel resource usage")
    ns = ap.parse_args(argv)
    if ns.lib:
        build_lib(ns.resources)
    else:
        build_lib(ns.resources)
        build_cli()
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
