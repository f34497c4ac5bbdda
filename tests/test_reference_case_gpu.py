"""The reference's only end-to-end parity case, run through the drop-in CLI.

analysis/correctness/ of the reference generates a 768x384x384 channel with a
solid cube (generate.py:6-42; restated by cases.config_correctness), runs the
solver twice with the argv of run.sh:13-37 -- plain pull with
blockDim [32,4,2] / gridDim [24,96,192], then in place + dynamic blocking with
innerLoop 9, --useRingDDFBuf and blockDim [96,1,1] / gridDim [1,94,86] --
and compares the four *_1539.dat dumps byte for byte (compare.sh:6-9).  Here
the same argv drives culbm_amd/bin/single_dev_expt_main; each dump is reduced
to its SHA-256 digest as soon as it is written, so the case needs about 2.7 GB
of scratch disk instead of 4.5 GB."""
from __future__ import annotations

import hashlib
import re
import subprocess
import time
from pathlib import Path

import pytest

from culbm_amd import cases

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "culbm_amd" / "bin" / "single_dev_expt_main"
FIELDS = ("rho", "vx", "vy", "vz")
STEPS = 1539

COMMON = ["--devId", "0", "--domainDim", "[768,384,384]", "--invTau", "0.2857142857142857",
          "--dstep", str(STEPS), "--nstep", str(STEPS), "--dumpRho", "--dumpVx", "--dumpVy", "--dumpVz"]
PLAIN = ["--blockDim", "[32,4,2]", "--gridDim", "[24,96,192]", "--streamPolicy", "0", "--optPolicy", "0"]
TEMPORAL = ["--blockDim", "[96,1,1]", "--gridDim", "[1,94,86]", "--innerLoop", "9", "--streamPolicy", "1",
            "--optPolicy", "3", "--useRingDDFBuf"]


def run(case, out, extra):
    t = time.time()
    p = subprocess.run([str(CLI), *COMMON, "--initStateFolder", str(case), "--dumpFolder", str(out), *extra],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    speeds = re.findall(r"^speed = ([0-9.]+) \(MLUPS\)$", p.stdout, re.M)
    print(f"{' '.join(extra)}: {time.time() - t:.1f} s, speed {speeds}", flush=True)
    assert len(speeds) == 1
    digests = {}
    for name in FIELDS:
        f = Path(out) / f"{name}_{STEPS}.dat"
        data = f.read_bytes()
        assert len(data) == 768 * 384 * 384 * 4, name
        digests[name] = hashlib.sha256(data).hexdigest()
        f.unlink()
    return digests


def test_reference_correctness_case_plain_equals_temporal(tmp_path):
    st = cases.config_correctness()
    case = tmp_path / "case"
    cases.write_state(case, st, fields=("vx",))  # the reference case folder: flag.dat, vx.dat
    del st
    plain = run(case, tmp_path / "plain", PLAIN)
    temporal = run(case, tmp_path / "temporal", TEMPORAL)
    for name in FIELDS:
        assert plain[name] == temporal[name], f"{name}_{STEPS}.dat differs between plain and temporal"
