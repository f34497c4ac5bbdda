"""The drop-in CLI end to end on the GPU: init files in, `speed = ... (MLUPS)`
per window, `<field>_<step>.dat` dumps out (simulator_expt_platform.cu:572-606,
638-659), checked bit for bit against the FP32 oracle, and the in-place +
blocked policy against plain pull through the files alone."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from culbm_amd import cases

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "culbm_amd" / "bin" / "single_dev_expt_main"
N = 32
FIELDS = ("rho", "vx", "vy", "vz")
PLAIN_GRID = ("--blockDim", "[32,16,2]", "--gridDim", f"[{N // 32},{N // 16},{N // 2}]")  # covers the domain


def run_cli(init, out, *args):
    cmd = [str(CLI), "--domainDim", f"[{N},{N},{N}]", "--invTau", "0.5", "--initStateFolder", str(init),
           "--dumpFolder", str(out), "--dumpRho", "--dumpVx", "--dumpVy", "--dumpVz", *map(str, args)]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    speeds = [float(m) for m in re.findall(r"^speed = ([0-9.]+) \(MLUPS\)$", p.stdout, re.M)]
    return p.stdout, speeds


@pytest.fixture(scope="module")
def state(tmp_path_factory):
    st = cases.config_small(seed=4, n=N)
    st.flag[10:14, 12:16, 9:13] = cases.WALL
    folder = tmp_path_factory.mktemp("init")
    cases.write_state(folder, st)
    return st, folder


def test_windows_and_dumps_match_the_oracle(state, tmp_path):
    from oracle import pyoracle as po
    st, folder = state
    stdout, speeds = run_cli(folder, tmp_path, "--streamPolicy", 0, "--optPolicy", 0, "--nstep", 40, "--dstep", 20,
                             *PLAIN_GRID)
    assert len(speeds) == 2 and all(s > 0 for s in speeds)
    assert "Init flag successfully, from file:" in stdout
    sim = po.Sim(st, 0.5, np.float32)  # the reference starts every run from rest
    for step in (20, 40):
        sim.step(20)
        want = sim.dump()
        for name, w in zip(FIELDS, want):
            got = cases.read_dump(tmp_path, name, step, st.domain)
            assert np.array_equal(got.view(np.uint32), np.asarray(w, np.float32).view(np.uint32)), (name, step)
    # cells without the DUMP bits keep their initial values
    rho40 = cases.read_dump(tmp_path, "rho", 40, st.domain)
    walls = (st.flag & cases.DUMP_RHO) == 0
    assert np.array_equal(rho40[walls], st.rho[walls])


def test_blocked_in_place_windows_overshoot_and_match_plain_pull(state, tmp_path):
    st, folder = state
    out_b, out_p = tmp_path / "blocked", tmp_path / "plain"
    _, speeds = run_cli(folder, out_b, "--streamPolicy", 1, "--optPolicy", 3, "--innerLoop", 4,
                        "--nstep", 40, "--dstep", 10)
    # windows advance in whole passes of innerLoop steps: 12, 24, 36, 48
    assert len(speeds) == 4
    assert sorted(p.name for p in out_b.glob("rho_*.dat")) == [f"rho_{s}.dat" for s in (12, 24, 36, 48)]
    run_cli(folder, out_p, "--streamPolicy", 0, "--optPolicy", 0, "--nstep", 48, "--dstep", 48, *PLAIN_GRID)
    for name in FIELDS:
        a = (out_b / f"{name}_48.dat").read_bytes()
        b = (out_p / f"{name}_48.dat").read_bytes()
        assert a == b, name
