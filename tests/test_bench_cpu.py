"""bench.py's host-side contract, without a GPU: every per-form traffic entry
under profiles/ carries the field the bench reads, every entry names a bench
form, and a multi-GPU request exits with a message before touching a device."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_traffic_entries_carry_bytes_per_cell_update():
    t = json.loads((ROOT / "profiles" / "traffic_288x272x280.json").read_text())
    assert bench.HEADLINE in t
    for form, e in t.items():
        assert form in bench.FORMS, form
        assert 100.0 < e["bytes_per_cell"] < 400.0, (form, e)


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr) or "1 GPU" in (p.stdout + p.stderr)
