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
        # per cell-update, against one step's algorithmic bytes (27 reads + 27 writes);
        # a launch that makes two steps (mixed policy) can go down to half of them
        alg = 54 * (4 if bench.FORMS[form][2] == "fp32" else 2)
        assert 0.5 * alg <= e["bytes_per_cell"] < 2.0 * alg, (form, e)


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr) or "1 GPU" in (p.stdout + p.stderr)


def test_fastest_forms_picks_per_storage():
    head = {"mlups": 100.0}
    forms = {"fp32_mixed": {"mlups": 140.0}, "fp32_plain_pull": {"mlups": 90.0}, "fp16_mixed": {"mlups": 210.0},
             "fp16_plain_pull": {"mlups": 190.0}, "bf16_inplace_blocked": {"mlups": 170.0}}
    best = bench.fastest_forms("fp32_inplace_blocked", head, forms)
    assert best["fp32"] == {"form": "fp32_mixed", "mlups": 140.0, "x_headline": 1.4}
    assert best["fp16"]["form"] == "fp16_mixed" and best["bf16"]["form"] == "bf16_inplace_blocked"
    assert bench.fastest_forms("fp32_inplace_blocked", head, {})["fp32"]["form"] == "fp32_inplace_blocked"


def test_fp16_claim_forms_are_stochastic_fp16():
    """The >= 1.7x at <= 1e-3 claim may only take an FP16 form with stochastic
    rounding (round to nearest misses the error bar); they all store the same
    bits, which the GPU variant tests pin."""
    assert bench.FP16_CLAIMS
    for form in bench.FP16_CLAIMS:
        spec = bench.FORMS[form]
        assert spec[2] == "fp16" and len(spec) > 4 and spec[4] == "stochastic", form
