"""The C ABI and the drop-in CLI without a GPU: the library loads and exports
every symbol include/culbm.h declares; configuration errors are reported
before any device call, with the reference's messages where it has one
(simulator_expt_platform.cu:365-368, 523-534, 105-115)."""
from __future__ import annotations

import ctypes
import json
import re
import subprocess
from pathlib import Path

import pytest

from culbm_amd import LatticeConfig, validate
from culbm_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "culbm_amd" / "bin" / "single_dev_expt_main"


def header_symbols():
    text = (ROOT / "include" / "culbm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(culbm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_struct_layout_matches_header():
    # culbm_config: 3 + 1 + 3 + 3 + 3 + 3 + 1 + 1 + 7 four-byte fields
    assert ctypes.sizeof(_lib.Config) == 4 * 25
    assert ctypes.sizeof(_lib.Info) == 8 + 8 + 4 * 4 + 4 * 8


def test_defaults_are_the_reference_cli_defaults():
    c = _lib.Config()
    _lib.lib().culbm_config_default(ctypes.byref(c))
    assert (c.nx, c.ny, c.nz) == (256, 256, 256)
    assert c.inv_tau == pytest.approx(0.5)
    assert list(c.block_dim) == [32, 16, 2] and list(c.grid_dim) == [2, 3, 19]
    assert (c.stream_policy, c.opt_policy, c.inner_loop) == (0, 0, 5)


@pytest.mark.parametrize("kw,msg", [
    (dict(domain=(0, 4, 4)), "domain dimensions"),
    (dict(stream_policy=2), "streamPolicy"),
    (dict(opt_policy=5), "optPolicy"),
    (dict(stream_policy=1, opt_policy=2), "optPolicy 2"),
    (dict(use_ring=True), "--useRingDDFBuf is unimplemented unless --streamPolicy 1 --optPolicy 3."),
    (dict(opt_policy=3, inner_loop=0), "innerLoop"),
    (dict(storage="fp16", storage_scale=3.0), "power of two"),
    (dict(storage="fp32", storage_scale=2.0), "no scale"),
    (dict(domain=(300, 64, 64), opt_policy=3, block_dim=(2, 1, 1), grid_dim=(1, 1, 1), inner_loop=4),
     "<Domain Dimension, Blocking Dimension and InnerLoop> don't match!"),
    (dict(inv_tau=float("nan")), "finite"),
    (dict(storage="fp32", rounding="stochastic"), "stochastic rounding applies to FP16 or BF16"),
    # one FP32 direction array above 4 GiB: padded in place (1088 x 1025 x 1025 cells), dense in pull
    (dict(domain=(1024, 1024, 1024), stream_policy=1), "32-bit offsets"),
    (dict(domain=(1100, 1100, 1100)), "32-bit offsets"),
])
def test_validation_errors(kw, msg):
    err = validate(LatticeConfig(**kw))
    assert err is not None and msg in err


def test_largest_domains_that_fit_the_32_bit_offsets_are_accepted():
    # FP32 pull at 1024^3: 2^30 cells x 4 B, exactly 4 GiB per direction, fits (the last
    # element starts at 2^32 - 4); one more plane does not; FP16 in place at 1024^3 fits
    assert validate(LatticeConfig(domain=(1024, 1024, 1024))) is None
    assert "32-bit offsets" in validate(LatticeConfig(domain=(1024, 1024, 1025)))
    assert validate(LatticeConfig(domain=(1024, 1024, 1024), stream_policy=1, storage="fp16")) is None


def test_valid_configs():
    for sp, op in [(0, 0), (1, 0), (0, 1), (1, 1), (0, 2), (0, 3), (1, 3)]:
        assert validate(LatticeConfig(domain=(64, 64, 64), stream_policy=sp, opt_policy=op)) is None
    assert validate(LatticeConfig(domain=(64, 64, 64), stream_policy=1, opt_policy=3, use_ring=True)) is None
    for storage in ("fp16", "bf16"):
        assert validate(LatticeConfig(domain=(64, 64, 64), storage=storage, rounding="stochastic")) is None


def test_create_reports_validation_error_without_device():
    c = LatticeConfig(domain=(8, 8, 8), stream_policy=7).to_c()
    h = ctypes.c_void_p()
    rc = _lib.lib().culbm_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == _lib.ERR_INVALID and h.value is None
    assert b"streamPolicy" in _lib.lib().culbm_last_error(None)
    assert _lib.lib().culbm_error_string(rc) == b"CULBM_ERR_INVALID"
    assert _lib.lib().culbm_destroy(None) == 0


# ---- the drop-in CLI -----------------------------------------------------------

def run_cli(*args):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not CLI.exists(), reason="CLI not built")
class TestCli:
    def test_help(self):
        r = run_cli("--help")
        assert r.returncode == 0 and "--streamPolicy" in r.stdout

    @pytest.mark.parametrize("vec", [["[64,48,38]"], ["64,48,38"], ["64", "48", "38"]])
    def test_vector_syntaxes(self, vec):
        r = run_cli("--domainDim", *vec, "--dryRun")
        assert r.returncode == 0, r.stderr
        assert json.loads(r.stdout)["domainDim"] == [64, 48, 38]

    def test_parsed_knobs(self):
        r = run_cli("--domainDim=[288,272,280]", "--blockDim", "[32,16,2]", "--gridDim", "[2,6,38]",
                    "--innerLoop", "5", "--streamPolicy", "1", "--optPolicy", "3", "--invTau", "0.2857142857142857",
                    "--nstep", "1539", "--dstep", "1539", "--dumpRho", "--dumpVz", "--useRingDDFBuf",
                    "--dumpFolder", "out", "--dryRun")
        assert r.returncode == 0, r.stderr
        c = json.loads(r.stdout)
        assert c["streamPolicy"] == 1 and c["optPolicy"] == 3 and c["useRingDDFBuf"] is True
        assert c["nstep"] == 1539 and c["dump"] == [1, 0, 0, 1] and c["dumpFolder"] == "out"
        assert abs(c["invTau"] - 0.2857142857142857) < 1e-7

    def test_unknown_option(self):
        r = run_cli("--bogus")
        assert r.returncode == 2 and "--bogus" in r.stderr

    def test_ring_rule(self):
        r = run_cli("--useRingDDFBuf")
        assert r.returncode == 1
        assert "--useRingDDFBuf is unimplemented unless --streamPolicy 1 --optPolicy 3." in r.stderr

    def test_plain_grid_must_cover_the_domain(self):
        r = run_cli("--domainDim", "[64,48,40]", "--streamPolicy", "0", "--optPolicy", "0")
        assert r.returncode == 1 and "Dimension of domain doesn't match computation grid!" in r.stderr
        r = run_cli("--domainDim", "[64,48,40]", "--gridDim", "[2,3,20]", "--dryRun")
        assert r.returncode == 0, r.stderr

    def test_missing_flag_file(self, tmp_path):
        r = run_cli("--initStateFolder", str(tmp_path), "--domainDim", "[8,8,8]")
        assert r.returncode == 1 and f"File: {tmp_path}/flag.dat doesn't exist!" in r.stderr


@pytest.mark.parametrize("mode,nbytes,iters", [(7, 1 << 20, 1), (-1, 1 << 20, 1), (0, 100, 1), (5, 1 << 20, 0),
                                               (6, (64 << 30) + 16, 1), (3, (8 << 30) + 16, 1),
                                               (5, 16 << 10, 1), (0x200, 1 << 20, 1)])
def test_bandwidth_probe_rejects_bad_arguments_before_touching_a_device(mode, nbytes, iters):
    g = ctypes.c_float()
    rc = _lib.lib().culbm_probe_bandwidth(0, mode, nbytes, iters, ctypes.byref(g))
    assert rc == 1  # CULBM_ERR_INVALID
