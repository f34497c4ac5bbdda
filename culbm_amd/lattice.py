"""Python handle on one simulation context of the C ABI.

`Lattice` is the object the reference keeps behind `Simulator::Data`
(src/simulator/single_dev_expt/simulator_expt_platform.cu:23-554): device
buffers, the stream, the policy, the step counter.  Arrays cross the boundary
as dense x-fastest numpy arrays of shape (nz, ny, nx) (populations:
(27, nz, ny, nx)), the layout of the reference's .dat files.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib

STORAGE = {"fp32": 0, "fp16": 1, "bf16": 2}
INIT = {"rest": 0, "feq": 1}


@dataclass
class LatticeConfig:
    """The reference CLI knobs (simulator_expt_platform.cu:312-354) + storage."""

    domain: Tuple[int, int, int] = (256, 256, 256)  # --domainDim (x, y, z)
    inv_tau: float = 0.5                            # --invTau
    stream_policy: int = 0                          # --streamPolicy
    opt_policy: int = 0                             # --optPolicy
    inner_loop: int = 5                             # --innerLoop
    block_dim: Tuple[int, int, int] = (32, 16, 2)   # --blockDim
    grid_dim: Tuple[int, int, int] = (2, 3, 19)     # --gridDim
    use_ring: bool = False                          # --useRingDDFBuf
    device: int = 0                                 # --devId
    storage: str = "fp32"
    storage_scale: float = 0.0
    wave_depth: int = 0

    def to_c(self) -> _lib.Config:
        c = _lib.Config()
        c.nx, c.ny, c.nz = (int(v) for v in self.domain)
        c.inv_tau = float(self.inv_tau)
        c.stream_policy = int(self.stream_policy)
        c.opt_policy = int(self.opt_policy)
        c.inner_loop = int(self.inner_loop)
        for k in range(3):
            c.block_dim[k] = int(self.block_dim[k])
            c.grid_dim[k] = int(self.grid_dim[k])
        c.use_ring = int(bool(self.use_ring))
        c.device = int(self.device)
        if self.storage not in STORAGE:
            raise ValueError(f"storage must be one of {sorted(STORAGE)}, got {self.storage!r}")
        c.storage = STORAGE[self.storage]
        c.storage_scale = float(self.storage_scale)
        c.wave_depth = int(self.wave_depth)
        return c


def validate(cfg: LatticeConfig) -> Optional[str]:
    """Host-only validation (no device call).  Returns None or the reason."""
    c = cfg.to_c()
    buf = ctypes.create_string_buffer(512)
    rc = _lib.lib().culbm_validate(ctypes.byref(c), buf, len(buf))
    return None if rc == _lib.OK else buf.value.decode()


class Lattice:
    def __init__(self, cfg: Optional[LatticeConfig] = None, **kw):
        self.cfg = cfg if cfg is not None else LatticeConfig(**kw)
        self._ctx = None
        c = self.cfg.to_c()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().culbm_create(ctypes.byref(c), ctypes.byref(h)))
        self._ctx = h

    # -- lifetime ----------------------------------------------------------
    def close(self) -> None:
        if self._ctx is not None:
            _lib.lib().culbm_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -----------------------------------------------------------
    @property
    def shape(self) -> Tuple[int, int, int]:
        nx, ny, nz = self.cfg.domain
        return (nz, ny, nx)

    @property
    def ncell(self) -> int:
        nz, ny, nx = self.shape
        return nx * ny * nz

    def _check(self, rc: int) -> None:
        _lib.check(rc, self._ctx)

    def _field(self, a, dtype) -> Optional[np.ndarray]:
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.size != self.ncell:
            raise ValueError(f"field has {a.size} values, domain has {self.ncell}")
        return a

    # -- state -------------------------------------------------------------
    def set_state(self, flag=None, rho=None, vx=None, vy=None, vz=None) -> None:
        arrs = [self._field(flag, np.uint32)] + [self._field(v, np.float32) for v in (rho, vx, vy, vz)]
        self._check(_lib.lib().culbm_set_state(self._ctx, *[_lib.ptr(a) for a in arrs]))

    def init_ddf(self, mode: str = "rest") -> None:
        self._check(_lib.lib().culbm_init_ddf(self._ctx, INIT[mode]))

    def get_ddf(self) -> np.ndarray:
        out = np.empty((27,) + self.shape, dtype=np.float32)
        self._check(_lib.lib().culbm_get_ddf(self._ctx, out))
        return out

    def set_ddf(self, ddf) -> None:
        a = np.ascontiguousarray(ddf, dtype=np.float32)
        if a.size != 27 * self.ncell:
            raise ValueError("ddf must hold 27 * ncell values")
        self._check(_lib.lib().culbm_set_ddf(self._ctx, a))

    # -- stepping ----------------------------------------------------------
    def step(self, n: int = 1) -> None:
        self._check(_lib.lib().culbm_step(self._ctx, int(n)))

    def time_steps(self, n: int) -> float:
        ms = ctypes.c_float()
        self._check(_lib.lib().culbm_time_steps(self._ctx, int(n), ctypes.byref(ms)))
        return ms.value

    def profile_steps(self, n: int) -> dict:
        """Run n steps with an event pair around every launch."""
        ms = ctypes.c_float()
        cnt = ctypes.c_int32()
        self._check(_lib.lib().culbm_profile_steps(self._ctx, int(n), ctypes.byref(ms), ctypes.byref(cnt)))
        return {"steps": int(n), "kernel_ms": ms.value, "launches": cnt.value}

    def run_window(self, dstep: int) -> Tuple[int, float]:
        steps = ctypes.c_int32()
        ms = ctypes.c_float()
        self._check(_lib.lib().culbm_run_window(self._ctx, int(dstep), ctypes.byref(steps), ctypes.byref(ms)))
        return steps.value, ms.value

    # -- output ------------------------------------------------------------
    def dump(self) -> None:
        self._check(_lib.lib().culbm_dump(self._ctx))

    def macros(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        out = [np.empty(self.shape, dtype=np.float32) for _ in range(4)]
        self._check(_lib.lib().culbm_get_macros(self._ctx, *[_lib.ptr(a) for a in out]))
        return tuple(out)

    def info(self) -> dict:
        i = _lib.Info()
        self._check(_lib.lib().culbm_get_info(self._ctx, ctypes.byref(i)))
        return {
            "step": i.step,
            "device_bytes": i.device_bytes,
            "wave_depth": i.wave_depth,
            "launches_per_pass": i.launches_per_pass,
            "bytes_per_pop": i.bytes_per_pop,
            "storage_scale": i.storage_scale,
        }


# ---- single-cell and codec probes (device) --------------------------------
def probe_collide(f: np.ndarray, inv_tau: float, device: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(f, dtype=np.float32).copy()
    n = a.size // 27
    _lib.check(_lib.lib().culbm_probe_cells(device, 0, float(inv_tau), _lib.ptr(a), None, None, n))
    return a


def probe_equilibrium(in4: np.ndarray, device: int = 0) -> np.ndarray:
    q = np.ascontiguousarray(in4, dtype=np.float32)
    n = q.size // 4
    out = np.zeros((n, 27), dtype=np.float32)
    _lib.check(_lib.lib().culbm_probe_cells(device, 1, 0.0, _lib.ptr(out), _lib.ptr(q), None, n))
    return out


def probe_moments(f: np.ndarray, device: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(f, dtype=np.float32)
    n = a.size // 27
    out = np.zeros((n, 4), dtype=np.float32)
    _lib.check(_lib.lib().culbm_probe_cells(device, 2, 0.0, _lib.ptr(a), None, _lib.ptr(out), n))
    return out


def probe_codec(x: np.ndarray, storage: str, scale: float = 1.0, device: int = 0):
    a = np.ascontiguousarray(x, dtype=np.float32)
    enc = np.zeros(a.shape, dtype=np.uint16)
    dec = np.zeros(a.shape, dtype=np.float32)
    _lib.check(_lib.lib().culbm_probe_codec(device, STORAGE[storage], float(scale), a,
                                            _lib.ptr(enc), dec, a.size))
    return enc, dec


BW_MODES = {"copy": 0, "read": 1, "write": 2, "copy_scalar": 3, "copy_shifted": 4}


def probe_bandwidth(mode: str, nbytes: int, iters: int = 20, device: int = 0) -> float:
    """Achieved GB/s of a streaming probe over `nbytes` per buffer."""
    g = ctypes.c_float()
    _lib.check(_lib.lib().culbm_probe_bandwidth(device, BW_MODES[mode], int(nbytes), int(iters),
                                                ctypes.byref(g)))
    return g.value
