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
ROUNDING = {"nearest": 0, "stochastic": 1}
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
    rounding: str = "nearest"                       # 16-bit stores: "nearest" or "stochastic"

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
        if self.rounding not in ROUNDING:
            raise ValueError(f"rounding must be one of {sorted(ROUNDING)}, got {self.rounding!r}")
        c.rounding = ROUNDING[self.rounding]
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

    def dump_deferred(self) -> None:
        """Arm the dump of the current state; the next step writes it (fused),
        or the next macros()/dump()/set_state()/... runs it first."""
        self._check(_lib.lib().culbm_dump_deferred(self._ctx))

    def macros(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        out = [np.empty(self.shape, dtype=np.float32) for _ in range(4)]
        self._check(_lib.lib().culbm_get_macros(self._ctx, *[_lib.ptr(a) for a in out]))
        return tuple(out)

    def moment_sums(self, flag_mask: int = 0) -> Tuple[float, float, float, float]:
        """FP64 (sum rho, sum rho u_x, sum rho u_y, sum rho u_z) of the stored
        populations over the cells whose flag shares a bit with flag_mask
        (every cell for 0): the mass-conservation check."""
        out = (ctypes.c_double * 4)()
        self._check(_lib.lib().culbm_moment_sums(self._ctx, int(flag_mask), out))
        return tuple(out)

    # -- discrete adjoint ----------------------------------------------------
    def adjoint(self, weight, nsteps: int, window: Tuple[int, int], checkpoint_every: int = 0,
                mem_cap: int = 0, want_f0: bool = False, timing: bool = False,
                max_slots: int = 0) -> "AdjointResult":
        """Objective J = sum_{t in window} sum_x weight(x) u_x(x, t) over the next
        `nsteps` steps and its gradients (the context ends `nsteps` steps on).

        `window` = (t_begin, t_end), inclusive, in steps from the current state.
        checkpoint_every = k > 0 keeps every k-th state; 0 picks the binomial
        schedule that fits `mem_cap` device bytes (0: 90% of free memory) and
        holds at most `max_slots` states (0: no limit)."""
        w = np.ascontiguousarray(weight, dtype=np.float64)
        if w.size != self.ncell:
            raise ValueError(f"weight has {w.size} values, domain has {self.ncell}")
        spec = _lib.AdjointSpec()
        spec.nsteps = int(nsteps)
        spec.t_begin, spec.t_end = int(window[0]), int(window[1])
        spec.checkpoint_every = int(checkpoint_every)
        spec.want_f0 = int(bool(want_f0))
        spec.mem_cap = int(mem_cap)
        spec.timing = int(bool(timing))
        spec.max_slots = int(max_slots)
        st = _lib.AdjointStats()
        g = [np.empty(self.shape, dtype=np.float64) for _ in range(4)]
        f0 = np.empty((27,) + self.shape, dtype=np.float32) if want_f0 else None
        self._check(_lib.lib().culbm_adjoint(self._ctx, ctypes.byref(spec), _lib.ptr(w), ctypes.byref(st),
                                             *[_lib.ptr(a) for a in g], _lib.ptr(f0)))
        return AdjointResult(
            objective=st.objective, grad_rho=g[0], grad_ux=g[1], grad_uy=g[2], grad_uz=g[3],
            grad_inv_tau=st.grad_inv_tau, grad_f0=f0,
            stats={"forward_steps": st.forward_steps, "device_bytes": st.device_bytes, "slots": st.slots,
                   "launches": st.launches, "ms_sweep": st.ms_sweep, "ms_recompute": st.ms_recompute,
                   "ms_reverse": st.ms_reverse, "ms_total": st.ms_total})

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


@dataclass
class AdjointResult:
    """J and its gradients; grad_* are (nz, ny, nx) float64, zero outside EQU cells."""
    objective: float
    grad_rho: np.ndarray
    grad_ux: np.ndarray
    grad_uy: np.ndarray
    grad_uz: np.ndarray
    grad_inv_tau: float
    grad_f0: Optional[np.ndarray]
    stats: dict


def adjoint_plan(nsteps: int, checkpoint_every: int = 0, slots: int = 1) -> Tuple[int, int]:
    """(forward steps, state slots) of a checkpoint schedule; host only."""
    fwd = ctypes.c_int64()
    used = ctypes.c_int32()
    _lib.check(_lib.lib().culbm_adjoint_plan(int(nsteps), int(checkpoint_every), int(slots),
                                             ctypes.byref(fwd), ctypes.byref(used)))
    return fwd.value, used.value


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


BW_MODES = {"copy": 0, "read": 1, "write": 2, "copy_scalar": 3, "copy_shifted": 4, "pop27": 5,
            "pop27_shifted": 6}


PROBE_CHECK_WRONG = 0x100


def probe_bandwidth(mode: str, nbytes: int, iters: int = 20, device: int = 0, check_wrong: bool = False) -> float:
    """Achieved GB/s of a streaming probe over `nbytes` per buffer.  The device
    checks afterwards that the probe moved exactly the data it claims (raises
    CulbmError otherwise); check_wrong=True checks against a deliberately wrong
    expectation, which must raise."""
    g = ctypes.c_float()
    m = BW_MODES[mode] | (PROBE_CHECK_WRONG if check_wrong else 0)
    _lib.check(_lib.lib().culbm_probe_bandwidth(device, m, int(nbytes), int(iters), ctypes.byref(g)))
    return g.value
