"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference-header binding (oracle/_ref/libculbm_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

Every lattice function takes dense arrays shaped (nz, ny, nx) (populations
(27, nz, ny, nx)).  Parity pins: see tests/test_oracle.py (closed forms and
textbook identities), tests/test_ref_parity.py (reference headers compiled on
the host) and tests/golden/ (vectors produced from those headers).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libculbm_ref.so"

_F64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_U16 = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i = ctypes.c_int
_i64 = ctypes.c_int64

_SIG = {
    "orc_velset": [_I32P, _I32P, _I32P, _F64],
    "orc_equ_f64": [ctypes.c_double] * 4 + [_F64],
    "orc_equ_f32": [ctypes.c_float] * 4 + [_F32],
    "orc_moments_f64": [_F64, _F64],
    "orc_moments_f32": [_F32, _F32],
    "orc_collide_f64": [ctypes.c_double, _F64],
    "orc_collide_f32": [ctypes.c_float, _F32],
    "orc_collide_batch_f32": [ctypes.c_float, _F32, _i64],
    "orc_equ_batch_f32": [_F32, _F32, _F32, _F32, _F32, _i64],
    "orc_moments_batch_f32": [_F32, _F32, _i64],
    "orc_step_f64": [_i, _i, _i, _U32, _F32, _F32, _F32, _F32, ctypes.c_double, _F64, _F64, _i],
    "orc_step_f32": [_i, _i, _i, _U32, _F32, _F32, _F32, _F32, ctypes.c_float, _F32, _F32, _i],
    "orc_step_s16": [_i, _i, _i, _U32, _F32, _F32, _F32, _F32, ctypes.c_float, _i, ctypes.c_float,
                     _U16, _U16, _i],
    "orc_step_s16_sr": [_i, _i, _i, _U32, _F32, _F32, _F32, _F32, ctypes.c_float, _i, ctypes.c_float,
                        ctypes.c_uint32, _U16, _U16, _i],
    "orc_encode16_sr": [_i, _F32, _U16, _i64, _i64, ctypes.c_uint32],
    "orc_dump_f32": [_i, _i, _i, _U32, _F32, _F32, _F32, _F32, _F32, _i],
    "orc_dump_s16": [_i, _i, _i, _U32, _i, ctypes.c_float, _U16, _F32, _F32, _F32, _F32, _i],
    "orc_dump_f64": [_i, _i, _i, _U32, _F64, _F64, _F64, _F64, _F64, _i],
    "orc_feq_soa_f64": [_F32, _F32, _F32, _F32, _F64, _i64],
    "orc_feq_soa_f32": [_F32, _F32, _F32, _F32, _F32, _i64],
    "orc_encode16": [_i, _F32, _U16, _i64],
    "orc_decode16": [_i, _U16, _F32, _i64],
    "orc_blk_valid": [_i, _i, _i],
    "orc_blk_num": [_i, _i, _i],
    "orc_valid_prev": [_i, _i, _i, _i, _i],
    "orc_max_threads": [],
}
_RES = {"orc_blk_valid": _i, "orc_blk_num": _i, "orc_valid_prev": _i, "orc_max_threads": _i}

_REF_SIG = {
    "ref_velset": [_I32P, _I32P, _I32P, _F32],
    "ref_equ": [ctypes.c_float] * 4 + [_F32],
    "ref_moments": [_F32, _F32],
    "ref_collide": [ctypes.c_float, _F32],
    "ref_collide_moments": [ctypes.c_float, _F32, _F32],
    "ref_collide_batch": [ctypes.c_float, _F32, _i64],
    "ref_equ_batch": [_F32, _F32, _F32, _F32, _F32, _i64],
    "ref_moments_batch": [_F32, _F32, _i64],
    "ref_blk_valid": [ctypes.c_uint32] * 3,
    "ref_blk_num": [ctypes.c_uint32] * 3,
    "ref_valid_prev": [ctypes.c_int32] * 5,
}
_REF_RES = {"ref_blk_valid": _i, "ref_blk_num": ctypes.c_uint32, "ref_valid_prev": ctypes.c_int32}

_o = None
_r = None


def lib():
    global _o
    if _o is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(str(ORACLE_PATH))
        for name, args in _SIG.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = _RES.get(name, None)
        _o = h
    return _o


def ref_available() -> bool:
    return REF_PATH.exists()


def ref():
    """The reference's own headers compiled on the host (absent where the
    reference checkout is not)."""
    global _r
    if _r is None:
        if not REF_PATH.exists():
            raise RuntimeError(f"{REF_PATH} missing (needs the reference checkout)")
        h = ctypes.CDLL(str(REF_PATH))
        for name, args in _REF_SIG.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = _REF_RES.get(name, None)
        _r = h
    return _r


def threads(n=None) -> int:
    if n:
        return int(n)
    return int(os.environ.get("OMP_NUM_THREADS", 0)) or lib().orc_max_threads()


def velset():
    cx = np.zeros(27, np.int32)
    cy = np.zeros(27, np.int32)
    cz = np.zeros(27, np.int32)
    w = np.zeros(27, np.float64)
    lib().orc_velset(cx, cy, cz, w)
    return cx, cy, cz, w


def equ(rho, ux, uy, uz, dtype=np.float64):
    f = np.zeros(27, dtype)
    (lib().orc_equ_f64 if dtype == np.float64 else lib().orc_equ_f32)(rho, ux, uy, uz, f)
    return f


def moments(f):
    f = np.ascontiguousarray(f)
    out = np.zeros(4, f.dtype)
    (lib().orc_moments_f64 if f.dtype == np.float64 else lib().orc_moments_f32)(f, out)
    return out


def collide(f, inv_tau):
    f = np.array(f, copy=True)
    (lib().orc_collide_f64 if f.dtype == np.float64 else lib().orc_collide_f32)(inv_tau, f)
    return f


def _fields(state):
    return [np.ascontiguousarray(getattr(state, k), np.float32) for k in ("rho", "vx", "vy", "vz")]


class Sim:
    """Double-buffered plain pull stepping of one state on the CPU.

    dtype float64: accuracy oracle; float32: bit-exact FP32 restatement;
    'fp16'/'bf16': FP32 arithmetic over 16-bit shifted storage."""

    def __init__(self, state, inv_tau, dtype=np.float64, scale=1.0, nthreads=None, rounding="nearest"):
        self.flag = np.ascontiguousarray(state.flag, np.uint32)
        self.nz, self.ny, self.nx = self.flag.shape
        self.fields = _fields(state)
        self.inv_tau = inv_tau
        self.kind = dtype if isinstance(dtype, str) else np.dtype(dtype).name
        self.scale = float(scale)
        self.nthreads = threads(nthreads)
        shp = (27, self.nz, self.ny, self.nx)
        st = np.uint16 if self.kind in ("fp16", "bf16") else np.dtype(dtype)
        self.src = np.zeros(shp, st)
        self.dst = np.zeros(shp, st)
        self.step_count = 0
        self.stochastic = rounding == "stochastic"

    # populations as float arrays (decoded)
    def ddf(self) -> np.ndarray:
        if self.kind in ("fp16", "bf16"):
            out = np.zeros(self.src.shape, np.float32)
            lib().orc_decode16(1 if self.kind == "fp16" else 2, self.src.reshape(-1), out.reshape(-1), out.size)
            return out * np.float32(1.0 / self.scale)
        return self.src.copy()

    def set_ddf(self, f) -> None:
        if self.kind in ("fp16", "bf16"):
            enc = np.zeros(self.src.shape, np.uint16)
            scaled = np.ascontiguousarray(f, np.float32) * np.float32(self.scale)
            lib().orc_encode16(1 if self.kind == "fp16" else 2, scaled.reshape(-1), enc.reshape(-1), enc.size)
            self.src[...] = enc
        else:
            self.src[...] = f
        self.dst[...] = self.src

    def init_feq(self) -> None:
        """Populations = feq(rho, u) of the fields (same values in both buffers)."""
        rho, vx, vy, vz = (a.reshape(-1) for a in self.fields)
        n = rho.size
        if self.kind == "float64":
            f = np.zeros(27 * n, np.float64)
            lib().orc_feq_soa_f64(rho, vx, vy, vz, f, n)
        else:
            f = np.zeros(27 * n, np.float32)
            lib().orc_feq_soa_f32(rho, vx, vy, vz, f, n)
        self.set_ddf(f.reshape(27, self.nz, self.ny, self.nx))

    def step(self, n: int = 1) -> None:
        rho, vx, vy, vz = self.fields
        for _ in range(n):
            if self.kind == "float64":
                lib().orc_step_f64(self.nx, self.ny, self.nz, self.flag, rho, vx, vy, vz,
                                   float(self.inv_tau), self.src, self.dst, self.nthreads)
            elif self.kind == "float32":
                lib().orc_step_f32(self.nx, self.ny, self.nz, self.flag, rho, vx, vy, vz,
                                   float(self.inv_tau), self.src, self.dst, self.nthreads)
            elif self.stochastic:
                lib().orc_step_s16_sr(self.nx, self.ny, self.nz, self.flag, rho, vx, vy, vz,
                                      float(self.inv_tau), 1 if self.kind == "fp16" else 2, self.scale,
                                      self.step_count + 1, self.src, self.dst, self.nthreads)
            else:
                lib().orc_step_s16(self.nx, self.ny, self.nz, self.flag, rho, vx, vy, vz,
                                   float(self.inv_tau), 1 if self.kind == "fp16" else 2, self.scale,
                                   self.src, self.dst, self.nthreads)
            self.src, self.dst = self.dst, self.src
            self.step_count += 1

    def dump(self):
        """rho/u fields after a dump (DUMP-flagged cells overwritten)."""
        if self.kind == "float64":
            out = [np.ascontiguousarray(f, np.float64) for f in self.fields]
            lib().orc_dump_f64(self.nx, self.ny, self.nz, self.flag, self.src, *out, self.nthreads)
            self.fields = [o.astype(np.float32) for o in out]
            return tuple(out)
        out = [f.copy() for f in self.fields]
        if self.kind == "float32":
            lib().orc_dump_f32(self.nx, self.ny, self.nz, self.flag, self.src, *out, self.nthreads)
        else:
            lib().orc_dump_s16(self.nx, self.ny, self.nz, self.flag, 1 if self.kind == "fp16" else 2,
                               self.scale, self.src, *out, self.nthreads)
        self.fields = out
        return tuple(out)


_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def adjoint_f64(state, inv_tau, nsteps, t0, t1, weight, f0, rho=None, ux=None, uy=None, uz=None):
    """FP64 forward + discrete adjoint on the CPU (small grids: every state is
    kept).  Returns (J, dict of gradients: rho, ux, uy, uz (per cell EQU inputs),
    inv_tau, f0)."""
    fn = lib().orc_adjoint_f64
    fn.restype = ctypes.c_double
    fn.argtypes = [_i, _i, _i, _U32, _F64P, _F64P, _F64P, _F64P, ctypes.c_double, _i, _i, _i,
                   _F64P, _F64P, _F64P, _F64P, _F64P, _F64P, ctypes.POINTER(ctypes.c_double), _F64P]
    nz, ny, nx = state.flag.shape
    n = nx * ny * nz
    fields = [np.ascontiguousarray(v if v is not None else getattr(state, k), np.float64).reshape(-1)
              for v, k in ((rho, "rho"), (ux, "vx"), (uy, "vy"), (uz, "vz"))]
    g = {k: np.zeros(n) for k in ("rho", "ux", "uy", "uz")}
    gf0 = np.zeros(27 * n)
    gt = ctypes.c_double()
    J = fn(nx, ny, nz, np.ascontiguousarray(state.flag, np.uint32).reshape(-1), *fields, float(inv_tau),
           int(nsteps), int(t0), int(t1), np.ascontiguousarray(weight, np.float64).reshape(-1),
           np.ascontiguousarray(f0, np.float64).reshape(-1), g["rho"], g["ux"], g["uy"], g["uz"],
           ctypes.byref(gt), gf0)
    g = {k: v.reshape(nz, ny, nx) for k, v in g.items()}
    g["inv_tau"] = gt.value
    g["f0"] = gf0.reshape(27, nz, ny, nx)
    return J, g


def _adj_fields(state, rho, ux, uy, uz):
    return [np.ascontiguousarray(v if v is not None else getattr(state, k), np.float64).reshape(-1)
            for v, k in ((rho, "rho"), (ux, "vx"), (uy, "vy"), (uz, "vz"))]


def objective_f64(state, inv_tau, nsteps, t0, t1, weight, f0, rho=None, ux=None, uy=None, uz=None,
                  nthreads=None, f_end=False):
    """FP64 forward run of nsteps steps from f0 and J = sum_{t in [t0, t1]}
    sum_x w(x) u_x(x, t) (EQU inputs as float64, so they can be perturbed).
    With f_end=True returns (J, f^nsteps)."""
    fn = lib().orc_objective_f64
    fn.restype = ctypes.c_double
    fn.argtypes = [_i, _i, _i, _U32, _F64P, _F64P, _F64P, _F64P, ctypes.c_double, _i, _i, _i,
                   _F64P, _F64P, ctypes.c_void_p, _i]
    nz, ny, nx = state.flag.shape
    out = np.empty(27 * nx * ny * nz) if f_end else None
    J = fn(nx, ny, nz, np.ascontiguousarray(state.flag, np.uint32).reshape(-1),
           *_adj_fields(state, rho, ux, uy, uz), float(inv_tau), int(nsteps), int(t0), int(t1),
           np.ascontiguousarray(weight, np.float64).reshape(-1), np.ascontiguousarray(f0, np.float64).reshape(-1),
           None if out is None else out.ctypes.data_as(ctypes.c_void_p), threads(nthreads))
    return (J, out.reshape(27, nz, ny, nx)) if f_end else J


def adjoint_ckpt_f64(state, inv_tau, nsteps, t0, t1, weight, f0, every, rho=None, ux=None, uy=None,
                     uz=None, nthreads=None):
    """adjoint_f64 with states kept every `every` steps, on `nthreads` threads
    (same results, bit for bit)."""
    fn = lib().orc_adjoint_ckpt_f64
    fn.restype = ctypes.c_double
    fn.argtypes = [_i, _i, _i, _U32, _F64P, _F64P, _F64P, _F64P, ctypes.c_double, _i, _i, _i,
                   _F64P, _F64P, _i, _F64P, _F64P, _F64P, _F64P, ctypes.POINTER(ctypes.c_double), _F64P, _i]
    nz, ny, nx = state.flag.shape
    n = nx * ny * nz
    g = {k: np.zeros(n) for k in ("rho", "ux", "uy", "uz")}
    gf0 = np.zeros(27 * n)
    gt = ctypes.c_double()
    J = fn(nx, ny, nz, np.ascontiguousarray(state.flag, np.uint32).reshape(-1),
           *_adj_fields(state, rho, ux, uy, uz), float(inv_tau), int(nsteps), int(t0), int(t1),
           np.ascontiguousarray(weight, np.float64).reshape(-1), np.ascontiguousarray(f0, np.float64).reshape(-1),
           int(every), g["rho"], g["ux"], g["uy"], g["uz"], ctypes.byref(gt), gf0, threads(nthreads))
    g = {k: v.reshape(nz, ny, nx) for k, v in g.items()}
    g["inv_tau"] = gt.value
    g["f0"] = gf0.reshape(27, nz, ny, nx)
    return J, g
