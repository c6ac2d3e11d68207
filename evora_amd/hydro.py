"""Python host mirror of the reference's hydro routines over libevora_hip.

`Hydro` exposes the step-level routines under the reference's names --
primitive (muscl.f90:26), bound (boundary.f90:26), hydrostep
(timestep.f90:26), compevo (muscl.f90:45), sweep (muscl.f90:171) -- with the
same call order, argument meaning and error behaviour (a reference `stop`
or `cleanstop` becomes an `EvoraError` carrying the reference's message and
status code).  The state lives on the GPU; `upload_q`/`download_q` move the
reference-layout q(6, nx, ny, nz) (numpy shape (nz, ny, nx, 6)).

There is no CPU fallback: constructing a `Hydro` without the built library
or without a HIP device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import Derived, HydroParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libevora_hip.so")

FP64, FP32_MIXED = 0, 1
MODE_FUSED, MODE_STAGED, MODE_FUSED_EXACT = 0, 1, 2
STAGE_SWEEP, STAGE_CONDUCTION, STAGE_GRAVPOTENTIAL, STAGE_GRAVINTEGRATION = 0, 1, 2, 3

ERR_NAMES = {
    1: "negative pressure", 2: "riemann unphysical pressure", 3: "hll", 4: "hllc", 5: "dt",
    6: "solver", 7: "limiter", 8: "config", 9: "hip", 10: "state",
}

# symbols of include/evora_hip.h (checked by tests/test_abi.py)
EXPORTS = (
    "evh_create", "evh_destroy", "evh_set_mode", "evh_upload_q", "evh_download_q",
    "evh_primitive", "evh_bound", "evh_hydrostep", "evh_compevo", "evh_sweep",
    "evh_thermalstep", "evh_thermalconduction",
    "evh_step", "evh_run", "evh_sums", "evh_download_w", "evh_download_evo_s",
    "evh_last_error", "evh_event_record", "evh_event_elapsed", "evh_synchronize",
    "evh_set_tuning", "evh_get_tuning", "evh_set_timing", "evh_kernel_stats", "evh_conduction_stats", "evh_stage_stats",
    "evh_gravpotential", "evh_gravintegration", "evh_download_phi",
    "evh_slab", "evh_slab_faces", "evh_slab_temp", "evh_slab_temp_faces",
    "evh_device_bytes", "evh_math_eval",
    "evh_batch_create", "evh_batch_destroy", "evh_batch_upload_q", "evh_batch_download_q",
    "evh_batch_set_time", "evh_batch_run", "evh_batch_times", "evh_batch_set_tuning",
    "evh_batch_get_tuning", "evh_batch_autotune", "evh_batch_last_error", "evh_batch_run_noalloc",
)


class EvhConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("boundtype", C.c_int), ("solver", C.c_int), ("limiter", C.c_int),
        ("secondorder", C.c_int), ("dualenergy", C.c_int),
        ("onedimension", C.c_int), ("twodimension", C.c_int),
        ("precision", C.c_int),
        ("g", C.c_double), ("g1", C.c_double), ("g1inv", C.c_double), ("gg1", C.c_double),
        ("cfl", C.c_double), ("de_thresh1", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dxinv", C.c_double), ("dyinv", C.c_double), ("dzinv", C.c_double),
        ("volume", C.c_double),
        ("mach_fp64", C.c_double),
        ("conduction", C.c_int), ("simplekappa", C.c_int),
        ("c_cond", C.c_double), ("temp0", C.c_double), ("kappa0", C.c_double),
        ("gravity", C.c_int), ("supercomoving", C.c_int),
        ("grav", C.c_double), ("fb", C.c_double), ("h2", C.c_double),
        ("kfx2", C.c_double), ("kfy2", C.c_double), ("kfz2", C.c_double),
        ("nyqx", C.c_int), ("nyqy", C.c_int), ("nyqz", C.c_int),
        ("dx2inv", C.c_double), ("dy2inv", C.c_double), ("dz2inv", C.c_double),
        ("hydro_off", C.c_int),
    ]


class EvoraError(RuntimeError):
    def __init__(self, code: int, message: str, cell=(0, 0, 0)):
        super().__init__(f"{message} [status {code}: {ERR_NAMES.get(code, '?')}, cell {tuple(cell)}]")
        self.code = code
        self.cell = tuple(cell)
        self.message = message


_lib = None


def load_library():
    """Load libevora_hip.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m evora_amd.build`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        sig = {
            "evh_create": [C.POINTER(EvhConfig), C.POINTER(C.c_void_p)],
            "evh_destroy": [P], "evh_set_mode": [P, C.c_int],
            "evh_upload_q": [P, D], "evh_download_q": [P, D],
            "evh_primitive": [P], "evh_bound": [P], "evh_hydrostep": [P, D],
            "evh_compevo": [P, C.c_double], "evh_sweep": [P, C.c_double],
            "evh_thermalstep": [P, D], "evh_thermalconduction": [P, C.c_double],
            "evh_step": [P, D, C.c_double, D, C.POINTER(C.c_int)],
            "evh_run": [P, C.c_int, D, C.c_double, D, C.POINTER(C.c_int)],
            "evh_sums": [P, D], "evh_download_w": [P, D], "evh_download_evo_s": [P, D, D],
            "evh_last_error": [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int],
            "evh_event_record": [P, C.c_int],
            "evh_event_elapsed": [P, C.c_int, C.c_int, C.POINTER(C.c_float)],
            "evh_synchronize": [P],
            "evh_set_timing": [P, C.c_int],
            "evh_set_tuning": [P, C.c_int, C.c_int, C.c_int],
            "evh_get_tuning": [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "evh_kernel_stats": [P, D, C.POINTER(C.c_longlong)],
            "evh_conduction_stats": [P, D, C.POINTER(C.c_longlong)],
            "evh_stage_stats": [P, C.c_int, D, C.POINTER(C.c_longlong)],
            "evh_gravpotential": [P, C.c_double],
            "evh_gravintegration": [P, C.c_double],
            "evh_download_phi": [P, D],
            "evh_slab": [P, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
            "evh_slab_faces": [P, C.c_void_p, C.c_void_p],
            "evh_slab_temp": [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
            "evh_slab_temp_faces": [P, C.c_void_p, C.c_void_p],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.evh_device_bytes.argtypes = [P]
        L.evh_device_bytes.restype = C.c_longlong
        L.evh_math_eval.argtypes = [C.c_int, C.c_longlong, D, D, D]
        L.evh_math_eval.restype = C.c_int
        _lib = L
    return _lib


def make_config(d: Derived, precision: int = FP64, mach_fp64: float = 5.0) -> EvhConfig:
    c = EvhConfig()
    for name, _ in EvhConfig._fields_:
        if name == "precision":
            c.precision = precision
        elif name == "mach_fp64":
            c.mach_fp64 = mach_fp64
        else:
            setattr(c, name, getattr(d, name))
    return c


MATH_POW, MATH_POW_G1, MATH_RPOW_G1, MATH_DIV, MATH_RCP, MATH_SQRT, MATH_POWF = range(7)


def math_eval(fn: int, x, y=None) -> np.ndarray:
    """The fast forms' device arithmetic f(x, y) elementwise (evh_math_eval)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    code = load_library().evh_math_eval(fn, x.size, _dptr(x), _dptr(y), _dptr(out))
    if code:
        raise EvoraError(code, "evh_math_eval failed")
    return out


def _dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Hydro:
    """One grid's hydro state on the GPU, driven by the reference's calls."""

    def __init__(self, params: HydroParams | Derived, precision: int = FP64, mode: int | None = None,
                 mach_fp64: float = 5.0):
        self.L = load_library()
        self.d = params.derive() if isinstance(params, HydroParams) else params
        self.precision = precision
        self.cfg = make_config(self.d, precision, mach_fp64)
        h = C.c_void_p()
        code = self.L.evh_create(C.byref(self.cfg), C.byref(h))
        if code:
            raise EvoraError(code, "evh_create failed")
        self.h = h
        if mode is not None:
            self._check(self.L.evh_set_mode(self.h, mode))

    # --------------------------------------------------------------- plumbing
    def _check(self, code: int):
        if code:
            c = C.c_int()
            cell = (C.c_int * 3)()
            msg = C.create_string_buffer(512)
            self.L.evh_last_error(self.h, C.byref(c), cell, msg, 512)
            raise EvoraError(code, msg.value.decode(), tuple(cell))

    def close(self):
        if getattr(self, "h", None):
            self.L.evh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape_q(self):
        return (self.d.nz, self.d.ny, self.d.nx, 6)

    def set_mode(self, mode: int):
        self._check(self.L.evh_set_mode(self.h, mode))

    def upload_q(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.shape_q:
            raise ValueError(f"q has shape {q.shape}, expected {self.shape_q}")
        self._check(self.L.evh_upload_q(self.h, _dptr(q)))

    def download_q(self) -> np.ndarray:
        q = np.empty(self.shape_q)
        self._check(self.L.evh_download_q(self.h, _dptr(q)))
        return q

    # --------------------------------------------------------------- reference routines
    def primitive(self):
        """muscl.f90:26-43"""
        self._check(self.L.evh_primitive(self.h))

    def bound(self):
        """boundary.f90:26-113"""
        self._check(self.L.evh_bound(self.h))

    def hydrostep(self, dt: float = 1e99) -> float:
        """timestep.f90:26-57: returns min(cfl*v, dt)."""
        x = C.c_double(dt)
        self._check(self.L.evh_hydrostep(self.h, C.byref(x)))
        return x.value

    def compevo(self, dt: float):
        """muscl.f90:45-113"""
        self._check(self.L.evh_compevo(self.h, dt))

    def sweep(self, dt: float):
        """muscl.f90:171-274"""
        self._check(self.L.evh_sweep(self.h, dt))

    def thermalstep(self, dt: float = 1e99) -> float:
        """timestep.f90:72-104: returns min(c_cond dx^2 min(rho/kappa)/temp0, dt)"""
        x = C.c_double(dt)
        self._check(self.L.evh_thermalstep(self.h, C.byref(x)))
        return x.value

    def thermalconduction(self, dt: float):
        """conduction.f90:26-123 (with the primitive and bound before it)"""
        self._check(self.L.evh_thermalconduction(self.h, dt))

    def gravpotential(self, cosmo: float = 1.0):
        """gravitation.f90:26-52: the potential of the current state (on the device)."""
        self._check(self.L.evh_gravpotential(self.h, cosmo))

    def gravintegration(self, dt: float):
        """gravitation.f90:172-254 with the last potential."""
        self._check(self.L.evh_gravintegration(self.h, dt))

    def download_phi(self) -> np.ndarray:
        """The last potential, shape (nz, ny, nx) == phi(nx, ny, nz)."""
        phi = np.empty((self.d.nz, self.d.ny, self.d.nx))
        self._check(self.L.evh_download_phi(self.h, _dptr(phi)))
        return phi

    # --------------------------------------------------------------- z-slabs
    def slab(self, lo: bool, hi: bool):
        """This grid as one rank's z-slab (parallel.f90:26-66): the ghost planes
        below (lo) and above (hi) come from the neighbours' planes, placed in
        the returned device halo buffers (addresses, None for a side that
        keeps bound's z rules)."""
        a, b = C.c_void_p(), C.c_void_p()
        self._check(self.L.evh_slab(self.h, int(bool(lo)), int(bool(hi)), C.byref(a), C.byref(b)))
        return a.value, b.value

    def slab_faces(self, lo_dst, hi_dst):
        """The two lowest and two highest planes of q into device buffers
        (addresses or None): the planes communicate sends
        (parallel.f90:317-381)."""
        self._check(self.L.evh_slab_faces(self.h, lo_dst, hi_dst))

    def slab_temp(self):
        """With conduction on: the device halo buffers of T for the planes next
        to the slab (addresses, None for a side without a neighbour)."""
        a, b = C.c_void_p(), C.c_void_p()
        self._check(self.L.evh_slab_temp(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def slab_temp_faces(self, lo_dst, hi_dst):
        """T of the slab's lowest and highest planes into device buffers: what
        the neighbours need before thermalconduction (main.f90:170-175)."""
        self._check(self.L.evh_slab_temp_faces(self.h, lo_dst, hi_dst))

    def step(self, t: float, tend: float):
        """One pass of main.f90:109-155; returns (t, dt, go)."""
        tt, dt, go = C.c_double(t), C.c_double(0.0), C.c_int(1)
        self._check(self.L.evh_step(self.h, C.byref(tt), tend, C.byref(dt), C.byref(go)))
        return tt.value, dt.value, bool(go.value)

    def run(self, nsteps: int, t: float = 0.0, tend: float = 1e99):
        """Up to nsteps passes; returns (t, dt, steps_done)."""
        tt, dt, done = C.c_double(t), C.c_double(0.0), C.c_int(0)
        self._check(self.L.evh_run(self.h, nsteps, C.byref(tt), tend, C.byref(dt), C.byref(done)))
        return tt.value, dt.value, done.value

    def sums(self) -> np.ndarray:
        """snapshot.f90:169-172 sums of q1, q5, q6 (no volume factor)."""
        out = np.zeros(3)
        self._check(self.L.evh_sums(self.h, _dptr(out)))
        return out

    # --------------------------------------------------------------- staged views
    def download_w(self) -> np.ndarray:
        w = np.empty((self.d.nz + 4, self.d.ny + 4, self.d.nx + 4, 5))
        self._check(self.L.evh_download_w(self.h, _dptr(w)))
        return w

    def download_evo_s(self):
        evo = np.empty((self.d.nz + 2, self.d.ny + 2, self.d.nx + 2, 5))
        s = np.empty((self.d.nz + 2, self.d.ny + 2, self.d.nx + 2, 3, 5))
        self._check(self.L.evh_download_evo_s(self.h, _dptr(evo), _dptr(s)))
        return evo, s

    # --------------------------------------------------------------- timing
    def event_record(self, slot: int):
        self._check(self.L.evh_event_record(self.h, slot))

    def event_elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(self.L.evh_event_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(self.L.evh_synchronize(self.h))

    def set_tuning(self, tile_ty: int = 12, wpe_hi: bool = True, zchunk: int = 0):
        """Launch shape of the fused kernel: tile height, occupancy target, z chunk
        (0: automatic, shorter chunks on small grids)."""
        self._check(self.L.evh_set_tuning(self.h, tile_ty, int(bool(wpe_hi)), zchunk))

    def get_tuning(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.evh_get_tuning(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, bool(b.value), c.value

    def set_timing(self, on: bool):
        self._check(self.L.evh_set_timing(self.h, int(bool(on))))

    def kernel_stats(self):
        """(total ms, launches) of the sweep kernels since set_timing(True)."""
        ms, n = C.c_double(), C.c_longlong()
        self._check(self.L.evh_kernel_stats(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def conduction_stats(self):
        """(total ms, launches) of the conduction kernel since set_timing(True)."""
        ms, n = C.c_double(), C.c_longlong()
        self._check(self.L.evh_conduction_stats(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def stage_stats(self, stage: int):
        """(total ms, launches) of a stage (STAGE_*) since set_timing(True)."""
        ms, n = C.c_double(), C.c_longlong()
        self._check(self.L.evh_stage_stats(self.h, stage, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def device_bytes(self) -> int:
        return int(self.L.evh_device_bytes(self.h))
