"""Python mirror of the batched form (`evh_batch_*` in include/evora_hip.h).

`Batch` advances many independent 3-D boxes of uneven sizes together: each
box is its own evora run (own t, dt from its own CFL limit, own error stop),
and one step of the ensemble is a device-side dt kernel plus one persistent
sweep kernel over a (box, tile) work list, optionally replayed from a
hipGraph so that the host never synchronises between steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .hydro import FP64, EvhConfig, EvoraError, _dptr, load_library, make_config
from .params import HydroParams

_SIG = None


def _lib():
    global _SIG
    L = load_library()
    if _SIG is None:
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        sig = {
            "evh_batch_create": [C.POINTER(EvhConfig), C.c_int, C.POINTER(C.c_void_p)],
            "evh_batch_destroy": [P],
            "evh_batch_upload_q": [P, C.c_int, D],
            "evh_batch_download_q": [P, C.c_int, D],
            "evh_batch_set_time": [P, D, D],
            "evh_batch_run": [P, C.c_int, C.c_int, C.POINTER(C.c_float)],
            "evh_batch_times": [P, D, D, I, I],
            "evh_batch_set_tuning": [P, C.c_int, C.c_int, C.c_int, C.c_int],
            "evh_batch_get_tuning": [P, I, I, I, I],
            "evh_batch_autotune": [P, C.c_int, C.c_char_p, C.c_int],
            "evh_batch_last_error": [P, I, I, I, C.c_char_p, C.c_int],
            "evh_batch_run_noalloc": [C.POINTER(EvhConfig), C.c_int, P, P, I, D, D, I, I, C.c_int, C.c_char_p,
                                      C.c_int],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _SIG = True
    return L


class Batch:
    """Many boxes on one GPU; box i is params[i] (HydroParams or Derived)."""

    def __init__(self, params, precision: int = FP64, mach_fp64: float = 5.0):
        self.L = _lib()
        self.d = [p.derive() if isinstance(p, HydroParams) else p for p in params]
        cfgs = (EvhConfig * len(self.d))(*[make_config(d, precision, mach_fp64) for d in self.d])
        h = C.c_void_p()
        code = self.L.evh_batch_create(cfgs, len(self.d), C.byref(h))
        if code:
            raise EvoraError(code, "evh_batch_create failed")
        self.h = h
        self.n = len(self.d)

    def _check(self, code):
        if code:
            c, box = C.c_int(), C.c_int()
            cell = (C.c_int * 3)()
            msg = C.create_string_buffer(256)
            self.L.evh_batch_last_error(self.h, C.byref(c), C.byref(box), cell, msg, 256)
            err = EvoraError(code, msg.value.decode(), tuple(cell))
            err.box = box.value
            raise err

    def close(self):
        if getattr(self, "h", None):
            self.L.evh_batch_destroy(self.h)
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

    def upload_q(self, box: int, q: np.ndarray):
        d = self.d[box]
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != (d.nz, d.ny, d.nx, 6):
            raise ValueError(f"box {box}: q has shape {q.shape}")
        self._check(self.L.evh_batch_upload_q(self.h, box, _dptr(q)))

    def download_q(self, box: int) -> np.ndarray:
        d = self.d[box]
        q = np.empty((d.nz, d.ny, d.nx, 6))
        self._check(self.L.evh_batch_download_q(self.h, box, _dptr(q)))
        return q

    def set_time(self, t=None, tend=None):
        tt = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
        te = None if tend is None else np.ascontiguousarray(tend, dtype=np.float64)
        self._check(self.L.evh_batch_set_time(self.h, None if tt is None else _dptr(tt),
                                              None if te is None else _dptr(te)))

    def run(self, nsteps: int, graph: bool = True) -> float:
        """Advance every running box nsteps steps; returns device milliseconds."""
        ms = C.c_float()
        self._check(self.L.evh_batch_run(self.h, nsteps, int(bool(graph)), C.byref(ms)))
        return ms.value

    def times(self):
        """Per-box (t, last dt, steps done, go)."""
        t, dt = np.empty(self.n), np.empty(self.n)
        steps, go = np.empty(self.n, dtype=np.int32), np.empty(self.n, dtype=np.int32)
        self._check(self.L.evh_batch_times(self.h, _dptr(t), _dptr(dt), steps.ctypes.data_as(C.POINTER(C.c_int)),
                                           go.ctypes.data_as(C.POINTER(C.c_int))))
        return t, dt, steps, go

    def set_tuning(self, tile_ty=12, wpe_hi=True, zchunk=64, blocks_per_cu=2):
        self._check(self.L.evh_batch_set_tuning(self.h, tile_ty, int(bool(wpe_hi)), zchunk, blocks_per_cu))

    def get_tuning(self):
        v = [C.c_int() for _ in range(4)]
        self._check(self.L.evh_batch_get_tuning(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, trial_steps: int = 3) -> str:
        """Time every launch shape on a snapshot, keep the fastest; returns the log."""
        buf = C.create_string_buffer(1 << 16)
        self._check(self.L.evh_batch_autotune(self.h, trial_steps, buf, 1 << 16))
        return buf.value.decode()


def run_noalloc(params, q0_ptr: int, q1_ptr: int, side, t, tend, steps, go, nsteps: int,
                precision: int = FP64, mach_fp64: float = 5.0) -> int:
    """evh_batch_run_noalloc: nsteps batched steps on the caller's device
    buffers (addresses q0_ptr, q1_ptr: box i's SoA state at element offset
    6 * sum of the earlier boxes' cells, in buffer side[i]), with nothing
    allocated on the device.  side, t, steps, go are updated in place
    (numpy int32 / float64 arrays); returns the status (0 or the first
    box's error code) and raises on argument errors."""
    L = _lib()
    d = [p.derive() if isinstance(p, HydroParams) else p for p in params]
    cfgs = (EvhConfig * len(d))(*[make_config(x, precision, mach_fp64) for x in d])
    I = C.POINTER(C.c_int)
    err = C.create_string_buffer(256)
    for a, dt_ in ((side, np.int32), (t, np.float64), (tend, np.float64), (steps, np.int32), (go, np.int32)):
        if a.dtype != dt_ or not a.flags.c_contiguous or a.size != len(d):
            raise ValueError("side/steps/go: int32, t/tend: float64, one value per box")
    code = L.evh_batch_run_noalloc(cfgs, len(d), C.c_void_p(q0_ptr), C.c_void_p(q1_ptr), side.ctypes.data_as(I),
                                   _dptr(t), _dptr(tend), steps.ctypes.data_as(I), go.ctypes.data_as(I), nsteps,
                                   err, 256)
    if code in (8, 9, 10):
        raise EvoraError(code, err.value.decode())
    return code
