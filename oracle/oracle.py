"""Python access to the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps

* liboracle.so -- the C restatement of the hydro step (evora_oracle.c), and
* _ref/evref   -- the reference's own Fortran hot path linked with
                  ref_driver.f90 (built from the reference sources when they
                  are present; the built binary travels with the tree).

Arrays use the reference's Fortran layouts, seen from numpy in C order:
  q   (nz, ny, nx, 6)            = q(nq, nx, ny, nz)
  w   (nz+4, ny+4, nx+4, 5)      = w(nw, -1:nx+2, -1:ny+2, -1:nz+2)
  evo (nz+2, ny+2, nx+2, 5)      = evo(nw, 0:nx+1, ...)
  s   (nz+2, ny+2, nx+2, 3, 5)   = s(nw, 3, 0:nx+1, ...)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "evref")

ERRORS = {
    0: "ok",
    1: "negative pressure",
    2: "riemann unphysical pressure",
    3: "hll",
    4: "hllc",
    5: "dt",
    6: "solver",
    7: "limiter",
    8: "config",
}


class EvParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("boundtype", C.c_int), ("solver", C.c_int), ("limiter", C.c_int),
        ("secondorder", C.c_int), ("dualenergy", C.c_int),
        ("onedimension", C.c_int), ("twodimension", C.c_int),
        ("g", C.c_double), ("g1", C.c_double), ("g1inv", C.c_double), ("gg1", C.c_double),
        ("cfl", C.c_double), ("de_thresh1", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dxinv", C.c_double), ("dyinv", C.c_double), ("dzinv", C.c_double),
        ("volume", C.c_double),
        ("conduction", C.c_int), ("simplekappa", C.c_int),
        ("c_cond", C.c_double), ("temp0", C.c_double), ("kappa0", C.c_double),
        ("gravity", C.c_int), ("supercomoving", C.c_int),
        ("grav", C.c_double), ("fb", C.c_double), ("h2", C.c_double),
        ("kfx2", C.c_double), ("kfy2", C.c_double), ("kfz2", C.c_double),
        ("nyqx", C.c_int), ("nyqy", C.c_int), ("nyqz", C.c_int),
        ("dx2inv", C.c_double), ("dy2inv", C.c_double), ("dz2inv", C.c_double),
        ("hydro_off", C.c_int),
    ]


def ev_params(d) -> EvParams:
    """EvParams from an evora_amd.params.Derived."""
    p = EvParams()
    for name, _ in EvParams._fields_:
        setattr(p, name, getattr(d, name))
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(EvParams)
        D = C.POINTER(C.c_double)
        for name, args in {
            "ev_primitive": [P, D, D],
            "ev_bound": [P, D],
            "ev_hydrostep": [P, D, D],
            "ev_compevo": [P, D, D, D, C.c_double],
            "ev_sweep": [P, D, D, D, D, C.c_double],
            "ev_riemannsolver": [P, D, D, D, C.c_int],
            "ev_comp_slope": [P, D, D, D, D],
            "ev_q2w": [P, D, D],
            "ev_step": [P, D, D, D, D, D, C.c_double, D, C.POINTER(C.c_int)],
            "ev_run": [P, D, C.c_int, D, C.c_double, D, C.POINTER(C.c_int)],
            "ev_thermalstep": [P, D, D, D],
            "ev_thermalconduction": [P, D, D, C.c_double],
            "ev_gravpotential": [P, D, D, C.c_double],
            "ev_gravintegration": [P, D, D, C.c_double],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.ev_sums.argtypes = [P, D, D]
        L.ev_sums.restype = None
        for name in ("ev_dft3", "ev_multigreen"):
            fn = getattr(L, name)
            fn.argtypes = [P, D, D] + ([C.c_int] if name == "ev_dft3" else [])
            fn.restype = None
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    def __init__(self, code):
        cell = tuple((C.c_int * 3).in_dll(lib(), "ev_err_cell"))
        super().__init__(f"oracle error {code} ({ERRORS.get(code, '?')}) at cell {cell}")
        self.code = code
        self.cell = cell


def _check(code):
    if code:
        raise OracleError(code)


def new_w(d):
    return np.zeros((d.nz + 4, d.ny + 4, d.nx + 4, 5))


def new_evo(d):
    return np.zeros((d.nz + 2, d.ny + 2, d.nx + 2, 5))


def new_s(d):
    return np.zeros((d.nz + 2, d.ny + 2, d.nx + 2, 3, 5))


class Oracle:
    """The reference's five step-level routines on host arrays."""

    def __init__(self, d):
        self.d = d
        self.p = ev_params(d)
        self.L = lib()

    def primitive(self, w, q):
        _check(self.L.ev_primitive(C.byref(self.p), _ptr(w), _ptr(q)))

    def bound(self, w):
        _check(self.L.ev_bound(C.byref(self.p), _ptr(w)))

    def hydrostep(self, w, dt=1e99):
        x = C.c_double(dt)
        _check(self.L.ev_hydrostep(C.byref(self.p), _ptr(w), C.byref(x)))
        return x.value

    def compevo(self, w, evo, s, dt):
        _check(self.L.ev_compevo(C.byref(self.p), _ptr(w), _ptr(evo), _ptr(s), dt))

    def sweep(self, w, q, evo, s, dt):
        _check(self.L.ev_sweep(C.byref(self.p), _ptr(w), _ptr(q), _ptr(evo), _ptr(s), dt))

    def riemann(self, wl, wr, dir):
        f = np.zeros(6)
        _check(self.L.ev_riemannsolver(C.byref(self.p), _ptr(np.ascontiguousarray(wl, dtype=np.float64)),
                                       _ptr(np.ascontiguousarray(wr, dtype=np.float64)), _ptr(f), dir))
        return f

    def comp_slope(self, wl, w, wr):
        s = np.zeros(5)
        _check(self.L.ev_comp_slope(C.byref(self.p), _ptr(np.ascontiguousarray(wl, dtype=np.float64)),
                                    _ptr(np.ascontiguousarray(w, dtype=np.float64)),
                                    _ptr(np.ascontiguousarray(wr, dtype=np.float64)), _ptr(s)))
        return s

    def stage(self, q, t=0.0, tend=1e99):
        """One pass of the step; returns w, evo, s (time n), q^{n+1}, dt."""
        q = q.copy()
        w, evo, s = new_w(self.d), new_evo(self.d), new_s(self.d)
        self.primitive(w, q)
        self.bound(w)
        dt = self.hydrostep(w, 1e99)
        if tend - t <= dt:
            dt = tend - t
        if self.d.secondorder:
            self.compevo(w, evo, s, dt)
        w_out, evo_out, s_out = w.copy(), evo.copy(), s.copy()
        self.sweep(w, q, evo, s, dt)
        return w_out, evo_out, s_out, q, dt

    def run(self, q, nsteps, t=0.0, tend=1e99):
        """nsteps passes of main.f90:109-155 in place on q; returns (t, dt, steps)."""
        tt = C.c_double(t)
        dt = C.c_double(0.0)
        done = C.c_int(0)
        _check(self.L.ev_run(C.byref(self.p), _ptr(q), nsteps, C.byref(tt), tend, C.byref(dt), C.byref(done)))
        return tt.value, dt.value, done.value

    def sums(self, q):
        out = np.zeros(3)
        self.L.ev_sums(C.byref(self.p), _ptr(q), _ptr(out))
        return out

    # gravity (gravitation.f90); phi and complex fields have shape (nz, ny, nx)
    def dft3(self, a, sign):
        """FFTW's unnormalised transform of fft.f90 (sign -1 forward, +1 backward)."""
        re = np.ascontiguousarray(a.real, dtype=np.float64)
        im = np.ascontiguousarray(a.imag, dtype=np.float64)
        self.L.ev_dft3(C.byref(self.p), _ptr(re), _ptr(im), sign)
        return re + 1j * im

    def multigreen(self, a):
        re = np.ascontiguousarray(a.real, dtype=np.float64)
        im = np.ascontiguousarray(a.imag, dtype=np.float64)
        self.L.ev_multigreen(C.byref(self.p), _ptr(re), _ptr(im))
        return re + 1j * im

    def gravpotential(self, q, cosmo=1.0):
        phi = np.zeros((self.d.nz, self.d.ny, self.d.nx))
        _check(self.L.ev_gravpotential(C.byref(self.p), _ptr(q), _ptr(phi), cosmo))
        return phi

    def gravintegration(self, q, phi, dt):
        _check(self.L.ev_gravintegration(C.byref(self.p), _ptr(q), _ptr(np.ascontiguousarray(phi)), dt))


# ---------------------------------------------------------------- reference run
def have_ref() -> bool:
    return os.path.exists(REF_PATH)


def run_ref(params, mode: str, data: np.ndarray, nsteps: int = 1):
    """Run the reference's own hot path (oracle/_ref/evref).

    Returns (out_bytes_as_float64, stdout).  `params` is an
    evora_amd.params.HydroParams; its namelist is what the reference reads.
    """
    if not have_ref():
        raise RuntimeError("oracle/_ref/evref is not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        para = os.path.join(td, "para")
        with open(para, "w") as fh:
            fh.write(params.to_namelist())
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        np.ascontiguousarray(data, dtype=np.float64).tofile(fin)
        r = subprocess.run([REF_PATH, para, mode, fin, fout, str(nsteps)], capture_output=True, text=True,
                           cwd=td, timeout=600)
        out = np.fromfile(fout, dtype=np.float64) if os.path.exists(fout) else None
        return out, r.stdout + r.stderr


def ref_stage(params, q):
    """One reference pass: w, evo, s (time n), q^{n+1}, dt."""
    d = params.derive()
    out, log = run_ref(params, "stage", q, 1)
    if out is None or "ERROR" in log:
        raise RuntimeError("reference stage failed:\n" + log)
    nw_ = 5 * (d.nx + 4) * (d.ny + 4) * (d.nz + 4)
    ne = 5 * (d.nx + 2) * (d.ny + 2) * (d.nz + 2)
    nq = 6 * d.ncells
    o = 0
    w = out[o:o + nw_].reshape(d.nz + 4, d.ny + 4, d.nx + 4, 5); o += nw_
    evo = out[o:o + ne].reshape(d.nz + 2, d.ny + 2, d.nx + 2, 5); o += ne
    s = out[o:o + 3 * ne].reshape(d.nz + 2, d.ny + 2, d.nx + 2, 3, 5); o += 3 * ne
    q1 = out[o:o + nq].reshape(d.nz, d.ny, d.nx, 6); o += nq
    dt = out[o]
    return w, evo, s, q1, dt


def ref_green(params, a):
    """The reference's multigreen on the complex field a (nz, ny, nx)."""
    raw = np.empty(a.shape + (2,))
    raw[..., 0], raw[..., 1] = a.real, a.imag
    out, log = run_ref(params, "green", raw, 1)
    if out is None:
        raise RuntimeError("reference green failed:\n" + log)
    out = out.reshape(a.shape + (2,))
    return out[..., 0] + 1j * out[..., 1]


def ref_gravint(params, q, phi, dt):
    """The reference's gravintegration of q (nz, ny, nx, 6) with potential phi and step dt."""
    data = np.concatenate([np.ravel(q), np.ravel(phi), [dt]])
    out, log = run_ref(params, "gravint", data, 1)
    if out is None:
        raise RuntimeError("reference gravint failed:\n" + log)
    return out.reshape(q.shape)


def ref_run(params, q, nsteps):
    """nsteps reference passes; returns (q_final, [(t, dt), ...], log)."""
    out, log = run_ref(params, "run", q, nsteps)
    steps = []
    for line in log.splitlines():
        if line.startswith("STEP"):
            parts = line.split()
            steps.append((float(parts[2]), float(parts[3])))
    qf = None if out is None else out.reshape(q.shape)
    return qf, steps, log
