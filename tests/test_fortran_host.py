"""Fortran host pieces on the CPU: the iso_c_binding module and dispatch
compile against the reference's module global, and the ic(q) plugin reads
the reference dump format written by evora_amd.dump."""
import os
import subprocess

import numpy as np
import pytest

from evora_amd import dump

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = os.path.join(ROOT, "oracle", "_ref")
FC = "/opt/rocm/lib/llvm/bin/flang"


def need_ref_objects():
    if not (os.path.exists(os.path.join(REF, "global.mod")) and os.path.exists(FC)):
        pytest.skip("reference objects (make -C oracle ref) or flang absent")


def test_dispatch_and_binding_compile(tmp_path):
    need_ref_objects()
    for src in ("evora_hip.f90", "hydro_dispatch.f90"):
        r = subprocess.run([FC, "-cpp", "-O2", "-c", "-I", REF, "-J", str(tmp_path),
                            os.path.join(ROOT, "evora_amd", "fortran", src), "-o", str(tmp_path / (src + ".o"))],
                           capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 0, r.stderr


def test_ic_readdump_reads_reference_format(tmp_path):
    need_ref_objects()
    exe = tmp_path / "readdump_check"
    r = subprocess.run([FC, "-cpp", "-O2", "-I", REF, os.path.join(ROOT, "tests", "fortran", "readdump_check.f90"),
                        os.path.join(ROOT, "evora_amd", "fortran", "ic_readdump.f90"), os.path.join(REF, "global.o"),
                        "-o", str(exe), "-Wl,--unresolved-symbols=ignore-all"], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    q = np.random.default_rng(3).standard_normal((5, 4, 7, 6))
    dump.write_dump(str(tmp_path / "ic.dump"), q)
    out = tmp_path / "q.bin"
    r = subprocess.run([str(exe), str(out), "7", "4", "5"], env=dict(os.environ, EVORA_IC_DUMP=str(tmp_path / "ic.dump")),
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and "ERROR" not in r.stdout, r.stdout
    assert np.array_equal(np.fromfile(out).reshape(q.shape), q)
    qq, hdr = dump.read_dump(str(tmp_path / "ic.dump"))
    assert np.array_equal(qq, q) and hdr == dict(nn=6, nx=7, ny=4, nz=5, lz=1, mz=5)
