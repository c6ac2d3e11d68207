"""The reference's binary dump format (source/write.f90:47-65, reader 179-214).

A dump is two Fortran sequential unformatted records, each framed by a 4-byte
little-endian length marker before and after:
  record 1: nn, nx, ny, nz, lz, mz    six int32
  record 2: a(nn, 1:nx, 1:ny, lz:mz)  float64, component fastest

Files are named data/<name>.<rank:3> (write.f90:38-41).  Records are written
whole, so this reader and writer cover dumps below 2 GiB per record; larger
records would need the compiler's sub-record framing and are rejected.
"""
from __future__ import annotations

import struct

import numpy as np

_MAX_RECORD = 2**31 - 1


def write_dump(path: str, q: np.ndarray, lz: int = 1, mz: int | None = None, nz_global: int | None = None) -> None:
    """Write q (numpy shape (mz-lz+1, ny, nx, nn)) as one rank's dump."""
    q = np.ascontiguousarray(q, dtype="<f8")
    nz, ny, nx, nn = q.shape
    if mz is None:
        mz = lz + nz - 1
    nbytes = q.nbytes
    if nbytes > _MAX_RECORD:
        raise ValueError(f"record of {nbytes} bytes exceeds one Fortran record")
    with open(path, "wb") as fh:
        head = struct.pack("<6i", nn, nx, ny, nz_global if nz_global is not None else mz, lz, mz)
        fh.write(struct.pack("<i", len(head)) + head + struct.pack("<i", len(head)))
        fh.write(struct.pack("<i", nbytes))
        fh.write(q.tobytes())
        fh.write(struct.pack("<i", nbytes))


def read_dump(path: str):
    """Return (q, header) with q shaped (mz-lz+1, ny, nx, nn)."""
    with open(path, "rb") as fh:
        (m1,) = struct.unpack("<i", fh.read(4))
        if m1 != 24:
            raise ValueError(f"{path}: first record is {m1} bytes, expected 24")
        nn, nx, ny, nz, lz, mz = struct.unpack("<6i", fh.read(24))
        (m1b,) = struct.unpack("<i", fh.read(4))
        (m2,) = struct.unpack("<i", fh.read(4))
        nloc = mz - lz + 1
        want = 8 * nn * nx * ny * nloc
        if m1b != 24 or m2 != want:
            raise ValueError(f"{path}: record markers {m1b}/{m2} do not match header {nn,nx,ny,lz,mz}")
        q = np.frombuffer(fh.read(want), dtype="<f8").reshape(nloc, ny, nx, nn).copy()
        (m2b,) = struct.unpack("<i", fh.read(4))
        if m2b != want:
            raise ValueError(f"{path}: trailing marker {m2b} != {want}")
    return q, dict(nn=nn, nx=nx, ny=ny, nz=nz, lz=lz, mz=mz)
