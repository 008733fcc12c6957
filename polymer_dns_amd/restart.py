"""Fortran unformatted sequential files of the reference: ``restart`` /
``last-restart`` (main_dns.f90:776-781 write, 1742-1747 read), the scalar
records of ``c-restart`` (782-793, 1910-1916) and the geometry file
(main_dns.f90:1120-1122).  Records carry a 4-byte little-endian length before
and after the payload.

Arrays use the reference layout (see kmm.py): complex (nxh, nz, nyp),
mean real (nyp,), integrals real (mx, mz, nyp), geometry int32 (mx, mz, nyp).
"""
from __future__ import annotations

import struct

import numpy as np

REC1 = ["v", "fn", "fnm1", "omy", "gn", "gnm1"]
REC1_MEAN = ["u0", "h1n", "h1nm1", "w0", "h3n", "h3nm1"]
REC2 = ["fxintg", "fyintg", "fzintg"]


def write_record(f, arrays):
    payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)
    if len(payload) > 0x7FFFFFFF:
        raise ValueError("record larger than 2 GiB")
    m = struct.pack("<I", len(payload))
    f.write(m)
    f.write(payload)
    f.write(m)


def read_record(f):
    h = f.read(4)
    if len(h) != 4:
        raise EOFError("missing record marker")
    (n,) = struct.unpack("<I", h)
    data = f.read(n)
    t = f.read(4)
    if len(data) != n or len(t) != 4 or struct.unpack("<I", t)[0] != n:
        raise ValueError("truncated or malformed record")
    return data


def write_restart(path, st, nx, ny, nz):
    nxh, nyp, mx, mz = nx // 2, ny + 1, 3 * nx // 2, 3 * nz // 2
    with open(path, "wb") as f:
        arrs = [np.asarray(st[k], dtype=np.complex128).reshape(nxh, nz, nyp) for k in REC1]
        arrs += [np.asarray(st[k], dtype=np.float64).reshape(nyp) for k in REC1_MEAN]
        write_record(f, arrs)
        zero = np.zeros((mx, mz, nyp))
        write_record(f, [np.asarray(st.get(k, zero), dtype=np.float64).reshape(mx, mz, nyp) for k in REC2])


def read_restart(path, nx, ny, nz):
    nxh, nyp, mx, mz = nx // 2, ny + 1, 3 * nx // 2, 3 * nz // 2
    out = {}
    with open(path, "rb") as f:
        d = read_record(f)
        S = nxh * nz * nyp
        if len(d) != 16 * S * 6 + 8 * nyp * 6:
            raise ValueError("restart record 1 does not match the grid")
        off = 0
        for k in REC1:
            out[k] = np.frombuffer(d, dtype=np.complex128, count=S, offset=off).reshape(nxh, nz, nyp).copy()
            off += 16 * S
        for k in REC1_MEAN:
            out[k] = np.frombuffer(d, dtype=np.float64, count=nyp, offset=off).copy()
            off += 8 * nyp
        d = read_record(f)
        Pn = mx * mz * nyp
        if len(d) != 8 * Pn * 3:
            raise ValueError("restart record 2 does not match the grid")
        for i, k in enumerate(REC2):
            out[k] = np.frombuffer(d, dtype=np.float64, count=Pn, offset=8 * Pn * i).reshape(mx, mz, nyp).copy()
    return out


def write_crestart_scalar(path, scalar, scn, scnm1):
    with open(path, "wb") as f:
        for a in (scalar, scn, scnm1):
            write_record(f, [np.asarray(a, dtype=np.complex128)])


def read_crestart_scalar(path, nx, ny, nz):
    nxh, nyp = nx // 2, ny + 1
    out = []
    with open(path, "rb") as f:
        for _ in range(3):
            d = read_record(f)
            out.append(np.frombuffer(d, dtype=np.complex128).reshape(nxh, nz, nyp).copy())
    return dict(zip(("scalar", "scn", "scnm1"), out))


def write_geometry(path, imatrix):
    """Geometry file read by main_dns.f90:1120-1122 as integer(nyp,mz,mx)."""
    with open(path, "wb") as f:
        write_record(f, [np.asarray(imatrix, dtype=np.int32)])


def read_geometry(path, nx, ny, nz):
    mx, mz, nyp = 3 * nx // 2, 3 * nz // 2, ny + 1
    with open(path, "rb") as f:
        d = read_record(f)
    if len(d) != 4 * mx * mz * nyp:
        raise ValueError("geometry record does not match (nyp,mz,mx)")
    return np.frombuffer(d, dtype=np.int32).reshape(mx, mz, nyp).copy()


def rib_geometry(nx, ny, nz, y_cells, x_center, x_width):
    """Spanwise rib of solid (imatrix=1) cells on the index-1 (y=+1) wall:
    y indices 0..y_cells-1, x indices centred at x_center, all z."""
    mx, mz, nyp = 3 * nx // 2, 3 * nz // 2, ny + 1
    im = np.zeros((mx, mz, nyp), dtype=np.int32)
    x0 = x_center - x_width // 2
    im[x0:x0 + x_width, :, :y_cells] = 1
    return im
