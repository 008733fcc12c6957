"""Geometry (imatrix) generation and the geometry file, restating
FFTW_DNS/code/src/geometry/Geom.f90:1-177.

Cell types written: 1 wall plane (y index kwall), 6 buffer region (x indices
2..bftail), 4 brick (geomtype 3: 3-D brick; geomtype 2: 2-D brick), 7 suction
plane (flow_select 40).  Geom.f90 is built by the reference's preconfigure
step with the compiler's default REAL (4 bytes, no -r8 flag,
preconfigure.py:84,144), so coordinates and bounds are evaluated in float32
here as well.

The file is one Fortran unformatted sequential record of int32.  Geom.f90
declares imatrix(nyp, mz+1, mx+1) and leaves the j = mz+1 and i = mx+1 planes
unset, while the solver reads imatrix(nyp, mz, mx) from the same record
(main_dns.f90:1120-1122): the solver sees the first nyp*mz*mx values of the
larger array, a shifted copy of the generated cells.  ``write_geometry(...,
q8=True)`` reproduces that file (unset planes written as 0); the default
writes the (nyp, mz, mx) array the solver is meant to see.

Arrays here are C-ordered (mx, mz, nyp), i.e. Fortran (nyp, mz, mx), the
layout ``Channel.set_geometry`` takes.
"""
from __future__ import annotations

import struct

import numpy as np

f32 = np.float32


def config_geometry(path):
    """The dns.config values Geom.f90 reads (lines 13-16, 20, 32, 34):
    bftail, yl, zl, xl, kwall, geomtype, flow_select."""
    with open(path) as f:
        lines = f.read().splitlines()

    def num(n, typ):
        tok = lines[n - 1].split()[0].replace("d", "e").replace("D", "e")
        return typ(float(tok)) if typ is int else typ(tok)

    return dict(bftail=num(13, int), yl=num(14, float), zl=num(15, float), xl=num(16, float),
                kwall=num(20, int), geomtype=num(32, int), flow_select=num(34, int))


def make_geometry(nx, ny, nz, yl, zl, xl, kwall=0, bftail=0, geomtype=0, flow_select=1):
    """imatrix of Geom.f90 for the 3/2 grid (mx = 3 nx/2, mz = 3 nz/2),
    C-ordered (mx, mz, nyp); Fortran indices i, j, k are 1-based there,
    0-based here."""
    mx, mz, nyp = 3 * nx // 2, 3 * nz // 2, ny + 1
    pi = f32(2.0) * np.arccos(f32(0.0))
    k1 = np.arange(nyp, dtype=np.float32)   # float(k-1)
    ycoord = (f32(1.0) - np.cos(k1 * pi / f32(ny))) * (f32(yl) / f32(2.0))
    delxm = f32(xl) / f32(mx)
    delzm = f32(zl) / f32(mz)
    bfhead = 2
    bfwidth = bftail - bfhead
    im = np.zeros((mx, mz, nyp), dtype=np.int32)
    # wall plane
    if 1 <= kwall <= nyp:
        im[:, :, kwall - 1] = 1
    # buffer region: i = bfhead..bftail, where not already set
    for i in range(bfhead, bftail + 1):
        if 1 <= i <= mx:
            col = im[i - 1]
            col[col == 0] = 6
    if geomtype == 3:   # 3-D brick
        xmin = f32(xl) * f32(0.1)
        xmax = xmin + f32(0.1) * f32(xl)
        zmin = f32(zl) * f32(0.2)
        zmax = f32(zl) * f32(0.8)
        ygeom = f32(yl) / f32(2.0) * f32(0.9)
        kbuff = kwall
        if not 1 <= kbuff <= nyp:
            raise ValueError("geomtype 3 measures the brick from the wall plane: kwall must be in 1..ny+1")
        ybuff = ycoord[kbuff - 1]
        ywidth = f32(0.25) * f32(yl)
        ymax = ygeom
        ymin = ygeom - ywidth
        for i in range(bftail + 1, mx + 1):
            xloc = f32(i - 1) * delxm
            if xloc <= xmax and xloc >= xmin:
                for j in range(1, mz + 1):
                    zloc = f32(j - 1) * delzm
                    if zloc <= zmax and zloc >= zmin:
                        for k in range(kbuff, nyp + 1):
                            yloc = ycoord[k - 1] - ybuff
                            if yloc <= ymax and yloc >= ymin:
                                im[i - 1, j - 1, k - 1] = 4
    if geomtype == 2:   # 2-D brick (the first of the two kmin2 entries only, as the loop runs n = 1,1)
        geomstart = f32(0.4) * f32(xl)
        geomwidth = f32(0.1) * f32(yl)
        imin = int(np.floor(geomstart / delxm)) + 1
        imax = int(np.floor((geomstart + geomwidth) / delxm)) + 1
        kmin = int(np.floor(f32(0.1) * f32(nyp)))
        kmax = kmin + int(np.floor(f32(0.1) * f32(nyp)))
        for i in range(imin, imax + 1):
            for k in range(kmin, kmax + 1):
                if 1 <= i <= mx and 1 <= k <= nyp:
                    im[i - 1, :, k - 1] = 4
    if flow_select == 40:   # suction plane
        imin = bfwidth // 2 + 1 if bfwidth >= 0 else -((-bfwidth) // 2) + 1
        k = int(np.floor(f32(0.703125) * f32(ny)))
        if 1 <= k <= nyp:
            for i in range(max(imin, 1), mx + 1):
                im[i - 1, :, k - 1] = 7
    return im


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_geometry(path, imatrix, q8=False):
    """One unformatted record of int32.  q8: Geom.f90's (nyp, mz+1, mx+1)
    array (the extra planes 0) instead of the (nyp, mz, mx) one."""
    im = np.ascontiguousarray(imatrix, dtype=np.int32)
    if q8:
        mx, mz, nyp = im.shape
        big = np.zeros((mx + 1, mz + 1, nyp), dtype=np.int32)
        big[:mx, :mz, :] = im
        im = big
    with open(path, "wb") as f:
        f.write(_record(im.astype("<i4").tobytes()))


def read_geometry(path, nx, ny, nz):
    """What the solver reads (main_dns.f90:1120-1122): the first
    nyp*mz*mx int32 of the record as imatrix(nyp, mz, mx), returned
    C-ordered (mx, mz, nyp)."""
    mx, mz, nyp = 3 * nx // 2, 3 * nz // 2, ny + 1
    with open(path, "rb") as f:
        head = f.read(4)
        if len(head) != 4:
            raise ValueError(f"{path}: empty geometry file")
        (nbytes,) = struct.unpack("<i", head)
        need = 4 * nyp * mz * mx
        if nbytes < need:
            raise ValueError(f"{path}: record of {nbytes} bytes, the solver reads {need}")
        data = np.frombuffer(f.read(need), dtype="<i4")
    return data.reshape(mx, mz, nyp).astype(np.int32)
