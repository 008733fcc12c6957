"""The geometry generator (Geom.f90) and the geometry file: cell types by
closed form, the dns.config lines it reads, and the solver's view of the
generator's (nyp, mz+1, mx+1) record (quirk Q8)."""
import math
import os

import numpy as np
import pytest

from polymer_dns_amd import geometry as G

HERE = os.path.dirname(os.path.abspath(__file__))


def test_config_lines_read_from_the_golden_config():
    g = G.config_geometry(os.path.join(HERE, "golden", "dns.config"))
    assert g == dict(bftail=0, yl=2.0, zl=4.1887902, xl=12.566371, kwall=0, geomtype=0, flow_select=1)


def test_wall_buffer_and_suction_planes():
    nx, ny, nz = 16, 32, 8
    im = G.make_geometry(nx, ny, nz, 2.0, 2.0, 4.0, kwall=5, bftail=6, flow_select=40)
    mx, mz, nyp = im.shape
    assert (mx, mz, nyp) == (24, 12, 33)
    assert (im[:, :, 4] == 1).all()                        # wall plane k = kwall
    # buffer x indices 2..bftail (1-based) are 6 off the wall and suction planes
    ks = int(math.floor(0.703125 * ny)) - 1                # suction plane (0-based)
    others = [k for k in range(nyp) if k not in (4, ks)]
    assert (im[1:6][:, :, others] == 6).all()
    assert (im[0][:, others] == 0).all() and (im[6:][:, :, others] == 0).all()
    # suction from x index bfwidth/2 + 1 = 3 (1-based) on, over the buffer too
    assert (im[2:, :, ks] == 7).all() and (im[1, :, ks] == 6).all() and (im[0, :, ks] == 0).all()


def test_two_d_brick_closed_form():
    nx, ny, nz, xl, yl = 64, 32, 16, 4.0 * math.pi, 2.0
    im = G.make_geometry(nx, ny, nz, yl, 2.0, xl, geomtype=2)
    mx, nyp = 3 * nx // 2, ny + 1
    dx = xl / mx
    imin, imax = math.floor(0.4 * xl / dx) + 1, math.floor((0.4 * xl + 0.1 * yl) / dx) + 1
    kmin = math.floor(0.1 * nyp)
    kmax = kmin + math.floor(0.1 * nyp)
    want = np.zeros_like(im)
    want[imin - 1:imax, :, kmin - 1:kmax] = 4
    assert np.array_equal(im, want)


def test_three_d_brick_inside_its_bounds():
    """Cells of type 4 lie in 0.1 xl <= x <= 0.2 xl, 0.2 zl <= z <= 0.8 zl and
    ygeom - yl/4 <= y(k) - y(kwall) <= ygeom with y(k) = (1 - cos((k-1) pi/ny)) yl/2,
    ygeom = 0.45 yl, and every cell clearly inside those bounds is type 4."""
    nx, ny, nz, xl, yl, zl, kwall = 32, 32, 32, 4.0 * math.pi, 2.0, 2.0 * math.pi, 3
    im = G.make_geometry(nx, ny, nz, yl, zl, xl, kwall=kwall, geomtype=3)
    mx, mz, nyp = im.shape
    x = np.arange(mx) * xl / mx
    z = np.arange(mz) * zl / mz
    y = (1.0 - np.cos(np.arange(nyp) * math.pi / ny)) * yl / 2.0 - (1.0 - math.cos((kwall - 1) * math.pi / ny)) * yl / 2.0
    ins = lambda v, lo, hi, e: (v >= lo - e) & (v <= hi + e)
    X, Z, Y = np.meshgrid(x, z, y, indexing="ij")
    K = np.broadcast_to(np.arange(nyp)[None, None, :], im.shape)
    loose = ins(X, 0.1 * xl, 0.2 * xl, 1e-5) & ins(Z, 0.2 * zl, 0.8 * zl, 1e-5) & \
        ins(Y, 0.45 * yl - 0.25 * yl, 0.45 * yl, 1e-5) & (K >= kwall - 1)
    tight = ins(X, 0.1 * xl, 0.2 * xl, -1e-5) & ins(Z, 0.2 * zl, 0.8 * zl, -1e-5) & \
        ins(Y, 0.45 * yl - 0.25 * yl, 0.45 * yl, -1e-5) & (K >= kwall - 1)
    brick = im == 4
    assert brick.any()
    assert not (brick & ~loose).any()
    assert brick[tight].all()
    assert (im[:, :, kwall - 1][~brick[:, :, kwall - 1]] == 1).all()


def test_file_round_trip_and_the_solvers_view_of_a_generator_file(tmp_path):
    nx, ny, nz = 8, 8, 8
    rng = np.random.default_rng(4)
    im = rng.integers(0, 8, size=(12, 12, 9)).astype(np.int32)
    G.write_geometry(tmp_path / "g", im)
    assert np.array_equal(G.read_geometry(tmp_path / "g", nx, ny, nz), im)
    # Q8: the generator's array is (nyp, mz+1, mx+1) in Fortran order; the solver
    # reads the first nyp*mz*mx values as (nyp, mz, mx): solver (k, j, i) is file
    # element L = k + nyp (j + mz i) = generator (L mod nyp, (L div nyp) mod (mz+1),
    # L div (nyp (mz+1)))
    G.write_geometry(tmp_path / "q8", im, q8=True)
    got = G.read_geometry(tmp_path / "q8", nx, ny, nz)
    mx, mz, nyp = im.shape
    i, j, k = np.meshgrid(np.arange(mx), np.arange(mz), np.arange(nyp), indexing="ij")
    L = k + nyp * (j + mz * i)
    kk, jj, ii = L % nyp, (L // nyp) % (mz + 1), L // (nyp * (mz + 1))
    big = np.zeros((mx + 1, mz + 1, nyp), dtype=np.int32)
    big[:mx, :mz] = im
    assert np.array_equal(got, big[ii, jj, kk])
    assert not np.array_equal(got, im)
    with pytest.raises(ValueError):
        G.read_geometry(tmp_path / "g", 16, 16, 16)
