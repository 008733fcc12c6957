"""kmm_solve_noalloc: the slab solve on the library's fixed device scratch,
called through the C ABI and compared with kmm_solve on the same inputs
(identical kernels, so the results agree bit for bit) and with the oracle's
solve (solvers.f90:174-283).  Two calls in a row on one GPU, the second on a
smaller slab, so the second call runs on the array the first one left.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import kmm_oracle as O
from polymer_dns_amd import kmm
from tests._cases import params, rel

pytestmark = pytest.mark.gpu


def _slab(p, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((p.nz, p.nyp)) + 1j * rng.standard_normal((p.nz, p.nyp))
    bt = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    bb = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    return f, bt, bb


def _call(lib, fn, h, f, x, g, dyde, bt, bb, ib):
    u = np.ascontiguousarray(f, dtype=np.complex128).copy()
    bt = np.ascontiguousarray(bt, dtype=np.complex128)
    bb = np.ascontiguousarray(bb, dtype=np.complex128)
    dp = C.POINTER(C.c_double)
    r = getattr(lib, fn)(h, u.ctypes.data_as(dp), C.c_double(x), C.c_double(g), C.c_double(dyde),
                         bt.ctypes.data_as(dp), bb.ctypes.data_as(dp), C.c_int(ib))
    assert r == kmm.KMM_OK, lib.kmm_status_string(r)
    return u


@pytest.mark.parametrize("ib", [0, 1])
def test_noalloc_solve_matches_solve_twice(ib):
    lib = kmm.load_library()
    lib.kmm_solve_noalloc.argtypes = lib.kmm_solve.argtypes
    cases = [(64, 64, 32, 11), (32, 32, 16, 12)]   # second call: a smaller slab
    for nx, ny, nz, seed in cases:
        p = params(nx, ny, nz)
        f, bt, bb = _slab(p, seed + ib)
        x = p.wavx[3] ** 2
        g = 0.0 if ib else 7.5
        if ib:
            x = 0.0   # kx = 0: the Neumann zero mode (kz = 0) takes the guard path
        with kmm.Channel(kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt,
                                         r_tau=p.R_tau)) as ch:
            ref = _call(lib, "kmm_solve", ch.h, f, x, g, p.dyde, bt, bb, ib)
            got = _call(lib, "kmm_solve_noalloc", ch.h, f, x, g, p.dyde, bt, bb, ib)
        assert np.array_equal(got, ref), f"{nx}x{ny + 1}x{nz}: noalloc differs from kmm_solve"
        if not ib:
            want = O.solve(f, p.wavz ** 2 + x + g, p.dyde, bt, bb, ib)
            assert rel(got, want) < 1e-11


def test_noalloc_solve_rejects_oversized_slab():
    lib = kmm.load_library()
    lib.kmm_solve_noalloc.argtypes = lib.kmm_solve.argtypes
    p = params(16, 128, 1024)   # nyp * nz = 132 096 > 131 072
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau)
    with kmm.Channel(cfg) as ch:
        f = np.zeros((p.nz, p.nyp), dtype=np.complex128)
        b = np.zeros(p.nz, dtype=np.complex128)
        dp = C.POINTER(C.c_double)
        r = lib.kmm_solve_noalloc(ch.h, f.ctypes.data_as(dp), 0.0, 1.0, p.dyde, b.ctypes.data_as(dp),
                                  b.ctypes.data_as(dp), 0)
    assert r == kmm.KMM_E_ARG
