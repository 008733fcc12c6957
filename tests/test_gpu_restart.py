"""Restart files through the C ABI: a restart written at half the streamwise
resolution read with irstrt = 2 (main_dns.f90:1748-1809) fills the low kx
modes and the first mx/2 columns of the IB integrals of the full-resolution
channel, bit for bit, with the rest zero; the channel then writes the
full-resolution file (main_dns.f90:776-781) that restart.read_restart reads
back unchanged."""
import numpy as np
import pytest

from oracle import kmm_oracle as O
from polymer_dns_amd import restart
from tests._cases import params, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

SPEC = ("v", "fn", "fnm1", "omy", "gn", "gnm1")
MEAN = ("u0", "h1n", "h1nm1", "w0", "h3n", "h3nm1")
INTG = ("fxintg", "fyintg", "fzintg")


def test_half_resolution_restart(tmp_path):
    nx, ny, nz = 64, 32, 32
    lo = params(nx // 2, ny, nz)
    rng = np.random.default_rng(21)
    st = seeded(lo, 22)
    for k in ("fn", "fnm1", "gn", "gnm1"):
        st[k] = O.norm(rng.standard_normal(st["v"].shape) + 1j * rng.standard_normal(st["v"].shape))
    for k in MEAN[1:3] + MEAN[4:]:
        st[k] = rng.standard_normal(lo.nyp)
    for k in INTG:
        st[k] = rng.standard_normal((lo.mx, lo.mz, lo.nyp))
    src = tmp_path / "restart_lo"
    restart.write_restart(src, st, nx // 2, ny, nz)
    hi = params(nx, ny, nz)
    cfg = kmm.make_config(nx, ny, nz, xl=hi.xl, zl=hi.zl, re=hi.re, dt=hi.dt, r_tau=hi.R_tau, irstrt=2)
    out = tmp_path / "restart_hi"
    with kmm.Channel(cfg) as ch:
        ch.read_restart(src)
        for k in SPEC:
            want = np.zeros((hi.nxh, hi.nz, hi.nyp), dtype=np.complex128)
            want[: lo.nxh] = st[k]
            assert np.array_equal(ch.get(k), want), k
        for k in MEAN:
            assert np.array_equal(ch.get(k), st[k]), k
        ch.write_restart(out)
    back = restart.read_restart(out, nx, ny, nz)
    for k in INTG:
        want = np.zeros((hi.mx, hi.mz, hi.nyp))
        want[: lo.mx] = st[k]
        assert np.array_equal(back[k], want), k
    assert np.array_equal(back["omy"][: lo.nxh], st["omy"])
    assert not np.any(back["v"][lo.nxh:])
