"""The reference's new-run initial fields (init_flow.f90) and the conversion
of physical fields to the spectral state (main_dns.f90:1923-2005): closed
forms of the profiles and a synthesis round trip."""
import math

import numpy as np
import pytest

from oracle import kmm_oracle as O
from polymer_dns_amd import ic
from tests._cases import params, seeded


def test_channel_and_couette_profiles_closed_forms():
    ny, re, R_tau = 16, 2800.0, math.sqrt(3.0 * 2800.0)
    y = ic.solver_ycoord(ny)
    u, v, w = ic.init_flow(1, 8, ny, 8, 2.0, re, R_tau)
    u0, Uc, _ = ic.laminar_u0(ny, re, R_tau)
    assert np.abs(u[3, 5] - Uc * (1.0 - y ** 2)).max() < 1e-12 * Uc
    assert np.abs(u[0, 0] - np.polynomial.chebyshev.chebval(y, u0)).max() < 1e-12 * Uc
    assert not v.any() and not w.any()
    u, _, _ = ic.init_flow(1, 8, ny, 8, 2.0, re, R_tau, kwall=2)
    assert not u[:, :, :2].any() and not u[:, :, -3:].any() and u[:, :, 2:-3].all()
    u, _, _ = ic.init_flow(2, 8, ny, 8, 2.0, re, R_tau, uinf=1.5)
    assert np.abs(u[1, 2] - 1.5 * (y + 1.0) / 2.0).max() < 1e-14
    for fs in (0, -1, 5):
        assert not any(a.any() for a in ic.init_flow(fs, 8, ny, 8, 2.0, re, R_tau))
    for fs in (3, 11, 12):
        with pytest.raises(ValueError):
            ic.init_flow(fs, 8, ny, 8, 2.0, re, R_tau)
    for fs in (4, 10):
        with pytest.raises(NotImplementedError):
            ic.init_flow(fs, 8, ny, 8, 2.0, re, R_tau)


def _synthesize(a, p):
    """Physical values on the unpadded grid (nx, nz, nyp) of a spectral field
    by direct summation: Chebyshev in y, Fourier in x and z (kx > 0 counted
    with its conjugate)."""
    x = np.arange(p.nx) * p.xl / p.nx
    z = np.arange(p.nz) * p.zl / p.nz
    T = np.cos(np.outer(np.arange(p.nyp), np.arccos(np.clip(p.ycoord / (p.yl / 2), -1, 1))))  # [n, k]
    out = np.zeros((p.nx, p.nz, p.nyp))
    for kx in range(p.nxh):
        ex = np.exp(1j * p.wavx[kx] * x)[:, None, None]
        for j in range(p.nz):
            c = a[kx, j] @ T                                  # physical y profile of this mode
            if not np.any(c):
                continue
            ez = np.exp(1j * p.wavz[j] * z)[None, :, None]
            term = (ex * ez * c[None, None, :]).real
            out += term if kx == 0 else 2.0 * term
    return out


def test_new_run_state_inverts_a_synthesized_state():
    """Spectral state -> velocity (veloc) -> physical values by direct
    summation -> new_run_state returns the state."""
    p = params(8, 8, 8)
    st = seeded(p, 7, amp=0.2, kmax=2)
    u, v, w = O.veloc(st["v"], st["omy"], st["u0"], st["w0"], p)
    phys = [_synthesize(a, p) for a in (u, v, w)]
    got = ic.new_run_state(*phys, p.xl, p.zl)
    for k in ("v", "omy"):
        assert np.abs(got[k] - st[k]).max() < 1e-12 * max(1.0, np.abs(st[k]).max()), k
    for k in ("u0", "w0"):
        assert np.abs(got[k] - np.real(st[k])).max() < 1e-12 * max(1.0, np.abs(st[k]).max()), k
