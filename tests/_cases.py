"""Shared parity cases: seeded states and oracle runs (test helpers)."""
import math

import numpy as np

from oracle import kmm_oracle as O
from polymer_dns_amd import ic


def params(nx, ny, nz, scalar=False, **kw):
    d = dict(xl=4.0 * math.pi, zl=2.0 * math.pi, re=2800.0, dt=0.005,
             R_tau=math.sqrt(3.0 * 2800.0))
    d.update(kw)
    return O.Params(nx=nx, ny=ny, nz=nz, scalar=scalar, **d)


def seeded(p, seed, amp=0.05, kmax=8):
    st = ic.seeded_state(p.nx, p.ny, p.nz, p.xl, p.zl, p.re, p.R_tau, seed, amp=amp,
                         kmax=min(kmax, p.nxh - 1, p.nzh - 1), yl=p.yl)
    st = {k: np.array(v) for k, v in st.items()}
    if p.scalar:
        rng = np.random.default_rng(seed)
        sc = np.zeros((p.nxh, p.nz, p.nyp), dtype=np.complex128)
        sc[0, 0, 0] = 0.5
        sc[1, 2, :4] = 0.01 * (rng.standard_normal(4) + 1j * rng.standard_normal(4))
        sc[0, 3, :4] = 0.01 * rng.standard_normal(4)
        sc[0, p.nz - 3, :4] = np.conj(sc[0, 3, :4])
        st["scalar"] = sc
    return st


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def poly_state(p, seed, amp=1e-3, kmax=3, nmax=6, mean=None):
    """Conformation tensor = identity + a smooth seeded perturbation of
    amplitude amp (low Fourier modes times low Chebyshev polynomials), as the
    six independent spectral components {"c11": ..., ...}.  `mean` adds
    physical y-profiles (length nyp) to components, {"11": ..., "12": ...}."""
    rng = np.random.default_rng(seed)
    x = (np.arange(p.nx) * p.xl / p.nx)[:, None, None]
    z = (np.arange(p.nz) * p.zl / p.nz)[None, :, None]
    yy = (p.ycoord / (p.yl / 2))[None, None, :]
    T = [np.ones_like(yy), yy]
    for n in range(2, nmax + 1):
        T.append(2 * yy * T[-1] - T[-2])
    out = {}
    for ij in O.POLY:
        f = np.zeros((p.nx, p.nz, p.nyp))
        for _ in range(6):
            kx, kz, n = rng.integers(0, kmax + 1), rng.integers(-kmax, kmax + 1), rng.integers(0, nmax + 1)
            ph = rng.uniform(0, 2 * np.pi)
            f = f + rng.standard_normal() * np.cos(kx * p.alpha * x + kz * p.beta * z + ph) * T[n]
        f = amp * f / max(1e-30, np.abs(f).max())
        if ij in ("11", "22", "33"):
            f = f + 1.0
        if mean is not None and ij in mean:
            f = f + np.asarray(mean[ij])[None, None, :]
        out["c" + ij] = O.xyzfft(f, p)
        O.norm(out["c" + ij])
    return out


def shear_conformation_mean(p, u0, lam):
    """Mean conformation of a steady simple shear at the local rate of the
    mean flow u0 (Chebyshev coefficients), Oldroyd-B limit: C12 = lam U',
    C11 = 1 + 2 (lam U')^2 as added profiles for poly_state(mean=...)."""
    d = np.polynomial.chebyshev.chebder(np.real(np.asarray(u0))) * (2.0 / p.yl)
    g = lam * np.polynomial.chebyshev.chebval(p.ycoord / (p.yl / 2), d)
    return {"12": g, "11": 2.0 * g * g}
