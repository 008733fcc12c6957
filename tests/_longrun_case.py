"""The reference's CPU-runnable case run long: 64 x 65 x 64 channel,
Lx = 4 pi, Lz = 2 pi, re = 2800, constant mass flux, seeded 5 % perturbation
of laminar flow, dt = 0.005, 2000 steps; statistics sampled every 20 steps
(tests/golden/longrun_c1_stats.npz holds the oracle's averages)."""
import numpy as np

from oracle import kmm_oracle as O
from tests._cases import params, seeded

NSTEPS, EVERY, SEED = 2000, 20, 20261202
PROFILES = ("umean", "urms", "vrms", "wrms", "rss")


def inputs():
    p = params(64, 64, 64, dt=0.005, cmf=True)
    return p, seeded(p, SEED, amp=0.05)


def spectra(st, p):
    """Energy spectra of u, v, w from the state: E(kx) (nxh) and E(kz)
    (nz), summing |coefficient|^2 over the other wavenumber and the
    Chebyshev index; kx > 0 modes count twice (their conjugates)."""
    u, v, w = O.veloc(st["v"], st["omy"], st["u0"], st["w0"], p)
    wx = np.where(np.arange(p.nxh) > 0, 2.0, 1.0)[:, None, None]
    out = {}
    for name, f in (("u", u), ("v", v), ("w", w)):
        e = wx * np.abs(f) ** 2
        out["ex_" + name] = e.sum(axis=(1, 2))
        out["ez_" + name] = e.sum(axis=(0, 2))
    return out
