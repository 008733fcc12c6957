"""The Re_tau 395 viscoelastic workload at full size (256 x 193 x 256, the
FENE-P parameters of the 192 x 129 x 192 case: Wi_tau 50, beta 0.9, L^2 900,
Sc 0.5, conformation at the mean-shear equilibrium plus seeded 1e-3 noise),
FP64: inputs rebuilt from seeds, and the sampling the fixture
tests/golden/c5_fenep.npz uses.  Re_bulk 6875 in half-height units is the
bulk Reynolds number of a Re_tau 395 channel (13 750 on the full height)."""
import math

import numpy as np

from polymer_dns_amd import ic
from tests._cases import params

NX, NY, NZ = 256, 192, 256
XL, ZL = 4.0 * math.pi, 4.0 * math.pi / 3.0
RE, RTAU, DT = 6875.0, 395.0, 0.001
WI_TAU, QBETA, ZLMAX, SC = 50.0, 0.9, 30.0, 0.5
SEED = 20261202 + 5
NSTEPS = 2
SAMPLE_SEED = 9
NRANDOM, NLARGEST = 256, 32
FIELDS = (["v", "omy", "fn", "gn", "u0", "w0"] + ["c" + q for q in ("11", "12", "13", "22", "23", "33")]
          + ["c%sn" % q for q in ("11", "12", "13", "22", "23", "33")]
          + ["str%sn" % q for q in ("11", "12", "13", "22", "23", "33")])


def tpoly():
    utau = RTAU / RE
    return WI_TAU / (RE * utau * utau)   # Wi_tau = lambda u_tau^2 / nu


def inputs():
    """(oracle Params, initial state dict)."""
    tp = tpoly()
    p = params(NX, NY, NZ, xl=XL, zl=ZL, re=RE, dt=DT, R_tau=RTAU, polymer=True, zlmax=ZLMAX, tpoly=tp,
               diffpoly=SC, qbeta=QBETA, ipeter=1, crstrt=1, src_start=0)
    st = ic.seeded_state(NX, NY, NZ, XL, ZL, RE, RTAU, SEED, amp=0.05, kmax=8)
    mean = ic.shear_equilibrium(st["u0"], tp, ZLMAX, ipeter=1)
    st.update(ic.conformation_state(NX, NY, NZ, SEED + 1, amp=1e-3, mean=mean))
    return p, st


def sample_index(f, rng):
    """NRANDOM seeded random flat indices of f plus its NLARGEST largest."""
    flat = np.abs(np.asarray(f).ravel())
    r = rng.choice(flat.size, min(NRANDOM, flat.size), replace=False)
    top = np.argsort(flat)[-NLARGEST:]
    return np.unique(np.concatenate([r, top]))
