"""CPU restatement (numpy, FP64) of the Kim-Moin-Moser channel time step of the
Polymer_DNS reference (FFTW_DNS/code/src/dns/*.f90).

TEST INFRASTRUCTURE ONLY.  This module is the checker: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.  The
product path (polymer_dns_amd) never imports, links or falls back to it.

Array layout: every spectral field is a numpy array ``a[kx, kz, n]`` of
complex128 with shape (nxh, nz, nyp).  In C order this is byte-identical to
the reference's Fortran column-major ``complex(nyp,nz,nxh)``: y (Chebyshev
index n, n=0 is y=+1) fastest, then kz in FFT order, then kx >= 0 (no x
Nyquist).  Physical 3/2 fields are ``p[k, j, i]`` with shape (mx, mz, nyp),
byte-identical to Fortran ``real(nyp,mz,mx)``.

Every function cites the reference routine it restates.  The restatement
follows the reference's equations and elimination order; floating-point
operation order inside an expression is not reproduced (parity is to 1e-10
relative, far above FP64 rounding).

Pinning (no oracle/_ref build is possible: the reference needs FFTW3, which is
absent from the image, and a stand-in library is not allowed):
  * cderiv / c1derbw / c1dertw -- closed forms T_n' = n U_{n-1},
    T_n'(+-1) = (+-1)^(n+1) n^2 (tests/test_oracle_pins.py)
  * solve / penta / psolve1d -- dense tau-matrix LU built independently from
    the Chebyshev recurrences, plus manufactured polynomial solutions and BC
    residuals
  * inverse/forward transforms -- scipy.fft.dct(type=1) == FFTW REDFT00 and
    numpy.fft == FFTW unnormalised DFTs, single-mode cos/sin fields, round trip
  * the full step -- laminar Poiseuille fixed point under constant pressure
    gradient, divergence-free velocity, Hermitian symmetry, and the linear
    Orr-Sommerfeld / Squire decay of a small perturbation
  * the nonlinear path (tests/test_oracle_nonlinear_pins.py) -- the
    Tollmien-Schlichting wave at Re 7500, alpha 1 grows at Orszag's (1971)
    rate through the full step; an oblique Orr-Sommerfeld mode with its
    forced Squire response follows its collocation eigenvalue in v and
    omega_y; a scalar mode follows its advection-diffusion eigenvalue;
    u . (u x a) = 0 per plane; u x omega and -u . grad(theta) of analytic
    modes; nonlin annihilates gradient fields; vort is the curl and veloc
    inverts it; gradv, gradscl, derivscji against analytic derivatives; the
    NL filter is exp(-3 eta^10) with eta = 1 at kz = nz/2, n = ny, and the
    state is not filtered.  Sign flips in any term of the cross product,
    nonlin, vort, gradscl or gradv, the smoothing normaliser, or the AB2
    weights each fail at least one of these pins.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import scipy.fft as sfft

# threads of the x/z/y transforms (scipy.fft workers); the CPU baseline sets
# it to the host's core count, the tests leave it at 1
WORKERS = 1


# --------------------------------------------------------------------------- #
#  Parameters (dns.config semantics, FFTW_DNS/code/src/dns/main_dns.f90:829-1171)
# --------------------------------------------------------------------------- #
@dataclasses.dataclass
class Params:
    nx: int
    ny: int
    nz: int
    xl: float = 4.0 * math.pi
    yl: float = 2.0
    zl: float = 2.0 * math.pi
    re: float = 2800.0
    dt: float = 0.005
    theta: float = 0.5
    gain: float = 0.0
    ugain: float = 0.0
    R_tau: float = math.sqrt(3.0 * 2800.0)
    flow_select: int = 1
    uinf: float = 0.0         # Uinf: the moving wall's velocity of the Couette cases
    perturbtime: int = 0
    kwall: int = 0            # KWALL: laminar start zero in the kwall rows next to each wall
    bftail: int = 0           # BFTAIL: last x index of the buffer region (imatrix 6, 8)
    xstart: float = 0.0       # XSTART: x origin of the suction profile (imatrix 7)
    # scalar (SCALAR build); scl_flag 6 = uniform volumetric source of rate
    # deltaT (extension, not a reference mode)
    scalar: bool = False
    diff: float = 1.0
    scl_flag: int = 0
    deltaT: float = 0.0
    # scl_flag 2: Gaussian blobs at the particles (dns.config line 65)
    sigmax: float = 0.0
    sigmay: float = 0.0
    sigmaz: float = 0.0
    # polymer (POLYMER build, FENE-P; dns.config lines 35-37, 55-62, 66)
    polymer: bool = False
    zlmax: float = 30.0       # L (maximum extensibility); Peterlin uses L^2
    tpoly: float = 0.1        # relaxation time lambda
    alpha_poly: float = 0.0
    diffpoly: float = 0.5     # artificial diffusivity 1/(re*diffpoly)
    c11z: float = 1.0
    c22z: float = 1.0
    c33z: float = 1.0
    qbeta: float = 0.9        # solvent viscosity ratio beta
    itarget: int = 0          # 0: uniform beta = qbeta
    ipeter: int = 1           # 1: FENE-P Peterlin function, 0: Oldroyd-B
    src_start: int = 0
    src_stop: int = 10 ** 9   # last step of the scl_flag 2 source window
    irstrt: int = 1
    crstrt: int = 0
    # constant mass flux (extension, no reference counterpart): dPdx is
    # corrected every step so that the bulk velocity stays at ubulk
    # (ubulk <= 0: the bulk velocity of the state at startup)
    cmf: bool = False
    ubulk: float = 0.0

    def __post_init__(self):
        assert self.ny % 2 == 0 and self.nz % 2 == 0 and self.nx % 2 == 0
        self.nyp = self.ny + 1
        self.nyh = self.ny // 2
        self.nyhp = self.nyh + 1
        self.nyhm = self.nyh - 1
        self.nxh = self.nx // 2
        self.nzh = self.nz // 2
        self.mx = 3 * self.nx // 2
        self.mz = 3 * self.nz // 2
        pi = 2.0 * math.acos(0.0)
        self.alpha = 2.0 * pi / self.xl
        self.beta = 2.0 * pi / self.zl
        # main_dns.f90:1137-1148
        self.wavx = np.arange(self.nxh, dtype=np.float64) * self.alpha
        kz = np.arange(self.nz, dtype=np.float64)
        kz[self.nzh:] -= self.nz
        self.wavz = kz * self.beta
        # main_dns.f90:1151-1155
        self.c = np.ones(self.nyp)
        self.c[0] = 2.0
        self.c[-1] = 2.0
        self.dyde = 2.0 / self.yl
        # main_dns.f90:1112-1117 (index 0 is y=+yl/2)
        n = np.arange(self.nyp, dtype=np.float64)
        self.ycoord = np.cos(n * pi / self.ny) * (self.yl / 2.0)
        self.ycoord[0] = self.yl / 2.0
        seg = np.zeros(self.nyp)
        seg[1:] = np.abs(self.ycoord[1:] - self.ycoord[:-1])
        seg[0] = seg[-1]
        self.seght = seg
        # init_flow.f90:117 (constant pressure gradient)
        self.dPdx = -2.0 / self.yl * (2.0 * self.R_tau / (self.yl * self.re)) ** 2
        # setbcs main_dns.f90:1240-1265 (channel), 1268-1330 (2-D Couette:
        # u = Uinf at the top wall; 3-D Couette: u = w = Uinf) and 1406-1409
        self.atu, self.btu, self.gtu = 1.0, 0.0, 0.0
        self.atw, self.btw, self.gtw = 1.0, 0.0, 0.0
        if self.flow_select in (2, 3, 12):
            self.gtu = self.uinf
        if self.flow_select == 3:
            self.gtw = self.uinf
        self.abu, self.bbu, self.gbu = 1.0, 0.0, 0.0
        self.abw, self.bbw, self.gbw = 1.0, 0.0, 0.0
        self.nt = 1
        self.nb = 1
        self.p1b, self.p1t, self.p2b, self.p2t = 1.0, -1.0, -1.0, 0.5

    @property
    def kx2(self):
        return (self.wavx ** 2)[:, None]

    @property
    def k2(self):
        return self.wavx[:, None] ** 2 + self.wavz[None, :] ** 2


# --------------------------------------------------------------------------- #
#  Chebyshev derivatives (FFTW_DNS/code/src/dns/derivatives.f90)
# --------------------------------------------------------------------------- #
def cderiv(f, dyde):
    """derivatives.f90:258-314 -- backward recurrence along the last axis."""
    ny = f.shape[-1] - 1
    df = np.zeros_like(f)
    df[..., ny - 1] = 2.0 * ny * f[..., ny] * dyde
    for n in range(ny - 2, -1, -1):
        df[..., n] = df[..., n + 2] + 2.0 * (n + 1) * f[..., n + 1] * dyde
    df[..., 0] = 0.5 * df[..., 0]
    return df


def c1derbw(f, dyde):
    """derivatives.f90:370-427 -- d/dy at y=-1: sum (-1)^(n+1) n^2 a_n dyde."""
    n = np.arange(f.shape[-1], dtype=np.float64)
    w = n * n * np.where(n % 2 == 1, 1.0, -1.0) * dyde
    return (f * w).sum(axis=-1)


def c1dertw(f, dyde):
    """derivatives.f90:541-595 -- d/dy at y=+1: sum n^2 a_n dyde."""
    n = np.arange(f.shape[-1], dtype=np.float64)
    return (f * (n * n * dyde)).sum(axis=-1)


# --------------------------------------------------------------------------- #
#  Tau systems (FFTW_DNS/code/src/dns/solvers.f90:174-1061)
#  All arrays below are 1-based along the first axis (index 0 unused) so the
#  elimination reads like the recurrences; trailing axes batch over (kx, kz).
# --------------------------------------------------------------------------- #
def _evnc(wn, dyde, nyh):
    """solvers.f90:324-351 -- even-mode tridiagonal coefficients a(i,1..3), i=1..nyh."""
    shp = (nyh + 2,) + wn.shape
    a1 = np.zeros(shp)
    a2 = np.zeros(shp)
    a3 = np.zeros(shp)
    for i in range(1, nyh + 1):
        j = 2 * i
        rj1 = 1.0 / (4 * j * (j - 1))
        rj2 = 1.0 / (2 * (j * j - 1))
        rj3 = 1.0 / (4 * j * (j + 1))
        a1[i] = wn * rj1
        a2[i] = -dyde * dyde - wn * rj2
        a3[i] = wn * rj3
    a1[1] = 2.0 * a1[1]
    a3[nyh - 1] = 0.0
    a2[nyh] = -dyde * dyde
    a3[nyh] = 0.0
    return a1, a2, a3


def _oddc(wn, dyde, nyh):
    """solvers.f90:355-383 -- odd-mode tridiagonal coefficients, i=1..nyh-1."""
    nyhm = nyh - 1
    shp = (nyh + 2,) + wn.shape
    a1 = np.zeros(shp)
    a2 = np.zeros(shp)
    a3 = np.zeros(shp)
    for i in range(1, nyhm + 1):
        j = 2 * i + 1
        rj1 = 1.0 / (4 * j * (j - 1))
        rj2 = 1.0 / (2 * (j * j - 1))
        rj3 = 1.0 / (4 * j * (j + 1))
        a1[i] = wn * rj1
        a2[i] = -dyde * dyde - wn * rj2
        a3[i] = wn * rj3
    a3[nyhm - 1] = 0.0
    a2[nyhm] = -dyde * dyde
    a3[nyhm] = 0.0
    return a1, a2, a3


def _evns(s, ny, c):
    """solvers.f90:387-440 (rows 2..nyhp; row 1 is the BC row set by the caller).
    s: Chebyshev coefficients with n on axis 0."""
    nyh = ny // 2
    nyhm = nyh - 1
    w = np.zeros((nyh + 2,) + s.shape[1:], dtype=s.dtype)
    for i in range(2, nyhm + 1):
        j = 2 * i - 2
        rj1 = 1.0 / (4 * j * (j - 1))
        rj2 = 1.0 / (2 * (j * j - 1))
        rj3 = 1.0 / (4 * j * (j + 1))
        w[i] = -c[j - 2] * s[j - 2] * rj1 + s[j] * rj2 - s[j + 2] * rj3
    j = ny - 2
    w[nyh] = -s[j - 2] / (4 * j * (j - 1)) + s[j] / (2 * (j * j - 1))
    j = ny
    w[nyh + 1] = -s[j - 2] / (4 * j * (j - 1))
    return w


def _odds(s, ny):
    """solvers.f90:444-496 (rows 2..nyh)."""
    nyh = ny // 2
    nyhm = nyh - 1
    w = np.zeros((nyh + 2,) + s.shape[1:], dtype=s.dtype)
    for i in range(2, nyhm):
        j = 2 * i - 1
        rj1 = 1.0 / (4 * j * (j - 1))
        rj2 = 1.0 / (2 * (j * j - 1))
        rj3 = 1.0 / (4 * j * (j + 1))
        w[i] = -s[j - 2] * rj1 + s[j] * rj2 - s[j + 2] * rj3
    j = ny - 3
    w[nyhm] = -s[j - 2] / (4 * j * (j - 1)) + s[j] / (2 * (j * j - 1))
    j = ny - 1
    w[nyh] = -s[j - 2] / (4 * j * (j - 1))
    return w


def _eqsolve(a1, a2, a3, t, w, nrank):
    """solvers.f90:500-553 -- quasi-tridiagonal elimination with a full BC row.
    t: BC row weights t[1..nrank] (1-based, shape (nrank+1,) + batch)."""
    m = nrank - 1
    d = np.zeros_like(a2)
    d[m] = a2[m]
    w = w.copy()
    for i in range(m - 1, 0, -1):
        d[i] = a2[i] - a3[i] * a1[i + 1] / d[i + 1]
        w[i + 1] = w[i + 1] - a3[i] * w[i + 2] / d[i + 1]
    f = t[m] - t[nrank] * a1[m] / d[m]
    w[1] = w[1] - t[nrank] * w[nrank] / d[m]
    for i in range(m - 1, 0, -1):
        w[1] = w[1] - f * w[i + 1] / d[i]
        f = t[i] - f * a1[i] / d[i]
    w[1] = w[1] / f
    for i in range(2, nrank + 1):
        w[i] = (w[i] - a1[i - 1] * w[i - 1]) / d[i - 1]
    return w


def solve(f, wn, dyde, bctop, bcbot, ib):
    """solvers.f90:174-283 -- (dyde^2 D^2 - wn) u = f, Dirichlet (ib=0) or
    Neumann (ib!=0) rows.  f: (..., nyp) Chebyshev RHS; wn, bctop, bcbot
    broadcast over the leading axes.  Returns u (..., nyp)."""
    nyp = f.shape[-1]
    ny = nyp - 1
    nyh = ny // 2
    nyhp = nyh + 1
    c = np.ones(nyp)
    c[0] = c[-1] = 2.0
    s = np.moveaxis(f, -1, 0)
    wn = np.broadcast_to(np.asarray(wn, dtype=np.float64), s.shape[1:]).copy()
    kzero = None
    if ib != 0:
        kzero = wn <= 1.0e-20
        wn[kzero] = 1.0
    bshape = (1,) * len(s.shape[1:])
    # even
    a1, a2, a3 = _evnc(wn, dyde, nyh)
    w = _evns(s, ny, c)
    if ib == 0:
        w[1] = 0.5 * (bctop + bcbot)
        te = np.ones(nyhp + 1)
    else:
        w[1] = 0.5 * (bctop - bcbot)
        te = np.array([0.0] + [dyde * (2 * (j - 1)) ** 2 for j in range(1, nyhp + 1)])
    we = _eqsolve(a1, a2, a3, te.reshape((-1,) + bshape), w, nyhp)
    # odd
    a1, a2, a3 = _oddc(wn, dyde, nyh)
    w = _odds(s, ny)
    if ib == 0:
        w[1] = 0.5 * (bctop - bcbot)
        to = np.ones(nyh + 1)
    else:
        w[1] = 0.5 * (bctop + bcbot)
        to = np.array([0.0] + [dyde * (2 * (j - 1) + 1) ** 2 for j in range(1, nyh + 1)])
    wo = _eqsolve(a1, a2, a3, to.reshape((-1,) + bshape), w, nyh)
    u = np.zeros_like(s)
    u[0::2] = we[1:nyhp + 1]
    u[1::2] = wo[1:nyh + 1]
    if kzero is not None:
        u[:, kzero] = 0.0
    return np.moveaxis(u, 0, -1)


def _mixed_rows(at, ab, bt, bb, dyde, nyp, bottom_sign):
    """mixedbc (solvers.f90:777-803, bottom_sign=+1, quirk: bottom derivative
    sign flipped) and mixedbc1d (solvers.f90:1249-1269, bottom_sign=-1)."""
    j1 = np.arange(nyp, dtype=np.float64)
    tt = at + dyde * bt * j1 * j1
    tb = ab + bottom_sign * dyde * bb * j1 * j1
    tb[1::2] = -tb[1::2]
    return tt, tb


def _pntsolve_pair(ae, ao, we, wo, tt, tb, gtop, gbot, ny):
    """pntsolve (solvers.f90:906-1017) for the even (ip=1) then odd (ip=2)
    systems sharing the two Robin rows; returns u (n on axis 0)."""
    nyh = ny // 2
    res = {}
    s = [None, np.array(gtop, dtype=np.complex128), np.array(gbot, dtype=np.complex128)]
    for ip, (a, w, nrank) in ((1, (ae, we, nyh + 1)), (2, (ao, wo, nyh))):
        a1, a2, a3 = a
        n = nrank
        m = n - 1
        cs = a1.copy()
        gs = w.copy()
        ds = np.zeros_like(a2)
        ds[m] = a2[m]
        for i in range(m - 1, 0, -1):
            ds[i] = a2[i] - a3[i] * cs[i + 1] / ds[i + 1]
            gs[i + 1] = gs[i + 1] - a3[i] * gs[i + 2] / ds[i + 1]
        f = {}
        for js, tv in ((2, tb), (1, tt)):
            # fillt solvers.f90:1021-1061
            if ip == 1:
                t = np.concatenate([[0.0], tv[0::2][:nrank]])
            else:
                t = np.concatenate([[0.0], tv[1::2][:nrank]])
            fj = t[m] - t[n] * cs[m] / ds[m]
            s[js] = s[js] - t[n] * gs[n] / ds[m]
            for i in range(m - 1, 0, -1):
                s[js] = s[js] - fj * gs[i + 1] / ds[i]
                fj = t[i] - fj * cs[i] / ds[i]
            f[js] = fj
        res[ip] = (cs, ds, gs, f)
    (cse, dse, gse, fe), (cso, dso, gso, fo) = res[1], res[2]
    det = fe[1] * fo[2] - fe[2] * fo[1]
    u0 = (s[1] * fo[2] - s[2] * fo[1]) / det
    u1 = (s[2] * fe[1] - s[1] * fe[2]) / det
    nyp = ny + 1
    u = np.zeros((nyp,) + np.shape(u0), dtype=np.complex128)
    u[0] = u0
    u[1] = u1
    for i in range(2, nyh + 2):          # even, ii = 2i-1 (1-based) -> n = 2i-2
        u[2 * i - 2] = (gse[i] - cse[i - 1] * u[2 * i - 4]) / dse[i - 1]
    for i in range(2, nyh + 1):          # odd, ii = 2i -> n = 2i-1
        u[2 * i - 1] = (gso[i] - cso[i - 1] * u[2 * i - 3]) / dso[i - 1]
    return u


def penta(f, wn, dyde, ib, at, bt, gtop, ab, bb, gbot):
    """solvers.f90:680-773 -- (dyde^2 D^2 - wn) u = f with Robin rows
    at*u + bt*u' = gtop (y=+1), ab*u +/- bb*u' = gbot (y=-1, quirk Q9)."""
    nyp = f.shape[-1]
    ny = nyp - 1
    nyh = ny // 2
    c = np.ones(nyp)
    c[0] = c[-1] = 2.0
    s = np.moveaxis(f, -1, 0)
    wn = np.broadcast_to(np.asarray(wn, dtype=np.float64), s.shape[1:]).copy()
    kzero = None
    if ib != 0:
        kzero = wn <= 1.0e-20
        wn[kzero] = 1.0
    tt, tb = _mixed_rows(at, ab, bt, bb, dyde, nyp, +1.0)
    ae = _evnc(wn, dyde, nyh)
    we = _evns(s, ny, c)
    we[1] = 0.0
    ao = _oddc(wn, dyde, nyh)
    wo = _odds(s, ny)
    wo[1] = 0.0
    gt = np.broadcast_to(gtop, s.shape[1:])
    gb = np.broadcast_to(gbot, s.shape[1:])
    u = _pntsolve_pair(ae, ao, we, wo, tt, tb, gt, gb, ny)
    if kzero is not None:
        u[:, kzero] = 0.0
    return np.moveaxis(u, 0, -1)


def psolve1d(f, g, dyde, at, bt, gtop, ab, bb, gbot):
    """solvers.f90:1180-1245 (+ p1d* 1249-1493): 1-D Robin Helmholtz for the
    mean mode, correct bottom-derivative sign (mixedbc1d)."""
    nyp = f.shape[-1]
    ny = nyp - 1
    nyh = ny // 2
    c = np.ones(nyp)
    c[0] = c[-1] = 2.0
    s = f.astype(np.complex128)
    wn = np.asarray(g, dtype=np.float64)
    tt, tb = _mixed_rows(at, ab, bt, bb, dyde, nyp, -1.0)
    ae = _evnc(wn, dyde, nyh)
    we = _evns(s, ny, c)
    we[1] = 0.0
    ao = _oddc(wn, dyde, nyh)
    wo = _odds(s, ny)
    wo[1] = 0.0
    u = _pntsolve_pair(ae, ao, we, wo, tt, tb, gtop, gbot, ny)
    return u.real.copy()


# --------------------------------------------------------------------------- #
#  Spectral-space operators (solvers.f90, helpers.f90)
# --------------------------------------------------------------------------- #
def norm(a):
    """helpers.f90:91-158 -- reality/symmetry projection (in place)."""
    nz = a.shape[1]
    a[0, 0, :] = a[0, 0, :].real
    j = np.arange(1, nz // 2)
    jp = nz - j
    a[0, j, :] = 0.5 * (a[0, j, :] + np.conj(a[0, jp, :]))
    a[0, jp, :] = np.conj(a[0, j, :])
    a[:, nz // 2, :] = 0.0
    return a


def phirhs(v, fn, fnm1, var, p: Params):
    """solvers.f90:555-676 -- CN/AB2 right-hand side.  Returns the RHS that
    the reference leaves in ``v`` (the zeroed bctop/bcbot are implicit)."""
    re, dt, th = p.re, p.dt, p.theta
    wrkc = -re * var / (dt * th) * v - re * var / (2.0 * th) * (3.0 * fn - fnm1)
    if th > 0.99:
        v = wrkc.copy()
    wrkc = wrkc + (1.0 - th) * p.k2[:, :, None] * v / th
    d2 = cderiv(cderiv(v, p.dyde), p.dyde)
    return wrkc - (1.0 - th) * d2 / th


def uzero(u0, h1, h1l, uflag, p: Params):
    """solvers.f90:1065-1176 -- mean (kx=kz=0) streamwise/spanwise velocity."""
    re, dt, th = p.re, p.dt, p.theta
    wrk1 = -re * u0 / (dt * th) - re * (3.0 * h1 - h1l) / (2.0 * th)
    if th > 0.99:
        rhs = wrk1
    else:
        d2 = cderiv(cderiv(u0, p.dyde), p.dyde)
        rhs = wrk1 - (1.0 - th) * d2 / th
    g = re / (th * dt)
    if uflag == 1:
        return psolve1d(rhs, g, p.dyde, p.atu, p.btu, p.gtu, p.abu, p.bbu, p.gbu)
    # quirk Q10: btu passed for w (solvers.f90:1173)
    return psolve1d(rhs, g, p.dyde, p.atw, p.btu, p.gtw, p.abw, p.bbw, p.gbw)


def bulk(a):
    """Mean over y in [-yl/2, yl/2] of a Chebyshev series a[..., n]:
    (1/2) int_{-1}^{1} T_n = 1/(1-n^2) for even n, 0 for odd n."""
    n = np.arange(a.shape[-1])
    w = np.zeros(a.shape[-1])
    w[n % 2 == 0] = 1.0 / (1.0 - n[n % 2 == 0] ** 2.0)
    return (np.asarray(a) * w).sum(axis=-1)


def cmf_unit_response(p: Params):
    """Mean-flow response phi to a unit streamwise forcing held over one step:
    the uzero system (solvers.f90:1065-1176) with right-hand side -re/theta
    (the forcing term -re (3 f - f)/(2 theta) for f = 1) and homogeneous
    wall data."""
    rhs = np.zeros(p.nyp)
    rhs[0] = -p.re / p.theta
    g = p.re / (p.theta * p.dt)
    return psolve1d(rhs, g, p.dyde, p.atu, p.btu, 0.0, p.abu, p.bbu, 0.0)


def cmf_correct(u0, p: Params):
    """Constant mass flux: u0 += f phi with f = (ubulk - bulk(u0))/bulk(phi);
    the pressure gradient of the following steps becomes dPdx - f."""
    phi = cmf_unit_response(p)
    f = (p.ubulk - bulk(u0)) / bulk(phi)
    return u0 + f * phi, f


def veloc(v, omy, u0, w0, p: Params):
    """solvers.f90:1576-1659 -- u, w from continuity and omega_y."""
    dv = cderiv(v, p.dyde)
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    w2 = (kx ** 2 + kz ** 2).copy()
    w2[0, 0, 0] = 1.0
    u = -(-1j * kx * dv + 1j * kz * omy) / w2
    w = -(-1j * kz * dv - 1j * kx * omy) / w2
    u[0, 0, :] = u0
    w[0, 0, :] = w0
    v = v.copy()
    v[0, 0, :] = 0.0
    return u, v, w


def vort(u, v, w, p: Params):
    """solvers.f90:1663-1722 -- omega_x, omega_z."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    omz = 1j * kx * v - cderiv(u, p.dyde)
    omx = cderiv(w, p.dyde) - 1j * kz * v
    return omx, omz


def nonlin(fn, omz, gn, p: Params):
    """solvers.f90:1497-1572 -- fn (4th-order v system) and gn (omega_y)."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    wrkc = 1j * (kz * omz + kx * gn)
    wrk1 = cderiv(wrkc, p.dyde)
    fn_new = -(kx ** 2 + kz ** 2) * fn - wrk1
    gn_new = 1j * (kz * gn - kx * omz)
    return fn_new, gn_new


def gradscl(sc, p: Params):
    """derivatives.f90:90-144."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    return 1j * kx * sc, cderiv(sc, p.dyde), 1j * kz * sc


def gradv(u, v, w, p: Params):
    """derivatives.f90:6-86 -- 9 gradients (u_ij = d u_i / d x_j) + 3 Laplacians."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    k2 = kx ** 2 + kz ** 2
    d = lambda a: cderiv(a, p.dyde)
    u12, u22, u32 = d(u), d(v), d(w)
    return dict(u11=1j * kx * u, u12=u12, u13=1j * kz * u,
                u21=1j * kx * v, u22=u22, u23=1j * kz * v,
                u31=1j * kx * w, u32=u32, u33=1j * kz * w,
                Lu=d(u12) - k2 * u, Lv=d(u22) - k2 * v, Lw=d(u32) - k2 * w)


def eig_cardano(g):
    """helpers.f90:197-319 -- the three eigenvalues of the velocity-gradient
    tensor g[i][k] (arrays) by Cardano's closed form in complex arithmetic
    with principal powers: A = tr^2/9 - (sum of principal 2x2 minors)/3,
    B = tr^3/27 - tr m2/6 + det/2, S = (B + (B^2 - A^3)^(1/2))^(1/3),
    l1 = tr/3 + A/S + S, l2,3 = tr/3 -+ i sqrt(3) (A/S - S)/2 - A/(2S) - S/2."""
    u11, u12, u13 = (np.asarray(g[0][k], dtype=np.complex128) for k in range(3))
    u21, u22, u23 = (np.asarray(g[1][k], dtype=np.complex128) for k in range(3))
    u31, u32, u33 = (np.asarray(g[2][k], dtype=np.complex128) for k in range(3))
    tr = u11 + u22 + u33
    A = tr ** 2 / 9 - (u11 * u22) / 3 + (u12 * u21) / 3 - (u11 * u33) / 3 + (u13 * u31) / 3 \
        - (u22 * u33) / 3 + (u23 * u32) / 3
    B = tr ** 3 / 27 - (tr * (u11 * u22 - u12 * u21 + u11 * u33 - u13 * u31 + u22 * u33 - u23 * u32)) / 6 \
        + (u11 * u22 * u33) / 2 - (u11 * u23 * u32) / 2 - (u12 * u21 * u33) / 2 + (u12 * u23 * u31) / 2 \
        + (u13 * u21 * u32) / 2 - (u13 * u22 * u31) / 2
    with np.errstate(all="ignore"):
        S = (B + (B ** 2 - A ** 3) ** 0.5) ** (1.0 / 3.0)
        l1 = tr / 3 + A / S + S
        l2 = tr / 3 - (math.sqrt(3.0) * (A / S - S) * 1j) / 2 - A / (2 * S) - S / 2
        l3 = tr / 3 + (math.sqrt(3.0) * (A / S - S) * 1j) / 2 - A / (2 * S) - S / 2
    return l1, l2, l3


def calcswirl(g):
    """helpers.f90:360-404 -- sqrt(sum Im(lambda)^2); 0 where that is 0 or NaN."""
    l1, l2, l3 = eig_cardano(g)
    with np.errstate(all="ignore"):
        sw = np.sqrt(l1.imag ** 2 + l2.imag ** 2 + l3.imag ** 2)
        bad = ~(np.floor(sw / sw) == 1)
    return np.where(bad, 0.0, sw)


def swirl_eig(g):
    """sqrt(2) |Im| of the complex eigenvalue pair of g[i][k] from LAPACK
    (numpy.linalg.eigvals): the swirl the closed form of calcswirl computes,
    without its cancellation near repeated roots."""
    M = np.stack([np.stack([np.asarray(g[i][k]) for k in range(3)], axis=-1) for i in range(3)], axis=-2)
    return math.sqrt(2.0) * np.abs(np.linalg.eigvals(M).imag).max(axis=-1)


def smoothing_factors(p: Params):
    """main_dns.f90:710-745 -- exp(-3 (|kz|/nzh)^10) * exp(-3 (n/ny)^10)."""
    j = np.arange(p.nz)
    jj = np.where(j < p.nzh, j, p.nz - j)
    zs = np.exp(-3.0 * (jj / p.nzh) ** 10)
    ys = np.exp(-3.0 * (np.arange(p.nyp) / p.ny) ** 10)
    return zs[None, :, None] * ys[None, None, :]


# --------------------------------------------------------------------------- #
#  Transforms (solvers.f90:2241-2879 inverse, 3147-3416 forward)
# --------------------------------------------------------------------------- #
def to_physical(a, p: Params):
    """Spectral (nxh, nz, nyp) -> physical 3/2 grid (mx, mz, nyp).
    Chebyshev scale c_n/2 then DCT-I (REDFT00) along y; zero-pad kz to mz
    (negative kz on top) and kx to mx/2+1; BACKWARD c2c along z then c2r along
    x, all unnormalised."""
    fac = p.c / 2.0
    ay = a * fac
    ayr = sfft.dct(ay.real, type=1, axis=-1, workers=WORKERS)
    ayi = sfft.dct(ay.imag, type=1, axis=-1, workers=WORKERS)
    ay = ayr + 1j * ayi                                  # (nxh, nz, nyp), physical y
    pad = np.zeros((p.nyp, p.mz, p.mx // 2 + 1), dtype=np.complex128)
    t = np.transpose(ay, (2, 1, 0))                      # (nyp, nz, nxh)
    pad[:, :p.nzh, :p.nxh] = t[:, :p.nzh, :]
    pad[:, p.mz - p.nzh:, :p.nxh] = t[:, p.nzh:, :]
    zt = sfft.ifft(pad, axis=1, workers=WORKERS) * p.mz                 # BACKWARD along z
    phys = sfft.irfft(zt, n=p.mx, axis=2, workers=WORKERS) * p.mx       # c2r along x
    return np.ascontiguousarray(np.transpose(phys, (2, 1, 0)))   # (mx, mz, nyp)


def to_spectral(ph, p: Params):
    """Physical 3/2 (mx, mz, nyp) -> spectral (nxh, nz, nyp): r2c along x,
    FORWARD c2c along z, keep nxh x nz modes, / (mx mz); DCT-I, / ny,
    halve n=0 and n=ny."""
    t = np.transpose(ph, (2, 1, 0))                      # (nyp, mz, mx)
    xt = sfft.rfft(t, axis=2, workers=WORKERS)
    zt = sfft.fft(xt, axis=1, workers=WORKERS)
    sp = np.empty((p.nyp, p.nz, p.nxh), dtype=np.complex128)
    sp[:, :p.nzh, :] = zt[:, :p.nzh, :p.nxh]
    sp[:, p.nzh:, :] = zt[:, p.mz - p.nzh:, :p.nxh]
    sp /= float(p.mx * p.mz)
    sp = np.transpose(sp, (2, 1, 0))                     # (nxh, nz, nyp)
    r = sfft.dct(sp.real, type=1, axis=-1, workers=WORKERS)
    i = sfft.dct(sp.imag, type=1, axis=-1, workers=WORKERS)
    out = (r + 1j * i) / float(p.ny)
    out[..., 0] /= 2.0
    out[..., -1] /= 2.0
    return np.ascontiguousarray(out)


def xyzfft(b, p: Params):
    """helpers.f90:6-87 (is=-1) -- physical (nx, nz, nyp) unpadded -> spectral."""
    t = sfft.dct(b, type=1, axis=-1) / float(p.ny)
    t[..., 0] /= 2.0
    t[..., -1] /= 2.0
    t = np.transpose(t, (2, 1, 0))                       # (nyp, nz, nx)
    xt = np.fft.rfft(t, axis=2)[:, :, :p.nxh] / float(p.nx)
    zt = np.fft.fft(xt, axis=1) / float(p.nz)
    return np.ascontiguousarray(np.transpose(zt, (2, 1, 0)))


# --------------------------------------------------------------------------- #
#  Immersed boundary (solvers.f90:2996-3108, main_dns.f90:1526-1558)
# --------------------------------------------------------------------------- #
def fspread_matrix():
    """main_dns.f90:1550-1558 -- fspread(j,i) = exp(-2 dsq) if dsq < 13."""
    fs = np.zeros((10, 10))   # [jdif-1, idif-1]
    for i in range(10):
        for j in range(10):
            dsq = float(i * i + j * j)
            if dsq < 13:
                fs[j, i] = math.exp(-2.0 * dsq)
    return fs


def kmaxsurf(imatrix):
    """main_dns.f90:1526-1535 -- largest 1-based x index of a solid cell (quirk Q7)."""
    solid = (imatrix == 1) | (imatrix == 3) | (imatrix == 4)
    ks = np.nonzero(solid.any(axis=(1, 2)))[0]
    return int(ks.max()) + 1 if ks.size else 0


def forcing_constants(p: Params):
    """forcepre (main_dns.f90:1490-1546): buffer gains bfgain/bfugain (1200
    entries, 1-based index i -> [i-1]), the spanwise wall slope length, the
    suction velocity vdes (mx entries, x spacing xl/(mx-1)) and the initial
    streamwise profile initu(y) of init_flow (channel: laminar re dPdx
    (d^2/2 - yl d/2), zero in the kwall rows; 2-D Couette: Uinf d/yl)."""
    bfhead = 2
    bfwidth = p.bftail - bfhead
    bfwh = int(bfwidth / 2) if bfwidth >= 0 else -int(-bfwidth / 2)
    bfcntr = bfwh + 1
    bfgain = np.zeros(1200)
    bfugain = np.zeros(1200)
    for i in range(1, bfwidth + 2):
        ib = i - bfcntr
        if abs(ib) <= int(bfwh / 3):
            bfgain[i - 1] = 20.0 * math.exp(-60.0 * (ib / bfwh) ** 2)
        bfugain[i - 1] = 20.0 * math.exp(-6.0 * (ib / bfwh) ** 2)
    slope = bfwidth / 2.0
    delxm = p.xl / float(p.mx - 1)
    xsuction = delxm * (p.bftail - 1)
    x = delxm * np.arange(p.mx)
    with np.errstate(divide="ignore", invalid="ignore"):
        vdes = np.where(x <= xsuction, 0.0,
                        (-0.8604 * p.uinf) / np.sqrt(p.re * p.uinf * (x + p.xstart - xsuction)))
    d = np.abs(p.ycoord - p.ycoord[-1])
    initu = np.zeros(p.nyp)
    if p.flow_select == 1:
        initu = p.re * p.dPdx * (0.5 * d ** 2 - p.yl / 2.0 * d)
        i1 = np.arange(1, p.nyp + 1)
        initu[(i1 <= p.kwall) | ((p.nyp - i1) <= p.kwall)] = 0.0
    elif p.flow_select == 2:
        initu = p.uinf * d / abs(p.ycoord[0] - p.ycoord[-1])
    return dict(bfhead=bfhead, bfwidth=bfwidth, bfgain=bfgain, bfugain=bfugain, slope=slope, vdes=vdes,
                initu=initu)


def apply_ib_plane(i, up, vp, wp, vw, drag_src, fxi, fyi, fzi, imat, it, p: Params,
                   gain, ugain, kmax):
    """Goldstein feedback at y-plane i (0-based), solvers.f90:2996-3108, with
    the reference's order-dependent 5x5 spreading (quirks Q5, Q6, Q7).

    Cell types: 1 (and 4 while it <= perturbtime) solid surface, 3
    spanwise damping, 8 spanwise moving wall at wdes(x), 11 constant x-force
    -- these set the drag that is spread; 6 buffer zone and 7 suction region
    add their force at the cell itself.  initv = initw = 0 (x- and
    z-independent starts).

    up/vp/wp/vw*: (mx, mz) arrays for this plane, x index first; the
    reference's (mz, mx) plane is column-major so the linear address of
    (j,k) (1-based) is (k-1)*mz + j, which is also the loop visit order.
    vw = [vwx, vwy, vwz] already hold the assigned products (incl. -dPdx).
    fxi/fyi/fzi: (mx, mz) integral slices, updated in place."""
    mx, mz = up.shape
    fc = forcing_constants(p) if np.any(np.isin(imat, (6, 7, 8))) else None
    dragx = np.zeros((mx, mz))
    dragy = np.zeros((mx, mz))
    dragz = np.zeros((mx, mz))
    flat = [a.reshape(-1) for a in vw]          # linear index = k*mz + j (0-based)
    srcs = []
    for k, j in np.argwhere(imat != 0):
        t = int(imat[k, j])
        u, v, w = up[k, j], vp[k, j], wp[k, j]
        ls = k * mz + j
        if t == 1 or (t == 4 and it <= p.perturbtime):
            fxi[k, j] += u
            fyi[k, j] += v
            fzi[k, j] += w
            dragx[k, j] = -ugain * u - gain * fxi[k, j]
            dragy[k, j] = -ugain * v - gain * fyi[k, j]
            dragz[k, j] = -ugain * w - gain * fzi[k, j]
        elif t == 3:
            fzi[k, j] += w
            dragz[k, j] = -ugain * w - gain * fzi[k, j]
        elif t == 8:
            k1 = k + 1
            bw = fc["bfwidth"]
            if k1 <= bw * 3 + 1:
                wdes = (0.3 * p.uinf) / fc["slope"] * (k1 - (bw * 3 + 1 - fc["slope"]))
            else:
                wdes = 0.3 * p.uinf
            fxi[k, j] += u
            fyi[k, j] += v
            fzi[k, j] += w - wdes
            dragx[k, j] = -ugain * u - gain * fxi[k, j]
            dragy[k, j] = -ugain * v - gain * fyi[k, j]
            dragz[k, j] = -ugain * (w - wdes) - gain * fzi[k, j]
        elif t == 6:
            b = (k + 1) - fc["bfhead"] + 1
            gu, gg = fc["bfugain"][b - 1], fc["bfgain"][b - 1]
            iu = fc["initu"][i]
            fxi[k, j] += u - iu
            fyi[k, j] += v
            fzi[k, j] += w
            flat[0][ls] += -gu * (u - iu) - gg * fxi[k, j]
            flat[1][ls] += -gu * v - gg * fyi[k, j]
            flat[2][ls] += -gu * w - gg * fzi[k, j]
        elif t == 7:
            vd = fc["vdes"][i]          # indexed by the y index (solvers.f90:3034)
            fyi[k, j] += v - vd
            flat[1][ls] += -gain * fyi[k, j] - ugain * (v - vd)
            fxi[k, j] += u - p.uinf
            flat[0][ls] += -gain * fxi[k, j] - ugain * (u - p.uinf)
            fzi[k, j] += w
            flat[2][ls] += -gain * fzi[k, j] - ugain * w
        elif t == 11:
            dragx[k, j] = 1.0
        elif t != 4:
            raise NotImplementedError(f"imatrix type {t}")
        if t in (1, 3, 4, 8, 11):
            srcs.append((k, j))
    if (i + 1) > kmax or not srcs:
        return
    fs = fspread_matrix()
    for k, j in srcs:
        ls = k * mz + j
        for ii in range(-2, 3):
            ipii = k + ii
            if ipii < 0:
                ipii += mx
            if ipii >= mx:
                ipii -= mx
            for jj in range(-2, 3):
                w = fs[abs(jj), abs(ii)]
                lt = ipii * mz + (j + jj)
                if lt < 0 or lt >= mx * mz:
                    raise ValueError("IB spreading leaves the plane (quirk Q5)")
                if lt <= ls:       # target visited at or before the source
                    flat[0][lt] += w * dragx[k, j]
                    flat[1][lt] += w * dragy[k, j]
                    flat[2][lt] += w * dragz[k, j]


# --------------------------------------------------------------------------- #
#  Nonlinear term (solvers.f90:1726-3538, NS + scalar + IB subset)
# --------------------------------------------------------------------------- #
class CFLError(RuntimeError):
    pass


def scalar_blob_source(part, p: Params):
    """solvers.f90:2078-2118 -- scsource(i,j,k) = sum over particles of
    deltaT exp(-(betax + betay + betaz)), betax = (x_k - x_n)^2/(2 sigmax^2)
    with x_k = (k-1) delxm (no periodic wrap), betaz likewise in z, betay with
    ycoord(i); a term counts only where betax < 20, betax + betaz < 20 and
    betax + betay + betaz < 18.  Returns the (mx, mz, nyp) physical field."""
    delxm = p.xl / p.mx
    delzm = p.zl / p.mz
    xk = np.arange(p.mx) * delxm
    zj = np.arange(p.mz) * delzm
    src = np.zeros((p.mx, p.mz, p.nyp))
    for xc, yc, zc in zip(part["x"], part["y"], part["z"]):
        bx = (xk - xc) ** 2 / (2.0 * p.sigmax ** 2)
        bz = (zj - zc) ** 2 / (2.0 * p.sigmaz ** 2)
        by = (p.ycoord - yc) ** 2 / (2.0 * p.sigmay ** 2)
        ok = (bx[:, None] < 20.0) & ((bx[:, None] + bz[None, :]) < 20.0)
        b = (bx[:, None, None] + by[None, None, :]) + bz[None, :, None]
        keep = ok[:, :, None] & (b < 18.0)
        src += np.where(keep, p.deltaT * np.exp(-np.where(keep, b, 0.0)), 0.0)
    return src


def vcw3dp(st, p: Params, it, fields, particles=None):
    """Returns (gn, fn, omz, scn) = spectral (u x omega)_x, _y, _z and the
    scalar nonlinear term, and updates the IB integrals in st.  With
    scl_flag 2 inside [src_start, src_stop] the blob source at `particles`
    enters as scsource/dt (solvers.f90:3061-3063); outside the window the
    reference reads an uninitialised local array, restated here as no
    source."""
    u, v, w, omx, omy, omz = (fields[k] for k in ("u", "v", "w", "omx", "omy", "omz"))
    up, vp, wp = to_physical(u, p), to_physical(v, p), to_physical(w, p)
    wxp, wyp, wzp = to_physical(omx, p), to_physical(omy, p), to_physical(omz, p)
    vwx = vp * wzp - wp * wyp
    vwy = -up * wzp + wp * wxp
    vwz = up * wyp - vp * wxp
    # CFL (solvers.f90:2985-2986, 3503-3514)
    delxm = p.xl / p.mx
    delzm = p.zl / p.mz
    cfl = (np.abs(up) / delxm + np.abs(vp) / p.seght[None, None, :]
           + np.abs(wp) / delzm) * p.dt
    cflmax = float(cfl.max())
    st["cflmax"] = cflmax
    vc = None
    if p.scalar:
        cxp = to_physical(fields["sclx"], p)
        cyp = to_physical(fields["scly"], p)
        czp = to_physical(fields["sclz"], p)
        vc = -(up * cxp + vp * cyp + wp * czp)
        if p.scl_flag == 6:
            vc = vc + p.deltaT
        if p.scl_flag == 2 and particles is not None and p.src_start <= it <= p.src_stop:
            vc = vc + scalar_blob_source(particles, p) / p.dt
    imat = st.get("imatrix")
    if p.flow_select == 1:
        if imat is None:
            vwx = vwx - p.dPdx
        else:
            vwx = vwx - p.dPdx * (imat == 0)
    if imat is not None and np.any(imat != 0):
        kmax = kmaxsurf(imat)
        for i in range(p.nyp):
            plane = imat[:, :, i]
            if not np.any(plane):
                continue
            vwp = [vwx[:, :, i].copy(), vwy[:, :, i].copy(), vwz[:, :, i].copy()]
            fxi = st["fxintg"][:, :, i].copy()
            fyi = st["fyintg"][:, :, i].copy()
            fzi = st["fzintg"][:, :, i].copy()
            apply_ib_plane(i, up[:, :, i], vp[:, :, i], wp[:, :, i], vwp, None,
                           fxi, fyi, fzi, plane, it, p, p.gain, p.ugain, kmax)
            vwx[:, :, i], vwy[:, :, i], vwz[:, :, i] = vwp
            st["fxintg"][:, :, i] = fxi
            st["fyintg"][:, :, i] = fyi
            st["fzintg"][:, :, i] = fzi
    gn = to_spectral(vwx, p)
    fn = to_spectral(vwy, p)
    omz_o = to_spectral(vwz, p)
    scn = to_spectral(vc, p) if vc is not None else None
    phys = dict(up=up, vp=vp, wp=wp, wxp=wxp, wyp=wyp, wzp=wzp)
    if "C" in fields:
        gv = fields["gradv"]
        gup = [[to_physical(gv[f"u{i}{k}"], p) for k in (1, 2, 3)] for i in (1, 2, 3)]
        cp = {ij: to_physical(fields["C"][ij], p) for ij in POLY}
        dcp = {ij: tuple(to_physical(a, p) for a in fields["dC"][ij]) for ij in POLY}
        cn, strp, qpp = polymer_products(cp, dcp, (up, vp, wp), gup, p)
        phys["poly"] = dict(cn={k: to_spectral(v, p) for k, v in cn.items()},
                            str={k: to_spectral(v, p) for k, v in strp.items()},
                            qp={k: to_spectral(v, p) for k, v in qpp.items()},
                            p12=qpp["12"], trC=cp["11"] + cp["22"] + cp["33"])
    return gn, fn, omz_o, scn, phys


# --------------------------------------------------------------------------- #
#  FENE-P conformation tensor (POLYMER build)
# --------------------------------------------------------------------------- #
POLY = ("11", "12", "13", "22", "23", "33")     # independent components (c21=c12, ...)
_CIDX = {(0, 0): "11", (0, 1): "12", (0, 2): "13", (1, 0): "12", (1, 1): "22",
         (1, 2): "23", (2, 0): "13", (2, 1): "23", (2, 2): "33"}
_PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def derivscji(C, p: Params):
    """derivatives.f90:148-254 -- dC_ij/dx_k for the independent components:
    {ij: (i kx C, D C, i kz C)} (dc_ij1 = i kx C_ij, dc_ij2 = cderiv, dc_ij3 = i kz)."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    return {ij: (1j * kx * C[ij], cderiv(C[ij], p.dyde), 1j * kz * C[ij]) for ij in POLY}


def polymer_products(cp, dcp, up3, gup, p: Params):
    """solvers.f90:2901-2983, pointwise on the 3/2 grid:
      c_ij^n = sum_k C_ik du_j/dx_k + sum_k C_kj du_i/dx_k - u . grad C_ij
    (the literal c11np..c33np lines: e.g. c12np = c11 u21 + c12 u22 + c13 u23
    + c12 u11 + c22 u12 + c32 u13 - (u dc121 + v dc122 + w dc123)),
    trC, Peterlin factor (L^2-3)/(L^2-trC) (ipeter=1) or 1, str = f C - I,
    qp = (1-beta)/(re beta lambda) str with beta = qbeta (itarget 0).
    cp: {ij: phys}, dcp: {ij: (dx, dy, dz) phys}, up3 = (u, v, w),
    gup[i][k] = du_i/dx_k phys.  Returns ({ij: c_n}, {ij: str}, {ij: qp})."""
    c = lambda i, j: cp[_CIDX[(i, j)]]
    cn = {}
    for i, j in _PAIRS:
        s = c(i, 0) * gup[j][0] + c(i, 1) * gup[j][1] + c(i, 2) * gup[j][2] \
            + c(0, j) * gup[i][0] + c(1, j) * gup[i][1] + c(2, j) * gup[i][2]
        d = dcp[_CIDX[(i, j)]]
        cn[_CIDX[(i, j)]] = s - (up3[0] * d[0] + up3[1] * d[1] + up3[2] * d[2])
    tr = cp["11"] + cp["22"] + cp["33"]
    if p.ipeter == 0:
        fac = 1.0
    else:
        L2 = p.zlmax * p.zlmax
        fac = (L2 - 3.0) / (L2 - tr)
    st = {ij: fac * cp[ij] for ij in POLY}
    for ij in ("11", "22", "33"):
        st[ij] = st[ij] - 1.0
    if p.itarget != 0:
        raise NotImplementedError("itarget 1/2 (scalar-dependent beta) is not restated")
    zbeta1 = (1.0 - p.qbeta) / (p.re * p.qbeta * p.tpoly)
    qp = {ij: zbeta1 * st[ij] for ij in POLY}
    return cn, st, qp


def writeoutputs(up, vp, wp, u12, swirl, beta, p12, trC, p: Params):
    """helpers.f90:1227-1436 with the mean u of solvers.f90:3487-3495, on the
    3/2 physical grid.  Arrays are (mx, mz, nyp) with y index 0 at y = +1
    (ycoord order); beta, p12, trC may be None (zeros).  Returns the y
    profiles umean, urms, vrms, wrms, rss, txy, betay, swirly, pss and the
    scalars massflux, twss, bwss, tpss, bpss, ckk."""
    nyp, mx, mz = p.nyp, p.mx, p.mz
    zero = np.zeros_like(up)
    beta = zero if beta is None else beta
    p12 = zero if p12 is None else p12
    trC = zero if trC is None else trC
    Lx, Lz = p.xl / mx, p.zl / mz
    volume = p.xl * p.yl * p.zl
    umean = up.sum(axis=1).sum(axis=0) / (mz * mx)            # uzmean over z, then over x
    upr = up - umean[None, None, :]
    y = p.ycoord
    Ly = np.empty(nyp)
    Ly[0] = abs(y[1] - y[0]) / 2.0
    Ly[-1] = abs(y[-1] - y[-2]) / 2.0
    Ly[1:-1] = abs(y[2:] - y[1:-1]) / 2.0 + abs(y[1:-1] - y[:-2]) / 2.0
    n = mx * mz
    prof = dict(
        umean=umean,
        urms=np.sqrt((upr ** 2).sum(axis=(0, 1)) / n),
        vrms=np.sqrt((vp ** 2).sum(axis=(0, 1)) / n),
        wrms=np.sqrt((wp ** 2).sum(axis=(0, 1)) / n),
        rss=(upr * vp).sum(axis=(0, 1)) / n,
        txy=(u12 / p.re).sum(axis=(0, 1)) / n,
        betay=(beta * Lx * Lz).sum(axis=(0, 1)) / (p.xl * p.zl),
        swirly=(swirl * Lx * Lz).sum(axis=(0, 1)) / (p.xl * p.zl),
        pss=p12.sum(axis=(0, 1)) / n,
    )
    mf = sum(1.0 / p.yl * 0.5 * (umean[i + 1] + umean[i]) * abs(y[i + 1] - y[i]) for i in range(nyp - 1))
    prof.update(massflux=mf,
                twss=float((u12[:, :, 0] / p.re).sum() / n), bwss=float((u12[:, :, -1] / p.re).sum() / n),
                tpss=float(p12[:, :, 0].sum() / n), bpss=float(p12[:, :, -1].sum() / n),
                ckk=float((trC * Lx * Ly[None, None, :] * Lz).sum() / volume))
    return prof


def conf_from_stress(st, p: Params):
    """Inverse of the stress map of solvers.f90:2925-2949 (str = f C - I,
    f = (L^2-3)/(L^2-trC) for ipeter 1, 1 for ipeter 0), the restart of the
    transport in the learned closure's audit steps: with s = tr str,
    f = (L^2 + s)/L^2 and C = (str + I)/f.  st: {ij: phys}; returns {ij: C}."""
    tr = st["11"] + st["22"] + st["33"]
    if p.ipeter == 0:
        inv = 1.0
    else:
        L2 = p.zlmax * p.zlmax
        inv = L2 / (L2 + tr)
    out = {ij: st[ij] * inv for ij in POLY}
    for ij in ("11", "22", "33"):
        out[ij] = out[ij] + inv
    return out


def polyNL(cn, cnm1, strn, strnm1, p: Params):
    """helpers.f90:610-674 -- 3 (c_n - str_n/lambda) - (c_nm1 - str_nm1/lambda)."""
    ti = 1.0 / p.tpoly
    return {ij: 3.0 * (cn[ij] - ti * strn[ij]) - (cnm1[ij] - ti * strnm1[ij]) for ij in POLY}


def polymer_implicit(C, cNL, p: Params):
    """main_dns.f90:240-321 -- phirhs(C, cNL, 2 cNL, diffpoly) (AB2 through
    phirhs's 3 fn - fnm1 = cNL) and the Neumann penta solve with
    g = re diffpoly/(dt theta); returns the new independent components."""
    g = p.re * p.diffpoly / (p.dt * p.theta)
    out = {}
    for ij in POLY:
        rhs = phirhs(C[ij], cNL[ij], 2.0 * cNL[ij], p.diffpoly, p)
        c = penta(rhs, p.k2 + g, p.dyde, 0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0)
        norm(c)
        out[ij] = c
    return out


def polyforce(qp, p: Params):
    """helpers.f90:498-560 -- t = div(qp): t1 = i(kx q11 + kz q13) + D q12,
    t2 = i(kx q12 + kz q23) + D q22, t3 = i(kx q13 + kz q33) + D q23."""
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    d = lambda a: cderiv(a, p.dyde)
    t1 = 1j * (kx * qp["11"] + kz * qp["13"]) + d(qp["12"])
    t2 = 1j * (kx * qp["12"] + kz * qp["23"]) + d(qp["22"])
    t3 = 1j * (kx * qp["13"] + kz * qp["33"]) + d(qp["23"])
    return t1, t2, t3


def polymer_init(p: Params):
    """main_dns.f90:206-236 -- C = diag(c11z, c22z, c33z) through xyzfft."""
    out = {}
    for ij in POLY:
        val = {"11": p.c11z, "22": p.c22z, "33": p.c33z}.get(ij, 0.0)
        out[ij] = xyzfft(np.full((p.nx, p.nz, p.nyp), val), p)
    return out


def polymer_flags(it, p: Params):
    """Time-line of the POLYMER branches of the main loop (main_dns.f90:206,
    240, 637, 674, 682) and of vcw3dp (solvers.f90:2900): init C at the start
    of the step; polyNL + implicit solve; polymer terms in vcw3dp; polymer
    force; the n -> n-1 copy before smoothing (nested inside the force
    branch, so its it == src_start-1 arm never fires)."""
    init = (it == p.src_start - 1) or (it == p.irstrt and p.crstrt == 0)
    solve = it >= p.src_start
    evaluate = it >= p.src_start - 1
    force = it >= p.src_start
    copy = force and init
    return dict(init=init, solve=solve, evaluate=evaluate, force=force, copy=copy)


# --------------------------------------------------------------------------- #
#  Influence matrix (solvers.f90:5-172, main_dns.f90:151-161, 433-466)
# --------------------------------------------------------------------------- #
def gfcn(pb, pt, p: Params):
    """Green's function for phi(+1)=pt, phi(-1)=pb, v(+-1)=0; returns
    (gf, dnvb, dnvt)."""
    g = p.re / (p.theta * p.dt)
    bct = np.full((p.nxh, p.nz), complex(pt, pt))
    bcb = np.full((p.nxh, p.nz), complex(pb, pb))
    bct[0, p.nz // 2:] = np.conj(bct[0, p.nz // 2:])
    bcb[0, p.nz // 2:] = np.conj(bcb[0, p.nz // 2:])
    bct[0, 0] = complex(pt, 0.0)
    bcb[0, 0] = complex(pb, 0.0)
    zero = np.zeros((p.nxh, p.nz, p.nyp), dtype=np.complex128)
    phi = solve(zero, p.k2 + g, p.dyde, bct, bcb, 0)
    gf = solve(phi, p.k2 + 0.0, p.dyde, 0.0, 0.0, 0)
    gf[0, 0, :] = gf[0, 0, :].real
    j = np.arange(1, p.nz // 2)
    jp = p.nz - j
    gf[0, j, :] = 0.5 * (gf[0, j, :] + np.conj(gf[0, jp, :]))
    gf[0, jp, :] = np.conj(gf[0, j, :])
    if p.nb == 1:
        dnvb = c1derbw(gf, p.dyde)
    else:
        dnvb = c1derbw(cderiv(gf, p.dyde), p.dyde)
    if p.nt == 1:
        dnvt = c1dertw(gf, p.dyde)
    else:
        dnvt = c1dertw(cderiv(gf, p.dyde), p.dyde)
    return gf, dnvb, dnvt


class Solver:
    """State + one reference time step.  State keys mirror the restart file:
    v, fn, fnm1, omy, gn, gnm1 (spectral), u0, h1n, h1nm1, w0, h3n, h3nm1
    (nyp), fxintg/fyintg/fzintg (mx, mz, nyp), optional scalar/scn/scnm1,
    optional imatrix (mx, mz, nyp) int32."""

    def __init__(self, p: Params, imatrix=None):
        self.p = p
        self.gf1, self.dnv1b, self.dnv1t = gfcn(p.p1b, p.p1t, p)
        self.gf2, self.dnv2b, self.dnv2t = gfcn(p.p2b, p.p2t, p)
        self.denom = self.dnv1t * self.dnv2b - self.dnv2t * self.dnv1b
        self.smooth = smoothing_factors(p)
        self.imatrix = imatrix
        self.it = 0
        self.particles = None
        self.part_state = 0
        self.ap = 0.0
        self.substep = 1
        self.part_swirl = True
        self.swirl_fn = calcswirl
        # learned-closure steps (no reference counterpart): a physical stress
        # {ij: (mx, mz, nyp)} supplied from outside replaces the FENE-P
        # transport and evaluation and feeds only the polymer force
        self.closure_stress = None

    def set_particles(self, x, y, z, ap=0.0, cd_switch=0, ratio=0.0, particle_flag=0, c_mu=1.0, gravity=0.0):
        """Particles initialised at the next evaluation (init_part): tracers
        (particle_flag 0) or inertial Maxey-Riley particles (> 0)."""
        self.particles = dict(x=np.asarray(x, float).copy(), y=np.asarray(y, float).copy(),
                              z=np.asarray(z, float).copy(), u=np.zeros(len(x)), v=np.zeros(len(x)),
                              w=np.zeros(len(x)))
        self.part_state = 1
        self.ap = ap
        self.substep = substep_check(self.p, cd_switch, ratio, ap)
        self.particle_flag = particle_flag
        self.mp = dict(ap=ap, ratio=ratio, c_mu=c_mu, cd_switch=cd_switch, gravity=gravity)
        self.mr_old = None
        self.umax1 = 0.0   # saved across steps (part_track.f90:104)

    # main_dns.f90:1602-2231 (new-run branch, from spectral u/v/w)
    def startup(self, st):
        """Given v, omy, u0, w0 (and scalar), fill the derived fields and the
        AB2 history exactly as ``initial`` does at it=0."""
        p = self.p
        if p.cmf and p.ubulk <= 0.0:
            p.ubulk = float(bulk(np.asarray(st["u0"], dtype=np.float64)))
        for k in ("v", "omy"):
            norm(st[k])
        u, v, w = veloc(st["v"], st["omy"], st["u0"], st["w0"], p)
        norm(u); norm(v); norm(w)
        st["v"] = v
        st["omy"][0, 0, :] = 0.0
        st.setdefault("fxintg", np.zeros((p.mx, p.mz, p.nyp)))
        st.setdefault("fyintg", np.zeros((p.mx, p.mz, p.nyp)))
        st.setdefault("fzintg", np.zeros((p.mx, p.mz, p.nyp)))
        if p.polymer:
            self.polymer_defaults(st)
        self._nonlinear(st, u, v, w, it=0, smooth=False)
        st["fnm1"] = st["fn"].copy()
        st["gnm1"] = st["gn"].copy()
        if p.polymer:
            # initial(): the n-1 polymer terms start equal to the n terms
            for ij in POLY:
                st["c%snm1" % ij] = st["c%sn" % ij].copy()
                st["str%snm1" % ij] = st["str%sn" % ij].copy()
        st["h1nm1"] = st["h1n"].copy()
        st["h3nm1"] = st["h3n"].copy()
        if p.scalar:
            st["scnm1"] = st["scn"].copy()
        return st

    def _nonlinear(self, st, u, v, w, it, smooth):
        p = self.p
        omx, omz = vort(u, v, w, p)
        norm(omx); norm(omz)
        fields = dict(u=u, v=v, w=w, omx=omx, omy=st["omy"], omz=omz)
        if p.scalar:
            norm(st["scalar"])
            sx, sy, sz = gradscl(st["scalar"], p)
            norm(sx); norm(sy); norm(sz)
            fields.update(sclx=sx, scly=sy, sclz=sz)
        pf = polymer_flags(it, p) if p.polymer else None
        closure = p.polymer and pf["evaluate"] and self.closure_stress is not None
        if p.polymer and pf["evaluate"] and not closure:
            C = {ij: st["c" + ij] for ij in POLY}
            fields["C"] = C
            fields["dC"] = derivscji(C, p)
            fields["gradv"] = gradv(u, v, w, p)
        if self.imatrix is not None:
            st["imatrix"] = self.imatrix
        gn, fn, omz_o, scn, phys = vcw3dp(st, p, it, fields, self.particles)
        st.pop("imatrix", None)
        self.last_phys = phys
        if self.particles is not None:
            # part_track (part_track.f90:131-168): initialise at the first
            # evaluation, Euler-advance afterwards
            uvw = (phys["up"], phys["vp"], phys["wp"])
            inertial = getattr(self, "particle_flag", 0) > 0
            fl = mr_fields(u, v, w, phys, p) if inertial else None
            if self.part_state == 1:
                self.particles = tracer_init(self.particles, uvw, p)
            elif inertial:
                # part_track.f90:131-150: substep RK4 sub-steps of dt/substep
                dts = p.dt / float(self.substep)
                for _ in range(self.substep):
                    self.particles, umax = mr_rk4(self.particles, fl, self.mr_old, p, self.mp, dts)
                    self.umax1 = max(self.umax1, umax)
                # particle CFL (part_track.f90:211-221)
                if self.umax1 * p.dt / self.ap >= 10.0:
                    raise ParticleCFLError(f"particle cfl failure at time step {it}")
            else:
                self.particles = tracer_update(self.particles, uvw, p, self.ap, self.substep)
            if inertial:
                # u_old, Lu_old for the next step (solvers.f90:3468-3474)
                self.mr_old = dict(u=fl["u"], v=fl["v"], w=fl["w"], L=fl["L"])
            self.part_state = 2
            if self.part_swirl:
                # swirl_part = swirl_3d at the particles (solvers.f90:3479-3481)
                gv = gradv(u, v, w, p)
                gp = [[to_physical(gv[f"u{i}{k}"], p) for k in (1, 2, 3)] for i in (1, 2, 3)]
                sw = self.swirl_fn(gp)
                pt = self.particles
                (pt["swirl"],) = fluid_interp(pt["x"], pt["y"], pt["z"], [sw], p)
                self.last_swirl = sw
        self.last_fields = fields
        if st["cflmax"] > 1.0:
            raise CFLError(f"CFL failure at time step {it}: {st['cflmax']}")
        for a in (fn, gn, omz_o):
            norm(a)
        if scn is not None:
            norm(scn)
        if closure:
            # qp = zbeta1 str transformed like the FENE-P stress, then polyforce
            zb = (1.0 - p.qbeta) / (p.re * p.qbeta * p.tpoly)
            qp = {ij: to_spectral(zb * np.asarray(self.closure_stress[ij], dtype=np.float64), p) for ij in POLY}
            for ij in POLY:
                norm(qp[ij])
            t1, t2, t3 = polyforce(qp, p)
            for a in (t1, t2, t3):
                norm(a)
            gn = gn + t1
            fn = fn + t2
            omz_o = omz_o + t3
        elif p.polymer and pf["evaluate"]:
            pol = phys["poly"]
            for ij in POLY:
                for key, d in (("cn", "c%sn"), ("str", "str%sn")):
                    norm(pol[key][ij])
                    st[d % ij] = pol[key][ij]
                norm(pol["qp"][ij])
            if pf["force"]:
                t1, t2, t3 = polyforce(pol["qp"], p)
                for a in (t1, t2, t3):
                    norm(a)
                gn = gn + t1
                fn = fn + t2
                omz_o = omz_o + t3
                if pf["copy"]:
                    for ij in POLY:
                        st["c%snm1" % ij] = st["c%sn" % ij].copy()
                        st["str%snm1" % ij] = st["str%sn" % ij].copy()
        if smooth:
            omz_o *= self.smooth
            gn *= self.smooth
            fn *= self.smooth
            if scn is not None:
                scn *= self.smooth
            if p.polymer and pf["evaluate"] and not closure:
                for ij in POLY:
                    st["c%sn" % ij] = st["c%sn" % ij] * self.smooth
                    st["str%sn" % ij] = st["str%sn" % ij] * self.smooth
        st["h1n"] = gn[0, 0, :].real.copy()
        st["h3n"] = omz_o[0, 0, :].real.copy()
        for a in (fn, gn, omz_o):
            norm(a)
        fn2, gn2 = nonlin(fn, omz_o, gn, p)
        norm(fn2); norm(gn2)
        st["fn"] = fn2
        st["gn"] = gn2
        if scn is not None:
            norm(scn)
            st["scn"] = scn

    # main_dns.f90:198-811 (NS + scalar)
    def polymer_defaults(self, st):
        """Zero conformation / history arrays that the caller did not supply."""
        z = lambda: np.zeros((self.p.nxh, self.p.nz, self.p.nyp), dtype=np.complex128)
        for ij in POLY:
            for key in ("c%s", "c%sn", "c%snm1", "str%sn", "str%snm1"):
                st.setdefault(key % ij, z())

    def step(self, st):
        p = self.p
        self.it += 1
        it = self.it
        g = p.re / (p.dt * p.theta)
        if p.polymer and self.closure_stress is None:
            self.polymer_defaults(st)
            pf = polymer_flags(it, p)
            if pf["init"]:
                for ij, c in polymer_init(p).items():
                    st["c" + ij] = c
            if pf["solve"]:
                get = lambda key: {ij: st[key % ij] for ij in POLY}
                cNL = polyNL(get("c%sn"), get("c%snm1"), get("str%sn"), get("str%snm1"), p)
                for ij in POLY:
                    st["c%snm1" % ij] = st["c%sn" % ij].copy()
                    st["str%snm1" % ij] = st["str%sn" % ij].copy()
                newc = polymer_implicit(get("c%s"), cNL, p)
                for ij in POLY:
                    st["c" + ij] = newc[ij]
        if p.scalar:
            rhs = phirhs(st["scalar"], st["scn"], st["scnm1"], p.diff, p)
            st["scnm1"] = st["scn"].copy()
            gs = p.re * p.diff / (p.dt * p.theta)
            st["scalar"] = penta(rhs, p.k2 + gs, p.dyde, 0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0)
        v = st["v"]
        phi = cderiv(cderiv(v, p.dyde), p.dyde) - p.k2[:, :, None] * v
        rhs = phirhs(phi, st["fn"], st["fnm1"], 1.0, p)
        st["fnm1"] = st["fn"].copy()
        phi = solve(rhs, p.k2 + g, p.dyde, 0.0, 0.0, 0)
        vp = solve(phi, p.k2 + 0.0, p.dyde, 0.0, 0.0, 0)
        orhs = phirhs(st["omy"], st["gn"], st["gnm1"], 1.0, p)
        st["gnm1"] = st["gn"].copy()
        vb = c1derbw(vp, p.dyde)
        vt = c1dertw(vp, p.dyde)
        c1 = (vb * self.dnv2t - vt * self.dnv2b) / self.denom
        c2 = (vt * self.dnv1b - vb * self.dnv1t) / self.denom
        v = vp + c1[:, :, None] * self.gf1 + c2[:, :, None] * self.gf2
        omy = penta(orhs, p.k2 + g, p.dyde, 0, p.atu, p.btu, 0.0, p.abu, p.bbu, 0.0)
        st["u0"] = uzero(st["u0"], st["h1n"], st["h1nm1"], 1, p)
        if p.cmf:
            st["u0"], fc = cmf_correct(st["u0"], p)
            p.dPdx = p.dPdx - fc
        st["w0"] = uzero(st["w0"], st["h3n"], st["h3nm1"], 3, p)
        omy[0, 0, :] = 0.0
        st["h1nm1"] = st["h1n"].copy()
        st["h3nm1"] = st["h3n"].copy()
        norm(v); norm(omy)
        st["omy"] = omy
        u, v, w = veloc(v, omy, st["u0"], st["w0"], p)
        norm(u); norm(w)
        st["v"] = v
        self._nonlinear(st, u, v, w, it=it, smooth=True)
        self.last_u = u
        self.last_w = w
        return st


# --------------------------------------------------------------------------- #
#  Seeded initial condition (new; no reference counterpart)
# --------------------------------------------------------------------------- #
class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) + Box-Muller, FP64."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------------------- #
#  Lagrangian tracers (FFTW_DNS/code/src/dns/part_track.f90)
# --------------------------------------------------------------------------- #
def interpolate3(x, y, z, x0, x1, y0, y1, z0, z1, c000, c100, c001, c101, c010, c110, c011, c111):
    """part_track.f90:1184-1222 -- trilinear; a degenerate axis (x1 <= x0,
    the periodic wrap cell) gets weight 0, i.e. the lower-corner value."""
    xd = np.where(x1 > x0, (x - x0) / np.where(x1 > x0, x1 - x0, 1.0), 0.0)
    yd = np.where(y1 > y0, (y - y0) / np.where(y1 > y0, y1 - y0, 1.0), 0.0)
    zd = np.where(z1 > z0, (z - z0) / np.where(z1 > z0, z1 - z0, 1.0), 0.0)
    c00 = c000 * (1.0 - xd) + c100 * xd
    c01 = c001 * (1.0 - xd) + c101 * xd
    c10 = c010 * (1.0 - xd) + c110 * xd
    c11 = c011 * (1.0 - xd) + c111 * xd
    c0 = c00 * (1.0 - yd) + c10 * yd
    c1 = c01 * (1.0 - yd) + c11 * yd
    return c0 * (1.0 - zd) + c1 * zd


def tracer_cells(xp, yp, zp, p: Params):
    """part_track.f90:537-553 -- 0-based cell indices on the 3/2 grid:
    imin/imax (x), jmin/jmax (z), kmax/kmin (y, kmin = kmax + 1)."""
    pi = 2.0 * math.acos(0.0)
    delxm = p.xl / p.mx
    delzm = p.zl / p.mz
    imin = np.mod(np.floor(xp / delxm).astype(np.int64) + p.mx, p.mx)
    imax = np.mod(imin + 1, p.mx)
    jmin = np.mod(np.floor(zp / delzm).astype(np.int64) + p.mz, p.mz)
    jmax = np.mod(jmin + 1, p.mz)
    kmax = np.floor(1.0 + float(p.ny) / pi * np.arccos(2.0 * yp / p.yl)).astype(np.int64) - 1
    kmin = kmax + 1
    return imin, imax, jmin, jmax, kmin, kmax


def fluid_interp(xp, yp, zp, fields, p: Params):
    """part_track.f90:518-613 -- trilinear interpolation of physical fields
    (each (mx, mz, nyp)) at the particle positions."""
    imin, imax, jmin, jmax, kmin, kmax = tracer_cells(xp, yp, zp, p)
    if np.any(kmin > p.ny) or np.any(kmax < 0):
        raise IndexError("particle outside the y range of the grid")
    delxm = p.xl / p.mx
    delzm = p.zl / p.mz
    xmin, xmax = imin * delxm, imax * delxm
    zmin, zmax = jmin * delzm, jmax * delzm
    ymin, ymax = p.ycoord[kmin], p.ycoord[kmax]
    out = []
    for u in fields:
        out.append(interpolate3(xp, yp, zp, xmin, xmax, ymin, ymax, zmin, zmax,
                                u[imin, jmin, kmin], u[imax, jmin, kmin], u[imin, jmax, kmin], u[imax, jmax, kmin],
                                u[imin, jmin, kmax], u[imax, jmin, kmax], u[imin, jmax, kmax], u[imax, jmax, kmax]))
    return out


def tracer_update(part, phys, p: Params, ap, substep=1):
    """part_track.f90:360-433 -- one explicit Euler step with dt/substep
    (part_track.f90:45-50 divides dt by substep and calls tracer_update
    once), then the velocity is re-interpolated at the new position."""
    dt = p.dt / float(substep)
    height = p.ycoord[0] - ap
    # Fortran MOD of reals truncates toward zero, like fmod
    x = np.fmod(p.xl + part["x"] + part["u"] * dt, p.xl)
    y = np.maximum(np.minimum(part["y"] + part["v"] * dt, height), ap + p.ycoord[-1])
    z = np.fmod(p.zl + part["z"] + part["w"] * dt, p.zl)
    u, v, w = fluid_interp(x, y, z, phys, p)
    return dict(x=x, y=y, z=z, u=u, v=v, w=w)


def substep_check(p: Params, cd_switch, ratio, ap):
    """part_track.f90:335-350 -- number of particle sub-steps per flow step."""
    pi_f = 1.0 / (ratio + 0.5)
    if cd_switch == 0:
        tau = 1.0
    elif cd_switch == 1:
        tau = p.re * ap ** 2 / (6.0 * pi_f)
    else:
        tau = 2.0 * p.re * ap ** 2 / (9.0 * pi_f)
    return int(math.ceil(max(1.0, 2.0 * p.dt / tau)))


def tracer_init(part, phys, p: Params):
    """init_part (part_track.f90:229-292): particle velocity = fluid velocity."""
    u, v, w = fluid_interp(part["x"], part["y"], part["z"], phys, p)
    return dict(x=part["x"].copy(), y=part["y"].copy(), z=part["z"].copy(), u=u, v=v, w=w)


# --------------------------------------------------------------------------- #
#  Inertial particles: Maxey-Riley equations by RK4 (part_track.f90:617-1075)
# --------------------------------------------------------------------------- #
class ParticleCFLError(RuntimeError):
    pass


def mr_fields(u, v, w, phys, p: Params):
    """The physical fields part_track receives (solvers.f90:3464-3466): u, v,
    w and the vorticity of vcw3dp, the nine gradients u_ik = du_i/dx_k and
    the Laplacians of gradv (derivatives.f90:6-86, normed as in
    main_dns.f90:544-558), all on the 3/2 grid (mx, mz, nyp)."""
    gv = gradv(u, v, w, p)
    for a in gv.values():
        norm(a)
    return dict(u=phys["up"], v=phys["vp"], w=phys["wp"], ox=phys["wxp"], oy=phys["wyp"], oz=phys["wzp"],
                g=[to_physical(gv[f"u{i}{k}"], p) for i in (1, 2, 3) for k in (1, 2, 3)],
                L=[to_physical(gv[k], p) for k in ("Lu", "Lv", "Lw")])


def mr_forces(xp, yp, zp, un, vn, wn, fl, old, p: Params, mp, dt):
    """RKstage (part_track.f90:805-873) and SubDerivative (:896-1075): the
    Maxey-Riley acceleration (pressure gradient / added mass, drag with the
    Faxen term, lift, Faxen time derivative, buoyancy).  dt is the particle
    sub-step: the reference divides dt by substep before the integration,
    so the fluid time derivatives (u - u_old)/dt use it too."""
    ap, re = mp["ap"], p.re
    pi_f = 1.0 / (mp["ratio"] + 0.5)
    pi_p = mp["ratio"] / (mp["ratio"] + 0.5)
    c_l = 0.5
    uf, vf, wf = fluid_interp(xp, yp, zp, [fl["u"], fl["v"], fl["w"]], p)
    wxf, wyf, wzf = fluid_interp(xp, yp, zp, [fl["ox"], fl["oy"], fl["oz"]], p)
    luf, lvf, lwf = fluid_interp(xp, yp, zp, fl["L"], p)
    g = fluid_interp(xp, yp, zp, fl["g"], p)
    # corner values (u - u_old)/dt interpolated (part_track.f90:1001-1009)
    dudt = fluid_interp(xp, yp, zp, [(fl[k] - old[k]) / dt for k in ("u", "v", "w")], p)
    dldt = fluid_interp(xp, yp, zp, [(fl["L"][i] - old["L"][i]) / dt for i in range(3)], p)
    sdu = dudt[0] + uf * g[0] + vf * g[1] + wf * g[2]
    sdv = dudt[1] + uf * g[3] + vf * g[4] + wf * g[5]
    sdw = dudt[2] + uf * g[6] + vf * g[7] + wf * g[8]
    re_p = 2.0 * ap * re * np.sqrt((uf - un) ** 2 + (vf - vn) ** 2 + (wf - wn) ** 2)
    cd = mp["cd_switch"]
    if cd == 0:
        c_d = np.zeros_like(re_p)
    elif cd == 1:
        c_d = np.full_like(re_p, 16.0 / (2.0 * ap * re) * mp["c_mu"])
    elif cd == 2:
        c_d = 24.0 / (2.0 * ap * re) * (1.0 + 0.15 * re_p ** 0.687)
    else:
        with np.errstate(divide="ignore"):
            c_d = np.where(re_p > 1.0, 16.31 * re_p ** (-0.8331), 17.28 * re_p ** (-0.9711))
    a2 = ap * ap / 6.0
    ru = uf - un + a2 * luf
    rv = vf - vn + a2 * lvf
    rw = wf - wn + a2 * lwf
    dfac = c_d * 3.0 / (8.0 * ap)
    ffac = 3.0 / 20.0 * (ap * ap)
    ku = pi_f * (1.5 * sdu + dfac * ru + c_l * (rv * wzf - rw * wyf) + ffac * dldt[0])
    kv = pi_f * (1.5 * sdv + dfac * rv + c_l * (rw * wxf - ru * wzf) + ffac * dldt[1]) - mp["gravity"] * (pi_f - pi_p)
    kw = pi_f * (1.5 * sdw + dfac * rw + c_l * (ru * wyf - rv * wxf) + ffac * dldt[2])
    return ku, kv, kw


def mr_rk4(part, fl, old, p: Params, mp, dt):
    """RK4 (part_track.f90:617-755) of one sub-step dt for every particle:
    stages at dt/2, dt/2, dt from the start state, positions periodic in x
    and z and clamped to [ap + y_min, y_max - ap] in y, an elastic bounce of
    v at a clamped wall.  Returns the new state and max |u_f - u_p| at the new
    positions."""
    ap = mp["ap"]
    height = p.ycoord[0] - ap
    bottom = ap + p.ycoord[-1]
    x0, y0, z0 = part["x"], part["y"], part["z"]
    up, vp, wp = part["u"], part["v"], part["w"]
    xp, yp, zp, un, vn, wn = x0, y0, z0, up, vp, wp
    ks, vs = [], []
    for sn in (1, 2, 3, 4):
        ku, kv, kw = mr_forces(xp, yp, zp, un, vn, wn, fl, old, p, mp, dt)
        h = dt if sn == 3 else dt / 2.0
        un, vn, wn = up + h * ku, vp + h * kv, wp + h * kw
        xp = np.fmod(p.xl + x0 + h * un, p.xl)
        yp = np.maximum(np.minimum(y0 + h * vn, height), bottom)
        zp = np.fmod(p.zl + z0 + h * wn, p.zl)
        ks.append((ku, kv, kw))
        vs.append((un, vn, wn))
    (u2, v2, w2), (u3, v3, w3), (u4, v4, w4) = vs[:3]
    (k1u, k1v, k1w), (k2u, k2v, k2w), (k3u, k3v, k3w), (k4u, k4v, k4w) = ks
    x = np.fmod(p.xl + x0 + dt * (up / 6 + u2 / 3 + u3 / 3 + u4 / 6), p.xl)
    y = np.maximum(np.minimum(y0 + dt * (vp / 6 + v2 / 3 + v3 / 3 + v4 / 6), height), bottom)
    z = np.fmod(p.zl + z0 + dt * (wp / 6 + w2 / 3 + w3 / 3 + w4 / 6), p.zl)
    u = up + dt * (k1u / 6 + k2u / 3 + k3u / 3 + k4u / 6)
    vnew = vp + dt * (k1v / 6 + k2v / 3 + k3v / 3 + k4v / 6)
    w = wp + dt * (k1w / 6 + k2w / 3 + k3w / 3 + k4w / 6)
    bounce = ((y == height) & (vnew > 0.0)) | ((y == bottom) & (vnew < 0.0))
    v = np.where(bounce, -vnew, vnew)
    uf, vf, wf = fluid_interp(x, y, z, [fl["u"], fl["v"], fl["w"]], p)
    umax = float(np.sqrt((uf - u) ** 2 + (vf - v) ** 2 + (wf - w) ** 2).max()) if x.size else 0.0
    return dict(x=x, y=y, z=z, u=u, v=v, w=w), umax
