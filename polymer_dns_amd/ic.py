"""Seeded initial conditions (no reference counterpart).

The reference starts channel runs from a pure laminar profile
(init_flow.f90:82-139) and trips turbulence with temporary IB bricks; the
seeded, divergence-free perturbation here is new.  It is delivered to either
code through a ``restart`` file (IRSTRT=1), so both start from identical
bytes.

Recipe: laminar u0 = Uc (1 - y^2) (Uc = -re dPdx yl^2/8), and for
0 <= kx <= kmax, |kz| <= kmax, (kx,kz) != (0,0):
    v_hat   = (1 - y^2)^2 * sum_{n<=4} r_n  T_n(y)
    omy_hat = (1 - y^2)   * sum_{n<=4} r'_n T_n(y)
with r, r' complex N(0,1) drawn from SplitMix64 + Box-Muller.  v = v' = 0 and
omega_y = 0 at both walls, so u, w from continuity satisfy no slip.  The
amplitude is scaled so that max |u'| on the collocation grid is
``amp * Uc``.  kx = 0 modes are made Hermitian in kz.
"""
from __future__ import annotations

import math

import numpy as np
from numpy.polynomial import chebyshev as cheb


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with Box-Muller normals."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def laminar_u0(ny, re, R_tau, yl=2.0):
    """Chebyshev coefficients of the CPG laminar profile (init_flow.f90:117-138)."""
    dPdx = -2.0 / yl * (2.0 * R_tau / (yl * re)) ** 2
    Uc = -re * dPdx * yl * yl / 8.0
    u0 = np.zeros(ny + 1)
    u0[0] = 0.5 * Uc
    u0[2] = -0.5 * Uc
    return u0, Uc, dPdx


def _phys_max_speed(v, omy, u0, nx, nz, xl, zl, yl):
    """max over the collocation grid of |(u', v', w')| for the perturbation."""
    nxh, nzz, nyp = v.shape
    ny = nyp - 1
    dyde = 2.0 / yl
    alpha = 2.0 * math.pi / xl
    beta = 2.0 * math.pi / zl
    kx = (np.arange(nxh) * alpha)[:, None, None]
    kzi = np.arange(nz)
    kzi[nz // 2:] -= nz
    kz = (kzi * beta)[None, :, None]
    dv = np.zeros_like(v)
    for a in range(nxh):
        for b in range(nz):
            d = cheb.chebder(v[a, b], scl=dyde)
            dv[a, b, :d.size] = d
    k2 = kx ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    u = -(-1j * kx * dv + 1j * kz * omy) / k2
    w = -(-1j * kz * dv - 1j * kx * omy) / k2
    u[0, 0] = 0.0
    w[0, 0] = 0.0
    y = np.cos(np.arange(nyp) * math.pi / ny)
    T = np.cos(np.outer(np.arange(nyp), np.arccos(y)))  # T_n(y_k): [n, k]
    speed2 = 0.0
    comps = []
    for a in (u, v, w):
        py = a @ T                                       # (nxh, nz, nyp) physical y
        full = np.zeros((nz, nx // 2 + 1, nyp), dtype=np.complex128)
        full[:, :nxh, :] = np.transpose(py, (1, 0, 2))
        zt = np.fft.ifft(full, axis=0) * nz
        ph = np.fft.irfft(zt, n=nx, axis=1) * nx
        comps.append(ph)
    speed2 = comps[0] ** 2 + comps[1] ** 2 + comps[2] ** 2
    return float(np.sqrt(speed2.max()))


def seeded_state(nx, ny, nz, xl, zl, re, R_tau, seed, amp=0.05, kmax=8, yl=2.0):
    """Returns dict(v, omy, u0, w0) in reference layout (see kmm.py)."""
    nxh, nyp = nx // 2, ny + 1
    rng = SplitMix64(seed)
    w1 = np.array([0.5, 0.0, -0.5])
    w2 = cheb.chebmul(w1, w1)
    v = np.zeros((nxh, nz, nyp), dtype=np.complex128)
    omy = np.zeros_like(v)
    for kx in range(0, min(kmax, nxh - 1) + 1):
        for kzs in range(-kmax, kmax + 1):
            if kx == 0 and kzs <= 0:
                continue
            if abs(kzs) >= nz // 2:
                continue
            r = np.array([complex(rng.normal(), rng.normal()) for _ in range(5)])
            rp = np.array([complex(rng.normal(), rng.normal()) for _ in range(5)])
            cv = cheb.chebmul(w2, r.real) + 1j * cheb.chebmul(w2, r.imag)
            co = cheb.chebmul(w1, rp.real) + 1j * cheb.chebmul(w1, rp.imag)
            j = kzs % nz
            v[kx, j, :cv.size] = cv
            omy[kx, j, :co.size] = co
            if kx == 0:
                v[0, (-kzs) % nz, :cv.size] = np.conj(cv)
                omy[0, (-kzs) % nz, :co.size] = np.conj(co)
    u0, Uc, dPdx = laminar_u0(ny, re, R_tau, yl)
    smax = _phys_max_speed(v, omy, u0, nx, nz, xl, zl, yl)
    if smax > 0:
        s = amp * Uc / smax
        v *= s
        omy *= s
    return dict(v=v, omy=omy, u0=u0, w0=np.zeros(nyp))


POLY = ("11", "12", "13", "22", "23", "33")


def shear_equilibrium(u0, tpoly, zlmax, ipeter=1, yl=2.0):
    """Steady FENE-P conformation of a homogeneous simple shear at the local
    rate U'(y) of the mean profile u0 (Chebyshev coefficients), at the
    Gauss-Lobatto points y_k = cos(k pi/ny) (yl/2).  With W = tpoly U' the
    steady equations L C + C L^T = (f C - I)/tpoly give C22 = C33 = 1/f,
    C12 = W/f^2, C11 = 1/f + 2 W^2/f^3, and the Peterlin function
    f = (L^2 - 3)/(L^2 - tr C) (ipeter 1) reduces to the cubic
    L^2 f^2 (f - 1) = 2 W^2 (root f >= 1, by Newton); ipeter 0: f = 1.
    Returns {"11", "12", "22", "33"} profiles (length ny+1)."""
    u0 = np.real(np.asarray(u0, dtype=np.complex128))
    ny = u0.size - 1
    y = np.cos(np.arange(ny + 1) * math.pi / ny)
    W = tpoly * cheb.chebval(y, cheb.chebder(u0)) * (2.0 / yl)
    f = np.ones_like(W)
    if ipeter:
        L2 = zlmax * zlmax
        f = 1.0 + np.cbrt(2.0 * W * W / L2)          # above the root
        for _ in range(60):
            g = L2 * f * f * (f - 1.0) - 2.0 * W * W
            dg = L2 * (3.0 * f * f - 2.0 * f)
            f = f - g / dg
    return {"11": 1.0 / f + 2.0 * W * W / f ** 3, "12": W / f ** 2, "22": 1.0 / f, "33": 1.0 / f}


def _cheb_coeffs(vals):
    """Chebyshev coefficients of values at the Gauss-Lobatto points
    y_k = cos(k pi/N) (DCT-I, ends halved)."""
    N = vals.size - 1
    k = np.arange(N + 1)
    w = np.ones(N + 1)
    w[0] = w[N] = 0.5
    a = (2.0 / N) * (np.cos(np.outer(k, k) * math.pi / N) @ (w * vals))
    a[0] *= 0.5
    a[N] *= 0.5
    return a


def conformation_state(nx, ny, nz, seed, amp=1e-3, kmax=4, nmax=6, mean=None):
    """Conformation tensor C = I + seeded noise, as the six independent
    spectral components {"c11": ..., "c33": ...} (complex (nx/2, nz, ny+1)).
    With ``mean`` (profiles at the Gauss-Lobatto points, e.g. from
    ``shear_equilibrium``) the mean of each given component is that profile
    instead of the identity's.

    Noise: for 0 <= kx <= kmax, |kz| <= kmax and Chebyshev n <= nmax, complex
    N(0,1) coefficients from SplitMix64 scaled by ``amp / (1 + kx + |kz| + n)``;
    the kx = 0 plane is made Hermitian in kz and the (0,0) pencil real, so the
    physical field is real.  The identity enters as 1 in the (0,0,0)
    coefficient of c11, c22, c33."""
    rng = SplitMix64(seed)
    nxh, nyp = nx // 2, ny + 1
    out = {}
    for ij in POLY:
        a = np.zeros((nxh, nz, nyp), dtype=np.complex128)
        for kx in range(0, min(kmax, nxh - 1) + 1):
            for kzs in range(-kmax, kmax + 1):
                kzi = kzs % nz
                for n in range(0, min(nmax, ny) + 1):
                    r = complex(rng.normal(), rng.normal())
                    a[kx, kzi, n] = amp * r / (1.0 + kx + abs(kzs) + n)
        j = np.arange(1, nz // 2)
        a[0, nz - j, :] = np.conj(a[0, j, :])
        a[0, 0, :] = a[0, 0, :].real
        a[:, nz // 2, :] = 0.0
        if mean is not None and ij in mean:
            a[0, 0, :] += _cheb_coeffs(np.asarray(mean[ij], dtype=np.float64))
        elif ij in ("11", "22", "33"):
            a[0, 0, 0] += 1.0
        out["c" + ij] = a
    return out


# --------------------------------------------------------------------------- #
#  The reference's new-run initial fields (init_flow.f90) and the conversion of
#  physical fields to the spectral state (main_dns.f90:1923-2005)
# --------------------------------------------------------------------------- #
def solver_ycoord(ny, yl=2.0):
    """main_dns.f90:1112-1116: ycoord(1) = yl/2, ycoord(k) = cos((k-1) pi/ny) yl/2."""
    y = np.cos(np.arange(ny + 1) * math.pi / ny) * (yl / 2.0)
    y[0] = yl / 2.0
    return y


def init_flow(flow_select, nx, ny, nz, yl, re, R_tau, uinf=1.0, kwall=0):
    """Physical initial velocity of init_flow.f90:23-40 on the unpadded grid,
    arrays (nx, nz, nyp) C-ordered, y index 0 at ycoord(1) = yl/2:
    still (flow_select <= 0 or 5), channel (1, init_flow.f90:82-139: the
    laminar profile re dPdx (d^2/2 - yl d/2), d = |y - y_min|, zero in the
    kwall rows next to each wall), Couette_2D (2, :147-186: uinf d/yl).
    Refused: Couette_3D (3) and vortex_Couette (12) divide by ycoord(nyp) -
    ycoord(nyp) = 0 (init_flow.f90:226-228, 818); vortex_channel (11) leaves
    the vortex centre vyc, vzc unset (init_flow.f90:515-676, where
    vortex_only sets them); the Blasius (4) and vortex_only (10) starts come
    with free-slip / outflow walls (setbcs, main_dns.f90:1331-1390) that this
    solver does not run."""
    nyp = ny + 1
    y = solver_ycoord(ny, yl)
    u = np.zeros((nx, nz, nyp))
    v = np.zeros_like(u)
    w = np.zeros_like(u)
    if flow_select <= 0 or flow_select == 5:
        return u, v, w
    if flow_select == 1:
        dpdx = -2.0 / yl * (2.0 * R_tau / (yl * re)) ** 2
        d = np.abs(y - y[-1])
        prof = re * dpdx * (0.5 * d ** 2 - yl / 2.0 * d)
        i1 = np.arange(1, nyp + 1)
        prof[(i1 <= kwall) | ((nyp - i1) <= kwall)] = 0.0
        u[:] = prof
        return u, v, w
    if flow_select == 2:
        u[:] = uinf * (np.abs(y - y[-1]) / abs(y[0] - y[-1]))
        return u, v, w
    if flow_select in (3, 12):
        raise ValueError(f"flow_select {flow_select}: the reference's initial u divides by "
                         "ycoord(nyp) - ycoord(nyp) = 0 (init_flow.f90:226-228, 818)")
    if flow_select == 11:
        raise ValueError("flow_select 11: vortex_channel uses the vortex centre vyc, vzc without "
                         "setting it (init_flow.f90:515-676)")
    raise NotImplementedError(f"flow_select {flow_select}: free-slip / outflow walls (setbcs) are not run "
                              "by this solver, so its initial field is not built")


def new_run_state(u, v, w, xl, zl):
    """The new-run branch of initial (main_dns.f90:1923-2005) from physical
    u, v, w (nx, nz, nyp): forward transforms (xyzfft, helpers.f90:6-87:
    DCT-I along y over ny, ends halved; r2c along x over nx keeping nx/2
    modes; c2c along z over nz), omega_y = i kz u - i kx w with the (0,0)
    mode zero, the symmetry projection (norm, helpers.f90:91-158), u0 and w0
    = the real (0,0) pencils of u and w, v(0,0) = 0.  Returns dict(v, omy,
    u0, w0) in the state layout ((nx/2, nz, nyp) complex; u0, w0 (nyp,))."""
    from scipy.fft import dct
    nx, nz, nyp = np.shape(u)
    ny, nxh = nyp - 1, nx // 2

    def fwd(b):
        t = dct(np.asarray(b, dtype=np.float64), type=1, axis=-1) / float(ny)
        t[..., 0] /= 2.0
        t[..., -1] /= 2.0
        t = np.transpose(t, (2, 1, 0))                       # (nyp, nz, nx)
        xt = np.fft.rfft(t, axis=2)[:, :, :nxh] / float(nx)
        zt = np.fft.fft(xt, axis=1) / float(nz)
        return np.ascontiguousarray(np.transpose(zt, (2, 1, 0)))

    def norm(a):
        a[0, 0, :] = a[0, 0, :].real
        j = np.arange(1, nz // 2)
        a[0, j, :] = 0.5 * (a[0, j, :] + np.conj(a[0, nz - j, :]))
        a[0, nz - j, :] = np.conj(a[0, j, :])
        a[:, nz // 2, :] = 0.0
        return a

    us, vs, ws = fwd(u), fwd(v), fwd(w)
    kx = (np.arange(nxh) * (2.0 * math.pi / xl))[:, None, None]
    kzi = np.arange(nz)
    kzi[nz // 2:] -= nz
    kz = (kzi * (2.0 * math.pi / zl))[None, :, None]
    omy = 1j * kz * us - 1j * kx * ws
    omy[0, 0, :] = 0.0
    for a in (us, vs, ws, omy):
        norm(a)
    u0 = us[0, 0, :].real.copy()
    w0 = ws[0, 0, :].real.copy()
    vs[0, 0, :] = 0.0
    return dict(v=vs, omy=omy, u0=u0, w0=w0)


def ensemble_members(seed, n=16, sizes=(64, 96, 128), nys=(48, 64, 80, 96), rtau=(120.0, 240.0)):
    """Seeded ensemble of independent channels of uneven size: per member
    (nx, ny, nz, Re_tau, state seed) with nx = nz drawn from ``sizes``, ny
    from ``nys`` and Re_tau uniform in ``rtau`` (SplitMix64 of ``seed``).
    The first members cover the smallest (64 x 49 x 64) and the largest
    (128 x 97 x 128) grid so that every ensemble spans the whole range."""
    rng = SplitMix64(seed)
    out = []
    for m in range(n):
        if m == 0:
            nx, ny = sizes[0], nys[0]
        elif m == 1:
            nx, ny = sizes[-1], nys[-1]
        else:
            nx = sizes[int(rng.uniform() * len(sizes))]
            ny = nys[int(rng.uniform() * len(nys))]
        rt = rtau[0] + (rtau[1] - rtau[0]) * rng.uniform()
        out.append((nx, ny, nx, rt, int(rng.next_u64() & 0x7FFFFFFF)))
    return out
