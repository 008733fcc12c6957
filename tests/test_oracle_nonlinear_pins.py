"""Pins of the oracle's nonlinear path that do not use the oracle's own
formulas.

* Linear stability of plane Poiseuille flow: the least stable Orr-Sommerfeld
  mode at centreline Re 7500, alpha 1 (Orszag 1971, J. Fluid Mech. 50:689,
  c = 0.24989154 + 0.00223497 i) and an oblique Orr-Sommerfeld mode with its
  forced Squire (omega_y) response at alpha = beta = 1, Re 2000.  The
  eigenvalues come from an independent Chebyshev collocation of the
  textbook equations (Schmid & Henningson 2001, eqs. 3.14-3.15); a small
  eigenmode on the laminar base, advanced by the oracle's full step (u x
  omega on the 3/2 grid, smoothing, nonlin, phirhs, the tau solves and the
  influence matrix, AB2/CN), must follow exp(sigma t).  This pins the signs of
  the base-flow terms of the cross product, nonlin's curl (the kx and kz
  terms of gn and fn), veloc/vort and the AB2 weighting.
* u . (u x a) = 0 on the 3/2 grid, read back through Parseval on the
  retained modes: pins every term of vcw3dp's cross product.
* nonlin annihilates gradients (curl grad = 0).
* gradv, gradscl and derivscji against analytic derivatives of single
  Fourier modes times polynomials in y.
* Smoothing: the NL filter is the exponential filter exp(-3 eta^10) with
  eta = 1 at the highest resolved mode (kz = nz/2, n = ny), and only the NL
  terms are filtered (a state mode near the z cut-off decays at its Stokes
  rate).
"""
import math

import numpy as np
import pytest
import scipy.fft as sfft
import scipy.linalg as sla
from numpy.polynomial import chebyshev as cheb

from oracle import kmm_oracle as O
from tests._cases import params


# ------------------------------------------------------ collocation (textbook)
def _cheb_mats(N):
    y = np.cos(np.pi * np.arange(N + 1) / N)
    T = np.zeros((N + 1, N + 1))
    T1, T2, T4 = np.zeros_like(T), np.zeros_like(T), np.zeros_like(T)
    for n in range(N + 1):
        e = np.zeros(N + 1)
        e[n] = 1.0
        T[:, n] = cheb.chebval(y, e)
        T1[:, n] = cheb.chebval(y, cheb.chebder(e))
        T2[:, n] = cheb.chebval(y, cheb.chebder(e, 2))
        T4[:, n] = cheb.chebval(y, cheb.chebder(e, 4))
    return y, T, T1, T2, T4


def os_squire_mode(N, alpha, beta, Re):
    """Least stable Orr-Sommerfeld mode of U = 1 - y^2 for a perturbation
    ~ exp(i(alpha x + beta z) + sigma t):
        sigma (D^2 - k^2) v = [-i alpha U (D^2 - k^2) + i alpha U'' + (D^2 - k^2)^2 / Re] v,
    v = v' = 0 at y = +-1, and its forced Squire response
        sigma eta = [-i alpha U + (D^2 - k^2) / Re] eta - i beta U' v,  eta(+-1) = 0,
    with eta = du/dz - dw/dx.  Returns sigma and the Chebyshev coefficients
    of v and eta."""
    y, T, T1, T2, T4 = _cheb_mats(N)
    k2 = alpha * alpha + beta * beta
    U, Up, Upp = 1.0 - y * y, -2.0 * y, -2.0 * np.ones_like(y)
    L = T2 - k2 * T
    L4 = T4 - 2.0 * k2 * T2 + k2 * k2 * T
    A = (-1j * alpha * U[:, None] * L + 1j * alpha * Upp[:, None] * T + L4 / Re).astype(complex)
    B = L.astype(complex)
    for r, row in ((0, T[0]), (1, T1[0]), (N - 1, T1[N]), (N, T[N])):
        A[r] = row
        B[r] = 0.0
    w, V = sla.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 1e6)
    i = int(np.argmax(np.where(ok, w.real, -np.inf)))
    s, v = w[i], V[:, i]
    M = (s * T + 1j * alpha * U[:, None] * T - L / Re).astype(complex)
    rhs = -1j * beta * Up * (T @ v)
    M[0], M[N] = T[0], T[N]
    rhs[0] = rhs[N] = 0.0
    return s, v, np.linalg.solve(M, rhs)


def _run_mode(Re, ny, nz, kz_index, sigma, v, eta, dt, nsteps, eps=1e-7):
    """Small eigenmode on the laminar base (centreline 1: R_tau^2 = 2 Re,
    so dPdx = -2/Re), advanced by the oracle; returns the relative
    deviations of v and omega_y of the mode from exp(sigma t)."""
    p = params(4, ny, nz, xl=2 * math.pi, zl=2 * math.pi, re=Re, dt=dt, R_tau=math.sqrt(2.0 * Re))
    z = np.zeros((p.nxh, p.nz, p.nyp), complex)
    st = dict(v=z.copy(), omy=z.copy(), u0=np.zeros(p.nyp), w0=np.zeros(p.nyp))
    st["u0"][0], st["u0"][2] = 0.5, -0.5          # 1 - y^2 = T0/2 - T2/2
    scale = eps / np.abs(v).max()
    st["v"][1, kz_index] = scale * v
    st["omy"][1, kz_index] = scale * eta
    sol = O.Solver(p)
    sol.startup(st)
    a0v, a0o = st["v"][1, kz_index].copy(), st["omy"][1, kz_index].copy()
    for _ in range(nsteps):
        sol.step(st)
    g = np.exp(sigma * dt * nsteps)
    dv = np.abs(st["v"][1, kz_index] - g * a0v).max() / np.abs(a0v).max()
    do = 0.0
    if np.abs(a0o).max() > 0:
        do = np.abs(st["omy"][1, kz_index] - g * a0o).max() / np.abs(a0o).max()
    lam = np.vdot(a0v, st["v"][1, kz_index]) / np.vdot(a0v, a0v)
    return dv, do, lam


def test_collocation_reproduces_orszag_eigenvalue():
    s, _, _ = os_squire_mode(64, 1.0, 0.0, 7500.0)
    c = 1j * s          # sigma = -i alpha c, alpha = 1
    assert abs(c.real - 0.24989154) < 2e-8 and abs(c.imag - 0.00223497) < 2e-8


def test_tollmien_schlichting_wave_grows_at_orszag_rate():
    s, v, eta = os_squire_mode(64, 1.0, 0.0, 7500.0)
    dv, _, lam = _run_mode(7500.0, 64, 4, 0, s, v, eta, dt=0.01, nsteps=100)
    # AB2/CN time error at dt = 0.01 is ~3e-6 here (2nd order: 8e-7 at 0.005)
    assert dv < 2e-5
    assert abs(lam) > 1.0    # the TS wave is unstable at Re 7500
    assert abs(math.log(abs(lam)) / 1.0 - s.real) < 0.02 * s.real


def test_oblique_orr_sommerfeld_squire_mode():
    s, v, eta = os_squire_mode(64, 1.0, 1.0, 2000.0)
    assert np.abs(eta).max() > 0.5 * np.abs(v).max()      # the Squire response is not negligible
    dv, do, _ = _run_mode(2000.0, 64, 16, 1, s, v, eta, dt=0.01, nsteps=100)
    assert dv < 3e-5 and do < 3e-5


# ---------------------------------------------------------- cross product
def _band_limited(p, rng, amp=1.0):
    a = amp * (rng.standard_normal((p.nxh, p.nz, p.nyp)) + 1j * rng.standard_normal((p.nxh, p.nz, p.nyp)))
    a *= np.exp(-0.2 * np.arange(p.nyp))[None, None, :]
    return O.norm(a)


def _phys_y(a, p):
    """Chebyshev coefficients -> values at the Gauss-Lobatto points (x/z
    stay spectral)."""
    f = a * (p.c / 2.0)
    return sfft.dct(f.real, type=1, axis=-1) + 1j * sfft.dct(f.imag, type=1, axis=-1)


def _plane_mean_dot(a, b):
    """Mean over the (x, z) plane of the product of two real fields given by
    their modes kx >= 0 (kx > 0 counted twice), per y."""
    w = np.full(a.shape[0], 2.0)
    w[0] = 1.0
    return np.einsum("i,ik->k", w, np.real(np.conj(a) * b).sum(axis=1))


def test_cross_product_is_orthogonal_to_velocity():
    """sum over each plane of u . (u x a) = 0 for any fields u, a."""
    p = params(8, 16, 12, R_tau=0.0)        # dPdx = 0
    rng = np.random.default_rng(41)
    f = {k: _band_limited(p, rng) for k in ("u", "v", "w", "omx", "omy", "omz")}
    st = dict(fxintg=np.zeros((p.mx, p.mz, p.nyp)), fyintg=np.zeros((p.mx, p.mz, p.nyp)),
              fzintg=np.zeros((p.mx, p.mz, p.nyp)))
    gn, fn, omz, _, _ = O.vcw3dp(st, p, 1, f)
    uy, vy, wy = (_phys_y(f[k], p) for k in ("u", "v", "w"))
    N = [_phys_y(a, p) for a in (gn, fn, omz)]
    s = _plane_mean_dot(uy, N[0]) + _plane_mean_dot(vy, N[1]) + _plane_mean_dot(wy, N[2])
    scale = _plane_mean_dot(uy, uy).max() * np.sqrt(_plane_mean_dot(_phys_y(f["omx"], p), _phys_y(f["omx"], p)).max())
    assert np.abs(s).max() < 1e-12 * scale
    # and each component alone is not orthogonal (the test has teeth)
    assert np.abs(_plane_mean_dot(uy, N[0])).max() > 1e-3 * scale


# ------------------------------------------------------------------ nonlin
def test_nonlin_annihilates_gradient_fields():
    """curl grad phi = 0: nonlin's fn = -(curl curl N)_y and gn = (curl N)_y
    vanish for N = grad phi."""
    p = params(8, 24, 12)
    rng = np.random.default_rng(5)
    phi = _band_limited(p, rng)
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    Nx, Ny, Nz = 1j * kx * phi, O.cderiv(phi, p.dyde), 1j * kz * phi
    fn, gn = O.nonlin(Ny, Nz, Nx, p)          # nonlin(fn, omz, gn): y, z, x components
    s = np.abs(Nx).max()
    assert np.abs(fn).max() < 1e-12 * s * np.abs(kx).max() ** 2
    assert np.abs(gn).max() < 1e-12 * s * np.abs(kx).max()
    # a solenoidal field is not annihilated
    fn2, gn2 = O.nonlin(np.zeros_like(Nx), Nx, Nz, p)
    assert np.abs(gn2).max() > 1e-2 * s


# --------------------------------------------------------------- gradients
def _mode_field(p, a, b, poly, phase):
    """Physical (nx, nz, nyp) field poly(y) cos(a alpha x + b beta z + phase)
    and the function computing its derivatives analytically on any grid."""
    def ev(x, z, y, d=None):
        arg = a * p.alpha * x + b * p.beta * z + phase
        if d is None:
            return cheb.chebval(y, poly) * np.cos(arg)
        if d == "x":
            return -a * p.alpha * cheb.chebval(y, poly) * np.sin(arg)
        if d == "z":
            return -b * p.beta * cheb.chebval(y, poly) * np.sin(arg)
        if d == "y":
            return cheb.chebval(y, cheb.chebder(poly)) * np.cos(arg)
        if d == "lap":
            k2 = (a * p.alpha) ** 2 + (b * p.beta) ** 2
            return (cheb.chebval(y, cheb.chebder(poly, 2)) - k2 * cheb.chebval(y, poly)) * np.cos(arg)
    return ev


def _grid(p, m):
    nx, nz = (p.mx, p.mz) if m else (p.nx, p.nz)
    X = (np.arange(nx) * p.xl / nx)[:, None, None]
    Z = (np.arange(nz) * p.zl / nz)[None, :, None]
    return X, Z, p.ycoord[None, None, :]


def test_gradv_gradscl_derivscji_on_analytic_modes():
    p = params(12, 20, 8)
    X, Z, Y = _grid(p, False)
    Xm, Zm, Ym = _grid(p, True)
    defs = {"u": (2, 1, [0.2, -0.5, 0.0, 0.3, 0.1], 0.4), "v": (1, -2, [0.0, 0.7, 0.2, 0.0, -0.4, 0.05], 1.1),
            "w": (3, 2, [1.0, 0.0, -0.6, 0.2], -0.3)}
    ev = {k: _mode_field(p, *d) for k, d in defs.items()}
    spec = {k: O.xyzfft(ev[k](X, Z, Y), p) for k in ev}
    g = O.gradv(spec["u"], spec["v"], spec["w"], p)
    dirs = ("x", "y", "z")
    for i, comp in enumerate(("u", "v", "w")):
        for k in range(3):
            got = O.to_physical(g[f"u{i + 1}{k + 1}"], p)
            ex = ev[comp](Xm, Zm, Ym, dirs[k])
            assert np.abs(got - ex).max() < 1e-11 * max(1.0, np.abs(ex).max()), (comp, dirs[k])
        got = O.to_physical(g["L" + comp], p)
        ex = ev[comp](Xm, Zm, Ym, "lap")
        assert np.abs(got - ex).max() < 1e-10 * np.abs(ex).max(), comp
    sx, sy, sz = O.gradscl(spec["v"], p)
    for got, d in ((sx, "x"), (sy, "y"), (sz, "z")):
        ex = ev["v"](Xm, Zm, Ym, d)
        assert np.abs(O.to_physical(got, p) - ex).max() < 1e-11 * np.abs(ex).max()
    C = {ij: spec["uvw"[n % 3]] * (n + 1) for n, ij in enumerate(O.POLY)}
    dC = O.derivscji(C, p)
    for n, ij in enumerate(O.POLY):
        for k, d in enumerate(dirs):
            ex = (n + 1) * ev["uvw"[n % 3]](Xm, Zm, Ym, d)
            assert np.abs(O.to_physical(dC[ij][k], p) - ex).max() < 1e-11 * max(1.0, np.abs(ex).max())


# --------------------------------------------------------------- smoothing
def test_nl_filter_is_exponential_filter_on_the_resolved_band():
    """Smoothing multiplies the NL terms by sigma(eta_z) sigma(eta_y) with
    sigma(eta) = exp(-3 eta^10) and eta = 1 at the highest resolved mode
    (|kz| = nz/2, n = ny).  gn = i(kz gn - kx omz) is diagonal per mode, so
    its ratio with and without smoothing is the filter."""
    p = params(8, 16, 16)
    rng = np.random.default_rng(17)
    sol = O.Solver(p)
    u, v, w = (_band_limited(p, rng, 0.01) for _ in range(3))
    out = []
    for smooth in (False, True):
        st = dict(omy=_band_limited(p, np.random.default_rng(3), 0.01),
                  fxintg=np.zeros((p.mx, p.mz, p.nyp)), fyintg=np.zeros((p.mx, p.mz, p.nyp)),
                  fzintg=np.zeros((p.mx, p.mz, p.nyp)))
        sol._nonlinear(st, u, v, w, it=1, smooth=smooth)
        out.append(st["gn"])
    kz_abs = np.abs(np.fft.fftfreq(p.nz, 1.0 / p.nz))
    eta_z, eta_y = kz_abs / (p.nz / 2), np.arange(p.nyp) / p.ny
    sig = np.exp(-3.0 * eta_z ** 10)[:, None] * np.exp(-3.0 * eta_y ** 10)[None, :]
    mask = np.abs(out[0]) > 1e-6 * np.abs(out[0]).max()
    ratio = np.where(mask, out[1] / np.where(mask, out[0], 1.0), 0.0)
    assert np.abs(np.where(mask, ratio - sig[None], 0.0)).max() < 1e-12
    # the highest resolved kz below the Nyquist is damped hard; low modes are not
    j = p.nz // 2 - 1
    assert np.all(np.abs(ratio[:, j, 0][mask[:, j, 0]] - math.exp(-3.0 * (j / (p.nz / 2)) ** 10)) < 1e-12)
    assert math.exp(-3.0 * (j / (p.nz / 2)) ** 10) < 0.5


def test_state_is_not_filtered():
    """A small Squire mode at the highest resolved |kz| below the Nyquist
    decays at the Crank-Nicolson factor of its Stokes eigenvalue: the
    smoothing touches only the NL terms (quirk Q3)."""
    p = params(4, 32, 16, R_tau=0.0, re=100.0, dt=0.01)
    kz = p.nz // 2 - 1
    k2 = p.wavx[1] ** 2 + p.wavz[kz] ** 2
    y, T, T1, T2, T4 = _cheb_mats(p.ny)
    A = (T2 - k2 * T) / p.re
    B = T.copy()
    A[0], A[-1], B[0], B[-1] = T[0], T[-1], 0.0, 0.0
    w, V = sla.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 1e8)
    i = int(np.argmax(np.where(ok, w.real, -np.inf)))
    s, vec = w[i].real, np.real(V[:, i])
    z = np.zeros((p.nxh, p.nz, p.nyp), complex)
    st = dict(v=z.copy(), omy=z.copy(), u0=np.zeros(p.nyp), w0=np.zeros(p.nyp))
    st["omy"][1, kz] = 1e-9 * vec
    sol = O.Solver(p)
    sol.startup(st)
    a0 = st["omy"][1, kz].copy()
    for _ in range(5):
        sol.step(st)
    g = (1 + 0.5 * s * p.dt) / (1 - 0.5 * s * p.dt)
    assert np.abs(st["omy"][1, kz] - g ** 5 * a0).max() < 1e-8 * np.abs(a0).max()


def test_vort_is_the_curl_and_veloc_inverts_it():
    """vort gives omega_x = dw/dy - dv/dz and omega_z = dv/dx - du/dy of
    analytic fields; veloc returns u, w with zero divergence whose
    du/dz - dw/dx is the omega_y it was given."""
    p = params(12, 20, 8)
    X, Z, Y = _grid(p, False)
    Xm, Zm, Ym = _grid(p, True)
    ev = {"u": _mode_field(p, 2, 1, [0.2, -0.5, 0.0, 0.3, 0.1], 0.4),
          "v": _mode_field(p, 1, -2, [0.0, 0.7, 0.2, 0.0, -0.4, 0.05], 1.1),
          "w": _mode_field(p, 3, 2, [1.0, 0.0, -0.6, 0.2], -0.3)}
    spec = {k: O.xyzfft(ev[k](X, Z, Y), p) for k in ev}
    omx, omz = O.vort(spec["u"], spec["v"], spec["w"], p)
    ex_x = ev["w"](Xm, Zm, Ym, "y") - ev["v"](Xm, Zm, Ym, "z")
    ex_z = ev["v"](Xm, Zm, Ym, "x") - ev["u"](Xm, Zm, Ym, "y")
    assert np.abs(O.to_physical(omx, p) - ex_x).max() < 1e-11 * np.abs(ex_x).max()
    assert np.abs(O.to_physical(omz, p) - ex_z).max() < 1e-11 * np.abs(ex_z).max()
    # veloc: random v, omega_y -> u, w
    rng = np.random.default_rng(8)
    v, omy = _band_limited(p, rng), _band_limited(p, rng)
    u0, w0 = rng.standard_normal(p.nyp), rng.standard_normal(p.nyp)
    u, v2, w = O.veloc(v, omy, u0, w0, p)
    kx = p.wavx[:, None, None]
    kz = p.wavz[None, :, None]
    dv = np.stack([np.stack([cheb.chebder(v2[a, b], scl=p.dyde) for b in range(p.nz)]) for a in range(p.nxh)])
    dv = np.concatenate([dv, np.zeros((p.nxh, p.nz, 1))], axis=-1)
    div = 1j * kx * u + dv + 1j * kz * w
    assert np.abs(div).max() < 1e-12 * np.abs(u).max() * np.abs(kx).max()
    cy = 1j * kz * u - 1j * kx * w
    cy[0, 0] = omy[0, 0]      # the mean modes come from u0, w0
    assert np.abs(cy - omy).max() < 1e-12 * np.abs(omy).max()
    assert np.all(u[0, 0] == u0) and np.all(w[0, 0] == w0)


def test_vcw3dp_products_of_analytic_modes():
    """u x omega and -u . grad(theta) of single Fourier modes times low
    polynomials: the products stay inside the retained band, so the
    dealiased output equals the analytic product on the 3/2 grid."""
    p = params(16, 20, 12, scalar=True, R_tau=0.0)
    X, Z, Y = _grid(p, False)
    Xm, Zm, Ym = _grid(p, True)
    modes = {"u": (2, 1, [0.2, -0.5, 0.3], 0.4), "v": (1, -1, [0.1, 0.7, 0.0, -0.2], 1.1),
             "w": (3, 0, [1.0, 0.0, -0.6], -0.3), "omx": (1, 2, [0.3, 0.2, 0.5], 0.2),
             "omy": (2, -1, [-0.4, 0.1, 0.0, 0.3], 2.0), "omz": (0, 1, [0.6, -0.3, 0.2], -1.0),
             "th": (2, 1, [0.5, 0.25, -0.3, 0.1], 0.7)}
    ev = {k: _mode_field(p, *d) for k, d in modes.items()}
    f = {k: O.norm(O.xyzfft(ev[k](X, Z, Y), p)) for k in modes if k != "th"}
    th = O.norm(O.xyzfft(ev["th"](X, Z, Y), p))
    f["sclx"], f["scly"], f["sclz"] = O.gradscl(th, p)
    st = dict(fxintg=np.zeros((p.mx, p.mz, p.nyp)), fyintg=np.zeros((p.mx, p.mz, p.nyp)),
              fzintg=np.zeros((p.mx, p.mz, p.nyp)))
    gn, fn, omz, scn, _ = O.vcw3dp(st, p, 1, f)
    P = {k: ev[k](Xm, Zm, Ym) for k in modes}
    ex = {"x": P["v"] * P["omz"] - P["w"] * P["omy"], "y": P["w"] * P["omx"] - P["u"] * P["omz"],
          "z": P["u"] * P["omy"] - P["v"] * P["omx"],
          "s": -(P["u"] * ev["th"](Xm, Zm, Ym, "x") + P["v"] * ev["th"](Xm, Zm, Ym, "y")
                 + P["w"] * ev["th"](Xm, Zm, Ym, "z"))}
    for got, k in ((gn, "x"), (fn, "y"), (omz, "z"), (scn, "s")):
        assert np.abs(O.to_physical(got, p) - ex[k]).max() < 1e-12 * np.abs(ex[k]).max(), k


def test_scalar_mode_advected_by_poiseuille_flow():
    """A small scalar mode theta(y) exp(i alpha x) on the laminar base
    follows the advection-diffusion eigenvalue
        sigma theta = -i alpha U theta + (D^2 - alpha^2) theta / (re diff),
    theta'(+-1) = 0 (zero-flux walls), computed by collocation."""
    re, diff, N = 500.0, 0.71, 32
    y, T, T1, T2, T4 = _cheb_mats(N)
    U = 1.0 - y * y
    A = (-1j * U[:, None] * T + (T2 - T) / (re * diff)).astype(complex)
    B = T.astype(complex)
    A[0], A[N], B[0], B[N] = T1[0], T1[N], 0.0, 0.0
    w, V = sla.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 1e6)
    i = int(np.argmax(np.where(ok, w.real, -np.inf)))
    s, vec = w[i], V[:, i]
    p = params(4, N, 4, scalar=True, diff=diff, xl=2 * math.pi, re=re, dt=0.005, R_tau=math.sqrt(2.0 * re))
    z = np.zeros((p.nxh, p.nz, p.nyp), complex)
    st = dict(v=z.copy(), omy=z.copy(), u0=np.zeros(p.nyp), w0=np.zeros(p.nyp), scalar=z.copy())
    st["u0"][0], st["u0"][2] = 0.5, -0.5
    st["scalar"][1, 0] = 1e-6 * vec / np.abs(vec).max()
    sol = O.Solver(p)
    sol.startup(st)
    a0 = st["scalar"][1, 0].copy()
    n = 100
    for _ in range(n):
        sol.step(st)
    g = np.exp(s * p.dt * n)
    assert np.abs(st["scalar"][1, 0] - g * a0).max() < 2e-5 * np.abs(a0).max()
