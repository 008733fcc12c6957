"""Per-routine entries of the C ABI against the oracle's restatements of the
same reference procedures, on seeded spectral fields (FP64, relative to the
largest magnitude of the oracle's result): phirhs, nonlin, veloc, vort,
gradscl, derivscji, polyforce, subforce, polyNL, norm, uzero, and vcw3dp
with the reference's argument order."""
import numpy as np
import pytest

from oracle import kmm_oracle as O
from tests._cases import params, rel, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

TOL = 1e-12


def rand_field(p, rng, scale=1.0):
    f = scale * (rng.standard_normal((p.nxh, p.nz, p.nyp)) + 1j * rng.standard_normal((p.nxh, p.nz, p.nyp)))
    # smooth in y so that derivatives stay O(1)
    f *= np.exp(-0.1 * np.arange(p.nyp))[None, None, :]
    return O.norm(f)


@pytest.fixture(scope="module", params=[(32, 32, 16), (64, 64, 32)])
def setup(request):
    nx, ny, nz = request.param
    p = params(nx, ny, nz, dt=0.004, polymer=True, tpoly=0.7)
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, polymer=1,
                          tpoly=p.tpoly, crstrt=1)
    ch = kmm.Channel(cfg)
    yield p, ch, np.random.default_rng(nx + ny)
    ch.close()


def test_phirhs(setup):
    p, ch, rng = setup
    v, fn, fnm1 = (rand_field(p, rng) for _ in range(3))
    for var in (1.0, 0.71):
        got = ch.phirhs(v, fn, fnm1, var)
        assert rel(got, O.phirhs(v, fn, fnm1, var, p)) < TOL


def test_nonlin_veloc_vort(setup):
    p, ch, rng = setup
    h2, h3, h1 = (rand_field(p, rng) for _ in range(3))
    fn, gn = ch.nonlin(h2, h3, h1)
    rfn, rgn = O.nonlin(h2, h3, h1, p)
    assert rel(fn, rfn) < TOL and rel(gn, rgn) < TOL
    st = seeded(p, 5, amp=0.3)
    u, v, w = ch.veloc(st["v"], st["omy"], st["u0"], st["w0"])
    ru, rv, rw = O.veloc(st["v"], st["omy"], st["u0"], st["w0"], p)
    for a, b in ((u, ru), (v, rv), (w, rw)):
        assert rel(a, b) < TOL
    omx, omz = ch.vort(ru, rv, rw)
    rx, rz = O.vort(ru, rv, rw, p)
    assert rel(omx, rx) < TOL and rel(omz, rz) < TOL


def test_gradscl_and_derivscji(setup):
    p, ch, rng = setup
    sc = rand_field(p, rng)
    for a, b in zip(ch.gradscl(sc), O.gradscl(sc, p)):
        assert rel(a, b) < TOL
    cij = [rand_field(p, rng) for _ in range(9)]
    got = ch.derivscji(cij)
    for i in range(3):
        for j in range(3):
            gx, gy, gz = O.gradscl(cij[3 * i + j], p)   # i kx c, D c, i kz c
            for k, ref in zip((1, 2, 3), (gx, gy, gz)):
                assert rel(got[f"{i + 1}{j + 1}{k}"], ref) < TOL, (i, j, k)


def test_polyforce_subforce_polynl(setup):
    p, ch, rng = setup
    s = {q: rand_field(p, rng) for q in O.POLY}
    t = ch.polyforce(s)
    for a, b in zip(t, O.polyforce(s, p)):
        assert rel(a, b) < TOL
    gn, fn, omz = (rand_field(p, rng) for _ in range(3))
    g2, f2, o2 = ch.subforce(gn, fn, omz, *t)
    assert rel(g2, gn + t[0]) < TOL and rel(f2, fn + t[1]) < TOL and rel(o2, omz + t[2]) < TOL
    d = [{q: rand_field(p, rng) for q in O.POLY} for _ in range(4)]
    got = ch.polyNL(*d)
    ref = O.polyNL(*d, p)
    for q in O.POLY:
        assert rel(got[q], ref[q]) < TOL


def test_norm(setup):
    p, ch, rng = setup
    a = rng.standard_normal((p.nxh, p.nz, p.nyp)) + 1j * rng.standard_normal((p.nxh, p.nz, p.nyp))
    got = ch.norm(a)
    ref = O.norm(a.copy())
    assert np.array_equal(got[:, p.nz // 2, :], np.zeros_like(got[:, p.nz // 2, :]))
    assert rel(got, ref) < 1e-15
    assert rel(ch.norm(got), got) == 0.0   # idempotent


def test_uzero(setup):
    p, ch, rng = setup
    st = seeded(p, 9)
    u0 = st["u0"]
    h1, h1l = 0.01 * rng.standard_normal(p.nyp), 0.01 * rng.standard_normal(p.nyp)
    for flag in (1, 3):
        got = ch.uzero(u0, h1, h1l, flag)
        assert rel(got, O.uzero(u0, h1, h1l, flag, p)) < 1e-11, flag


def test_vcw3dp_polymer_argument_list():
    """vcw3dp with the POLYMER argument list (solvers.f90:1726-1740, FENE-P
    products :2901-2983) on a seeded state and conformation, against the
    oracle's vcw3dp: u x omega and c_n, str_n, qp of every component."""
    from tests._cases import poly_state
    p = params(32, 32, 32, dt=0.005, polymer=True, tpoly=0.5, zlmax=30.0, qbeta=0.9, src_start=0)
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, polymer=1,
                          tpoly=p.tpoly, zlmax=p.zlmax, qbeta=p.qbeta, crstrt=1, src_start=0)
    st = seeded(p, 44, amp=0.3)
    C = poly_state(p, 45, amp=0.05)
    C = {ij: C["c" + ij] for ij in O.POLY}
    u, v, w = O.veloc(st["v"], st["omy"], st["u0"], st["w0"], p)
    omx, omz = O.vort(u, v, w, p)
    for a in (u, v, w, omx, omz):
        O.norm(a)
    gv = O.gradv(u, v, w, p)
    dC = O.derivscji(C, p)
    fields = dict(u=u, v=v, w=w, omx=omx, omy=st["omy"], omz=omz, C=C, dC=dC, gradv=gv)
    ost = dict(fxintg=np.zeros((p.mx, p.mz, p.nyp)), fyintg=np.zeros((p.mx, p.mz, p.nyp)),
               fzintg=np.zeros((p.mx, p.mz, p.nyp)))
    gn, fn, oz, _, phys = O.vcw3dp(ost, p, 1, fields)
    uij = [gv[f"u{i}{k}"] for i in (1, 2, 3) for k in (1, 2, 3)]
    full = {(i, j): C[O._CIDX[(i, j)]] for i in range(3) for j in range(3)}
    cij = [full[(i, j)] for i in range(3) for j in range(3)]
    dc = [None] * 27
    for i in range(3):
        for j in range(3):
            for k in range(3):
                dc[9 * j + 3 * i + k] = dC[O._CIDX[(i, j)]][k]
    with kmm.Channel(cfg) as ch:
        g2, f2, z2, cn, sn, qp = ch.vcw3dp_poly(u, v, w, omx, st["omy"], omz, uij, cij, dc)
    for a, b in ((g2, gn), (f2, fn), (z2, oz)):
        assert rel(a, O.norm(b.copy())) < 1e-12
    for ij in O.POLY:
        for got, key in ((cn, "cn"), (sn, "str"), (qp, "qp")):
            want = O.norm(phys["poly"][key][ij].copy())
            assert rel(got[ij], want) < 1e-11, (key, ij, rel(got[ij], want))
