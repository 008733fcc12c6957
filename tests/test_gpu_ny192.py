"""ny = 192 (193 Chebyshev points), the wall-normal resolution of the Re_tau
395 viscoelastic workload: each pencil holds its 96 coefficient pairs in two
segments of a wavefront, the Chebyshev transforms are 384-point FFTs and the
tau eliminations scan across both segments.  Against the oracle on seeded
inputs: cderiv, solve (Dirichlet and Neumann) and penta to the tolerances
of the ny <= 128 tests; the channel step with the scalar and the FENE-P
step at 32 x 193 x 32 (FP64, 1e-10 per spectral field); mixed precision at
64 x 193 x 64; and the full 256 x 193 x 256 FENE-P workload through
start-up and two steps against the sampled oracle fixture
tests/golden/c5_fenep.npz (tests/golden/make_c5_fixture.py)."""
import os

import numpy as np
import pytest

from oracle import kmm_oracle as O
from tests import _c5_case as c5
from tests._cases import params, poly_state, rel, seeded, shear_conformation_mean

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

NY = 192
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c5_fenep.npz")


def log(*a):
    print(*a, flush=True)


def cfg_for(p, **kw):
    return kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, yl=p.yl, re=p.re, dt=p.dt, theta=p.theta,
                           r_tau=p.R_tau, gain=p.gain, ugain=p.ugain, scalar=int(p.scalar), diff=p.diff,
                           scl_flag=p.scl_flag, deltaT=p.deltaT, perturbtime=p.perturbtime, **kw)


def test_cderiv_ny192():
    p = params(32, NY, 16)
    rng = np.random.default_rng(192)
    f = rng.standard_normal((p.nxh, p.nz, p.nyp)) + 1j * rng.standard_normal((p.nxh, p.nz, p.nyp))
    with kmm.Channel(cfg_for(p)) as ch:
        d = ch.cderiv(f)
    assert rel(d, O.cderiv(f, p.dyde)) < 1e-13


@pytest.mark.parametrize("g,ib", [(0.0, 0), (3.7, 0), (2.5e6, 0), (0.0, 1), (2.0, 1)])
def test_solve_ny192(g, ib):
    p = params(32, NY, 16)
    rng = np.random.default_rng(1920 + ib)
    f = rng.standard_normal((p.nz, p.nyp)) + 1j * rng.standard_normal((p.nz, p.nyp))
    bt = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    bb = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    x = p.wavx[3] ** 2
    with kmm.Channel(cfg_for(p)) as ch:
        u = ch.solve(f, x, g, p.dyde, bt, bb, ib)
    ref = O.solve(f, p.wavz ** 2 + x + g, p.dyde, bt, bb, ib)
    assert rel(u, ref) < 1e-11


@pytest.mark.parametrize("g,robin", [(5.0, (1.0, 0.0, 1.0, 0.0)), (1.1e6, (0.0, 1.0, 0.0, 1.0)),
                                     (0.0, (1.0, 0.3, 0.5, 0.2))])
def test_penta_ny192(g, robin):
    p = params(32, NY, 16)
    at, bt, ab, bb = robin
    rng = np.random.default_rng(1921)
    f = rng.standard_normal((p.nz, p.nyp)) + 1j * rng.standard_normal((p.nz, p.nyp))
    gt = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    gb = rng.standard_normal(p.nz) + 1j * rng.standard_normal(p.nz)
    x = p.wavx[2] ** 2
    with kmm.Channel(cfg_for(p)) as ch:
        u = ch.penta(f, x, g, p.dyde, 0, at, bt, gt, ab, bb, gb)
    ref = O.penta(f, p.wavz ** 2 + x + g, p.dyde, 0, at, bt, gt, ab, bb, gb)
    assert rel(u, ref) < 1e-11


FIELDS = ["v", "omy", "fn", "fnm1", "gn", "gnm1", "u0", "w0", "h1n", "h3n"]


def run_steps(p, nsteps, precision, tol, seed):
    st = seeded(p, seed)
    sol = O.Solver(p)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    with kmm.Channel(cfg_for(p, precision=precision)) as ch:
        for k in ("v", "omy", "u0", "w0"):
            ch.put(k, st[k])
        if p.scalar:
            ch.put("scalar", st["scalar"])
        ch.startup()
        names = FIELDS + (["scalar", "scn", "scnm1"] if p.scalar else [])
        for it in range(nsteps + 1):
            if it:
                sol.step(ost)
                ch.step(1)
            errs = {k: rel(ch.get(k), ost[k]) for k in names}
            log(f"  step {it}: worst {max(errs.values()):.3e}")
            bad = {k: v for k, v in errs.items() if not v < tol}
            assert not bad, (it, bad)
    return errs


def test_steps_ny192_with_scalar_fp64():
    p = params(32, NY, 32, dt=0.002, scalar=True, diff=0.71, scl_flag=6, deltaT=0.3)
    log("case: 32x193x32 channel + scalar, FP64")
    run_steps(p, 5, 0, 1e-10, 61)


def test_steps_ny192_mixed():
    p = params(64, NY, 64, dt=0.002)
    log("case: 64x193x64 channel, mixed")
    errs = run_steps(p, 3, 1, 1e-4, 62)
    assert max(errs.values()) > 1e-12   # the FP32 path really ran


def test_fenep_ny192_fp64():
    kw = dict(polymer=True, zlmax=30.0, tpoly=2.2, diffpoly=0.5, qbeta=0.9, crstrt=1, src_start=0)
    p = params(32, NY, 32, dt=0.001, re=6875.0, R_tau=395.0, **kw)
    log("case: FENE-P 32x193x32 FP64")
    st = seeded(p, 63, amp=0.2)
    st.update(poly_state(p, 64, amp=0.05, mean=shear_conformation_mean(p, st["u0"], p.tpoly)))
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, polymer=1,
                          zlmax=p.zlmax, tpoly=p.tpoly, diffpoly=p.diffpoly, qbeta=p.qbeta, ipeter=p.ipeter,
                          itarget=0, src_start=0, crstrt=1)
    sol = O.Solver(p)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    keys = ["v", "omy", "fn", "gn"] + ["c" + q for q in O.POLY] + ["str%sn" % q for q in O.POLY]
    with kmm.Channel(cfg) as ch:
        for k in ("v", "omy", "u0", "w0"):
            ch.put(k, st[k])
        for ij in O.POLY:
            ch.put("c" + ij, st["c" + ij])
        ch.startup()
        for _ in range(3):
            ch.step(1)
            sol.step(ost)
        errs = {k: rel(ch.get(k), ost[k]) for k in keys}
    bad = {k: v for k, v in errs.items() if not v < 1e-10}
    assert not bad, bad


def test_c5_workload_full_size_matches_oracle():
    """The Re_tau 395 viscoelastic workload (256 x 193 x 256, FENE-P), FP64:
    per checked spectral field and stage, the fixture's 288 sampled entries
    within 1e-10 of the field's largest magnitude, the largest magnitude and
    the sum of squared magnitudes within 1e-10 relative."""
    ref = np.load(FIX)
    p, st = c5.inputs()
    log("case: Re_tau 395 viscoelastic workload 256x193x256 FENE-P, FP64")
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, polymer=1,
                          zlmax=p.zlmax, tpoly=p.tpoly, diffpoly=p.diffpoly, qbeta=p.qbeta, ipeter=1, itarget=0,
                          crstrt=1, src_start=0, precision=0)
    tol = 1e-10
    with kmm.Channel(cfg) as ch:
        for k in ("v", "omy", "u0", "w0") + tuple("c" + q for q in O.POLY):
            ch.put(k, st[k])
        ch.startup()
        for stage in range(c5.NSTEPS + 1):
            if stage > 0:
                ch.step(1)
            for k in c5.FIELDS:
                flat = np.asarray(ch.get(k)).ravel()
                fmax = float(ref[f"{k}_{stage}_max"])
                err = np.abs(flat[ref[f"{k}_idx"]] - ref[f"{k}_{stage}_val"]).max() / max(fmax, 1e-300)
                assert err < tol, (stage, k, err)
                assert abs(np.abs(flat).max() - fmax) <= tol * fmax, (stage, k)
                ss = float(ref[f"{k}_{stage}_ss"])
                assert abs(np.sum(np.abs(flat) ** 2) - ss) <= 2 * tol * ss, (stage, k)
            log(f"  stage {stage}: fields match")
