"""The viscoelastic workload at its full size (192 x 129 x 192, FENE-P
Wi_tau 50, beta 0.9, L^2 900, Sc 0.5, conformation at the mean-shear
equilibrium, 100 000 tracers), FP64, through start-up and two steps against
the oracle's run of the same seeded inputs, kept as a sampled fixture
(tests/golden/make_c3_fixture.py): per checked spectral field, 288 sampled
entries within 1e-10 of the field's largest magnitude, the largest magnitude
and the sum of squared magnitudes within 1e-10 relative; tracer positions
and velocities of 2 000 sampled particles and the sums over all particles;
the cell indices of every tracer (part_track.f90:537-553) bit for bit.
"""
import os

import numpy as np
import pytest

from tests import _c3_case as c3

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_fenep_tracers.npz")
TOL = 1e-10


def test_viscoelastic_workload_full_size_matches_oracle():
    ref = np.load(FIX)
    p, st, (x, y, z) = c3.inputs()
    print("case: viscoelastic workload 192x129x192 FENE-P + 100k tracers, FP64", flush=True)
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, polymer=1,
                          zlmax=p.zlmax, tpoly=p.tpoly, diffpoly=p.diffpoly, qbeta=p.qbeta, ipeter=1, itarget=0,
                          crstrt=1, src_start=0, precision=0, ap=c3.AP)
    pick = ref["part_pick"]
    with kmm.Channel(cfg) as ch:
        for k in ("v", "omy", "u0", "w0") + tuple("c" + q for q in ("11", "12", "13", "22", "23", "33")):
            ch.put(k, st[k])
        ch.set_particles(x, y, z)
        ch.startup()
        for stage in range(c3.NSTEPS + 1):
            if stage > 0:
                ch.step(1)
            for k in c3.FIELDS:
                flat = np.asarray(ch.get(k)).ravel()
                fmax = float(ref[f"{k}_{stage}_max"])
                err = np.abs(flat[ref[f"{k}_idx"]] - ref[f"{k}_{stage}_val"]).max() / max(fmax, 1e-300)
                assert err < TOL, (stage, k, err)
                assert abs(np.abs(flat).max() - fmax) <= TOL * fmax, (stage, k)
                ss = float(ref[f"{k}_{stage}_ss"])
                assert abs(np.sum(np.abs(flat) ** 2) - ss) <= 2 * TOL * ss, (stage, k)
            P = ch.particles()
            for q in "xyzuvw":
                a, b = P[q][pick], ref[f"p{q}_{stage}_val"]
                assert np.abs(a - b).max() <= TOL * max(1.0, np.abs(b).max()), (stage, q)
                s = float(ref[f"p{q}_{stage}_sum"])
                assert abs(np.sum(P[q]) - s) <= TOL * np.abs(P[q]).sum(), (stage, q)
            _, cells = ch.fluid_interp(P["x"], P["y"], P["z"], [np.zeros((p.mx, p.mz, p.nyp))], return_cells=True)
            cells = cells.astype(np.int16)
            assert np.array_equal(cells[:, pick], ref[f"cells_{stage}_pick"]), stage
            assert c3.cells_digest(cells) == str(ref[f"cells_{stage}_sha256"]), stage
            print(f"  stage {stage}: fields, tracers and cells match", flush=True)
