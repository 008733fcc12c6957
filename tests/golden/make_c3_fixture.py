"""Writes tests/golden/c3_fenep_tracers.npz: the oracle's viscoelastic
workload at full size (192 x 129 x 192, FENE-P Wi_tau 50, beta 0.9, L^2 900,
Sc 0.5, the conformation at the mean-shear equilibrium plus seeded 1e-3
noise, 100 000 tracers), FP64, through start-up and two steps.

The full fields are 38 MB each, so the fixture keeps, for every checked
spectral field and stage, 256 seeded random entries plus the 32 largest,
the field's largest magnitude and its sum of squared magnitudes; for the
tracers, 2 000 seeded particles' positions, velocities and cell indices
(imin, jmin, kmax of part_track.f90:537-553), the sums over all particles,
and the SHA-256 digest of every particle's cell indices (int16, C order), so
that the GPU's indices can be checked bit for bit.  The inputs are rebuilt by the test from the same seeds
(polymer_dns_amd.ic, the benchmark's constants).

Run from the repository root: python tests/golden/make_c3_fixture.py
(about 10 minutes on one core).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from threadpoolctl import threadpool_limits  # noqa: E402

from oracle import kmm_oracle as O  # noqa: E402
from tests import _c3_case as c3_case  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_fenep_tracers.npz")


def main():
    p, st, xyz = c3_case.inputs()
    sol = O.Solver(p)
    sol.set_particles(*xyz, ap=c3_case.AP)
    st = {k: np.array(v) for k, v in st.items()}
    out = {}
    rng = np.random.default_rng(c3_case.SAMPLE_SEED)
    pick = rng.choice(c3_case.NPART, c3_case.NPSAMPLE, replace=False)
    out["part_pick"] = pick
    idx = {}
    with threadpool_limits(1):
        sol.startup(st)
        for stage in range(c3_case.NSTEPS + 1):
            if stage > 0:
                sol.step(st)
            for k in c3_case.FIELDS:
                f = np.asarray(st[k])
                if k not in idx:
                    idx[k] = c3_case.sample_index(f, rng)
                    out[f"{k}_idx"] = idx[k]
                flat = f.ravel()
                out[f"{k}_{stage}_val"] = flat[idx[k]]
                out[f"{k}_{stage}_max"] = np.abs(flat).max()
                out[f"{k}_{stage}_ss"] = np.sum(np.abs(flat) ** 2)
            P = sol.particles
            for q in "xyzuvw":
                out[f"p{q}_{stage}_val"] = P[q][pick]
                out[f"p{q}_{stage}_sum"] = np.sum(P[q])
            imin, _, jmin, _, _, kmax = O.tracer_cells(P["x"], P["y"], P["z"], p)
            cells = np.ascontiguousarray(np.stack([imin, jmin, kmax]).astype(np.int16))
            out[f"cells_{stage}_sha256"] = np.array(c3_case.cells_digest(cells))
            out[f"cells_{stage}_pick"] = cells[:, pick]
            print(f"stage {stage} done", flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
