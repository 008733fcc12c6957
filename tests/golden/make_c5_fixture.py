"""Writes tests/golden/c5_fenep.npz: the oracle's Re_tau 395 viscoelastic
workload at full size (256 x 193 x 256, FENE-P Wi_tau 50, beta 0.9, L^2 900,
Sc 0.5, the conformation at the mean-shear equilibrium plus seeded 1e-3
noise), FP64, through start-up and two steps.

As for the 192 x 129 x 192 fixture, every checked spectral field keeps, per
stage, 256 seeded random entries plus the 32 largest, the field's largest
magnitude and its sum of squared magnitudes.  The inputs are rebuilt by the
test from the same seeds (tests/_c5_case.py).

Run from the repository root: python tests/golden/make_c5_fixture.py
(about half an hour).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import kmm_oracle as O  # noqa: E402
from tests import _c5_case as c5_case  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_fenep.npz")


def main():
    p, st = c5_case.inputs()
    sol = O.Solver(p)
    st = {k: np.array(v) for k, v in st.items()}
    out = {}
    rng = np.random.default_rng(c5_case.SAMPLE_SEED)
    idx = {}
    sol.startup(st)
    for stage in range(c5_case.NSTEPS + 1):
        if stage > 0:
            sol.step(st)
        for k in c5_case.FIELDS:
            f = np.asarray(st[k])
            if k not in idx:
                idx[k] = c5_case.sample_index(f, rng)
                out[f"{k}_idx"] = idx[k]
            flat = f.ravel()
            out[f"{k}_{stage}_val"] = flat[idx[k]]
            out[f"{k}_{stage}_max"] = np.abs(flat).max()
            out[f"{k}_{stage}_ss"] = np.sum(np.abs(flat) ** 2)
        print(f"stage {stage} done", flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
