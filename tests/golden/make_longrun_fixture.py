"""Writes tests/golden/longrun_c1_stats.npz: time-averaged statistics of the
oracle's run of the reference's CPU-runnable case -- 64 x 65 x 64 channel,
Lx = 4 pi, Lz = 2 pi, re = 2800, constant mass flux, seeded 5 % perturbation
of laminar flow, dt = 0.005 -- over 2000 steps, sampled every 20 steps:
the writeoutputs profiles mean u, u/v/w rms and Reynolds shear stress
(helpers.f90:1227-1436, on the 3/2 physical grid) and the streamwise and
spanwise energy spectra of u, v, w (sum of |u_hat|^2 over the other
wavenumber and the Chebyshev coefficients, kx > 0 counted twice).

Run from the repository root: python tests/golden/make_longrun_fixture.py
(about half an hour on one core).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from threadpoolctl import threadpool_limits  # noqa: E402

from oracle import kmm_oracle as O  # noqa: E402
from tests import _longrun_case as L  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "longrun_c1_stats.npz")


def main():
    p, st = L.inputs()
    sol = O.Solver(p)
    st = {k: np.array(v) for k, v in st.items()}
    acc = {}
    with threadpool_limits(1):
        sol.startup(st)
        for it in range(1, L.NSTEPS + 1):
            sol.step(st)
            if it % L.EVERY == 0:
                ph = sol.last_phys
                z = np.zeros_like(ph["up"])
                prof = O.writeoutputs(ph["up"], ph["vp"], ph["wp"], z, z, None, None, None, p)
                sample = {k: prof[k] for k in L.PROFILES}
                sample.update(L.spectra(st, p))
                for k, v in sample.items():
                    acc[k] = acc.get(k, 0.0) + np.asarray(v)
                if it % 200 == 0:
                    print(f"step {it}: umean max {prof['umean'].max():.6f}", flush=True)
    n = L.NSTEPS // L.EVERY
    np.savez_compressed(OUT, **{k: v / n for k, v in acc.items()})
    print("wrote", OUT)


if __name__ == "__main__":
    main()
