"""The compiled CPU restatement (oracle/kmm_cpu.c, the benchmark's CPU leg)
against the numpy oracle: start-up by the numpy oracle, then three steps
each way on the same seeded state, channel + passive scalar (uniform
source) with and without a Goldstein rib, every state field within 1e-11 of
its largest magnitude.  Also the refusals: configurations outside its subset
raise instead of running."""
import numpy as np
import pytest

from oracle import kmm_cpu
from oracle import kmm_oracle as O
from polymer_dns_amd.restart import rib_geometry
from tests._cases import params, rel, seeded

FIELDS = ("v", "omy", "fn", "gn", "fnm1", "gnm1", "u0", "w0", "h1n", "h3n", "scalar", "scn", "scnm1", "fxintg",
          "fyintg", "fzintg")


@pytest.fixture(scope="module", autouse=True)
def built():
    kmm_cpu.build()


def pair(nx, ny, nz, ib, seed):
    p = params(nx, ny, nz, scalar=True, diff=0.71, scl_flag=6, deltaT=0.3, gain=-10.0, ugain=-1.0, dt=0.002)
    st = seeded(p, seed)
    im = rib_geometry(nx, ny, nz, y_cells=3, x_center=p.mx // 3, x_width=2) if ib else None
    a, b = O.Solver(p, imatrix=im), O.Solver(p, imatrix=im)
    sa = {k: v.copy() for k, v in st.items()}
    a.startup(sa)
    sb = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in sa.items()}
    return a, sa, b, sb


@pytest.mark.parametrize("grid,ib", [((32, 32, 32), False), ((32, 32, 32), True), ((48, 64, 32), True)])
def test_compiled_steps_match_numpy_oracle(grid, ib):
    a, sa, b, sb = pair(*grid, ib, sum(grid) + ib)
    for _ in range(3):
        a.step(sa)
    cfl = kmm_cpu.steps(b, sb, 3, threads=2)
    assert b.it == a.it == 3
    assert abs(cfl - sa["cflmax"]) <= 1e-12 * sa["cflmax"]
    for k in FIELDS:
        assert rel(sb[k], sa[k]) < 1e-11, k
    if ib:
        assert np.abs(sa["fxintg"]).max() > 0.0   # the feedback integrals really ran


def test_refuses_outside_its_subset():
    p = params(32, 32, 32, polymer=True)
    sol = O.Solver(p)
    with pytest.raises(NotImplementedError):
        kmm_cpu.steps(sol, {}, 1)
