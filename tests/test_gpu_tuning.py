"""The autotuner: occupancy variants of the pencil step kernel and the
x-plane kernel give the same step, the tuner times both on a twin of the
channel without advancing the channel itself, and keeps a valid choice."""
import numpy as np
import pytest

from tests._cases import params, rel, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

KEYS = ("v", "omy", "fn", "gn", "u0", "w0", "scalar", "scn")


def make(p, st, precision):
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, scalar=1,
                          diff=0.71, scl_flag=6, deltaT=0.3, precision=precision)
    ch = kmm.Channel(cfg)
    for k in ("v", "omy", "u0", "w0", "scalar"):
        ch.put(k, st[k])
    ch.startup()
    return ch


@pytest.mark.parametrize("precision", [0, 1])
def test_variants_agree_and_autotune_leaves_the_channel(precision):
    p = params(64, 64, 64, scalar=True)
    st = seeded(p, 91, amp=0.2)
    print(f"case: tuning variants, 64^3 scalar, precision {precision}", flush=True)
    a, b = make(p, st, precision), make(p, st, precision)
    a.set_tuning(0, 0)
    b.set_tuning(1, 1)
    a.step(3)
    b.step(3)
    errs = {k: rel(b.get(k), a.get(k)) for k in KEYS}
    print("variant 1 vs 0:", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-13
    before = {k: a.get(k) for k in KEYS}
    it0 = a.last_cfl()[1]
    res = a.autotune(4)
    print("autotune:", res, flush=True)
    assert all(t > 0 for t in res["pencil_ms"] + res["xplane_ms"])
    assert res["choice"]["pencil"] in (0, 1) and res["choice"]["xplane"] in (0, 1)
    assert a.last_cfl()[1] == it0
    for k in KEYS:
        assert np.array_equal(a.get(k), before[k]), k
    a.step(1)   # the tuned channel keeps stepping
    b.set_tuning(res["choice"]["pencil"], res["choice"]["xplane"])
    b.step(1)
    assert max(rel(b.get(k), a.get(k)) for k in KEYS) < 1e-13
    a.close()
    b.close()
