"""Audit steps and the residual gate of the learned closure: the conformation
tensor rebuilt from a FENE-P stress returns the transported tensor, an audit
window with a loose gate hands the run back to the network, and a gate of 0
leaves the run on the FENE-P transport (fallback)."""
import numpy as np
import pytest

from tests._cases import params, poly_state, rel, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

POLY = dict(polymer=1, zlmax=30.0, tpoly=0.5, diffpoly=0.5, qbeta=0.9, crstrt=1, src_start=0)
CFIELDS = ("c11", "c12", "c13", "c22", "c23", "c33")


def channel(p, st, precision=0):
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau,
                          precision=precision, **POLY)
    ch = kmm.Channel(cfg)
    for k, v in st.items():
        ch.put(k, v)
    ch.startup()
    return ch


def state(p):
    st = seeded(p, 61, amp=0.3)
    st.update(poly_state(p, 62, amp=0.05))
    return {k: v for k, v in st.items() if k in ("v", "omy", "u0", "w0") or k.startswith("c")}


@pytest.mark.parametrize("precision", [0, 1])
def test_conformation_rebuild_returns_the_transported_tensor(precision):
    p = params(32, 32, 32, dt=0.005, polymer=True, crstrt=1, src_start=0)
    print(f"case: rebuild, 32^3, precision {precision}", flush=True)
    ch = channel(p, state(p), precision)
    ch.set_poly_capture(True)
    ch.step(3)
    before = {f: ch.get(f) for f in CFIELDS}
    ch.conf_from_stress()
    after = {f: ch.get(f) for f in CFIELDS}
    errs = {f: rel(after[f], before[f]) for f in CFIELDS}
    print("rebuilt C vs transported C:", {f: f"{e:.2e}" for f, e in errs.items()}, flush=True)
    # the captured stress is FP32: the rebuild carries its rounding
    assert max(errs.values()) < 2e-5
    ch.close()


def test_audit_gate_resumes_then_falls_back():
    p = params(32, 32, 32, dt=0.005, polymer=True, crstrt=1, src_start=0)
    st = state(p)
    print("case: audit gate, 32^3 FP64", flush=True)
    ch = channel(p, st)
    ch.set_poly_capture(True)
    ch.step(2)
    Xs, Ts = [], []
    for s in range(20):
        ch.step(1)
        X, T = ch.poly_samples(2000, seed=100 + s)
        Xs.append(X)
        Ts.append(T)
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    net = kmm.Mlp(seed=3)
    net.set_norm(X.mean(0), X.std(0) + 1e-12, T.mean(0), T.std(0) + 1e-12)
    net.train(X, T, steps=1200, batch=512, lr=1e-3, seed=5)
    ref = channel(p, st)
    ref.step(22)

    it0 = 22                                 # steps done; the closure runs from step 23
    ch.set_closure(net)
    ch.set_closure_audit(5, 3, 0.5)          # first window: steps 28..30
    ch.set_graph(3)                          # replays must stop short of the windows
    ch.step(12)                              # steps 23..34; the closure resumes at 31
    s1 = ch.closure_status()
    print("after a loose gate:", s1, flush=True)
    assert s1["active"] and s1["audits"] == 1 and s1["fallback_it"] == -1
    assert 0.0 < s1["residual"] < 0.05
    ch.set_graph(0)
    ch.set_closure_audit(5, 3, 0.0)          # window 36..38 fails the gate
    ch.step(4)                               # steps 35..38
    s2 = ch.closure_status()
    print("after gate 0:", s2, flush=True)
    assert not s2["active"] and s2["audits"] == 2 and s2["fallback_it"] == it0 + 17
    ch.step(4)
    ref.step(20)
    v = rel(ch.get("v"), ref.get("v"))
    c12 = rel(ch.get("c12"), ref.get("c12"))
    print(f"fallback run vs FENE-P run at step {it0 + 20}: v rel {v:.3e}, c12 rel {c12:.3e}", flush=True)
    assert np.isfinite(ch.get("c11")).all()
    assert v < 5e-3 and c12 < 0.1
    ch.close()
    ref.close()
    net.close()
