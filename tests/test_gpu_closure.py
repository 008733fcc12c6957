"""The learned polymer-stress closure in the channel step: training data
drawn by the GPU from its own FENE-P run, GPU training, the held-out gain over
the closure without the network (zero increment, i.e. frozen stress) and over
the untrained network, and a closure run that tracks the FENE-P run."""
import numpy as np
import pytest

from tests._cases import params, poly_state, rel, seeded, shear_conformation_mean

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

POLY = dict(polymer=1, zlmax=30.0, tpoly=0.5, diffpoly=0.5, qbeta=0.9, crstrt=1, src_start=0)


def channel(p, st):
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, **POLY)
    ch = kmm.Channel(cfg)
    for k, v in st.items():
        ch.put(k, v)
    ch.startup()
    return ch


def collect(ch, steps, n, seed):
    Xs, Ts = [], []
    for s in range(steps):
        ch.step(1)
        X, T = ch.poly_samples(n, seed=seed + s)
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def test_closure_training_gain_and_run():
    p = params(32, 32, 32, dt=0.005, polymer=True, crstrt=1, src_start=0)
    st = seeded(p, 61, amp=0.3)
    # conformation at the shear equilibrium of the mean flow plus a seeded
    # perturbation: a developed mean polymer shear stress, as in a channel
    # past its start-up transient
    st.update(poly_state(p, 62, amp=0.05, mean=shear_conformation_mean(p, st["u0"], POLY["tpoly"])))
    st = {k: v for k, v in st.items() if k in ("v", "omy", "u0", "w0") or k.startswith("c")}
    ch = channel(p, st)
    ch.set_poly_capture(True)
    ch.step(2)
    Xtr, Ttr = collect(ch, 24, 2000, seed=100)
    Xte, Tte = collect(ch, 4, 2000, seed=900)
    mx, sx = Xtr.mean(0), Xtr.std(0) + 1e-12
    my, sy = Ttr.mean(0), Ttr.std(0) + 1e-12
    net = kmm.Mlp(seed=3)
    net.set_norm(mx, sx, my, sy)
    untrained = net.infer(Xte)
    losses = net.train(Xtr, Ttr, steps=1500, batch=512, lr=1e-3, seed=5)
    pred = net.infer(Xte)

    def mse(a):
        return float((((a - Tte) / sy) ** 2).mean())

    # targets are corrections to linear extrapolation: predicting 0 is the
    # closure without the network; -(str(n-1) - str(n-2)) freezes the stress
    m_net, m_extrap, m_untr = mse(pred), mse(np.zeros_like(Tte)), mse(untrained)
    m_frozen = mse(-Xte[:, 16:22])
    print(f"held-out normalised MSE: network {m_net:.3e}, linear extrapolation {m_extrap:.3e}, "
          f"frozen stress {m_frozen:.3e}, untrained {m_untr:.3e}", flush=True)
    assert np.isfinite(losses).all() and losses[-100:].mean() < 0.5 * losses[:20].mean()
    assert m_net < 0.5 * m_extrap
    assert m_net < 1e-3 * m_frozen

    # closure run vs the FENE-P run from the same state, with audit windows
    # of FENE-P transport (3 steps of every 25)
    ref = channel(p, st)
    ref.set_poly_capture(True)
    ref.step(2 + 24 + 4)
    ch.set_closure(net)
    ch.set_closure_audit(25, 3, 0.1)
    ch.step(6)
    ref.step(6)
    s12 = rel(ch.poly_stress(1), ref.poly_stress(1))
    v = rel(ch.get("v"), ref.get("v"))
    print(f"closure vs FENE-P after 6 steps: stress12 rel {s12:.3e}, v rel {v:.3e}", flush=True)
    assert s12 < 0.05 and v < 1e-3
    # mean velocity, Reynolds shear stress and polymer shear stress profiles
    # within 1 % of FENE-P after 60 closure steps (2 audits); without audits
    # the polymer stress profile error reaches ~2 % here
    ch.step(54)
    ref.step(54)
    st_c = ch.closure_status()
    print("audits:", st_c, flush=True)
    assert st_c["active"] and st_c["audits"] == 2
    a, b = ch.statistics(), ref.statistics()
    e = {k: float(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()) for k in ("umean", "rss", "pss", "urms")}
    print("closure vs FENE-P profiles after 60 steps:", {k: f"{x:.2e}" for k, x in e.items()}, flush=True)
    assert max(e.values()) < 1e-2, e
    ch.close()
    ref.close()
    net.close()


def test_closure_steps_match_the_oracle_fed_the_same_stress():
    """Steps with the learned closure against the oracle's closure hook fed
    the stress the GPU network predicted (FP64): the force of a closure
    stress, the first closure step finishing the last FENE-P evaluation, and
    the following closure-only steps."""
    from oracle import kmm_oracle as O
    p = params(32, 32, 32, dt=0.005, polymer=True, zlmax=30.0, tpoly=0.5, diffpoly=0.5, qbeta=0.9, crstrt=1,
               src_start=0)
    print("case: closure steps vs oracle, 32^3 FP64", flush=True)
    st = seeded(p, 81, amp=0.2)
    st.update(poly_state(p, 82, amp=0.05, mean=shear_conformation_mean(p, st["u0"], POLY["tpoly"])))
    st = {k: v for k, v in st.items() if k in ("v", "omy", "u0", "w0") or k.startswith("c")}
    ch = channel(p, st)
    sol = O.Solver(p)
    sol.polymer_defaults(st)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    ch.set_poly_capture(True)
    for _ in range(3):
        ch.step(1)
        sol.step(ost)
    net = kmm.Mlp(seed=4)   # untrained: small corrections to the extrapolation
    net.set_norm(np.zeros(22), np.ones(22), np.zeros(6), np.full(6, 1e-3))
    ch.set_closure(net)
    for k in range(4):
        ch.step(1)
        sol.closure_stress = {ij: ch.poly_stress(q) for q, ij in enumerate(O.POLY)}
        sol.step(ost)
        e = {f: rel(ch.get(f), ost[f]) for f in ("v", "omy")}
        e["u0"] = rel(ch.get("u0"), ost["u0"])
        print(f"closure step {k + 1}:", {f: f"{x:.2e}" for f, x in e.items()}, flush=True)
        assert max(e.values()) < 1e-10, (k, e)
    ch.close()
    net.close()
