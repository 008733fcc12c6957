"""Graph replay of steady time steps (kmm_set_graph) against ordinary
launches: the same seeded run advanced both ways must agree bit for bit
after every call, over several replays with host work in between."""
import numpy as np
import pytest

from tests._cases import params, poly_state, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")


def make(p, **kw):
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, **kw)
    return kmm.Channel(cfg)


def rib(p):
    from polymer_dns_amd.restart import rib_geometry
    return rib_geometry(p.nx, p.ny, p.nz, y_cells=3, x_center=p.mx // 3, x_width=2)


def fill(ch, st, keys):
    for k in keys:
        ch.put(k, st[k])


def check_equal(a, b, keys):
    for k in keys:
        assert np.array_equal(a.get(k), b.get(k)), k


@pytest.mark.parametrize("k", [4])
def test_graph_replay_scalar_ib_tracers_cmf(k):
    p = params(32, 32, 32, dt=0.005)
    st = seeded(p, 51, amp=0.2)
    st["scalar"] = np.zeros((p.nxh, p.nz, p.nyp), dtype=np.complex128)
    st["scalar"][0, 0, 0] = 0.5
    kw = dict(scalar=1, diff=0.71, scl_flag=6, deltaT=0.2, gain=2.0, ugain=5.0, cmf=1)
    chans = [make(p, **kw), make(p, **kw)]
    rng = np.random.default_rng(52)
    xyz = (rng.uniform(0, p.xl, 300), rng.uniform(-0.9, 0.9, 300), rng.uniform(0, p.zl, 300))
    for ch in chans:
        ch.set_geometry(rib(p))
        fill(ch, st, ("v", "omy", "u0", "w0", "scalar"))
        ch.set_particles(*xyz)
        ch.startup()
    chans[1].set_graph(k)
    for c in chans:
        c.step(2)
    keys = ["v", "omy", "fn", "gn", "u0", "scalar", "scn", "fxintg"]
    for call in range(4):
        for c in chans:
            c.step(k + 1)
        check_equal(chans[0], chans[1], keys)
        pa, pb = chans[0].particles(), chans[1].particles()
        for q in ("x", "y", "z", "u", "swirl"):
            assert np.array_equal(pa[q], pb[q]), (call, q)
        assert chans[0].dpdx() == chans[1].dpdx()
    for c in chans:
        c.close()


def test_graph_replay_polymer_mixed():
    p = params(32, 32, 32, dt=0.005, polymer=True, crstrt=1, src_start=0)
    st = seeded(p, 53, amp=0.2)
    st.update(poly_state(p, 54))
    kw = dict(polymer=1, zlmax=30.0, tpoly=0.5, diffpoly=0.5, qbeta=0.9, crstrt=1, src_start=0, precision=1)
    chans = [make(p, **kw), make(p, **kw)]
    for ch in chans:
        fill(ch, st, ("v", "omy", "u0", "w0") + tuple("c" + q for q in ("11", "12", "13", "22", "23", "33")))
        ch.startup()
    chans[1].set_graph(3)
    for c in chans:
        c.step(2)
    keys = ["v", "omy", "fn", "c11", "c12", "c33", "c12n", "str22n"]
    for call in range(3):
        for c in chans:
            c.step(7)
        check_equal(chans[0], chans[1], keys)
    for c in chans:
        c.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_learned_closure_with_audits(precision):
    """Closure steps replayed as graphs (3-step history period) between audit
    windows of FENE-P transport, against the same run launched step by step:
    velocity, stress and audit state agree bit for bit."""
    p = params(32, 32, 32, dt=0.005, polymer=True, crstrt=1, src_start=0)
    print("case: graph replay of closure steps, 32^3, precision", precision, flush=True)
    st = seeded(p, 57, amp=0.2)
    st.update(poly_state(p, 58, amp=0.05))
    kw = dict(polymer=1, zlmax=30.0, tpoly=0.5, diffpoly=0.5, qbeta=0.9, crstrt=1, src_start=0,
              precision=precision)
    net = kmm.Mlp(seed=6)   # untrained: small corrections to the extrapolation
    net.set_norm(np.zeros(22), np.ones(22), np.zeros(6), np.full(6, 1e-3))
    chans = [make(p, **kw), make(p, **kw)]
    for ch in chans:
        fill(ch, st, ("v", "omy", "u0", "w0") + tuple("c" + q for q in ("11", "12", "13", "22", "23", "33")))
        ch.startup()
        ch.set_poly_capture(True)
        ch.step(3)
        ch.set_closure(net)
        ch.set_closure_audit(12, 3, 0.5)
    chans[1].set_graph(3)
    for call in range(4):
        for c in chans:
            c.step(9)
        check_equal(chans[0], chans[1], ["v", "omy", "fn", "gn"])
        assert np.array_equal(chans[0].poly_stress(1), chans[1].poly_stress(1))
        assert chans[0].closure_status() == chans[1].closure_status()
    assert chans[0].closure_status()["audits"] >= 2
    for c in chans:
        c.close()
    net.close()


def test_ensemble_matches_members_run_alone():
    """Uneven members (ny 48/64/96, nx = nz 64/96) with seeded Re_tau advance
    through kmm_ensemble_step exactly as each does alone, and the first
    member (64 x 49 x 64, mixed precision) stays within 1e-4 of the FP64
    oracle over the 14 steps."""
    from oracle import kmm_oracle as O
    from tests._cases import rel
    rng = np.random.default_rng(55)
    grids = [(64, 48, 64), (96, 64, 96), (64, 96, 64), (96, 48, 96)]
    members, alone = [], []
    oracle = None
    for g in grids:
        rtau = float(rng.uniform(120.0, 240.0))
        p = params(*g, dt=0.004, R_tau=rtau)
        st = seeded(p, int(rng.integers(1 << 30)), amp=0.1)
        if oracle is None:
            sol = O.Solver(p)
            ost = {k: v.copy() for k, v in st.items()}
            sol.startup(ost)
            oracle = (sol, ost)
        pair = []
        for _ in range(2):
            ch = make(p, precision=1)
            fill(ch, st, ("v", "omy", "u0", "w0"))
            ch.startup()
            ch.set_graph(4)
            pair.append(ch)
        members.append(pair[0])
        alone.append(pair[1])
    for ch in members + alone:   # NL terms valid: the ensemble chunks replay graphs
        ch.step(1)
    kmm.ensemble_step(members, 13)
    for ch in alone:
        ch.step(13)
    sol, ost = oracle
    for _ in range(14):
        sol.step(ost)
    for k in ("v", "omy", "fn", "gn", "u0"):
        assert rel(members[0].get(k), ost[k]) < 1e-4, k
    for a, b in zip(members, alone):
        check_equal(a, b, ["v", "omy", "fn", "u0"])
        a.close()
        b.close()
