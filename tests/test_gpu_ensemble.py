"""The 16-member ensemble of independent channels of uneven size (nx = nz
in {64, 96, 128}, ny in {48, 64, 80, 96}, Re_tau uniform in [120, 240], all
seeded; mixed precision; steady steps replayed as HIP graphs) advanced
together by kmm_ensemble_step (members of one size in shared launches, one
per stage and step): every member equals the same channel run alone, bit
for bit, and the smallest (64 x 49 x 64) and the largest
(128 x 97 x 128) member stay within 1e-4 of the FP64 oracle."""
import math

import numpy as np
import pytest

from oracle import kmm_oracle as O
from polymer_dns_amd import ic
from tests._cases import params, rel, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")

SEED = 20261202 + 4
XL, ZL, DT = 4.0 * math.pi, 4.0 * math.pi / 3.0, 0.002


def member(nx, ny, nz, rt, seed):
    p = params(nx, ny, nz, xl=XL, zl=ZL, dt=DT, R_tau=rt)
    st = seeded(p, seed, amp=0.05)
    cfg = kmm.make_config(nx, ny, nz, xl=XL, zl=ZL, re=p.re, dt=DT, r_tau=rt, precision=1)
    ch = kmm.Channel(cfg)
    for k in ("v", "omy", "u0", "w0"):
        ch.put(k, st[k])
    ch.startup()
    ch.set_graph(4)
    return p, st, ch


def test_sixteen_member_ensemble():
    mem = ic.ensemble_members(SEED)
    assert len(mem) == 16 and {m[1] for m in mem} == {48, 64, 80, 96}
    print("case: 16-member ensemble", [m[:3] for m in mem], flush=True)
    ens = [member(*m) for m in mem]
    alone = [member(*m)[2] for m in mem]
    oracles = []
    for p, st, _ in ens[:2]:
        sol = O.Solver(p)
        ost = {k: v.copy() for k, v in st.items()}
        sol.startup(ost)
        oracles.append((sol, ost))
    chans = [e[2] for e in ens]
    for ch in chans + alone:
        ch.step(1)
    kmm.ensemble_step(chans, 12)
    for ch in alone:
        ch.step(12)
    # members sharing a shape ran in their group's shared launches
    shapes = [m[:3] for m in mem]
    for ch, sh in zip(chans, shapes):
        assert ch.batched_steps() == (12 if shapes.count(sh) > 1 else 0), sh
    for a, b in zip(chans, alone):
        for k in ("v", "omy", "fn", "gn", "u0"):
            assert np.array_equal(a.get(k), b.get(k)), k
    for (sol, ost), ch in zip(oracles, chans[:2]):
        for _ in range(13):
            sol.step(ost)
        for k in ("v", "omy", "fn", "gn", "u0", "w0"):
            e = rel(ch.get(k), ost[k])
            print(f"  {ch.nx}x{ch.nyp}x{ch.nz} {k}: rel {e:.3e}", flush=True)
            assert e < 1e-4, k
    for ch in chans + alone:
        ch.close()


def test_shared_launches_equal_own_streams():
    """Four channels of one size (64x49x64, their own Re_tau and seeds):
    kmm_ensemble_step with shared launches and with every member on its own
    stream give the same states bit for bit; only the members of the first
    run count their steps as shared."""
    mem = [(64, 48, 64, 120.0 + 30.0 * i, 100 + i) for i in range(4)]
    a = [member(*m)[2] for m in mem]
    b = [member(*m)[2] for m in mem]
    for ch in a + b:
        ch.step(1)
    kmm.ensemble_step(a, 9, batch=True)
    kmm.ensemble_step(b, 9, batch=False)
    assert [ch.batched_steps() for ch in a] == [9] * 4
    assert [ch.batched_steps() for ch in b] == [0] * 4
    for x, y in zip(a, b):
        for k in ("v", "omy", "fn", "gn", "u0", "w0"):
            assert np.array_equal(x.get(k), y.get(k)), k
    for ch in a + b:
        ch.close()
