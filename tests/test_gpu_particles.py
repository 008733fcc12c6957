"""GPU parity of the Lagrangian tracer path (part_track.f90) through the C ABI:
trilinear interpolation (fluid_interp / interpolate3) against the oracle on
the same fields and points, cell indices bit-exact; tracers carried by the
channel step against the oracle's step; particles.dat reading and the
part-t-XXXXXX.dat writer."""
import math
import os

import numpy as np
import pytest

from oracle import kmm_oracle as O
from tests._cases import params, rel, seeded

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")


def channel(p, **kw):
    cfg = kmm.make_config(p.nx, p.ny, p.nz, xl=p.xl, zl=p.zl, re=p.re, dt=p.dt, r_tau=p.R_tau, **kw)
    return kmm.Channel(cfg)


def random_points(p, n, rng, wall=0.999):
    return (rng.uniform(0, p.xl, n), rng.uniform(-wall, wall, n), rng.uniform(0, p.zl, n))


def test_fluid_interp_matches_oracle_and_cells_bit_exact():
    p = params(32, 32, 32)
    rng = np.random.default_rng(11)
    fields = [rng.standard_normal((p.mx, p.mz, p.nyp)) for _ in range(3)]
    x, y, z = random_points(p, 20000, rng)
    # include the periodic wrap cells and the last y interval at both walls
    x[:50] = p.xl * (1 - 0.5 / p.mx)
    z[50:100] = p.zl * (1 - 0.3 / p.mz)
    y[100:150] = rng.uniform(p.ycoord[1], 1.0, 50)
    y[150:200] = rng.uniform(-0.9999, p.ycoord[-2], 50)
    with channel(p) as ch:
        out, cells = ch.fluid_interp(x, y, z, fields, return_cells=True)
    ref = O.fluid_interp(x, y, z, fields, p)
    imin, _, jmin, _, _, kmax = O.tracer_cells(x, y, z, p)
    assert np.array_equal(cells[0], imin)
    assert np.array_equal(cells[1], jmin)
    assert np.array_equal(cells[2], kmax)
    for a, b in zip(out, ref):
        assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()


def test_fluid_interp_trilinear_exact_and_out_of_grid_error():
    p = params(16, 32, 16)
    x3 = (np.arange(p.mx) * (p.xl / p.mx))[:, None, None]
    z3 = (np.arange(p.mz) * (p.zl / p.mz))[None, :, None]
    y3 = p.ycoord[None, None, :]
    f = 1.0 + 0.5 * x3 - 2.0 * y3 + 0.75 * z3 + 0.3 * x3 * y3 * z3
    rng = np.random.default_rng(12)
    n = 4000
    x = rng.uniform(0, p.xl * (p.mx - 1) / p.mx, n)
    z = rng.uniform(0, p.zl * (p.mz - 1) / p.mz, n)
    y = rng.uniform(-0.99, 0.99, n)
    with channel(p) as ch:
        (g,) = ch.fluid_interp(x, y, z, [f])
        assert np.abs(g - (1.0 + 0.5 * x - 2.0 * y + 0.75 * z + 0.3 * x * y * z)).max() < 1e-12
        # y = -yl/2 exactly puts kmin past the last plane; y > yl/2 has no cell
        for bad_y in (-1.0, 1.5):
            with pytest.raises(kmm.KmmError) as e:
                ch.fluid_interp(np.array([0.1]), np.array([bad_y]), np.array([0.1]), [f])
            assert e.value.code == kmm.KMM_E_PARTICLE


@pytest.mark.parametrize("precision", [0, 1])
def test_tracers_follow_the_channel_step(precision):
    p = params(32, 32, 32, dt=0.01)
    st = seeded(p, 21, amp=0.2)
    rng = np.random.default_rng(22)
    n = 500
    x, y, z = random_points(p, n, rng, wall=0.98)
    ap = 0.005
    ch = channel(p, precision=precision, ap=ap, cd_switch=0)
    for k in ("v", "omy", "u0", "w0"):
        ch.put(k, st[k])
    ch.set_particles(x, y, z)
    ch.startup()
    sol = O.Solver(p)
    sol.swirl_fn = O.swirl_eig   # the GPU swirl has no cancellation near repeated roots
    sol.set_particles(x, y, z, ap=ap)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    tol = 1e-10 if precision == 0 else 1e-4
    got = ch.particles()
    for k in "uvw":
        assert rel(got[k], sol.particles[k]) < tol, k
    for step in range(6):
        ch.step(1)
        sol.step(ost)
        got = ch.particles()
        for k in "xyz":
            assert np.abs(got[k] - sol.particles[k]).max() < tol, (step, k)
        for k in "uvw":
            assert rel(got[k], sol.particles[k]) < tol, (step, k)
    # up3d / vp3d / wp3d of the last evaluation
    for c, key in enumerate(("up", "vp", "wp")):
        assert rel(ch.physical(c), sol.last_phys[key]) < tol
    # swirl_part: swirl_3d (from eigenvalues) interpolated at the particles
    scale = np.abs(sol.last_swirl).max()
    assert np.abs(got["swirl"] - sol.particles["swirl"]).max() < (1e-9 if precision == 0 else 1e-3) * scale
    ch.close()


def test_particles_set_after_startup_initialise_on_next_step():
    """A restart run (irstrt = 1) initialises particles at its first step."""
    p = params(32, 32, 32, dt=0.01)
    st = seeded(p, 23, amp=0.2)
    rng = np.random.default_rng(24)
    x, y, z = random_points(p, 64, rng, wall=0.9)
    with channel(p) as ch:
        for k in ("v", "omy", "u0", "w0"):
            ch.put(k, st[k])
        ch.startup()
        ch.set_particles(x, y, z)
        ch.step(1)
        got = ch.particles()
        assert np.array_equal(got["x"], x) and np.array_equal(got["y"], y) and np.array_equal(got["z"], z)
        u = ch.fluid_interp(x, y, z, [ch.physical(0), ch.physical(1), ch.physical(2)])
        for c, k in enumerate("uvw"):
            assert np.array_equal(got[k], u[c])
        ch.step(1)
        assert not np.array_equal(ch.particles()["x"], x)


def test_particles_dat_read_and_part_file_write(tmp_path):
    p = params(32, 32, 32, dt=0.01)
    st = seeded(p, 25, amp=0.1)
    pts = np.array([[0.5, 0.1, 0.25], [3.0, -0.7, 6.0], [12.0, 0.95, 1.0]])
    src = tmp_path / "particles.dat"
    with open(src, "w") as f:
        f.write("x y z u v w tp\n")
        for r in pts:
            f.write(f"{r[0]:.6e} {r[1]:.6e}, {r[2]:.6e} 0.0 0.0d0 1.0D-2 0.0\n")
    out = tmp_path / "part-t-000001.dat"
    with channel(p) as ch:
        for k in ("v", "omy", "u0", "w0"):
            ch.put(k, st[k])
        ch.read_particles(src)
        assert ch.particle_count() == 3
        ch.startup()
        ch.step(1)
        got = ch.particles()
        ch.write_particles(out, 1)
    lines = open(out).read().splitlines()
    assert lines[0].strip().startswith("variables")
    assert "solutiontime=" in lines[1]
    rows = [l for l in lines[2:] if l.strip()]
    assert len(rows) == 3
    for j, row in enumerate(rows):
        assert len(row) == 7 * 15
        fields = [row[15 * i:15 * i + 14] for i in range(7)]
        for i, s in enumerate(fields):
            mant = s.strip().lstrip("-")
            assert mant.startswith("0.") and mant[8] == "E", s
        vals = np.array([float(s) for s in fields])
        ref = np.array([got[k][j] for k in ("x", "y", "z", "u", "v", "w", "swirl")])
        assert np.allclose(vals, ref, rtol=5.1e-6, atol=0)  # 6 significant digits


def test_bad_particle_file_is_an_io_error(tmp_path):
    p = params(32, 32, 32)
    bad = tmp_path / "particles.dat"
    bad.write_text("header\n0.1 0.2 0.3 0 0 0\n")
    with channel(p) as ch:
        with pytest.raises(kmm.KmmError) as e:
            ch.read_particles(bad)
        assert e.value.code == kmm.KMM_E_IO


@pytest.mark.parametrize("precision", [0, 1])
def test_blob_scalar_source_matches_oracle(precision):
    """scl_flag 2: Gaussian blobs of rate deltaT at the tracers feed the
    scalar between src_start and src_stop (solvers.f90:2078-2118,
    3061-3063); the run crosses both ends of the window."""
    p = params(32, 32, 32, scalar=True, dt=0.01, diff=0.71, scl_flag=2, deltaT=0.4,
               sigmax=0.5, sigmay=0.08, sigmaz=0.3, src_start=2, src_stop=4)
    print(f"case: blob source, 32^3, precision {precision}", flush=True)
    st = seeded(p, 23, amp=0.2)
    rng = np.random.default_rng(24)
    x, y, z = random_points(p, 40, rng, wall=0.9)
    ch = channel(p, precision=precision, ap=0.005, cd_switch=0, scalar=1, diff=p.diff, scl_flag=2,
                 deltaT=p.deltaT, sigmax=p.sigmax, sigmay=p.sigmay, sigmaz=p.sigmaz,
                 src_start=p.src_start, src_stop=p.src_stop)
    for k in ("v", "omy", "u0", "w0", "scalar"):
        ch.put(k, st[k])
    ch.set_particles(x, y, z)
    ch.startup()
    sol = O.Solver(p)
    sol.swirl_fn = O.swirl_eig
    sol.set_particles(x, y, z, ap=0.005)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    tol = 1e-10 if precision == 0 else 1e-4
    for step in range(6):
        ch.step(1)
        sol.step(ost)
        errs = {k: rel(ch.get(k), ost[k]) for k in ("scalar", "scn", "v")}
        print(f"  step {step + 1}:", {k: f"{e:.2e}" for k, e in errs.items()}, flush=True)
        assert max(errs.values()) < tol, (step, errs)
    ch.close()


@pytest.mark.parametrize("case", [
    dict(precision=0, cd_switch=2, ap=0.01, ratio=2.0, gravity=0.5, c_mu=1.0),    # Schiller-Naumann, 1 sub-step
    dict(precision=0, cd_switch=1, ap=0.003, ratio=0.5, gravity=-0.2, c_mu=1.2),  # Stokes, 2 sub-steps
    dict(precision=1, cd_switch=2, ap=0.01, ratio=2.0, gravity=0.5, c_mu=1.0),
])
def test_inertial_particles_follow_the_channel_step(case):
    """Maxey-Riley particles (particle_flag 1, RK4 with sub-steps) carried by
    the channel step against the oracle's restatement of part_track/RK4/
    RKstage/SubDerivative on the same flow."""
    case = dict(case)
    precision = case.pop("precision")
    p = params(32, 32, 32, dt=0.01)
    print("case: inertial particles, 32^3,", case, "precision", precision, flush=True)
    st = seeded(p, 31, amp=0.2)
    rng = np.random.default_rng(32)
    n = 400
    x, y, z = random_points(p, n, rng, wall=0.95)
    ch = channel(p, precision=precision, particle_flag=1, **case)
    for k in ("v", "omy", "u0", "w0"):
        ch.put(k, st[k])
    ch.set_particles(x, y, z)
    ch.startup()
    sol = O.Solver(p)
    sol.swirl_fn = O.swirl_eig
    sol.set_particles(x, y, z, particle_flag=1, **case)
    ost = {k: v.copy() for k, v in st.items()}
    sol.startup(ost)
    tol = 1e-10 if precision == 0 else 1e-4
    print("sub-steps:", sol.substep, flush=True)
    for step in range(5):
        ch.step(1)
        sol.step(ost)
        got = ch.particles()
        e = {k: float(np.abs(got[k] - sol.particles[k]).max()) for k in "xyz"}
        e.update({k: rel(got[k], sol.particles[k]) for k in "uvw"})
        print(f"step {step + 1}:", {k: f"{v:.1e}" for k, v in e.items()}, flush=True)
        assert max(e.values()) < tol, (step, e)
    # the particles are inertial: their velocity differs from the fluid's
    uf = O.fluid_interp(got["x"], got["y"], got["z"], [sol.last_phys["up"]], p)[0]
    assert np.abs(got["u"] - uf).max() > 1e-6
    ch.close()


def test_inertial_particle_cfl_failure_is_reported():
    """A heavy particle (density ratio 1e6) under strong buoyancy leaves the
    fluid velocity by more than 10 a/dt in one step: the step returns the
    particle-CFL error (part_track.f90:216-221)."""
    p = params(32, 32, 32, dt=0.01)
    print("case: inertial particle CFL, 32^3", flush=True)
    st = seeded(p, 33, amp=0.1)
    ch = channel(p, particle_flag=1, cd_switch=2, ap=1e-4, ratio=1e6, gravity=100.0)
    for k in ("v", "omy", "u0", "w0"):
        ch.put(k, st[k])
    ch.set_particles([1.0, 2.0], [0.0, 0.2], [1.0, 3.0])
    ch.startup()
    with pytest.raises(kmm.KmmError) as ei:
        ch.step(1)
    assert ei.value.code == kmm.KMM_E_CFL and "particle cfl" in str(ei.value)
    ch.close()
