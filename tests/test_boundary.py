"""Drop-in boundary on the CPU: the C ABI library loads and exports every
entry point declared in include/kmm.h, the positional dns.config reader,
the Fortran unformatted restart/c-restart/geometry formats, and the seeded
initial condition.  No compute call touches a GPU here."""
import ctypes
import math
import os
import re
import struct

import numpy as np
import pytest

from polymer_dns_amd import ic, kmm, restart

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "kmm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int)\s+(kmm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(kmm.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(kmm.EXPORTS)


def test_status_strings():
    lib = kmm.load_library()
    assert lib.kmm_status_string(0) == b"ok"
    assert b"CFL" in lib.kmm_status_string(1)


def test_config_reader_positional():
    cfg = kmm.read_config(os.path.join(GOLDEN, "dns.config"))
    assert (cfg.nx, cfg.ny, cfg.nz) == (128, 128, 128)
    assert cfg.nsteps == 25 and cfg.iprnfrq == 500
    assert cfg.dt == pytest.approx(0.002) and cfg.theta == 0.5
    assert cfg.gain == 20 and cfg.ugain == 100
    assert cfg.zl == pytest.approx(4.1887902) and cfg.xl == pytest.approx(12.566371)
    assert cfg.re == 2800.0 and cfg.r_tau == 180.0 and cfg.perturbtime == 40
    assert cfg.irstrt == 1 and cfg.flow_select == 1 and cfg.scl_flag == 6
    assert cfg.diff == pytest.approx(0.71)
    assert cfg.deltaT == 1.0          # Fortran d exponent
    assert cfg.src_stop == cfg.nsteps  # src_stop < src_start -> nsteps
    assert cfg.cadence == 10
    # particle block (main_dns.f90:987, 1013-1016)
    assert cfg.particle_flag == 0 and cfg.cd_switch == 2
    assert cfg.ratio == 0.0 and cfg.ap == pytest.approx(0.0085) and cfg.c_mu == 1.0


def test_config_reader_rejects_short_file(tmp_path):
    p = tmp_path / "short.config"
    p.write_text("\n".join(["!"] * 20))
    with pytest.raises(kmm.KmmError):
        kmm.read_config(str(p))


def test_restart_round_trip_and_markers(tmp_path):
    nx, ny, nz = 16, 8, 8
    rng = np.random.default_rng(0)
    shp = (nx // 2, nz, ny + 1)
    st = {k: rng.standard_normal(shp) + 1j * rng.standard_normal(shp) for k in restart.REC1}
    st.update({k: rng.standard_normal(ny + 1) for k in restart.REC1_MEAN})
    st.update({k: rng.standard_normal((3 * nx // 2, 3 * nz // 2, ny + 1)) for k in restart.REC2})
    p = tmp_path / "restart"
    restart.write_restart(p, st, nx, ny, nz)
    raw = p.read_bytes()
    n1 = 6 * 16 * np.prod(shp) + 6 * 8 * (ny + 1)
    assert struct.unpack("<I", raw[:4])[0] == n1
    assert struct.unpack("<I", raw[4 + n1:8 + n1])[0] == n1
    back = restart.read_restart(p, nx, ny, nz)
    for k, v in st.items():
        assert np.array_equal(back[k], v)
    # Fortran column-major (nyp,nz,nxh): y fastest
    v = st["v"]
    first = np.frombuffer(raw[4:4 + 32], dtype=np.complex128)
    assert first[0] == v[0, 0, 0] and first[1] == v[0, 0, 1]


def test_truncated_restart_is_rejected(tmp_path):
    p = tmp_path / "r"
    p.write_bytes(b"\x10\x00\x00\x00" + b"\x00" * 8)
    with pytest.raises((EOFError, ValueError)):
        restart.read_restart(p, 16, 8, 8)


def test_geometry_round_trip(tmp_path):
    im = restart.rib_geometry(32, 32, 16, y_cells=3, x_center=20, x_width=3)
    p = tmp_path / "geometry"
    restart.write_geometry(p, im)
    assert np.array_equal(restart.read_geometry(p, 32, 32, 16), im)
    assert im.sum() == 3 * 3 * 24


def test_seeded_ic_deterministic_and_solenoidal():
    a = ic.seeded_state(32, 32, 32, 4 * math.pi, 2 * math.pi, 2800.0, math.sqrt(8400.0), 7)
    b = ic.seeded_state(32, 32, 32, 4 * math.pi, 2 * math.pi, 2800.0, math.sqrt(8400.0), 7)
    for k in a:
        assert a[k].tobytes() == b[k].tobytes()
    v, omy = a["v"], a["omy"]
    n = np.arange(33)
    # v = v' = 0 and omega_y = 0 at both walls
    for s in (1.0, -1.0):
        assert np.abs((v * s ** n).sum(-1)).max() < 1e-14
        assert np.abs((omy * s ** n).sum(-1)).max() < 1e-14
        assert np.abs((v * n * n * s ** (n + 1)).sum(-1)).max() < 1e-13
    # Hermitian kx = 0 plane
    for j in range(1, 16):
        assert np.allclose(v[0, j], np.conj(v[0, 32 - j]))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="host has a GPU")
def test_no_cpu_fallback():
    with pytest.raises(kmm.KmmError):
        kmm.Channel(kmm.make_config(32, 32, 32))
