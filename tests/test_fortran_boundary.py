"""The Fortran side of the drop-in boundary on the CPU: the iso_c_binding
module (include/kmm.f90) and the drop-in modules derivs/solvers/helpers
(include/kmm_dropin.f90) compile with ROCm flang, and a Fortran program
reading dns.config through kmm_config_read gets exactly the kmm_config
the ctypes binding gets (every field, integers equal, reals bit for bit)."""
import os
import shutil
import subprocess

import pytest

from polymer_dns_amd import kmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="no flang in this image")


def _flang(args, cwd):
    r = subprocess.run([FLANG] + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r


def test_fortran_modules_compile_and_read_config_like_ctypes(tmp_path):
    inc = os.path.join(ROOT, "include")
    _flang(["-c", os.path.join(inc, "kmm.f90"), "-o", "kmm.o"], tmp_path)
    _flang(["-c", os.path.join(inc, "kmm_dropin.f90"), "-o", "dropin.o"], tmp_path)
    for mod in ("kmm_c", "derivs", "solvers", "helpers"):
        assert (tmp_path / f"{mod}.mod").exists(), mod
    libdir = os.path.dirname(kmm.LIB_PATH)
    _flang(["-I.", os.path.join(ROOT, "tests", "fortran", "cfg_dump.f90"), "kmm.o", "-L" + libdir, "-lkmm",
            "-Wl,-rpath," + libdir, "-o", "cfg_dump"], tmp_path)
    path = os.path.join(ROOT, "tests", "golden", "dns.config")
    out = subprocess.run([str(tmp_path / "cfg_dump"), path], capture_output=True, text=True, check=True).stdout
    vals = dict(line.split(None, 1) for line in out.strip().splitlines())
    assert vals.pop("rc") == "0"
    cfg = kmm.read_config(path)
    names = [f for f, _ in kmm.KmmConfig._fields_]
    assert sorted(vals) == sorted(names)
    for name, ctype in kmm.KmmConfig._fields_:
        want = getattr(cfg, name)
        got = vals[name].strip()
        if isinstance(want, int):
            assert int(got) == want, name
        else:
            assert float(got.replace("D", "E")) == want, name
