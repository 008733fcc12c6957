"""Build libkmm.so (HIP kernels for gfx950 + C ABI) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkmm.so")
SOURCES = ["kmm_pencil.hip", "kmm_poly.hip", "kmm_plane.hip", "kmm_particles.hip", "kmm_mlp.hip", "kmm_audit.hip", "kmm_stats.hip", "kmm_routines.hip", "kmm_api.cpp", "kmm_mlp_api.cpp"]
ARCH = os.environ.get("KMM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]


def _obj(src: str) -> str:
    return os.path.join(CSRC, "_build", src + ".o")


def _compile(src: str) -> str:
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    out = _obj(src)
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out


ROOT = os.path.dirname(HERE)
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_BIN = os.path.join(ROOT, "tests", "fortran", "bin")


def build_fortran(verbose: bool = False) -> str | None:
    """The Fortran host programs of the drop-in tests (tests/fortran/dns_kmm.f90
    over include/kmm.f90, tests/fortran/dropin_routines.f90 over the drop-in
    modules), linked to libkmm.so by a path relative to the binary.  Skipped
    (None) when the image has no flang."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran host program not built")
        return None
    os.makedirs(FORTRAN_BIN, exist_ok=True)
    out = os.path.join(FORTRAN_BIN, "dns_kmm")
    rpath = "-Wl,-rpath,$ORIGIN/../../../polymer_dns_amd"
    steps = [
        [FLANG, "-c", os.path.join(ROOT, "include", "kmm.f90"), "-o", "kmm.o"],
        [FLANG, "-c", os.path.join(ROOT, "include", "kmm_dropin.f90"), "-o", "kmm_dropin.o"],
        [FLANG, "-I.", os.path.join(ROOT, "tests", "fortran", "dns_kmm.f90"), "kmm.o", "-L" + HERE, "-lkmm",
         rpath, "-o", out],
        [FLANG, "-I.", os.path.join(ROOT, "tests", "fortran", "dropin_routines.f90"), "kmm.o", "kmm_dropin.o",
         "-L" + HERE, "-lkmm", rpath, "-o", os.path.join(FORTRAN_BIN, "dropin_routines")],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, cwd=FORTRAN_BIN, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print("built", out)
    return out


def build(verbose: bool = False) -> str:
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
