"""ctypes binding of oracle/kmm_cpu.c, the compiled CPU restatement of the
reference's time step (channel + passive scalar + Goldstein IB cell types
1 and 4, FP64, OpenMP).

TEST INFRASTRUCTURE AND CPU BASELINE ONLY: the CPU tests check it against
the numpy oracle, and bench.py times it as its CPU leg.  It steps a state
dict of the numpy oracle's layout (``oracle.kmm_oracle.Solver``) in place,
starting from the solver's Green's functions and step counter.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "kmm_cpu.c")
LIB = os.path.join(HERE, "libkmm_cpu.so")
CFLAGS = ["-O3", "-fPIC", "-shared", "-fopenmp", "-fcx-limited-range", "-std=c11", "-D_DEFAULT_SOURCE"]


def build(verbose: bool = False) -> str:
    """gcc -O3 -fopenmp of kmm_cpu.c into oracle/libkmm_cpu.so."""
    cmd = ["gcc"] + CFLAGS + [SRC, "-o", LIB, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for kmm_cpu.c:\n{r.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


class _Params(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in ("nx", "ny", "nz")]
                + [(k, ctypes.c_double) for k in ("xl", "yl", "zl", "re", "dt", "theta", "dPdx", "gain", "ugain")]
                + [(k, ctypes.c_int) for k in ("flow_select", "perturbtime", "scalar", "scl_flag")]
                + [(k, ctypes.c_double) for k in ("diff", "deltaT")]
                + [(k, ctypes.c_double) for k in ("atu", "btu", "gtu", "abu", "bbu", "gbu",
                                                  "atw", "btw", "gtw", "abw", "bbw", "gbw")]
                + [("threads", ctypes.c_int)])


_P = ctypes.c_void_p


class _State(ctypes.Structure):
    _fields_ = [(k, _P) for k in ("v", "fn", "fnm1", "omy", "gn", "gnm1", "sc", "scn", "scnm1",
                                  "u0", "w0", "h1n", "h1nm1", "h3n", "h3nm1", "fxi", "fyi", "fzi", "imat",
                                  "gf1", "gf2", "dnv1b", "dnv1t", "dnv2b", "dnv2t")]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.kc_steps.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_State), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)]
        _lib.kc_steps.restype = ctypes.c_int
        _lib.kc_max_threads.restype = ctypes.c_int
    return _lib


class CFLError(RuntimeError):
    pass


def max_threads() -> int:
    return int(load().kc_max_threads())


def steps(sol, st: dict, nsteps: int, threads: int = 0) -> float:
    """Advances the oracle solver's state dict `st` by nsteps steps with the
    compiled restatement (in place); returns the last step's CFL number."""
    p = sol.p
    if p.polymer or p.cmf or sol.particles is not None:
        raise NotImplementedError("the compiled restatement covers the channel, scalar and IB types 1, 4")
    lib = load()
    pp = _Params(nx=p.nx, ny=p.ny, nz=p.nz, xl=p.xl, yl=p.yl, zl=p.zl, re=p.re, dt=p.dt, theta=p.theta,
                 dPdx=p.dPdx, gain=p.gain, ugain=p.ugain, flow_select=p.flow_select, perturbtime=p.perturbtime,
                 scalar=int(bool(p.scalar)), scl_flag=p.scl_flag, diff=p.diff, deltaT=p.deltaT,
                 atu=p.atu, btu=p.btu, gtu=p.gtu, abu=p.abu, bbu=p.bbu, gbu=p.gbu,
                 atw=p.atw, btw=p.btw, gtw=p.gtw, abw=p.abw, bbw=p.bbw, gbw=p.gbw, threads=threads)
    keep = {}

    def arr(key, dtype, src=None):
        a = np.ascontiguousarray(st[key] if src is None else src, dtype=dtype)
        keep[key] = a
        return a.ctypes.data

    s = _State()
    for k in ("v", "fn", "fnm1", "omy", "gn", "gnm1"):
        setattr(s, k, arr(k, np.complex128))
    if p.scalar:
        s.sc = arr("scalar", np.complex128)
        s.scn = arr("scn", np.complex128)
        s.scnm1 = arr("scnm1", np.complex128)
    for k in ("u0", "w0", "h1n", "h1nm1", "h3n", "h3nm1"):
        setattr(s, k, arr(k, np.float64))
    s.fxi = arr("fxintg", np.float64)
    s.fyi = arr("fyintg", np.float64)
    s.fzi = arr("fzintg", np.float64)
    if sol.imatrix is not None:
        s.imat = arr("_imat", np.int32, sol.imatrix)
    s.gf1 = arr("_gf1", np.complex128, sol.gf1)
    s.gf2 = arr("_gf2", np.complex128, sol.gf2)
    for k, v in (("dnv1b", sol.dnv1b), ("dnv1t", sol.dnv1t), ("dnv2b", sol.dnv2b), ("dnv2t", sol.dnv2t)):
        setattr(s, k, arr("_" + k, np.complex128, v))
    cfl = ctypes.c_double(0.0)
    r = lib.kc_steps(ctypes.byref(pp), ctypes.byref(s), int(sol.it), int(nsteps), ctypes.byref(cfl))
    # results back into the state dict (copies were made where the caller's
    # arrays were not contiguous of the right type)
    names = {"scalar": "scalar"}
    for k, a in keep.items():
        if not k.startswith("_"):
            st[names.get(k, k)] = a
    if r == 2:
        raise NotImplementedError("configuration outside the compiled restatement")
    if r == 3:
        raise ValueError("IB spreading leaves the plane (quirk Q5)")
    sol.it += nsteps if r == 0 else 0
    st["cflmax"] = cfl.value
    if r == 1:
        raise CFLError(f"CFL failure: {cfl.value}")
    return cfl.value
