"""ctypes binding of the FP64 CPU oracle (oracle/kpm_oracle.c).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker. The product package never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkpm_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

ADVANCED, RETARDED, PRINCIPAL, NONPRINCIPAL = 0, 1, 2, 3
TYPE_NAMES = {ADVANCED: "Advanced", RETARDED: "Retarded",
              PRINCIPAL: "Principal", NONPRINCIPAL: "NonPrincipal"}

# -O2 with contraction off and no -march: the same x86-64 SSE2 arithmetic the
# reference build (oracle/Makefile.ref) uses.
CFLAGS = ["-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=gnu11"]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "kpm_oracle.c")
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", *CFLAGS, src, "-o", LIB_PATH, "-lm"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.kpm_moments.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_double,
                                  ctypes.c_int64, ctypes.c_int, P, ctypes.c_int, P]
        L.kpm_lorentz.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, P]
        L.kpm_jackson.argtypes = [ctypes.c_int, P]
        L.kpm_lookup_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, P]
        L.kpm_gf_lookup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P]
        L.kpm_gf_direct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, P, P]
        L.kpm_moments_trace.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_double,
                                        ctypes.c_int, P, ctypes.c_int, P]
        L.kpm_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.kpm_uniform.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def moments(csr, scale, frm, to, M):
    """Raw moments c[K][M] (complex128) for to-basis list `to`, from-basis `frm`."""
    rp, col, val = (np.ascontiguousarray(x) for x in csr)
    val = np.ascontiguousarray(val, dtype=np.complex128)
    to = np.ascontiguousarray(np.atleast_1d(to), dtype=np.int64)
    out = np.zeros((to.size, M), dtype=np.complex128)
    lib().kpm_moments(rp.size - 1, _p(rp), _p(col), _p(val), scale, int(frm),
                      to.size, _p(to), M, _p(out))
    return out


def lorentz(c, broadening):
    c = np.array(c, dtype=np.complex128, order="C", ndmin=2)
    lib().kpm_lorentz(c.shape[0], c.shape[1], broadening, _p(c))
    return c


def jackson(M):
    g = np.zeros(M)
    lib().kpm_jackson(M, _p(g))
    return g


def lookup_table(M, E, lo, hi, scale):
    L = np.zeros((M, E), dtype=np.complex128)
    lib().kpm_lookup_table(M, E, lo, hi, scale, _p(L))
    return L


def gf_lookup(gtype, table, c):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    M, E = table.shape
    G = np.zeros(E, dtype=np.complex128)
    lib().kpm_gf_lookup(gtype, M, E, _p(np.ascontiguousarray(table)), _p(c), _p(G))
    return G


def gf_direct(gtype, c, lo, hi, E, scale):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    G = np.zeros(E, dtype=np.complex128)
    lib().kpm_gf_direct(gtype, c.size, E, lo, hi, scale, _p(c), _p(G))
    return G


def moments_trace(csr, scale, vecs, M):
    rp, col, val = (np.ascontiguousarray(x) for x in csr)
    val = np.ascontiguousarray(val, dtype=np.complex128)
    vecs = np.ascontiguousarray(vecs, dtype=np.complex128)
    mu = np.zeros(M, dtype=np.complex128)
    lib().kpm_moments_trace(rp.size - 1, _p(rp), _p(col), _p(val), scale,
                            vecs.shape[0], _p(vecs), M, _p(mu))
    return mu


def uniform(seed, counters):
    f = lib().kpm_uniform
    return np.array([f(seed, int(c)) for c in np.atleast_1d(counters)])


def energy_grid(lo, hi, E):
    dx = (hi - lo) / (E - 1)
    return lo + np.arange(E) * dx


def ldos_from_moments(c, lo, hi, E, scale, table=None):
    """LDOS[e] = Im(G_NonPrincipal[e])/pi summed over the rows of c, as
    PropertyExtractor::ChebyshevExpander::calculateLDOSCallback
    (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:661-684)."""
    c = np.atleast_2d(c)
    out = np.zeros(E)
    for row in c:
        if table is None:
            G = gf_direct(NONPRINCIPAL, row, lo, hi, E, scale)
        else:
            G = gf_lookup(NONPRINCIPAL, table, row)
        out += G.imag / np.pi
    return out


def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def run_ref(threads: int = 1, **kw):
    """Run the compiled reference driver (oracle/_ref/ref_driver).

    One OpenMP thread by default: the reference builds its lookup table inside
    the `omp parallel for` of PropertyExtractor::ChebyshevExpander::
    calculateGreensFunctions (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:
    278-292 -> ensureLookupTableIsReady), which races and can crash with
    several threads."""
    args = [REF_DRIVER] + [f"{k}={v}" for k, v in kw.items()]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    return subprocess.run(args, check=True, capture_output=True, text=True, env=env)
