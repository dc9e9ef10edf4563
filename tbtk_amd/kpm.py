"""Host-side helpers of the Chebyshev path (numpy; no GPU work here).

Damping kernels, the energy grid and the LDOS/DOS assembly that sit between
the engine's moments and its contraction. Arithmetic follows the reference
where the reference defines it:
  * Lorentz damping  Lib/src/Solver/ChebyshevExpander.cpp:190-196
  * Range            Lib/src/Utilities/Range.cpp:30-57, Range.h:130
  * LDOS             Lib/src/PropertyExtractor/ChebyshevExpander.cpp:661-684
"""
from __future__ import annotations

import functools
import math

import numpy as np


def energy_grid(lo: float, hi: float, res: int) -> np.ndarray:
    """Range(lo, hi, res)[e] = lo + e*dx, dx = (hi - lo)/(res - 1)."""
    dx = (hi - lo) / (res - 1)
    return lo + np.arange(res) * dx


@functools.lru_cache(maxsize=64)
def _lorentz_factors(M: int, broadening: float):
    """sinh(lambda (1 - n/M)) for n < M and sinh(lambda), lambda =
    broadening*M, with the C library's sinh (math.sinh)."""
    lam = broadening * M
    num = np.array([math.sinh(lam * (1 - n / float(M))) for n in range(M)])[None, :]
    num.setflags(write=False)
    return num, math.sinh(lam)


def lorentz(c: np.ndarray, broadening: float) -> np.ndarray:
    """c[k][n] * sinh(lambda (1 - n/M)) / sinh(lambda), lambda = broadening*M,
    evaluated with the C library's sinh in the reference's operation order."""
    c = np.array(c, dtype=np.complex128, ndmin=2)
    if broadening == 0:
        return c
    num, den = _lorentz_factors(c.shape[1], float(broadening))
    # Component by component, as std::complex<double> * double and / double
    # do (numpy's complex / real would multiply by 1/den instead).
    out = np.empty_like(c)
    out.real = (c.real * num) / den
    out.imag = (c.imag * num) / den
    return out


def jackson(M: int) -> np.ndarray:
    """Jackson kernel g_n (Rev. Mod. Phys. 78, 275 (2006), eq. 71)."""
    q = math.pi / (M + 1)
    n = np.arange(M)
    return ((M - n + 1) * np.cos(q * n) + np.sin(q * n) * math.cos(q) / math.sin(q)) / (M + 1)


def check_window(lo: float, hi: float, scale: float) -> None:
    """The strict containment Solver::ChebyshevExpander::setEnergyWindow
    asserts (Lib/include/TBTK/Solver/ChebyshevExpander.h:434-451)."""
    eps = np.finfo(float).eps
    if not (lo > -scale * (1 - 16 * eps) and hi < scale * (1 - 16 * eps)):
        raise ValueError(
            f"Invalid energy window [{lo}, {hi}]: not contained in (-{scale}, {scale})")


def ldos_from_gf(G_nonprincipal: np.ndarray) -> np.ndarray:
    """LDOS(E) = Im G_NonPrincipal(E) / pi, summed over rows."""
    return (np.atleast_2d(G_nonprincipal).imag / math.pi).sum(axis=0)


def gershgorin_scale(csr, margin: float = 1.05) -> float:
    """A scale factor from the Gershgorin bound max_i sum_j |H_ij| (FP64)."""
    rp, col, val = csr
    a = np.abs(np.asarray(val))
    rp = np.asarray(rp, dtype=np.int64)
    sums = np.add.reduceat(a, rp[:-1]) if a.size else np.zeros(rp.size - 1)
    sums = np.where(rp[1:] == rp[:-1], 0.0, sums)
    return float(sums.max()) * margin
