"""GPU tests of the device-side FP64 damping of the contraction's
coefficients (tbtkx_set_damping).

An identity lookup table (L[n][e] = 1 for e == n, else 0) makes the FP64
Retarded contraction return its coefficients exactly (c * (1, 0) and sums with
zeros are exact), so the damped coefficients themselves are compared with the
host's: Lorentz damping against the oracle's restatement of the reference
(Lib/src/Solver/ChebyshevExpander.cpp:190-196) and the Jackson kernel against
its closed form, each within a few ulp (the device sinh/sin/cos against the
host libm's). No damping, and Lorentz with broadening 0 as in the reference,
leave the coefficients bit for bit as given."""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm

from conftest import load_fixture

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


@pytest.fixture
def damping(engine):
    yield engine
    engine.set_damping(X.DAMP_NONE)


def identity_table(M):
    return np.eye(M, dtype=np.complex128)


def coefficients(K, M, seed):
    rng = np.random.default_rng(seed)
    decay = np.exp(-np.arange(M) / (0.3 * M))
    return (rng.normal(size=(K, M)) + 1j * rng.normal(size=(K, M))) * decay


def damped_on_device(engine, c, kind, broadening=0.0):
    engine.set_damping(kind, broadening)
    engine.lut_upload(identity_table(c.shape[1]))
    return engine.contract(c, X.RETARDED)


def within_ulps(got, ref, ulps):
    bound = ulps * EPS * np.maximum(np.abs(ref.real), np.abs(ref.imag)) + 1e-300
    return (np.abs(got.real - ref.real) <= bound).all() and \
        (np.abs(got.imag - ref.imag) <= bound).all()


def test_identity_table_returns_coefficients(damping):
    c = coefficients(3, 257, 1)
    got = damped_on_device(damping, c, X.DAMP_NONE)
    assert np.array_equal(got, c)


def test_lorentz_zero_broadening_is_identity(damping):
    c = coefficients(2, 130, 2)
    got = damped_on_device(damping, c, X.DAMP_LORENTZ, 0.0)
    assert np.array_equal(got, c)


@pytest.mark.parametrize("broadening", [1e-6, 1e-3, 0.02])
def test_lorentz_matches_oracle(damping, oracle_mod, broadening):
    K, M = 5, 1000
    c = coefficients(K, M, 3)
    got = damped_on_device(damping, c, X.DAMP_LORENTZ, broadening)
    ref = oracle_mod.lorentz(c, broadening)
    # the host helper the bench used before is bitwise the oracle
    assert np.array_equal(kpm.lorentz(c, broadening), ref)
    assert within_ulps(got, ref, 8)


@pytest.mark.parametrize("M", [1, 2, 77, 2000])
def test_jackson_matches_closed_form(damping, M):
    c = coefficients(4, M, 4)
    got = damped_on_device(damping, c, X.DAMP_JACKSON)
    ref = c * kpm.jackson(M)[None, :]
    # Near n = M the two terms of g_n cancel to O(q^2): a last-bit difference
    # of cos or sin scales with the terms' magnitude, not with g_n's.
    q = np.pi / (M + 1)
    n = np.arange(M)
    terms = (np.abs((M - n + 1) * np.cos(q * n))
             + np.abs(np.sin(q * n) * np.cos(q) / np.sin(q))) / (M + 1)
    bound = 8 * EPS * terms[None, :] * np.maximum(np.abs(c.real), np.abs(c.imag))
    assert (np.abs(got.real - ref.real) <= bound).all()
    assert (np.abs(got.imag - ref.imag) <= bound).all()
    # g_0 = 1 exactly in exact arithmetic; the kernel is positive and decays
    g = got[0] / c[0]
    assert abs(g[0].real - 1) < 1e-14
    assert (g.real > -1e-15).all() and np.all(np.diff(g.real) < 1e-15)


def test_lorentz_matches_reference_fixture(damping):
    """The reference's own Lorentz-damped moments (30x30 with spin, from
    (15,15,0), M = 1000, broadening 1e-6) from its raw moments, damped on the
    device."""
    fx = load_fixture("square30_spin")
    got = damped_on_device(damping, fx["coef_b0"], X.DAMP_LORENTZ, 1e-6)
    assert within_ulps(got, fx["coef_b1e-06"], 8)


def test_ldos_with_device_lorentz(damping, oracle_mod):
    """Config-1 LDOS at (15,15,0) (s = 10, M = 1000, E = 1000 on
    [-9.99, 9.99]) from the reference's raw moments with the Lorentz damping
    done on the device: equal to the oracle's LDOS within 1e-13 of its
    maximum."""
    fx = load_fixture("square30_spin")
    meta = fx["meta"]
    s, M, E, lo, hi = meta["scale"], meta["M"], 1000, -9.99, 9.99
    row = len(meta["to"]) - 1          # to = from = (15,15,0), last position
    assert meta["to"][row] == meta["frm"]
    c = fx["coef_b0"][row:row + 1]
    table = oracle_mod.lookup_table(M, E, lo, hi, s)
    ref = oracle_mod.ldos_from_moments(oracle_mod.lorentz(c, 1e-6), lo, hi, E, s, table)
    damping.lut_upload(table)
    damping.set_damping(X.DAMP_LORENTZ, 1e-6)
    got = kpm.ldos_from_gf(damping.contract(c, X.NONPRINCIPAL))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_mfma_with_device_jackson(damping):
    K, M, E, s = 200, 512, 300, 7.0
    c = coefficients(K, M, 5)
    damping.lut_generate(M, E, -0.95 * s, 0.95 * s, s)
    ref = damping.contract(c * kpm.jackson(M)[None, :], X.NONPRINCIPAL)
    damping.set_damping(X.DAMP_JACKSON)
    damping.coeffs_upload(c)
    got = damping.contract_mfma(X.NONPRINCIPAL)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 3e-5


def test_unknown_kind_is_an_error(damping):
    with pytest.raises(X.EngineError):
        damping.set_damping(7)
    with pytest.raises(X.EngineError):
        damping.set_damping(X.DAMP_LORENTZ, float("nan"))
