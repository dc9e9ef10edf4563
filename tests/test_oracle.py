"""Pins the FP64 CPU oracle (oracle/kpm_oracle.c) before anything is checked
against it:
  * bit-for-bit against golden vectors written by the compiled reference
    (tests/golden/make_fixtures.py);
  * against the closed forms of the reference's own unit tests
    (Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-565) and its LDOS
    sum rule (Test/include/TBTK/Test/PropertyExtractor/ChebyshevExpander.h:19-45);
  * against mathematics independent of the recurrence: moments from a dense
    eigendecomposition, c_n = sum_l <i|l><l|j> T_n(l/s), the trace from the
    same decomposition, and the Jackson kernel as a brute-force
    autocorrelation.
"""
import numpy as np
import pytest

from conftest import load_fixture
from tbtk_amd import lattice as L

EPS100 = 100 * np.finfo(float).eps

MOMENT_FIXTURES = {
    "periodic5": lambda: L.periodic_square(5),
    "square30_spin": L.config_plain_small,
    "square48_spin_disorder": lambda: L.square_spin(48, 48, t=1.0, J=0.5, disorder=1.0, seed=2),
    "nambu12": lambda: L.nambu_swave(12, 12, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.5, seed=3),
    "nambu12_b": lambda: L.nambu_swave(12, 12, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.5, seed=3),
    "honeycomb6x9": lambda: L.honeycomb_ribbon(6, 9, t=2.7, disorder=0.5, seed=4),
    "cubic8": lambda: L.cubic(8, 8, 8, t=1.0, disorder=1.0, seed=5),
}


def csr_digest(rp, col, val) -> str:
    """SHA-256 of rowptr (u32), col (u32) and val (complex128) bytes, as
    tests/golden/make_fixtures.py stores it for the larger models."""
    import hashlib
    h = hashlib.sha256()
    for a, dt in ((rp, "<u4"), (col, "<u4"), (val, "<c16")):
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", sorted(MOMENT_FIXTURES))
def test_csr_matches_reference_assembly(name):
    """The generators' CSR equals the reference's assembly of the same
    HoppingAmplitudes (getSparseMatrix): array for array where the fixture
    keeps it, else by the SHA-256 of its bytes."""
    fx = load_fixture(name)
    rp, col, val = MOMENT_FIXTURES[name]().csr()
    if "rowptr" not in fx:
        assert csr_digest(rp, col, val) == str(fx["csr_sha256"])
        return
    assert np.array_equal(rp, fx["rowptr"])
    assert np.array_equal(col, fx["col"])
    assert np.array_equal(val, fx["val"])


@pytest.mark.parametrize("name", sorted(MOMENT_FIXTURES))
def test_oracle_moments_bitexact_vs_reference(oracle_mod, name):
    O = oracle_mod
    fx = load_fixture(name)
    meta = fx["meta"]
    m = MOMENT_FIXTURES[name]()
    frm = m.basis_index(meta["frm"])
    to = [m.basis_index(t) for t in meta["to"]]
    raw = O.moments(m.csr(), meta["scale"], frm, to, meta["M"])
    for b in meta["broadenings"]:
        got = O.lorentz(raw, b) if b else raw
        assert np.array_equal(got, fx[f"coef_b{b:g}"]), (name, b)


def test_known_answers_periodic_square(oracle_mod):
    """c0, c1, c2 of the 5x5 periodic lattice, mu=-2, t=1, s=10
    (Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-277)."""
    O = oracle_mod
    m = L.periodic_square(5, t=1.0, mu=-2.0)
    mu, t, s = -2.0, 1.0, 10.0
    c = O.moments(m.csr(), s, m.basis_index((0, 0)),
                  [m.basis_index((0, 0)), m.basis_index((1, 0))], 100)
    assert abs(c[0, 0] - 1) < EPS100
    assert abs(c[0, 1] - (-mu / s)) < EPS100
    assert abs(c[0, 2] - (2 * (mu * mu + 4 * t * t) / (s * s) - 1)) < EPS100
    assert abs(c[1, 0]) < EPS100
    assert abs(c[1, 1] - (-t / s)) < EPS100
    assert abs(c[1, 2] - 4 * mu * t / (s * s)) < EPS100
    single = O.moments(m.csr(), s, m.basis_index((0, 0)), [m.basis_index((1, 0))], 100)
    assert np.array_equal(single[0], c[1])


def test_repeated_to_index_keeps_last_position(oracle_mod):
    O = oracle_mod
    m = L.periodic_square(5)
    c = O.moments(m.csr(), 10.0, 0, [0, 1, 0], 20)
    assert np.all(c[0] == 0)
    assert c[2, 0] == 1


def _dense(m):
    rp, col, val = m.csr()
    N = m.basis_size
    H = np.zeros((N, N), dtype=np.complex128)
    for r in range(N):
        H[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return H


@pytest.mark.parametrize("model", [
    L.nambu_swave(4, 5, disorder=0.5, seed=9),
    L.honeycomb_ribbon(3, 4, disorder=0.5, seed=1),
    L.square_spin(6, 5, disorder=1.0, seed=2),
])
def test_oracle_moments_vs_eigendecomposition(oracle_mod, model):
    O = oracle_mod
    H = _dense(model)
    assert np.allclose(H, H.conj().T)
    s = model.gershgorin_bound() * 1.1
    lam, V = np.linalg.eigh(H)
    M = 300
    theta = np.arccos(lam / s)
    Tn = np.cos(np.outer(np.arange(M), theta))          # [M, N]
    frm, to = 3, [3, 7, 11]
    exact = np.einsum("kl,l,nl->kn", V[to, :], V[frm, :].conj(), Tn)
    got = O.moments(model.csr(), s, frm, to, M)
    assert np.abs(got - exact).max() < 1e-11
    assert np.abs(got).max() <= 1 + 1e-12


@pytest.mark.parametrize("gname", ["c123", "crand"])
def test_gf_types_bitexact_vs_reference(oracle_mod, gname):
    O = oracle_mod
    fx = load_fixture("gf_types")
    lo, hi, E = fx["meta"]["windows"][gname]
    c = fx[f"{gname}_coef"]
    table = O.lookup_table(c.size, E, lo, hi, 10.0)
    for t, tname in O.TYPE_NAMES.items():
        assert np.array_equal(O.gf_direct(t, c, lo, hi, E, 10.0),
                              fx[f"{gname}_{tname}_lookup0"]), tname
        assert np.array_equal(O.gf_lookup(t, table, c),
                              fx[f"{gname}_{tname}_lookup1"]), tname


def test_gf_closed_forms(oracle_mod):
    """Test/include/TBTK/Test/Solver/ChebyshevExpander.h:371-505."""
    O = oracle_mod
    s = 10.0
    E = O.energy_grid(-5.0, 5.0, 10)
    c = np.array([1, 2, 3], dtype=np.complex128)
    th = np.arccos(E / s)
    n = np.arange(1, 3)[:, None]
    den = np.sqrt(s * s - E * E)
    ret = (-c[0] / 2 - (c[1:, None] * np.exp(-1j * n * th)).sum(0)) * 2j / den
    adv = (c[0] / 2 + (c[1:, None] * np.exp(1j * n * th)).sum(0)) * 2j / den
    pri = (c[1:, None] * np.sin(n * th)).sum(0) * 2 / den
    npr = (c[0] / 2 + (c[1:, None] * np.cos(n * th)).sum(0)) * 2j / den
    table = O.lookup_table(3, 10, -5.0, 5.0, s)
    for t, ref in [(O.RETARDED, ret), (O.ADVANCED, adv), (O.PRINCIPAL, pri), (O.NONPRINCIPAL, npr)]:
        assert np.abs(O.gf_direct(t, c, -5.0, 5.0, 10, s) - ref).max() < EPS100
        assert np.abs(O.gf_lookup(t, table, c) - ref).max() < EPS100


def test_ldos_bitexact_vs_reference(oracle_mod):
    O = oracle_mod
    fx = load_fixture("square30_spin_ldos")
    meta = fx["meta"]
    m = L.config_plain_small()
    lo, hi, E = meta["window"]
    rows = []
    for s in (0, 1):
        b = m.basis_index((15, 15, s))
        rows.append(O.lorentz(O.moments(m.csr(), 10.0, b, [b], 1000), 1e-6)[0])
    table = O.lookup_table(1000, E, lo, hi, 10.0)
    for lk in (0, 1):
        got = O.ldos_from_moments(np.array(rows), lo, hi, E, 10.0, table if lk else None)
        assert np.array_equal(got, fx[f"ldos_lookup{lk}"])
    dE = (hi - lo) / (E - 1)
    assert abs(fx["ldos_lookup1"].sum() * dE - 2) < 1e-3


def test_gf_pe_bitexact_vs_reference(oracle_mod):
    O = oracle_mod
    fx = load_fixture("square30_spin_ldos")
    g = fx["meta"]["gf"]
    m = L.config_plain_small()
    lo, hi, E = g["window"]
    c = O.moments(m.csr(), 10.0, m.basis_index(g["frm"]),
                  [m.basis_index(t) for t in g["to"]], g["M"])
    c = O.lorentz(c, 1e-6)
    table = O.lookup_table(g["M"], E, lo, hi, 10.0)
    for t, tname in O.TYPE_NAMES.items():
        got = np.array([O.gf_lookup(t, table, row) for row in c])
        assert np.array_equal(got, fx[f"gf_{tname}"]), tname


def test_ring_sum_rule(oracle_mod):
    """Test/include/TBTK/Test/PropertyExtractor/ChebyshevExpander.h:19-45."""
    O = oracle_mod
    fx = load_fixture("ring10_ldos")
    m = L.ring(10)
    c = O.lorentz(O.moments(m.csr(), 10.0, 0, [0], 1000), 1e-6)
    for frac in (0.9, 0.7):
        lo, hi = -10 * frac, 10 * frac
        ldos = O.ldos_from_moments(c, lo, hi, 1000, 10.0)
        assert np.array_equal(ldos, fx[f"ldos_{frac:g}"])
        assert abs(ldos.sum() * (hi - lo) / 999 - 1) < 1e-3


def test_jackson_kernel_is_normalised_autocorrelation(oracle_mod):
    """g_n = sum_v a_v a_{v+n} / sum_v a_v^2, a_v = sin(pi (v+1)/(M+1)),
    v = 0..M-1: the construction of Rev. Mod. Phys. 78, 275, Sec. II.C."""
    O = oracle_mod
    for M in (4, 17, 256):
        a = np.sin(np.pi * (np.arange(M) + 1) / (M + 1))
        brute = np.array([a[: M - n] @ a[n:] for n in range(M)]) / (a @ a)
        assert np.abs(O.jackson(M) - brute).max() < 1e-13


def test_trace_moments_vs_dense(oracle_mod):
    O = oracle_mod
    m = L.nambu_swave(3, 4, disorder=0.5, seed=1)
    N = m.basis_size
    s = m.gershgorin_bound() * 1.1
    R, seed, M = 3, 77, 64
    u = O.uniform(seed, np.arange(R * N)).reshape(R, N)
    vecs = np.exp(2j * np.pi * u)
    H = _dense(m)
    lam, V = np.linalg.eigh(H)
    Tn = np.cos(np.outer(np.arange(M), np.arccos(lam / s)))
    proj = np.abs(vecs @ V.conj()) ** 2      # [R, N] |<l|r>|^2
    exact = (proj.sum(0)[None, :] * Tn).sum(1)
    got = O.moments_trace(m.csr(), s, vecs, M)
    assert np.abs(got - exact).max() < 1e-10 * N * R
    # and the stochastic estimate tracks the exact trace
    tr = Tn.sum(1)
    assert abs(got[0].real / R - N) < 1e-9
    assert np.abs(got.real / R - tr).max() < 0.5 * N


def test_uniform_stream_statistics(oracle_mod):
    u = oracle_mod.uniform(5, np.arange(20000))
    assert 0 <= u.min() and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert len(np.unique(u)) == u.size


def test_host_lorentz_bitwise_vs_oracle(oracle_mod):
    """The host-side Lorentz damping the engine's LDOS uses (tbtk_amd.kpm)
    is the oracle's, bit for bit: c_n sinh(lambda (1 - n/M)) / sinh(lambda)
    per real and imaginary part, as std::complex<double> times and divided by
    a double (Lib/src/Solver/ChebyshevExpander.cpp:190-196). A complex
    division by the real denominator (which multiplies by its reciprocal)
    differs in the last bit for a share of the entries."""
    from tbtk_amd import kpm
    rng = np.random.default_rng(5)
    for K, M, b in ((1, 1000, 1e-6), (3, 257, 0.01), (2, 4000, 0.003)):
        c = rng.standard_normal((K, M)) + 1j * rng.standard_normal((K, M))
        got = kpm.lorentz(c, b)
        ref = oracle_mod.lorentz(c, b)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (K, M, b)
