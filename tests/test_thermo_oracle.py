"""CPU pin of the thermodynamic-equilibrium fixtures (no GPU).

tests/golden/square30_spin_thermo.npz holds the reference's Density,
Magnetization, SpinPolarizedLDOS and expectation values
(Lib/src/PropertyExtractor/ChebyshevExpander.cpp:297-453, 575-717) on the
30x30 spin model. Here they are recomputed from the FP64 oracle's Green's
functions (oracle/kpm_oracle.c: moments, Lorentz damping, lookup-table
contraction, each pinned to the reference elsewhere) with the textbook
definitions
    n_i     = sum_E f(E) Im G_ii(E) / pi dE,
    M_i^rc  = sum_E f(E) (-i) G_(i r),(i c)(E) / pi dE,
    <c_a^+ c_b> = -sum_E f(E) conj(i G_ab(E)) dE / pi,
with the Fermi-Dirac f at T = 300 K (k_B = 1.380649e-23 / 1.602176634e-19
eV/K) and the step function at T = 0, and dE = E_1 - E_0. The Ranges-format
blocks are checked against the same sums over the walked indices, the
Custom-format blocks against the IndexTree (sorted) order.
"""
import numpy as np
import pytest

from conftest import load_fixture
from tbtk_amd import lattice as L

KB = 1.380649e-23 / 1.602176634e-19
TOL = 1e-12


def fermi(E, mu, T):
    if T == 0:
        return np.where(E < mu, 1.0, np.where(E == mu, 0.5, 0.0))
    return 1.0 / (np.exp((E - mu) / (KB * T)) + 1.0)


@pytest.fixture(scope="module")
def setup(oracle_mod):
    fx = load_fixture("square30_spin_thermo")
    meta = fx["meta"]
    m = L.config_plain_small()
    csr = m.csr()
    lo, hi, E = meta["window"]
    s, M = meta["scale"], meta["M"]
    table = oracle_mod.lookup_table(M, int(E), lo, hi, s)
    Egrid = oracle_mod.energy_grid(lo, hi, int(E))

    def gf(to, frm):
        c = oracle_mod.moments(csr, s, m.basis_index(frm), [m.basis_index(to)], M)
        c = oracle_mod.lorentz(c, meta["broadening"])[0]
        return oracle_mod.gf_lookup(3, table, c)   # NonPrincipal

    return fx, meta, m, Egrid, gf


def density_of(gf, Egrid, f, idx):
    dE = Egrid[1] - Egrid[0]
    return float((f * gf(idx, idx).imag / np.pi * dE).sum())


def test_density_fixture_vs_oracle(setup):
    fx, meta, m, Eg, gf = setup
    for T in meta["T"]:
        f = fermi(Eg, meta["mu"], T)
        tag = f"T{T:g}"
        # Custom: IndexTree order of {15,15,+}, {3,4,*}, {0,0,1}
        exp = [density_of(gf, Eg, f, (0, 0, 1)),
               density_of(gf, Eg, f, (3, 4, 0)),
               density_of(gf, Eg, f, (3, 4, 1)),
               density_of(gf, Eg, f, (15, 15, 0)) + density_of(gf, Eg, f, (15, 15, 1))]
        assert np.allclose(fx[f"density_custom_{tag}"], exp, rtol=TOL, atol=0), tag
        # Ranges: {x, 15, +} over x < 4
        exp = [density_of(gf, Eg, f, (x, 15, 0)) + density_of(gf, Eg, f, (x, 15, 1))
               for x in range(4)]
        assert np.allclose(fx[f"density_ranges_{tag}"], exp, rtol=TOL, atol=0), tag


def spin_matrix(gf, Eg, f, site):
    dE = Eg[1] - Eg[0]
    out = []
    for n in range(4):
        g = gf((*site, n // 2), (*site, n % 2))
        out.append((f * (-1j) * g / np.pi * dE).sum())
    return out


def test_magnetization_fixture_vs_oracle(setup):
    fx, meta, m, Eg, gf = setup
    for T in meta["T"]:
        f = fermi(Eg, meta["mu"], T)
        tag = f"T{T:g}"
        exp = spin_matrix(gf, Eg, f, (2, 3)) + spin_matrix(gf, Eg, f, (15, 15))
        got = fx[f"mag_custom_{tag}"]
        assert np.abs(got - exp).max() <= TOL * np.abs(exp).max(), tag
        exp = sum((spin_matrix(gf, Eg, f, (x, 15)) for x in range(3)), [])
        got = fx[f"mag_ranges_{tag}"]
        assert np.abs(got - exp).max() <= TOL * np.abs(exp).max(), tag
        # a collinear Zeeman model has no spin-flip terms, and the diagonal
        # elements are the spin-resolved densities
        assert np.all(got.reshape(-1, 4)[:, 1:3] == 0)
        assert np.allclose(got.reshape(-1, 4)[0, 0].real,
                           density_of(gf, Eg, f, (0, 15, 0)), rtol=TOL)


def test_splodos_fixture_vs_oracle(setup):
    fx, meta, m, Eg, gf = setup

    def block(site):
        gs = [gf((*site, n // 2), (*site, n % 2)) for n in range(4)]
        return np.stack([-1j * g / np.pi for g in gs], axis=1)   # E x 4

    exp = block((15, 15)).ravel()
    got = fx["splodos_custom"]
    assert np.abs(got - exp).max() <= TOL * np.abs(exp).max()
    exp = np.concatenate([block((x, 14)).ravel() for x in range(2)])
    got = fx["splodos_ranges"]
    assert np.abs(got - exp).max() <= TOL * np.abs(exp).max()


def test_expectation_values_fixture_vs_oracle(setup):
    fx, meta, m, Eg, gf = setup
    dE = Eg[1] - Eg[0]
    for T in meta["T"]:
        f = fermi(Eg, meta["mu"], T)
        tag = f"T{T:g}"
        for name, (to, frm) in (("offdiag", ((15, 15, 0), (16, 15, 0))),
                                ("diag", ((15, 15, 1), (15, 15, 1)))):
            exp = -(f * np.conj(1j * gf(to, frm)) * dE / np.pi).sum()
            got = fx[f"expval_{name}_{tag}"][0]
            assert abs(got - exp) <= TOL * abs(exp), (name, tag)
        # <c^+ c> of one state equals its density
        assert abs(fx[f"expval_diag_{tag}"][0].real -
                   density_of(gf, Eg, f, (15, 15, 1))) < 1e-13
