"""GPU tests of the plain FP64 step's compressed CSR forms: 16-bit column
offsets (TBTKX_OPT_COL16: each column as an offset from the first row of its
row block, when every offset of the model fits in int16) and the value
palette (TBTKX_OPT_PALETTE: a 1-byte code per non-zero into the model's at
most 255 distinct off-diagonal values, the diagonal from a per-row array).
The step forms the same products from the same values in the same order in
every form, so the moments are bitwise equal to the 32-bit, full-value
path's and to the oracle's (the reference's CPU recurrence,
Lib/src/Solver/ChebyshevExpander.cpp:65-82, 157-186)."""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def col16(engine):
    yield engine
    engine.set_option(X.OPT_COL16, 1)
    engine.set_option(X.OPT_PALETTE, 1)
    engine.set_option(X.OPT_GRAPH, 0)


MODELS = {
    "square_spin_40x37": lambda: L.square_spin(40, 37, disorder=1.0, seed=8),
    "nambu_12x9": lambda: L.nambu_swave(12, 9, disorder=0.5, seed=3),
    "square_flux_20x20": lambda: L.square(20, 20, disorder=0.3, seed=2, flux=0.07),
    "cubic_9": lambda: L.cubic(9, 9, 9, disorder=1.0, seed=5),
}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_col16_bitwise_equal_to_32bit_and_oracle(col16, oracle_mod, name):
    m = MODELS[name]()
    csr = m.csr()
    s = kpm.gershgorin_scale(csr, 1.05)
    col16.set_csr(csr, s)
    frm = m.basis_size // 2
    to = [frm, 0, m.basis_size - 1, frm + 1]
    M = 300
    ref = oracle_mod.moments(csr, s, frm, to, M)
    col16.set_option(X.OPT_COL16, 1)
    got16 = col16.moments(frm, to, M)
    assert col16.get_option(X.OPT_COL16) == 1
    col16.set_option(X.OPT_COL16, 0)
    got32 = col16.moments(frm, to, M)
    assert col16.get_option(X.OPT_COL16) == 0
    assert np.array_equal(got16, got32)
    assert np.array_equal(got16, ref)


def test_col16_falls_back_on_wide_offsets(col16, oracle_mod):
    """A 70000-site ring: the wrap-around bond puts column N-1 in the first
    row block (offset 69999 > 32767), so the model keeps 32-bit columns."""
    m = L.ring(70000)
    csr = m.csr()
    col16.set_csr(csr, 2.5)
    to = [0, 69999, 35000]
    got = col16.moments(0, to, 40)
    assert col16.get_option(X.OPT_COL16) == 0
    assert np.array_equal(got, oracle_mod.moments(csr, 2.5, 0, to, 40))


def test_col16_device_built_lattice(col16):
    """H built on the device from the lattice description gets its 16-bit
    columns on the device too; moments equal the host-CSR model's."""
    m = L.nambu_swave(16, 16, disorder=0.5, seed=3)
    spec = L.nambu_swave_spec(16, 16, disorder=0.5, seed=3)
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    to = [5, 100, 1023]
    col16.set_csr(m.csr(), s)
    ref = col16.moments(100, to, 200)
    col16.set_lattice(spec, s)
    got = col16.moments(100, to, 200)
    assert col16.get_option(X.OPT_COL16) == 1
    assert np.array_equal(got, ref)


def test_col16_graph_replay_switches_format(col16, oracle_mod):
    m = L.square_spin(33, 31, disorder=1.0, seed=4)
    csr = m.csr()
    col16.set_csr(csr, 11.0)
    col16.set_option(X.OPT_GRAPH, 1)
    ref = oracle_mod.moments(csr, 11.0, 77, [77, 3], 129)
    for use in (1, 0, 1, 1):
        col16.set_option(X.OPT_COL16, use)
        got = col16.moments(77, [77, 3], 129)
        assert col16.get_option(X.OPT_COL16) == use
        assert np.array_equal(got, ref), use


@pytest.mark.parametrize("name", sorted(MODELS))
def test_palette_bitwise_equal_to_full_values_and_oracle(col16, oracle_mod, name):
    m = MODELS[name]()
    csr = m.csr()
    s = kpm.gershgorin_scale(csr, 1.05)
    col16.set_csr(csr, s)
    frm = m.basis_size // 3
    to = [frm, 1, m.basis_size - 2]
    M = 257
    ref = oracle_mod.moments(csr, s, frm, to, M)
    out = {}
    for pal in (1, 0):
        col16.set_option(X.OPT_PALETTE, pal)
        out[pal] = col16.moments(frm, to, M)
        assert col16.get_option(X.OPT_PALETTE) == pal
        assert col16.get_option(X.OPT_COL16) == 1
    assert np.array_equal(out[1], out[0])
    assert np.array_equal(out[1], ref)


def random_chain(n, seed):
    """Open chain with a different random hopping on every bond (more than
    255 distinct off-diagonal values) and random on-site energies."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.5, 1.5, n - 1)
    eps = rng.uniform(-1, 1, n)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, t[i - 1] if i > 0 else None), (i, eps[i]),
                     (i + 1, t[i] if i < n - 1 else None)):
            if v is not None:
                rows.append(i)
                cols.append(j)
                vals.append(v)
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(rp, np.array(rows) + 1, 1)
    return (np.cumsum(rp).astype(np.uint32), np.array(cols, dtype=np.uint32),
            np.array(vals, dtype=np.complex128))


def test_palette_falls_back_on_many_values(col16, oracle_mod):
    csr = random_chain(600, 11)
    col16.set_option(X.OPT_PALETTE, 1)
    col16.set_csr(csr, 4.0)
    got = col16.moments(300, [300, 0, 599], 120)
    assert col16.get_option(X.OPT_COL16) == 1
    assert col16.get_option(X.OPT_PALETTE) == 0
    assert np.array_equal(got, oracle_mod.moments(csr, 4.0, 300, [300, 0, 599], 120))


def test_palette_device_built_lattice(col16):
    m = L.square_spin(24, 20, disorder=1.0, seed=6)
    spec = L.square_spin_spec(24, 20, disorder=1.0, seed=6)
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    col16.set_option(X.OPT_PALETTE, 1)
    col16.set_csr(m.csr(), s)
    ref = col16.moments(480, [480, 7], 300)
    col16.set_lattice(spec, s)
    got = col16.moments(480, [480, 7], 300)
    assert col16.get_option(X.OPT_PALETTE) == 1
    assert np.array_equal(got, ref)


def test_palette_duplicated_diagonal_keeps_full_values(col16, oracle_mod):
    """A raw CSR that lists a row's diagonal twice (unmerged entries, which
    the plain step adds one after the other) must not go through the
    palette, which holds one diagonal value per row."""
    n = 50
    rows, cols, vals = [], [], []
    for i in range(n):
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(-1.0)
        rows.append(i); cols.append(i); vals.append(0.25 * (i % 3))
        if i % 7 == 0:
            rows.append(i); cols.append(i); vals.append(0.5)
        if i < n - 1:
            rows.append(i); cols.append(i + 1); vals.append(-1.0)
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(rp, np.array(rows) + 1, 1)
    csr = (np.cumsum(rp).astype(np.uint32), np.array(cols, dtype=np.uint32),
           np.array(vals, dtype=np.complex128))
    col16.set_csr(csr, 3.0)
    ref = oracle_mod.moments(csr, 3.0, 21, [21, 0, 14], 90)
    out = {}
    for pal in (1, 0):
        col16.set_option(X.OPT_PALETTE, pal)
        out[pal] = col16.moments(21, [21, 0, 14], 90)
        assert col16.get_option(X.OPT_PALETTE) == 0
    assert np.array_equal(out[1], ref)
    assert np.array_equal(out[0], ref)
