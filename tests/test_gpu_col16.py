"""GPU tests of the plain FP64 step with 16-bit column offsets
(TBTKX_OPT_COL16): the step reads each column as an offset from the first
row of its row block when every offset of the model fits in int16, else the
32-bit columns. Same products in the same order either way, so the moments
are bitwise equal to the 32-bit path's and to the oracle's (the reference's
CPU recurrence, Lib/src/Solver/ChebyshevExpander.cpp:65-82, 157-186)."""
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
