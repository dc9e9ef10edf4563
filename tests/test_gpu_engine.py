"""GPU parity tests of the HIP engine through the C ABI (tbtk_amd.engine).

Plain FP64 path: bitwise equal to the reference's CPU recurrence (golden
vectors from the compiled reference, and the oracle). Contraction with an
uploaded table: bitwise equal to the reference. Device-generated table:
within 1e-12 of the host table. Batched FP32 path: within the stated FP32
tolerance of the FP64 oracle; batched FP64 path: within 1e-12.
"""
import numpy as np
import pytest

from conftest import load_fixture
from test_oracle import MOMENT_FIXTURES
from tbtk_amd import engine as X
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


def _fixture_case(name):
    fx = load_fixture(name)
    meta = fx["meta"]
    m = MOMENT_FIXTURES[name]()
    return fx, meta, m


@pytest.mark.parametrize("kernel", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(MOMENT_FIXTURES))
def test_plain_moments_bitexact(engine, oracle_mod, name, kernel):
    fx, meta, m = _fixture_case(name)
    engine.set_option(X.OPT_PLAIN_KERNEL, kernel)
    engine.set_csr(m.csr(), meta["scale"])
    frm = m.basis_index(meta["frm"])
    to = [m.basis_index(t) for t in meta["to"]]
    raw = engine.moments(frm, to, meta["M"])
    for b in meta["broadenings"]:
        got = oracle_mod.lorentz(raw, b) if b else raw
        assert np.array_equal(got, fx[f"coef_b{b:g}"]), (name, b, kernel)
    engine.set_option(X.OPT_PLAIN_KERNEL, 2)


@pytest.mark.parametrize("graph", [0, 1])
def test_plain_graph_replay(engine, oracle_mod, graph):
    m = L.square_spin(40, 37, disorder=1.0, seed=8)
    engine.set_csr(m.csr(), 11.0)
    engine.set_option(X.OPT_GRAPH, graph)
    try:
        ref = oracle_mod.moments(m.csr(), 11.0, 100, [100, 5, 2000], 257)
        for rep in range(3):   # replays of one captured graph
            got = engine.moments(100, [100, 5, 2000], 257)
            assert np.array_equal(got, ref), rep
        got = engine.moments(7, [7], 64)    # new shape: recapture
        assert np.array_equal(got, oracle_mod.moments(m.csr(), 11.0, 7, [7], 64))
    finally:
        engine.set_option(X.OPT_GRAPH, 0)


def test_plain_long_rows_and_edge_cases(engine, oracle_mod):
    # A star graph: row 0 couples to every site (longer than one LDS block).
    N = 5000
    rows = np.concatenate([np.zeros(N - 1, int), np.arange(1, N), np.arange(N)])
    cols = np.concatenate([np.arange(1, N), np.zeros(N - 1, int), np.arange(N)])
    vals = np.concatenate([np.full(N - 1, 0.01 + 0.02j), np.full(N - 1, 0.01 - 0.02j),
                           np.linspace(-1, 1, N)]).astype(np.complex128)
    csr = L.coo_to_csr(rows, cols, vals, N)
    s = 3.0
    engine.set_csr(csr, s)
    for kernel in (0, 1):
        engine.set_option(X.OPT_PLAIN_KERNEL, kernel)
        for M in (1, 2, 3, 50):
            got = engine.moments(0, [0, 9, 0, N - 1], M)
            ref = oracle_mod.moments(csr, s, 0, [0, 9, 0, N - 1], M)
            assert np.array_equal(got, ref), (kernel, M)
            assert np.all(got[0] == 0)       # repeated to-index: last wins
    engine.set_option(X.OPT_PLAIN_KERNEL, 2)


def test_plain_rejects_bad_arguments(engine):
    m = L.periodic_square(5)
    engine.set_csr(m.csr(), 10.0)
    with pytest.raises(X.EngineError):
        engine.moments(25, [0], 10)
    with pytest.raises(X.EngineError):
        engine.moments(0, [0], 0)
    with pytest.raises(X.EngineError):
        engine.set_csr(m.csr(), 0.0)


def test_contract_uploaded_table_bitexact(engine, oracle_mod):
    O = oracle_mod
    fx = load_fixture("square30_spin_ldos")
    g = fx["meta"]["gf"]
    m = L.config_plain_small()
    engine.set_csr(m.csr(), 10.0)
    lo, hi, E = g["window"]
    c = engine.moments(m.basis_index(g["frm"]), [m.basis_index(t) for t in g["to"]], g["M"])
    c = O.lorentz(c, 1e-6)
    engine.lut_upload(O.lookup_table(g["M"], E, lo, hi, 10.0))
    for t, tname in O.TYPE_NAMES.items():
        assert np.array_equal(engine.contract(c, t), fx[f"gf_{tname}"]), tname
    engine.lut_destroy()
    with pytest.raises(X.EngineError):
        engine.contract(c, X.RETARDED)


def test_ldos_end_to_end_bitexact(engine, oracle_mod):
    O = oracle_mod
    fx = load_fixture("square30_spin_ldos")
    lo, hi, E = fx["meta"]["window"]
    m = L.config_plain_small()
    engine.set_csr(m.csr(), 10.0)
    engine.lut_upload(O.lookup_table(1000, E, lo, hi, 10.0))
    ldos = np.zeros(E)
    for s in (0, 1):
        b = m.basis_index((15, 15, s))
        c = O.lorentz(engine.moments(b, [b], 1000), 1e-6)
        ldos += engine.contract(c, X.NONPRINCIPAL)[0].imag / np.pi
    assert np.array_equal(ldos, fx["ldos_lookup1"])


def test_lut_generate_matches_host_table(engine, oracle_mod):
    M, E, lo, hi, s = 700, 333, -6.1, 7.3, 8.0
    engine.lut_generate(M, E, lo, hi, s)
    rng = np.random.default_rng(3)
    c = rng.normal(size=(3, M)) + 1j * rng.normal(size=(3, M))
    table = oracle_mod.lookup_table(M, E, lo, hi, s)
    for t in range(4):
        got = engine.contract(c, t)
        ref = np.array([oracle_mod.gf_lookup(t, table, row) for row in c])
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), t
    engine.lut_destroy()


BLOCK_MODELS = {
    "nambu": (lambda: L.nambu_swave(10, 9, disorder=0.5, seed=3), 1.05),
    "honeycomb": (lambda: L.honeycomb_ribbon(7, 11, disorder=0.5, seed=4), 1.05),
    "square_spin": (lambda: L.square_spin(20, 23, disorder=1.0, seed=2), 1.02),
}


@pytest.mark.parametrize("name", sorted(BLOCK_MODELS))
@pytest.mark.parametrize("flags", [0, X.DOUBLING, X.H_FP64 | X.DOUBLING,
                                   X.VEC_FP64 | X.H_FP64, X.VEC_FP64 | X.H_FP64 | X.DOUBLING])
def test_block_moments_vs_oracle(engine, oracle_mod, name, flags):
    make, margin = BLOCK_MODELS[name]
    m = make()
    s = m.gershgorin_bound() * margin
    engine.set_csr(m.csr(), s)
    rng = np.random.default_rng(5)
    sites = rng.choice(m.basis_size, 70, replace=False)   # 2 blocks of 64
    M = 301
    got = engine.moments_block(sites, M, flags)
    ref = np.array([oracle_mod.moments(m.csr(), s, b, [b], M)[0] for b in sites])
    err = np.abs(got - ref).max()
    tol = 1e-12 if flags & X.VEC_FP64 else 2e-5
    assert err < tol, (name, flags, err)


@pytest.mark.parametrize("width", [64, 128, 256])
def test_block_widths_agree(engine, width):
    m = L.nambu_swave(8, 8, disorder=0.5, seed=1)
    s = m.gershgorin_bound() * 1.05
    engine.set_csr(m.csr(), s)
    sites = np.arange(0, 256, 3)
    engine.set_option(X.OPT_BLOCK_WIDTH, width)
    try:
        got = engine.moments_block(sites, 120, X.VEC_FP64 | X.H_FP64)
    finally:
        engine.set_option(X.OPT_BLOCK_WIDTH, 0)
    engine.set_option(X.OPT_BLOCK_WIDTH, 64)
    ref = engine.moments_block(sites, 120, X.VEC_FP64 | X.H_FP64)
    engine.set_option(X.OPT_BLOCK_WIDTH, 0)
    assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("flags,tol", [(X.VEC_FP64 | X.H_FP64, 1e-10), (0, 3e-5)])
def test_trace_vs_oracle(engine, oracle_mod, flags, tol):
    m = L.cubic(6, 5, 4, disorder=1.0, seed=5)
    N = m.basis_size
    s = m.gershgorin_bound() * 1.05
    engine.set_csr(m.csr(), s)
    R, seed, M = 70, 1234, 200
    u = oracle_mod.uniform(seed, np.arange(R * N)).reshape(R, N)
    vecs = np.exp(2j * np.pi * u)
    ref = oracle_mod.moments_trace(m.csr(), s, vecs, M)
    got = engine.moments_trace(R, seed, M, flags)
    assert np.abs(got - ref).max() < tol * R * N, np.abs(got - ref).max()


def test_block_width_autotuner_and_graph(engine):
    """The width autotuner picks one of the kernel widths and gives the same
    moments as a fixed width (FP32 vectors are column-independent; only the
    FP64 dot-sum order differs); the block recurrence replayed as a hipGraph
    matches eager launches on every replay, with eager work in between."""
    m = L.nambu_swave(12, 11, disorder=0.5, seed=4)
    s = m.gershgorin_bound() * 1.05
    engine.set_csr(m.csr(), s)
    sites = np.random.default_rng(3).choice(m.basis_size, 150, replace=False)
    engine.set_option(X.OPT_BLOCK_WIDTH, 0)
    tuned = engine.moments_block(sites, 81, X.DOUBLING)
    width = engine.get_option(X.OPT_BLOCK_WIDTH)
    assert width in (64, 128, 256)
    engine.set_option(X.OPT_BLOCK_WIDTH, 64)
    fixed = engine.moments_block(sites, 81, X.DOUBLING)
    assert np.abs(tuned - fixed).max() < 1e-12
    gather = engine.moments_block(sites, 81, 0)
    engine.set_option(X.OPT_BLOCK_WIDTH, 0)
    engine.set_option(X.OPT_GRAPH, 1)
    try:
        for _ in range(3):
            g = engine.moments_block(sites, 81, X.DOUBLING)
            assert np.abs(g - tuned).max() < 1e-12
            g0 = engine.moments_block(sites, 81, 0)
            assert np.array_equal(g0, gather)
            engine.moments(int(sites[0]), [int(sites[0])], 7)   # eager work between replays
        # a larger call reallocates the block buffers; the cached graph of the
        # first call must not be replayed on the freed ones
        engine.set_option(X.OPT_GRAPH, 0)
        engine.moments_block(np.arange(400), 33, X.DOUBLING)
        engine.set_option(X.OPT_GRAPH, 1)
        g = engine.moments_block(sites, 81, X.DOUBLING)
        assert np.abs(g - tuned).max() < 1e-12
    finally:
        engine.set_option(X.OPT_GRAPH, 0)


@pytest.mark.parametrize("strip,share,rmul", [(-1, 1, 1), (7, 1, 1), (12, 2, 2), (40, 4, 1),
                                              (64, 1, 2), (10 ** 6, 2, 1)])
def test_block_strip_traversal_bitwise(engine, strip, share, rmul):
    """The L2 strip traversal only changes the order rows are visited in:
    gather-mode moments are bitwise equal to the natural order for strip
    widths that do not divide the band, exceed it, or leave partial layers."""
    m = L.nambu_swave(9, 13, disorder=0.5, seed=6)     # band 52, N = 468
    s = m.gershgorin_bound() * 1.05
    engine.set_csr(m.csr(), s)
    sites = np.random.default_rng(8).choice(m.basis_size, 90, replace=False)
    try:
        engine.set_option(X.OPT_BLOCK_WIDTH, 64)
        engine.set_option(X.OPT_STRIP_ROWS, -1)
        engine.set_option(X.OPT_GRID_SHARE, 1)
        engine.set_option(X.OPT_ROWS_IN_FLIGHT, 1)
        ref = engine.moments_block(sites, 47, 0)
        ref_d = engine.moments_block(sites, 47, X.DOUBLING)
        engine.set_option(X.OPT_STRIP_ROWS, strip)
        engine.set_option(X.OPT_GRID_SHARE, share)
        engine.set_option(X.OPT_ROWS_IN_FLIGHT, rmul)
        got = engine.moments_block(sites, 47, 0)
        got_d = engine.moments_block(sites, 47, X.DOUBLING)
        assert engine.get_option(X.OPT_STRIP_ROWS) == strip
    finally:
        engine.set_option(X.OPT_BLOCK_WIDTH, 0)
        engine.set_option(X.OPT_STRIP_ROWS, 0)
        engine.set_option(X.OPT_GRID_SHARE, 0)
        engine.set_option(X.OPT_ROWS_IN_FLIGHT, 0)
    assert np.array_equal(got, ref)
    assert np.abs(got_d - ref_d).max() < 1e-12
