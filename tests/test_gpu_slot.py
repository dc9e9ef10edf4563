"""GPU tests of the slot-major ELL form of the plain FP64 step
(TBTKX_OPT_PLAIN_KERNEL = 2): W slots per row, slot s of row r at s*N + r,
16-bit column offsets from the row, palette codes or full values. The slots
keep each row's CSR order, so the moments are bitwise equal to the oracle's
(the reference's CPU recurrence, Lib/src/Solver/ChebyshevExpander.cpp:
65-82, 157-186) and to the other kernels'. Models whose rows exceed 8
non-zeros or whose offsets exceed int16 fall back to the row-block kernel.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm
from tbtk_amd import lattice as L
from test_gpu_col16 import random_chain

pytestmark = pytest.mark.gpu


@pytest.fixture
def slot(engine):
    engine.set_option(X.OPT_PLAIN_KERNEL, 2)
    yield engine
    for opt, val in ((X.OPT_PLAIN_KERNEL, 2), (X.OPT_COL16, 1), (X.OPT_PALETTE, 1),
                     (X.OPT_GRAPH, 0), (X.OPT_REAL, 1), (X.OPT_ROW_PATTERNS, 1)):
        engine.set_option(opt, val)


MODELS = {
    "square_spin_40x37": (lambda: L.square_spin(40, 37, disorder=1.0, seed=8), 1),
    "nambu_12x9": (lambda: L.nambu_swave(12, 9, disorder=0.5, seed=3), 1),
    "square_flux_20x20": (lambda: L.square(20, 20, disorder=0.3, seed=2, flux=0.07), 1),
    "cubic_9": (lambda: L.cubic(9, 9, 9, disorder=1.0, seed=5), 1),
    "honeycomb_7x11": (lambda: L.honeycomb_ribbon(7, 11, t=2.7, disorder=0.5, seed=4), 1),
    "periodic_5": (lambda: L.periodic_square(5), 1),
}


@pytest.mark.parametrize("real", [1, 0])
@pytest.mark.parametrize("name", sorted(MODELS))
def test_slot_bitwise_vs_oracle_and_row_blocks(slot, oracle_mod, name, real):
    make, palette = MODELS[name]
    m = make()
    csr = m.csr()
    s = kpm.gershgorin_scale(csr, 1.05)
    slot.set_option(X.OPT_REAL, real)
    slot.set_csr(csr, s)
    frm = m.basis_size // 2
    to = [frm, 0, m.basis_size - 1, frm + 1, frm]
    M = 300
    ref = oracle_mod.moments(csr, s, frm, to, M)
    print(f"{name}: N={m.basis_size} real={real}", flush=True)
    for pal, pat in ((1, 1), (1, 2), (1, 0), (0, 1)):
        slot.set_option(X.OPT_PALETTE, pal)
        slot.set_option(X.OPT_ROW_PATTERNS, pat)
        got = slot.moments(frm, to, M)
        assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
        assert slot.get_option(X.OPT_PALETTE) == (pal and palette)
        assert slot.get_option(X.OPT_ROW_PATTERNS) == (pal and pat and palette)
        assert np.array_equal(got, ref), (name, pal, pat)
    slot.set_option(X.OPT_PLAIN_KERNEL, 0)
    assert np.array_equal(slot.moments(frm, to, M), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 0


def test_row_patterns_many_rows_fall_back(slot, oracle_mod):
    """A chain with 40 distinct hoppings in a random sequence over its 699
    bonds keeps the palette (40 values) but has more than 255 distinct rows
    (pairs of neighbouring hoppings), so it runs on the per-slot codes."""
    n = 700
    rng = np.random.default_rng(5)
    hop = rng.choice(np.linspace(-1.2, -0.8, 40), size=n - 1)
    rows, cols, vals = [], [], []
    for i in range(n):
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(hop[i - 1])
        rows.append(i); cols.append(i); vals.append(0.1 * (i % 3))
        if i < n - 1:
            rows.append(i); cols.append(i + 1); vals.append(hop[i])
    csr = L.coo_to_csr(np.array(rows), np.array(cols), np.array(vals, dtype=np.complex128), n)
    slot.set_csr(csr, 3.0)
    got = slot.moments(350, [350, 0, 699], 100)
    assert slot.get_option(X.OPT_PALETTE) == 1
    assert slot.get_option(X.OPT_ROW_PATTERNS) == 0
    assert np.array_equal(got, oracle_mod.moments(csr, 3.0, 350, [350, 0, 699], 100))


def test_slot_full_offset_range(slot, oracle_mod):
    """The 16x1024 strip of the 1024x1024 spin lattice (offsets from the row
    up to +-2048) with the palette, bitwise."""
    m = L.square_spin(16, 1024, t=1.0, J=0.5, disorder=1.0, seed=2)
    csr = m.csr()
    s, M = 12.0, 400
    centre = m.basis_index((8, 512, 0))
    to = [centre, centre + 1, 0, m.basis_size - 1, centre - 2048, centre + 2049]
    slot.set_csr(csr, s)
    got = slot.moments(centre, to, M)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
    assert slot.get_option(X.OPT_PALETTE) == 1
    assert slot.get_option(X.OPT_ROW_PATTERNS) == 1
    assert np.array_equal(got, oracle_mod.moments(csr, s, centre, to, M))


def test_slot_full_size(slot, oracle_mod):
    """The 1024x1024 spin lattice (2,097,152 states), 32 moments, bitwise."""
    m = L.config_plain_large()
    csr = m.csr()
    s, M = 12.0, 32
    frm = m.basis_index((512, 512, 0))
    to = [frm, m.basis_index((512, 512, 1)), m.basis_index((513, 512, 0)),
          m.basis_index((500, 520, 0))]
    slot.set_csr(csr, s)
    got = slot.moments(frm, to, M)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
    assert slot.get_option(X.OPT_ROW_PATTERNS) == 1
    assert np.array_equal(got, oracle_mod.moments(csr, s, frm, to, M))


def test_slot_many_values_full_values(slot, oracle_mod):
    """599 distinct hoppings: no palette, the slots carry full values."""
    csr = random_chain(600, 11)
    slot.set_csr(csr, 4.0)
    got = slot.moments(300, [300, 0, 599], 120)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
    assert slot.get_option(X.OPT_PALETTE) == 0
    assert np.array_equal(got, oracle_mod.moments(csr, 4.0, 300, [300, 0, 599], 120))


def test_slot_falls_back(slot, oracle_mod):
    """A 70000-site ring (wrap-around offset 69999) and a star graph (a row
    of 5000 non-zeros) run on the row-block kernel, bitwise."""
    ring = L.ring(70000).csr()
    slot.set_csr(ring, 2.5)
    got = slot.moments(0, [0, 1, 69999], 80)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 0
    assert np.array_equal(got, oracle_mod.moments(ring, 2.5, 0, [0, 1, 69999], 80))
    N = 5000
    rows = np.concatenate([np.zeros(N - 1, int), np.arange(1, N), np.arange(N)])
    cols = np.concatenate([np.arange(1, N), np.zeros(N - 1, int), np.arange(N)])
    vals = np.concatenate([np.full(N - 1, 0.01), np.full(N - 1, 0.01),
                           np.linspace(-1, 1, N)]).astype(np.complex128)
    star = L.coo_to_csr(rows, cols, vals, N)
    slot.set_csr(star, 3.0)
    got = slot.moments(0, [0, 9, N - 1], 50)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 0
    assert np.array_equal(got, oracle_mod.moments(star, 3.0, 0, [0, 9, N - 1], 50))


def test_slot_device_lattice_graph_and_noalloc(slot, oracle_mod):
    """A device-built lattice, graph replays and the allocation-free entry
    point on the slot form give the host model's oracle moments."""
    m = L.square_spin(24, 20, disorder=1.0, seed=6)
    spec = L.square_spin_spec(24, 20, disorder=1.0, seed=6)
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    ref = oracle_mod.moments(m.csr(), s, 480, [480, 7], 300)
    slot.set_lattice(spec, s)
    assert np.array_equal(slot.moments(480, [480, 7], 300), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
    slot.set_option(X.OPT_GRAPH, 1)
    for _ in range(3):
        assert np.array_equal(slot.moments(480, [480, 7], 300), ref)
    slot.set_option(X.OPT_GRAPH, 0)
    assert np.array_equal(slot.moments_noalloc(480, [480, 7], 300), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2


@pytest.mark.parametrize("rows", [1, 2, 4])
def test_row_patterns_uniform_wavefronts(slot, oracle_mod, rows):
    """Wavefronts whose 64-row chunks share one pattern read it through
    scalar loads (TBTKX_OPT_ROW_PATTERNS = 1); with 2 every row reads its
    pattern byte. Both are bitwise the oracle's, for 1, 2 and 4 rows per
    thread, on a lattice with long uniform runs (800 rows per column of the
    20x400 spin lattice), one whose runs are shorter than a chunk (cubic 17,
    4913 rows, not a multiple of 64) and a 1000-site open chain (every
    chunk uniform but the ends), and in a graph replay."""
    models = [L.square_spin(20, 400, disorder=1.0, seed=21),
              L.cubic(17, 17, 17, disorder=1.0, seed=22),
              L.square(1000, 1, disorder=0.5, seed=23)]
    try:
        slot.set_option(X.OPT_SLOT_ROWS, rows)
        for m in models:
            csr = m.csr()
            s = kpm.gershgorin_scale(csr, 1.05)
            slot.set_csr(csr, s)
            frm = m.basis_size // 2 + 1
            to = [frm, 0, m.basis_size - 1, 63, 64, frm - 64]
            M = 200
            ref = oracle_mod.moments(csr, s, frm, to, M)
            print(f"{m.name}: N={m.basis_size} rows={rows}", flush=True)
            for pat in (1, 2):
                slot.set_option(X.OPT_ROW_PATTERNS, pat)
                got = slot.moments(frm, to, M)
                assert slot.get_option(X.OPT_ROW_PATTERNS) == 1
                assert np.array_equal(got, ref), (m.name, rows, pat)
            if m.basis_size == 1000:
                # chunk 0 holds the end row 0; rows 960-999 fill no chunk
                assert slot.get_option(X.OPT_PATTERN_BYTE_ROWS) == 64 + 40
            slot.set_option(X.OPT_ROW_PATTERNS, 1)
            slot.set_option(X.OPT_GRAPH, 1)
            for _ in range(2):
                assert np.array_equal(slot.moments(frm, to, M), ref), (m.name, rows, "graph")
            slot.set_option(X.OPT_GRAPH, 0)
    finally:
        slot.set_option(X.OPT_SLOT_ROWS, 0)
        slot.set_option(X.OPT_GRAPH, 0)


def test_device_lattice_builds_plain_forms_on_first_plain_call(slot, oracle_mod):
    """A device-built model copies its CSR back for the palette, the row
    patterns and the slots only on its first plain recurrence: a batched
    call leaves them unbuilt, the allocation-free entry point runs the
    row-block form meanwhile, and every form gives the oracle's moments."""
    m = L.square_spin(6, 300, disorder=1.0, seed=9)      # columns of 600 rows
    spec = L.square_spin_spec(6, 300, disorder=1.0, seed=9)
    N = m.basis_size
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    to = [N // 2 + 1, 5, N - 1]
    ref = oracle_mod.moments(m.csr(), s, N // 2 + 1, to, 150)
    slot.set_lattice(spec, s)
    assert slot.get_option(X.OPT_PATTERN_BYTE_ROWS) == -1
    slot.moments_block([N // 2 + 1, 5], 20, 0)
    assert slot.get_option(X.OPT_PATTERN_BYTE_ROWS) == -1
    assert np.array_equal(slot.moments_noalloc(N // 2 + 1, to, 150), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 0
    assert np.array_equal(slot.moments(N // 2 + 1, to, 150), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
    assert slot.get_option(X.OPT_ROW_PATTERNS) == 1
    # most 64-row chunks lie inside one 600-row column: uniform
    assert 0 < slot.get_option(X.OPT_PATTERN_BYTE_ROWS) < N // 2
    assert np.array_equal(slot.moments_noalloc(N // 2 + 1, to, 150), ref)
    assert slot.get_option(X.OPT_PLAIN_KERNEL) == 2
