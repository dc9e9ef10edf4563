"""LatticeSpec builders against the COO generators (CPU).

tbtkx_set_lattice builds a row's entries as: the on-site amplitude, then
every term whose to-orbital is the row's orbital, at the column of the
partner cell (cell - d, wrapped along periodic axes, skipped outside open
ones), equal columns summed in that order. `spec_coo` below applies that rule
on the host; the resulting CSR must equal the same-named generator's CSR bit
for bit (values compared as raw 64-bit patterns, so signed zeros count): the
spec builders describe exactly the models the generators build, and the
device generator is then checked against these same CSRs on the GPU.
"""
import numpy as np
import pytest

from tbtk_amd import lattice as L


def spec_coo(spec):
    Lx, Ly, Lz = spec.cells
    O = spec.orbitals
    N = spec.basis_size
    rows, cols, vals = [], [], []
    if spec.onsite is not None:
        d = np.arange(N, dtype=np.int64)
        rows.append(d); cols.append(d); vals.append(np.asarray(spec.onsite, np.complex128))
    cell = np.indices((Lx, Ly, Lz)).reshape(3, -1).astype(np.int64)
    for dvec, to, frm, re, im in spec.terms:
        fc = cell - np.array(dvec, dtype=np.int64)[:, None]
        ok = np.ones(cell.shape[1], dtype=bool)
        for a in range(3):
            if spec.periodic[a]:
                fc[a] %= spec.cells[a]
            else:
                ok &= (fc[a] >= 0) & (fc[a] < spec.cells[a])
        tc, f = cell[:, ok], fc[:, ok]
        rows.append(((tc[0] * Ly + tc[1]) * Lz + tc[2]) * O + to)
        cols.append(((f[0] * Ly + f[1]) * Lz + f[2]) * O + frm)
        vals.append(np.full(tc.shape[1], complex(re, im), dtype=np.complex128))
    return L.coo_to_csr(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), N)


PAIRS = [
    (lambda: L.square_spin(7, 5, disorder=1.0, seed=2),
     lambda: L.square_spin_spec(7, 5, disorder=1.0, seed=2)),
    (lambda: L.square_spin(30, 30), lambda: L.square_spin_spec(30, 30)),
    (lambda: L.nambu_swave(6, 5, disorder=0.5, seed=3),
     lambda: L.nambu_swave_spec(6, 5, disorder=0.5, seed=3)),
    (lambda: L.honeycomb_ribbon(5, 4, disorder=0.5, seed=4),
     lambda: L.honeycomb_ribbon_spec(5, 4, disorder=0.5, seed=4)),
    (lambda: L.honeycomb_ribbon(3, 2), lambda: L.honeycomb_ribbon_spec(3, 2)),
    (lambda: L.square(6, 4, disorder=0.5, seed=1), lambda: L.square_spec(6, 4, disorder=0.5, seed=1)),
    (lambda: L.square(1, 1), lambda: L.square_spec(1, 1)),
    (lambda: L.cubic(4, 3, 5, disorder=1.0, seed=5), lambda: L.cubic_spec(4, 3, 5, disorder=1.0, seed=5)),
    (lambda: L.periodic_square(2), lambda: L.periodic_square_spec(2)),
    (lambda: L.periodic_square(5), lambda: L.periodic_square_spec(5)),
    (lambda: L.ring(10), lambda: L.ring_spec(10)),
    (lambda: L.ring(2), lambda: L.ring_spec(2)),
]


@pytest.mark.parametrize("i", range(len(PAIRS)))
def test_spec_matches_generator(i):
    make_model, make_spec = PAIRS[i]
    m, spec = make_model(), make_spec()
    assert spec.basis_size == m.basis_size
    rp, col, val = m.csr()
    srp, scol, sval = spec_coo(spec)
    assert np.array_equal(rp, srp)
    assert np.array_equal(col, scol)
    assert np.array_equal(val.view(np.uint64), sval.view(np.uint64))


def test_uneven_specs_match_uneven_models():
    a = L.config_uneven(seed=4, count=6, max_states=20_000)
    b = L.config_uneven_specs(seed=4, count=6, max_states=20_000)
    for (m, s, M), (spec, s2, M2) in zip(a, b):
        assert (m.basis_size, M) == (spec.basis_size, M2) and np.array_equal(s, s2)
        rp, col, val = m.csr()
        srp, scol, sval = spec_coo(spec)
        assert np.array_equal(rp, srp) and np.array_equal(col, scol)
        assert np.array_equal(val.view(np.uint64), sval.view(np.uint64))


# Models whose terms put two amplitudes on one (row, column): the two-cell
# periodic lattices, where the +1 and -1 neighbours coincide.
COINCIDING = {("periodic_square", 2), ("ring", 2)}


@pytest.mark.parametrize("i", range(len(PAIRS)))
def test_spec_gershgorin_and_row_length(i):
    """The spec's Gershgorin bound is the CSR row-sum bound (up to the order
    of the sum); where entries coincide it counts them separately, which
    only loosens it. Its fullest-row count bounds the CSR's longest row and
    equals it when no entries coincide and interior rows exist."""
    make_model, make_spec = PAIRS[i]
    m, spec = make_model(), make_spec()
    coinciding = (m.name, m.params.get("size")) in COINCIDING
    g_csr, g_spec = m.gershgorin_bound(), spec.gershgorin_bound()
    assert g_spec >= g_csr * (1 - 1e-14)
    if not coinciding:
        assert g_spec <= g_csr * (1 + 1e-14)
    longest = int(np.diff(m.csr()[0].astype(np.int64)).max())
    assert spec.max_row_entries() >= longest
    moved = {a for d, *_ in spec.terms for a in range(3) if d[a]}
    interior = all(spec.cells[a] >= 3 for a in moved if not spec.periodic[a])
    if interior and not coinciding:
        assert spec.max_row_entries() == longest
