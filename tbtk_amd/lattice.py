"""Seeded synthetic tight-binding models for the Chebyshev path, built as CSR.

Every generator returns a :class:`LatticeModel` whose amplitudes are exactly
what the corresponding TBTK model-building loop would add with
``model << HoppingAmplitude(amp, to, from)`` (the ``+ HC`` partner written out
explicitly), and whose basis numbering is TBTK's: depth-first over subindex
values (``Lib/src/Core/HoppingAmplitudeTree.cpp:596-616``), which for the
rectangular index sets used here is plain lexicographic order of the Index.
The CSR form follows ``HoppingAmplitudeSet::getSparseMatrix``
(``Lib/include/Core/TBTK/HoppingAmplitudeSet.h:397-424``): row = basis(to),
column = basis(from), columns ascending within a row, duplicates summed
(``Lib/include/Utilities/TBTK/SparseMatrix.h:1309-1420``). The generators
never emit duplicate (to, from) pairs: each state gets one combined on-site
amplitude, so the CSR values do not depend on a summation order.

Model idioms followed:
  * square lattice with spin, Zeeman ``-J*2*(1/2 - s)``: reference ``README.md``
  * Nambu s-wave superconductor: ``Templates/UnderDevelopment/
    SelfConsistentSuperconductivityChebyshev/src/main.cpp:188-214`` and its
    ``DeltaCallback`` (:147-170)
  * honeycomb 4-site cell: ``Templates/HexagonalLattice/src/main.cpp:58-103``
Open boundary conditions throughout (the ``if(x+1 < SIZE_X)`` idiom).
Random on-site disorder is drawn with ``numpy.random.default_rng(seed)``, one
value per lattice site in lexicographic site order.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np


@dataclass
class LatticeModel:
    name: str
    shape: tuple            # extent of each subindex (rectangular index set)
    rows: np.ndarray        # int64 [nnz]  basis(to)
    cols: np.ndarray        # int64 [nnz]  basis(from)
    vals: np.ndarray        # complex128 [nnz]
    params: dict = field(default_factory=dict)
    _csr: tuple | None = None

    @property
    def nsub(self) -> int:
        return len(self.shape)

    @property
    def basis_size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def nnz(self) -> int:
        return int(self.rows.size)

    def basis_index(self, index) -> int:
        """TBTK basis number of a physical Index (lexicographic order)."""
        b = 0
        for v, n in zip(index, self.shape):
            if not 0 <= v < n:
                raise IndexError(f"index {index} outside {self.shape}")
            b = b * n + int(v)
        return b

    def physical_index(self, b: int) -> tuple:
        out = []
        for n in reversed(self.shape):
            out.append(b % n)
            b //= n
        return tuple(reversed(out))

    def csr(self):
        """(rowptr u32[N+1], col u32[Z], val c128[Z]); cached."""
        if self._csr is None:
            self._csr = coo_to_csr(self.rows, self.cols, self.vals, self.basis_size)
        return self._csr

    def csr_c64(self):
        rp, col, val = self.csr()
        return rp, col, val.astype(np.complex64)

    def gershgorin_bound(self) -> float:
        """max_i sum_j |H_ij| in FP64: an upper bound of the spectral radius."""
        rp, col, val = self.csr()
        a = np.abs(val)
        row_sum = np.add.reduceat(a, rp[:-1].astype(np.int64)) if a.size else a
        # reduceat on empty rows returns the next element; guard them
        empty = rp[1:] == rp[:-1]
        row_sum = np.where(empty, 0.0, row_sum)
        return float(row_sum.max())


def coo_to_csr(rows, cols, vals, n):
    """Sort by (row, col), merge duplicates by summation, build CSR."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.complex128)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    uniq, start = np.unique(key, return_index=True)
    if uniq.size != key.size:
        merged = np.add.reduceat(vals, start)
    else:
        merged = vals
    r = uniq // n
    c = uniq % n
    counts = np.bincount(r, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(counts, out=rowptr[1:])
    return rowptr, c.astype(np.uint32), np.ascontiguousarray(merged)


def _grid(*shape):
    return np.indices(shape).reshape(len(shape), -1)


def _bonds(shape, axis, sub_slice=None):
    """(to_multi, from_multi) pairs for +1 steps along `axis` (open BC)."""
    idx = _grid(*shape)
    keep = idx[axis] + 1 < shape[axis]
    frm = idx[:, keep]
    to = frm.copy()
    to[axis] += 1
    return to, frm


def _flat(multi, shape):
    return np.ravel_multi_index(tuple(multi), shape).astype(np.int64)


def _site_disorder(lx, ly, disorder, seed):
    """U[-disorder, disorder] per lattice site (lexicographic), or zeros."""
    if disorder:
        return np.random.default_rng(seed).uniform(-disorder, disorder, size=(lx, ly))
    return np.zeros((lx, ly))


def _uniform_onsite(n, disorder, seed):
    """U[-disorder, disorder] per state, or zeros, as complex amplitudes."""
    w = np.zeros(n)
    if disorder:
        w = np.random.default_rng(seed).uniform(-disorder, disorder, size=n)
    return w.astype(np.complex128)


def _square_spin_onsite(lx, ly, J, mu, disorder, seed):
    site_w = _site_disorder(lx, ly, disorder, seed)
    g = _grid(lx, ly, 2)
    s = g[2]
    diag = -mu + (-J * 2 * (0.5 - s)) + site_w[g[0], g[1]]
    return diag.astype(np.complex128)


def _nambu_onsite(lx, ly, mu, J, disorder, seed):
    site_w = _site_disorder(lx, ly, disorder, seed)
    g = _grid(lx, ly, 4)
    s = g[2]
    w = site_w[g[0], g[1]]
    se = np.where(s < 2, s, s - 2)
    zee = -J * 2 * (0.5 - se)
    diag = np.where(s < 2, -mu + zee + w, -(-mu + zee + w))
    return diag.astype(np.complex128)


def square_spin(lx, ly, t=1.0, J=0.5, mu=0.0, disorder=0.0, seed=0):
    """{x, y, s}: -t nearest-neighbour hopping, Zeeman -J*2*(1/2-s), -mu,
    on-site disorder U[-disorder, disorder] per site (same for both spins)."""
    shape = (lx, ly, 2)
    rows, cols, vals = [], [], []
    for axis in (0, 1):
        to, frm = _bonds(shape, axis)
        a, b = _flat(to, shape), _flat(frm, shape)
        amp = np.full(a.size, -t, dtype=np.complex128)
        rows += [a, b]; cols += [b, a]; vals += [amp, np.conj(amp)]
    d = np.arange(np.prod(shape), dtype=np.int64)
    rows.append(d); cols.append(d); vals.append(_square_spin_onsite(lx, ly, J, mu, disorder, seed))
    return LatticeModel(
        "square_spin", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals),
        dict(lx=lx, ly=ly, t=t, J=J, mu=mu, disorder=disorder, seed=seed))


def nambu_swave(lx, ly, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.0, seed=0):
    """{x, y, s}, s in 0..3: electron block s=0,1 (-t hopping, -mu, Zeeman
    -J*2*(1/2-s), disorder w), hole block s=2,3 (negated), and the s-wave
    pair potential of the SC template's DeltaCallback (to={x,y,3-s},
    from={x,y,s}: conj(D) for s=0, -conj(D) for s=1; + HC)."""
    shape = (lx, ly, 4)
    rows, cols, vals = [], [], []
    for axis in (0, 1):
        to, frm = _bonds(shape, axis)
        a, b = _flat(to, shape), _flat(frm, shape)
        orb = frm[2]
        amp = np.where(orb < 2, -t, t).astype(np.complex128)
        rows += [a, b]; cols += [b, a]; vals += [amp, np.conj(amp)]
    d = np.arange(np.prod(shape), dtype=np.int64)
    rows.append(d); cols.append(d); vals.append(_nambu_onsite(lx, ly, mu, J, disorder, seed))
    # pair potential
    site = _grid(lx, ly)
    for so in (0, 1):
        frm = np.vstack([site, np.full(site.shape[1], so)])
        to = np.vstack([site, np.full(site.shape[1], 3 - so)])
        a, b = _flat(to, shape), _flat(frm, shape)
        D = complex(delta)
        amp = np.full(a.size, np.conj(D) if so == 0 else -np.conj(D), dtype=np.complex128)
        rows += [a, b]; cols += [b, a]; vals += [amp, np.conj(amp)]
    return LatticeModel(
        "nambu_swave", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals),
        dict(lx=lx, ly=ly, t=t, mu=mu, delta=delta, J=J, disorder=disorder, seed=seed))


def honeycomb_ribbon(lx, ly, t=2.7, disorder=0.0, seed=0):
    """{x, y, n}, n in 0..3: the 4-site rectangular cell of the honeycomb
    lattice (chain 0-1-2-3 inside the cell, 3 -> next cell's 0 along x,
    1 -> next cell's 0 and 2 -> next cell's 3 along y), -t hopping, on-site
    disorder U[-disorder, disorder] per atom."""
    shape = (lx, ly, 4)
    rows, cols, vals = [], [], []
    cell = _grid(lx, ly)
    ncell = cell.shape[1]

    def add(to_m, frm_m):
        a, b = _flat(to_m, shape), _flat(frm_m, shape)
        amp = np.full(a.size, -t, dtype=np.complex128)
        rows.extend([a, b]); cols.extend([b, a]); vals.extend([amp, np.conj(amp)])

    for n0, n1 in ((0, 1), (1, 2), (2, 3)):
        add(np.vstack([cell, np.full(ncell, n1)]), np.vstack([cell, np.full(ncell, n0)]))
    kx = cell[0] + 1 < lx
    c = cell[:, kx]
    add(np.vstack([c[0] + 1, c[1], np.zeros(c.shape[1], int)]),
        np.vstack([c, np.full(c.shape[1], 3)]))
    ky = cell[1] + 1 < ly
    c = cell[:, ky]
    add(np.vstack([c[0], c[1] + 1, np.zeros(c.shape[1], int)]),
        np.vstack([c, np.full(c.shape[1], 1)]))
    add(np.vstack([c[0], c[1] + 1, np.full(c.shape[1], 3)]),
        np.vstack([c, np.full(c.shape[1], 2)]))
    n = lx * ly * 4
    d = np.arange(n, dtype=np.int64)
    rows.append(d); cols.append(d); vals.append(_uniform_onsite(n, disorder, seed))
    return LatticeModel(
        "honeycomb_ribbon", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals), dict(lx=lx, ly=ly, t=t, disorder=disorder, seed=seed))


def square(lx, ly, t=1.0, disorder=0.0, seed=0, flux=0.0):
    """{x, y}: -t nearest-neighbour hopping, disorder U[-disorder, disorder].
    flux: magnetic flux per plaquette in flux quanta (Landau gauge: the hop
    (x, y) -> (x, y + 1) carries exp(2 pi i flux x)), which makes H complex."""
    shape = (lx, ly)
    rows, cols, vals = [], [], []
    for axis in (0, 1):
        to, frm = _bonds(shape, axis)
        a, b = _flat(to, shape), _flat(frm, shape)
        amp = np.full(a.size, -t, dtype=np.complex128)
        if axis == 1 and flux:
            amp = amp * np.exp(2j * np.pi * flux * frm[0])
        rows += [a, b]; cols += [b, a]; vals += [amp, np.conj(amp)]
    n = lx * ly
    d = np.arange(n, dtype=np.int64)
    rows.append(d); cols.append(d); vals.append(_uniform_onsite(n, disorder, seed))
    return LatticeModel(
        "square", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals),
        dict(lx=lx, ly=ly, t=t, disorder=disorder, seed=seed, flux=flux))


def cubic(lx, ly, lz, t=1.0, disorder=0.0, seed=0):
    """{x, y, z}: -t nearest-neighbour hopping, disorder U[-disorder, disorder]."""
    shape = (lx, ly, lz)
    rows, cols, vals = [], [], []
    for axis in (0, 1, 2):
        to, frm = _bonds(shape, axis)
        a, b = _flat(to, shape), _flat(frm, shape)
        amp = np.full(a.size, -t, dtype=np.complex128)
        rows += [a, b]; cols += [b, a]; vals += [amp, np.conj(amp)]
    n = lx * ly * lz
    d = np.arange(n, dtype=np.int64)
    rows.append(d); cols.append(d); vals.append(_uniform_onsite(n, disorder, seed))
    return LatticeModel(
        "cubic", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals), dict(lx=lx, ly=ly, lz=lz, t=t, disorder=disorder, seed=seed))


def periodic_square(size, t=1.0, mu=-2.0):
    """The 5x5 periodic square lattice of the reference's known-answer test
    (``Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-200``): {x, y},
    -mu on site, -t to (x+1)%SIZE and (y+1)%SIZE, + HC. For SIZE=2 the
    periodic partner coincides with the open one and duplicates are summed."""
    shape = (size, size)
    rows, cols, vals = [], [], []
    g = _grid(size, size)
    frm = _flat(g, shape)
    rows.append(frm); cols.append(frm); vals.append(np.full(frm.size, -mu, dtype=np.complex128))
    for axis in (0, 1):
        to = g.copy()
        to[axis] = (to[axis] + 1) % size
        a = _flat(to, shape)
        amp = np.full(a.size, -t, dtype=np.complex128)
        rows += [a, frm]; cols += [frm, a]; vals += [amp, np.conj(amp)]
    return LatticeModel(
        "periodic_square", shape, np.concatenate(rows), np.concatenate(cols),
        np.concatenate(vals), dict(size=size, t=t, mu=mu))


def ring(size, t=-1.0):
    """The 10-site ring of the reference's LDOS sum-rule test
    (``Test/include/TBTK/Test/PropertyExtractor/ChebyshevExpander.h:19-25``):
    HoppingAmplitude(-1, {n}, {(n+1)%SIZE}) + HC."""
    shape = (size,)
    n = np.arange(size, dtype=np.int64)
    to = n
    frm = (n + 1) % size
    amp = np.full(size, t, dtype=np.complex128)
    return LatticeModel(
        "ring", shape, np.concatenate([to, frm]), np.concatenate([frm, to]),
        np.concatenate([amp, np.conj(amp)]), dict(size=size, t=t))


@dataclass
class LatticeSpec:
    """A lattice for the device generator (``tbtkx_set_lattice``): the unit
    cell's terms, the cell counts, boundary conditions and every state's
    on-site amplitude. Term ``((dx, dy, dz), to, frm, re, im)`` puts
    ``re + i im`` on H[(c + d, to), (c, frm)] for every cell c whose partner
    lies inside (or wraps along a periodic axis). The ``*_spec`` builders give
    exactly the CSR of the same-named generator: same amplitudes (the
    reverse bond is ``conj(amp)`` as there), same on-site array, and the
    summation order onsite-then-terms of any duplicate entries."""
    name: str
    cells: tuple            # (Lx, Ly, Lz)
    orbitals: int
    periodic: tuple         # (px, py, pz)
    terms: list             # [((dx, dy, dz), to, frm, re, im), ...]
    onsite: np.ndarray | None
    shape: tuple            # Index extents of the equivalent LatticeModel
    params: dict = field(default_factory=dict)

    @property
    def basis_size(self) -> int:
        return int(np.prod(self.cells)) * int(self.orbitals)

    def max_row_entries(self) -> int:
        """Entries of the fullest row (on-site plus the terms of one
        to-orbital); rows at open edges have fewer."""
        per = [0 if self.onsite is None else 1] * self.orbitals
        for _, to, _, _, _ in self.terms:
            per[to] += 1
        return max(per)

    def gershgorin_bound(self) -> float:
        """max over rows of sum |H_ij| (an upper bound of the spectral
        radius), from the terms without building the CSR; entries that
        coincide are counted separately, which only loosens the bound."""
        cell = np.indices(self.cells).reshape(3, -1).astype(np.int64)
        Ly, Lz, O = self.cells[1], self.cells[2], self.orbitals
        s = np.zeros(self.basis_size) if self.onsite is None else np.abs(self.onsite)
        for dvec, to, frm, re, im in self.terms:
            fc = cell - np.array(dvec, dtype=np.int64)[:, None]
            ok = np.ones(cell.shape[1], dtype=bool)
            for a in range(3):
                if not self.periodic[a]:
                    ok &= (fc[a] >= 0) & (fc[a] < self.cells[a])
            tc = cell[:, ok]
            s[((tc[0] * Ly + tc[1]) * Lz + tc[2]) * O + to] += abs(complex(re, im))
        return float(s.max())


def _bond(terms, d, to, frm, amp):
    """Append a bond and its hermitian partner, the latter as conj(amp)."""
    a = np.complex128(amp)
    b = np.conj(a)
    terms.append((tuple(d), to, frm, float(a.real), float(a.imag)))
    terms.append((tuple(-x for x in d), frm, to, float(b.real), float(b.imag)))


_AXES = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def square_spin_spec(lx, ly, t=1.0, J=0.5, mu=0.0, disorder=0.0, seed=0):
    terms = []
    for e in _AXES[:2]:
        for s in (0, 1):
            _bond(terms, e, s, s, -t)
    return LatticeSpec("square_spin", (lx, ly, 1), 2, (0, 0, 0), terms,
                       _square_spin_onsite(lx, ly, J, mu, disorder, seed), (lx, ly, 2),
                       dict(lx=lx, ly=ly, t=t, J=J, mu=mu, disorder=disorder, seed=seed))


def nambu_swave_spec(lx, ly, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.0, seed=0):
    terms = []
    for e in _AXES[:2]:
        for s in range(4):
            _bond(terms, e, s, s, -t if s < 2 else t)
    D = complex(delta)
    for so in (0, 1):
        _bond(terms, (0, 0, 0), 3 - so, so, np.conj(D) if so == 0 else -np.conj(D))
    return LatticeSpec("nambu_swave", (lx, ly, 1), 4, (0, 0, 0), terms,
                       _nambu_onsite(lx, ly, mu, J, disorder, seed), (lx, ly, 4),
                       dict(lx=lx, ly=ly, t=t, mu=mu, delta=delta, J=J, disorder=disorder,
                            seed=seed))


def honeycomb_ribbon_spec(lx, ly, t=2.7, disorder=0.0, seed=0):
    terms = []
    for n0, n1 in ((0, 1), (1, 2), (2, 3)):
        _bond(terms, (0, 0, 0), n1, n0, -t)
    _bond(terms, (1, 0, 0), 0, 3, -t)
    _bond(terms, (0, 1, 0), 0, 1, -t)
    _bond(terms, (0, 1, 0), 3, 2, -t)
    return LatticeSpec("honeycomb_ribbon", (lx, ly, 1), 4, (0, 0, 0), terms,
                       _uniform_onsite(lx * ly * 4, disorder, seed), (lx, ly, 4),
                       dict(lx=lx, ly=ly, t=t, disorder=disorder, seed=seed))


def square_spec(lx, ly, t=1.0, disorder=0.0, seed=0):
    terms = []
    for e in _AXES[:2]:
        _bond(terms, e, 0, 0, -t)
    return LatticeSpec("square", (lx, ly, 1), 1, (0, 0, 0), terms,
                       _uniform_onsite(lx * ly, disorder, seed), (lx, ly),
                       dict(lx=lx, ly=ly, t=t, disorder=disorder, seed=seed, flux=0.0))


def cubic_spec(lx, ly, lz, t=1.0, disorder=0.0, seed=0):
    terms = []
    for e in _AXES:
        _bond(terms, e, 0, 0, -t)
    return LatticeSpec("cubic", (lx, ly, lz), 1, (0, 0, 0), terms,
                       _uniform_onsite(lx * ly * lz, disorder, seed), (lx, ly, lz),
                       dict(lx=lx, ly=ly, lz=lz, t=t, disorder=disorder, seed=seed))


def periodic_square_spec(size, t=1.0, mu=-2.0):
    terms = []
    for e in _AXES[:2]:
        _bond(terms, e, 0, 0, -t)
    return LatticeSpec("periodic_square", (size, size, 1), 1, (1, 1, 0), terms,
                       np.full(size * size, -mu, dtype=np.complex128), (size, size),
                       dict(size=size, t=t, mu=mu))


def ring_spec(size, t=-1.0):
    # H[n][n+1] = t (to-cell n, from-cell n+1: d = -1), then its partner
    terms = []
    _bond(terms, (-1, 0, 0), 0, 0, t)
    return LatticeSpec("ring", (size, 1, 1), 1, (1, 0, 0), terms, None, (size,),
                       dict(size=size, t=t))


def write_amplitudes(model: LatticeModel, path: str) -> None:
    """Binary HoppingAmplitude list: b"TBHA", int32 version=1, int32 nsub,
    int64 count, then per amplitude int32 to[nsub], int32 from[nsub],
    float64 re, float64 im (little endian)."""
    nsub = model.nsub
    to = np.stack(np.unravel_index(model.rows, model.shape), axis=1).astype(np.int32)
    frm = np.stack(np.unravel_index(model.cols, model.shape), axis=1).astype(np.int32)
    rec = np.dtype([("to", np.int32, (nsub,)), ("frm", np.int32, (nsub,)),
                    ("re", np.float64), ("im", np.float64)])
    arr = np.empty(model.nnz, dtype=rec)
    arr["to"] = to
    arr["frm"] = frm
    arr["re"] = model.vals.real
    arr["im"] = model.vals.imag
    with open(path, "wb") as f:
        f.write(b"TBHA")
        f.write(struct.pack("<iiq", 1, nsub, model.nnz))
        arr.tofile(f)


# --------------------------------------------------------------------------
# The workloads of the Chebyshev path, sized as used by the bench and tests.
# --------------------------------------------------------------------------

def config_plain_small():
    """30x30 with spin, t=1, J=0.5, mu=0 (1,800 states)."""
    return square_spin(30, 30, t=1.0, J=0.5, mu=0.0)


def config_plain_large(seed=2):
    """1024x1024 with spin, t=1, J=0.5, disorder U[-1,1] (2,097,152 states)."""
    return square_spin(1024, 1024, t=1.0, J=0.5, mu=0.0, disorder=1.0, seed=seed)


def config_nambu(l=512, seed=3):
    """l x l Nambu s-wave, t=1, mu=-1, Delta=0.3, J=0.5, disorder U[-1/2,1/2]."""
    return nambu_swave(l, l, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.5, seed=seed)


def config_cubic(l=128, seed=5):
    """l^3 cubic lattice, t=1, disorder U[-1,1]."""
    return cubic(l, l, l, t=1.0, disorder=1.0, seed=seed)


def config_uneven(seed=4, count=64, max_states=1_000_000):
    """Uneven batch of `count` independent models drawn from `seed`: each a
    honeycomb ribbon (t=2.7, 4-site cell) or a square lattice (t=1), width
    20-400 and length 50-2000 (the length cut so that a model has at most
    `max_states` states), on-site disorder U[-1/2, 1/2], 500-8000 moments and
    1-64 LDOS sites drawn uniformly. Returns [(model, sites, M), ...]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        honey = bool(rng.integers(2))
        width = int(rng.integers(20, 401))
        length = int(rng.integers(50, 2001))
        per = 4 if honey else 1
        length = max(1, min(length, max_states // (per * width)))
        M = int(rng.integers(500, 8001))
        B = int(rng.integers(1, 65))
        mseed = int(rng.integers(1 << 31))
        if honey:
            model = honeycomb_ribbon(width, length, t=2.7, disorder=0.5, seed=mseed)
        else:
            model = square(width, length, t=1.0, disorder=0.5, seed=mseed)
        sites = np.sort(rng.choice(model.basis_size, B, replace=False))
        out.append((model, sites, M))
    return out


def config_plain_large_spec(seed=2):
    """The 1024x1024 spin lattice of config_plain_large as a LatticeSpec."""
    return square_spin_spec(1024, 1024, t=1.0, J=0.5, mu=0.0, disorder=1.0, seed=seed)


def config_nambu_spec(l=512, seed=3):
    return nambu_swave_spec(l, l, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.5, seed=seed)


def config_cubic_spec(l=128, seed=5):
    return cubic_spec(l, l, l, t=1.0, disorder=1.0, seed=seed)


def config_uneven_specs(seed=4, count=64, max_states=1_000_000):
    """config_uneven as LatticeSpecs (the same models, sites and moment
    counts from the same draws), for tbtkx_multi_set_lattices."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        honey = bool(rng.integers(2))
        width = int(rng.integers(20, 401))
        length = int(rng.integers(50, 2001))
        per = 4 if honey else 1
        length = max(1, min(length, max_states // (per * width)))
        M = int(rng.integers(500, 8001))
        B = int(rng.integers(1, 65))
        mseed = int(rng.integers(1 << 31))
        if honey:
            spec = honeycomb_ribbon_spec(width, length, t=2.7, disorder=0.5, seed=mseed)
        else:
            spec = square_spec(width, length, t=1.0, disorder=0.5, seed=mseed)
        sites = np.sort(rng.choice(spec.basis_size, B, replace=False))
        out.append((spec, sites, M))
    return out
