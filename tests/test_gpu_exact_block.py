"""Exact FP64 blocks (tbtkx_moments_multi): many start vectors through one
SpMM per step, each column in the reference's operation order, so every
from-index's moments equal the oracle's plain recurrence
(Lib/src/Solver/ChebyshevExpander.cpp:65-82, 103-199) bit for bit; the same
call with plain recurrences gives the same bits. Models: real H (two real
vectors per block column), complex H (a square lattice with a Peierls
flux), a Nambu superconductor and a honeycomb ribbon; block widths 64, 128
and 256; to-lists with repeated and absent indices.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def multi(engine):
    yield engine
    engine.set_option(X.OPT_MULTI, 0)


def _queries(N, K, seed):
    """K from-indices and to-lists (lengths 0-4, a repeated entry, the
    from-index itself in most lists)."""
    rng = np.random.default_rng(seed)
    froms = rng.choice(N, size=K, replace=K > N)
    tos = []
    for j, f in enumerate(froms):
        n = int(rng.integers(0, 4))
        t = list(rng.integers(0, N, size=n))
        if j % 3 != 2:
            t.insert(0, int(f))
        if j % 7 == 0 and t:
            t.append(t[0])   # repeated to-index: only its last position is filled
        tos.append(t)
    return froms, tos


CASES = [
    ("square_spin", lambda: L.square_spin(15, 12, t=1.0, J=0.5, disorder=1.0, seed=2), 8.0, 40),
    ("square_flux", lambda: L.square(14, 13, t=1.0, disorder=0.5, seed=6, flux=0.05), 6.0, 300),
    ("nambu", lambda: L.nambu_swave(9, 8, disorder=0.5, seed=3), 7.0, 512),
    ("honeycomb", lambda: L.honeycomb_ribbon(6, 9, t=2.7, disorder=0.5, seed=4), 9.5, 97),
]


@pytest.mark.parametrize("name,make,s,K", CASES, ids=[c[0] for c in CASES])
def test_exact_blocks_bitwise_vs_oracle(multi, oracle_mod, name, make, s, K):
    m = make()
    csr = m.csr()
    M = 160
    froms, tos = _queries(m.basis_size, K, seed=K)
    multi.set_csr(csr, s)
    multi.set_option(X.OPT_MULTI, 2)
    print(f"{name}: N={m.basis_size} K={K} M={M}", flush=True)
    got = multi.moments_multi(froms, tos, M)
    assert multi.get_option(X.OPT_MULTI) == 2
    width = multi.get_option(X.OPT_BLOCK_WIDTH)
    multi.set_option(X.OPT_MULTI, 1)
    plain = multi.moments_multi(froms, tos, M)
    assert multi.get_option(X.OPT_MULTI) == 1
    for j, (f, t) in enumerate(zip(froms, tos)):
        if not t:
            continue
        ref = oracle_mod.moments(csr, s, int(f), [int(x) for x in t], M)
        assert np.array_equal(got[j], ref), (name, j, width)
        assert np.array_equal(plain[j], ref), (name, j)


def test_exact_block_widths(multi, oracle_mod):
    """512 diagonal queries of a real model fill one block of 256 packed
    columns; 40 fill a block of 64; both bitwise."""
    m = L.square_spin(16, 16, t=1.0, J=0.5, disorder=1.0, seed=9)
    csr, s, M = m.csr(), 7.0, 90
    multi.set_csr(csr, s)
    multi.set_option(X.OPT_MULTI, 2)
    for K, B in ((512, 256), (40, 64)):
        froms = np.arange(K) % m.basis_size
        got = multi.moments_multi(froms, [[int(f)] for f in froms], M)
        assert multi.get_option(X.OPT_BLOCK_WIDTH) == B
        for j in range(0, K, 37):
            ref = oracle_mod.moments(csr, s, int(froms[j]), [int(froms[j])], M)
            assert np.array_equal(got[j], ref), (K, j)


def test_auto_mode_picks_blocks_for_many_vectors(multi):
    """By bytes moved: one from-index runs as a plain recurrence, 300 of
    them as exact blocks."""
    m = L.square_spin(20, 20, t=1.0, J=0.5, disorder=1.0, seed=1)
    multi.set_csr(m.csr(), 8.0)
    multi.moments_multi([5], [[5]], 20)
    assert multi.get_option(X.OPT_MULTI) == 1
    multi.moments_multi(np.arange(300), [[i] for i in range(300)], 20)
    assert multi.get_option(X.OPT_MULTI) == 2
