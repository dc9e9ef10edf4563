"""CPU checks of the synthetic workloads (no GPU): Hermiticity, sizes and
seeded determinism of the models the parity tests and the bench use."""
import numpy as np

from tbtk_amd import lattice as L


def dense(csr, n):
    rp, col, val = csr
    H = np.zeros((n, n), dtype=np.complex128)
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            H[r, col[k]] += val[k]
    return H


def test_square_flux_is_hermitian_and_gauge_invariant_spectrum():
    m0 = L.square(6, 5, disorder=0.0)
    m1 = L.square(6, 5, disorder=0.0, flux=0.25)
    H0 = dense(m0.csr(), m0.basis_size)
    H1 = dense(m1.csr(), m1.basis_size)
    assert np.allclose(H1, H1.conj().T)
    assert np.abs(H1.imag).max() > 0.5
    # integer flux per plaquette is a pure gauge: same spectrum as no flux
    m2 = L.square(6, 5, disorder=0.0, flux=1.0)
    H2 = dense(m2.csr(), m2.basis_size)
    assert np.allclose(np.linalg.eigvalsh(H2), np.linalg.eigvalsh(H0), atol=1e-12)
    # 1/4 flux changes it
    assert not np.allclose(np.linalg.eigvalsh(H1), np.linalg.eigvalsh(H0), atol=1e-3)


def test_uneven_batch_ranges_and_determinism():
    a = L.config_uneven(seed=4, count=6, max_states=20_000)
    b = L.config_uneven(seed=4, count=6, max_states=20_000)
    for (ma, sa, Ma), (mb, sb, Mb) in zip(a, b):
        assert ma.name == mb.name and ma.basis_size == mb.basis_size
        assert np.array_equal(sa, sb) and Ma == Mb
        assert np.array_equal(ma.csr()[2], mb.csr()[2])
        assert ma.name in ("honeycomb_ribbon", "square")
        assert 1 <= len(sa) <= 64 and len(np.unique(sa)) == len(sa)
        assert 500 <= Ma <= 8000
        assert ma.basis_size <= 20_000
        assert sa.max() < ma.basis_size
        rp = ma.csr()[0]
        assert np.diff(rp).max() <= 5
        lx, ly = ma.params["lx"], ma.params["ly"]
        assert 20 <= lx <= 400 and 1 <= ly <= 2000
