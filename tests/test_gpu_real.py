"""Real arithmetic for a real Hamiltonian (TBTKX_OPT_REAL) on the GPU.

With a real H and unit start vectors the complex recurrence keeps every
imaginary part exactly zero, and its real parts come out of the same
operations in the same order as a real recurrence. The engine therefore runs
- the plain FP64 path on real vectors (12 Z + 28 N bytes per step instead of
  20 Z + 52 N), and
- the batched paths (blocks and the uneven batch) with two real vectors per
  complex column, one in the real and one in the imaginary part (half the
  vector bytes).
Checks: moments of the real runs are bitwise equal to those of the complex
runs (TBTKX_OPT_REAL = 0) wherever they are read from the vectors (plain
path, gather mode), including odd vector counts and block boundaries;
doubling-mode moments (the same vectors, FP64 dot sums grouped by a
different launch shape) agree to 1e-12 relative; the option reports what
the last call ran; a complex H and random-phase trace vectors run complex.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import lattice as L
from test_gpu_multi import real_batch, upload

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore(engine):
    yield
    for opt, val in ((X.OPT_REAL, 1), (X.OPT_BLOCK_WIDTH, 0), (X.OPT_PLAIN_KERNEL, 2),
                     (X.OPT_GRAPH, 0)):
        engine.set_option(opt, val)


def real_and_complex(engine, run):
    """run() with real arithmetic allowed, then forced complex; returns both
    results and what each call reported."""
    engine.set_option(X.OPT_REAL, 1)
    a = run()
    ra = engine.get_option(X.OPT_REAL)
    engine.set_option(X.OPT_REAL, 0)
    b = run()
    rb = engine.get_option(X.OPT_REAL)
    engine.set_option(X.OPT_REAL, 1)
    return a, b, ra, rb


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("graph", [0, 1])
def test_plain_real_bitwise(engine, restore, kernel, graph):
    m = L.square_spin(12, 9, disorder=1.0, seed=2)
    engine.set_csr(m.csr(), m.gershgorin_bound() * 1.05)
    engine.set_option(X.OPT_PLAIN_KERNEL, kernel)
    engine.set_option(X.OPT_GRAPH, graph)
    to = [0, 5, 77, m.basis_size - 1]
    a, b, ra, rb = real_and_complex(engine, lambda: engine.moments(5, to, 200))
    assert (ra, rb) == (1, 0)
    assert np.array_equal(a, b)
    assert np.all(a.imag == 0)


def test_complex_h_and_trace_run_complex(engine, restore):
    m = L.square(6, 5, disorder=0.5, seed=3, flux=0.13)
    engine.set_csr(m.csr(), m.gershgorin_bound() * 1.05)
    engine.moments(3, [3], 50)
    assert engine.get_option(X.OPT_REAL) == 0
    engine.moments_block(np.arange(10), 50, 0)
    assert engine.get_option(X.OPT_REAL) == 0
    m = L.square(6, 5, disorder=0.5, seed=3)
    engine.set_csr(m.csr(), m.gershgorin_bound() * 1.05)
    engine.moments_trace(8, 3, 40, 0)
    assert engine.get_option(X.OPT_REAL) == 0
    engine.moments_block(np.arange(10), 50, 0)
    assert engine.get_option(X.OPT_REAL) == 1


@pytest.mark.parametrize("flags", [0, X.H_FP64, X.VEC_FP64 | X.H_FP64])
@pytest.mark.parametrize("nsites,width", [(1, 64), (131, 64), (256, 128), (255, 0)])
def test_block_packed_gather_bitwise(engine, restore, flags, nsites, width):
    """131 sites at width 64 are one full block of 128 packed vectors and one
    block holding 3 (the last column has no imaginary-part vector)."""
    m = L.nambu_swave(10, 9, disorder=0.5, seed=3)
    engine.set_csr(m.csr(), m.gershgorin_bound() * 1.05)
    rng = np.random.default_rng(nsites)
    sites = rng.choice(m.basis_size, nsites, replace=False)
    engine.set_option(X.OPT_BLOCK_WIDTH, width)
    a, b, ra, rb = real_and_complex(engine, lambda: engine.moments_block(sites, 97, flags))
    assert (ra, rb) == (1, 0)
    assert np.array_equal(a, b)
    assert np.all(a.imag == 0)


@pytest.mark.parametrize("flags", [X.DOUBLING, X.VEC_FP64 | X.H_FP64 | X.DOUBLING])
def test_block_packed_doubling(engine, restore, flags):
    m = L.honeycomb_ribbon(9, 13, disorder=0.5, seed=4)
    engine.set_csr(m.csr(), m.gershgorin_bound() * 1.05)
    sites = np.random.default_rng(1).choice(m.basis_size, 131, replace=False)
    engine.set_option(X.OPT_BLOCK_WIDTH, 64)
    a, b, ra, rb = real_and_complex(engine, lambda: engine.moments_block(sites, 300, flags))
    assert (ra, rb) == (1, 0)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    assert np.all(a.imag == 0)


@pytest.mark.parametrize("flags", [0, X.DOUBLING])
def test_multi_packed(engine, restore, flags):
    batch = real_batch()
    upload(engine, batch)
    run = lambda: engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], flags)
    a, b, ra, rb = real_and_complex(engine, run)
    assert (ra, rb) == (1, 0)
    for (m, sites, M), x, y in zip(batch, a, b):
        assert x.shape == y.shape == (len(sites), M)
        if flags:
            assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max()), m.name
        else:
            assert np.array_equal(x, y), m.name
        assert np.all(x.imag == 0)
