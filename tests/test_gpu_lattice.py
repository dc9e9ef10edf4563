"""Lattice models built on the device (tbtkx_set_lattice,
tbtkx_multi_set_lattices) against the same models built on the host.

The LatticeSpec builders are pinned to the COO generators on the CPU
(tests/test_lattice_spec.py); here the device-built CSR (row pointer,
columns, H/s as raw 64-bit patterns) must equal the one tbtkx_set_csr makes
from the generator's host CSR, and every recurrence that reads the other
device layouts (FP32 copies, real copies, the diagonal-first ELL, the band
that sizes the L2 strips) must give bitwise equal moments: plain FP64, the
batched gather mode, the batched doubling mode at a fixed launch shape, and
the uneven batch.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu

CASES = {
    "square_spin": (lambda: L.square_spin(17, 13, disorder=1.0, seed=2),
                    lambda: L.square_spin_spec(17, 13, disorder=1.0, seed=2)),
    "nambu": (lambda: L.nambu_swave(12, 9, disorder=0.5, seed=3),
              lambda: L.nambu_swave_spec(12, 9, disorder=0.5, seed=3)),
    "honeycomb": (lambda: L.honeycomb_ribbon(9, 7, disorder=0.5, seed=4),
                  lambda: L.honeycomb_ribbon_spec(9, 7, disorder=0.5, seed=4)),
    "square": (lambda: L.square(23, 11, disorder=0.5, seed=1),
               lambda: L.square_spec(23, 11, disorder=0.5, seed=1)),
    "cubic": (lambda: L.cubic(7, 6, 5, disorder=1.0, seed=5),
              lambda: L.cubic_spec(7, 6, 5, disorder=1.0, seed=5)),
    "periodic2": (lambda: L.periodic_square(2), lambda: L.periodic_square_spec(2)),
    "periodic5": (lambda: L.periodic_square(5), lambda: L.periodic_square_spec(5)),
    "ring": (lambda: L.ring(10), lambda: L.ring_spec(10)),
    "one_cell": (lambda: L.square(1, 1), lambda: L.square_spec(1, 1)),
}


@pytest.fixture
def fixed_shape(engine):
    for opt, v in ((X.OPT_BLOCK_WIDTH, 64), (X.OPT_STRIP_ROWS, 16), (X.OPT_GRID_SHARE, 1),
                   (X.OPT_ROWS_IN_FLIGHT, 1)):
        engine.set_option(opt, v)
    yield
    for opt in (X.OPT_BLOCK_WIDTH, X.OPT_STRIP_ROWS, X.OPT_GRID_SHARE, X.OPT_ROWS_IN_FLIGHT):
        engine.set_option(opt, 0)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_csr_equals_host_csr(engine, name):
    make_model, make_spec = CASES[name]
    m, spec = make_model(), make_spec()
    s = m.gershgorin_bound() * 1.05 + 0.5
    engine.set_csr(m.csr(), s)
    host = engine.csr_copy()
    engine.set_lattice(spec, s)
    dev = engine.csr_copy()
    assert np.array_equal(host[0], dev[0])
    assert np.array_equal(host[1], dev[1])
    assert np.array_equal(bits(host[2]), bits(dev[2]))


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_model_moments_bitwise(engine, fixed_shape, name):
    make_model, make_spec = CASES[name]
    m, spec = make_model(), make_spec()
    s = m.gershgorin_bound() * 1.05 + 0.5
    rng = np.random.default_rng(7)
    sites = rng.choice(m.basis_size, min(m.basis_size, 37), replace=False)
    out = {}
    for how in ("host", "device"):
        if how == "host":
            engine.set_csr(m.csr(), s)
        else:
            engine.set_lattice(spec, s)
        plain = engine.moments(int(sites[0]), sites, 61)
        gather = engine.moments_block(sites, 61, 0)
        dbl = engine.moments_block(sites, 80, X.DOUBLING)
        out[how] = (plain, gather, dbl)
    for a, b in zip(out["host"], out["device"]):
        assert np.array_equal(bits(a), bits(b)), name


def test_device_uneven_batch_bitwise(engine):
    specs = [L.honeycomb_ribbon_spec(9, 13, disorder=0.5, seed=1),
             L.square_spec(31, 17, disorder=0.5, seed=2),
             L.honeycomb_ribbon_spec(3, 2, disorder=0.5, seed=3),
             L.square_spec(5, 40, disorder=0.5, seed=4)]
    models = [L.honeycomb_ribbon(9, 13, disorder=0.5, seed=1),
              L.square(31, 17, disorder=0.5, seed=2),
              L.honeycomb_ribbon(3, 2, disorder=0.5, seed=3),
              L.square(5, 40, disorder=0.5, seed=4)]
    rng = np.random.default_rng(3)
    sites = [np.sort(rng.choice(m.basis_size, b, replace=False))
             for m, b in zip(models, (5, 64, 3, 17))]
    Ms = [70, 33, 120, 51]
    scales = [m.gershgorin_bound() * 1.05 for m in models]
    engine.multi_set([m.csr() for m in models], scales)
    host = engine.multi_moments(sites, Ms, 0)
    engine.multi_set_lattices(specs, scales)
    dev = engine.multi_moments(sites, Ms, 0)
    for a, b in zip(host, dev):
        assert np.array_equal(bits(a), bits(b))


def test_complex_lattice_runs_complex(engine):
    spec = L.square_spec(6, 5)
    spec.terms = [(d, to, frm, re, 0.25 if d[1] == 1 else (-0.25 if d[1] == -1 else im))
                  for d, to, frm, re, im in spec.terms]
    engine.set_lattice(spec, 6.0)
    mu = engine.moments(3, [3, 4], 40)
    assert engine.get_option(X.OPT_REAL) == 0
    assert np.abs(mu.imag).max() > 0


def test_lattice_rejects_bad_input(engine):
    spec = L.square_spec(4, 4)
    bad = L.LatticeSpec("bad", (4, 4, 1), 1, (0, 0, 0), [((1, 0, 0), 1, 0, 1.0, 0.0)], None, (4, 4))
    with pytest.raises(X.EngineError, match="orbital out of range"):
        engine.set_lattice(bad, 4.0)
    many = L.LatticeSpec("many", (4, 4, 1), 1, (0, 0, 0),
                         [((k % 3 - 1, k // 3 - 1, 0), 0, 0, 1.0, 0.0) for k in range(9)] * 2,
                         np.zeros(16, np.complex128), (4, 4))
    with pytest.raises(X.EngineError, match="at most 16 entries"):
        engine.set_lattice(many, 4.0)
    with pytest.raises(X.EngineError, match="scale factor"):
        engine.set_lattice(spec, 0.0)
    wide = L.cubic_spec(3, 3, 3)
    wide.terms = wide.terms + [((1, 1, 0), 0, 0, 0.5, 0.0), ((-1, -1, 0), 0, 0, 0.5, 0.0)]
    with pytest.raises(X.EngineError, match="more than 8 non-zeros"):
        engine.multi_set_lattices([wide], [10.0])


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_gershgorin(engine, name):
    """Device max_r sum_c |H_rc| equals the host CSR row-sum bound (device
    hypot against numpy's abs: 1e-13 relative)."""
    make_model, make_spec = CASES[name]
    m, spec = make_model(), make_spec()
    g = engine.lattice_gershgorin(spec)
    assert abs(g - m.gershgorin_bound()) <= 1e-13 * m.gershgorin_bound()
