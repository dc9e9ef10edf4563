"""Direct oracle pins of the headline kernels at their real shapes.

* The plain FP64 step (`plain_step_lds` with 16-bit column offsets and the
  value palette, the forms the 1024x1024 spin lattice runs in) on a 16x1024
  strip of the same lattice: its column offsets from the first row of a
  256-row block span -2048 ... 2303, the full range of the 1024x1024 model,
  and the moments must equal the oracle's bit for bit
  (Lib/src/Solver/ChebyshevExpander.cpp:65-82, 157-186).
* The full 1024x1024 spin lattice (2,097,152 states) at 32 moments, bitwise
  against the oracle.
* The split-BF16 MFMA contraction, all four types, against the oracle's FP64
  lookup contraction (Lib/src/Solver/ChebyshevExpander.cpp:324-370) on the
  oracle's table (:201-270), on 96 coefficient vectors of a real model:
  within 3e-5 of max|G| (about 16 significant bits per operand).
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def plain(engine):
    yield engine
    for opt, val in ((X.OPT_COL16, 1), (X.OPT_PALETTE, 1), (X.OPT_GRAPH, 0),
                     (X.OPT_PLAIN_KERNEL, 2)):
        engine.set_option(opt, val)


def block_offsets(csr, rows=256, cap=1536):
    """Column offsets from the first row of each plain-kernel row block
    (blocks of at most `rows` rows and `cap` non-zeros), as min and max."""
    rp = np.asarray(csr[0], dtype=np.int64)
    col = np.asarray(csr[1], dtype=np.int64)
    N = rp.size - 1
    lo, hi, r = 0, 0, 0
    while r < N:
        start, nnz = r, 0
        while r < N and r - start < rows and nnz + rp[r + 1] - rp[r] <= cap:
            nnz += rp[r + 1] - rp[r]
            r += 1
        if r == start:
            r += 1
        d = col[rp[start]:rp[r]] - start
        if d.size:
            lo, hi = min(lo, int(d.min())), max(hi, int(d.max()))
    return lo, hi


def test_plain_step_full_offset_range_bitwise(plain, oracle_mod):
    m = L.square_spin(16, 1024, t=1.0, J=0.5, disorder=1.0, seed=2)
    csr = m.csr()
    lo, hi = block_offsets(csr)
    assert (lo, hi) == (-2048, 2303), (lo, hi)
    s, M = 12.0, 400
    centre = m.basis_index((8, 512, 0))
    to = [centre, centre + 1, 0, m.basis_size - 1, centre - 2048, centre + 2049]
    plain.set_csr(csr, s)
    ref = oracle_mod.moments(csr, s, centre, to, M)
    got = plain.moments(centre, to, M)
    assert plain.get_option(X.OPT_COL16) == 1
    assert plain.get_option(X.OPT_PALETTE) == 1
    assert plain.get_option(X.OPT_REAL) == 1
    assert np.array_equal(got, ref)
    # the 32-bit, full-value form of the same step gives the same bits
    plain.set_option(X.OPT_COL16, 0)
    assert np.array_equal(plain.moments(centre, to, M), ref)
    assert plain.get_option(X.OPT_COL16) == 0


def test_plain_full_size_bitwise(plain, oracle_mod):
    """The 1024x1024 spin lattice with disorder U[-1, 1] (seed 2), scale
    factor 12, from the centre site: 32 moments equal the oracle's."""
    m = L.config_plain_large()
    csr = m.csr()
    s, M = 12.0, 32
    frm = m.basis_index((512, 512, 0))
    to = [frm, m.basis_index((512, 512, 1)), m.basis_index((513, 512, 0)),
          m.basis_index((512, 513, 0)), m.basis_index((500, 520, 0))]
    plain.set_csr(csr, s)
    got = plain.moments(frm, to, M)
    assert plain.get_option(X.OPT_COL16) == 1
    assert plain.get_option(X.OPT_PALETTE) == 1
    ref = oracle_mod.moments(csr, s, frm, to, M)
    assert np.array_equal(got, ref)
    # the moments reach the far to-index only after 12 + 8 steps
    assert np.count_nonzero(ref[4]) > 0


def test_plain_complex_full_size_bitwise(plain, oracle_mod):
    """A complex H at full size: the 1024 x 1024 square lattice with a flux
    of 0.01 flux quanta per plaquette (Landau-gauge Peierls phases, so more
    than 255 distinct values and the palette off) and disorder U[-1, 1]
    (seed 7). 32 moments from the centre equal the oracle's bit for bit, in
    complex arithmetic on the slot-major step with full FP64 values, and on
    the row-block step."""
    m = L.square(1024, 1024, t=1.0, disorder=1.0, seed=7, flux=0.01)
    csr = m.csr()
    s, M = m.gershgorin_bound() * 1.05, 32
    frm = 512 * 1024 + 512
    to = [frm, frm + 1, frm + 1024, frm - 1, 0, m.basis_size - 1, frm + 5 * 1024 + 3]
    ref = oracle_mod.moments(csr, s, frm, to, M)
    plain.set_csr(csr, s)
    got = plain.moments(frm, to, M)
    assert plain.get_option(X.OPT_REAL) == 0
    assert plain.get_option(X.OPT_PALETTE) == 0
    assert plain.get_option(X.OPT_PLAIN_KERNEL) == 2
    assert np.array_equal(got, ref)
    plain.set_option(X.OPT_PLAIN_KERNEL, 0)
    assert np.array_equal(plain.moments(frm, to, M), ref)
    assert np.count_nonzero(ref[6]) > 0


@pytest.fixture(scope="module")
def chebyshev_coefficients():
    """96 Jackson-damped moment vectors (600 moments) of a Nambu model, from
    the oracle's FP64 recurrence: 48 diagonal and 48 off-diagonal."""
    from oracle import oracle as O
    m = L.nambu_swave(8, 7, disorder=0.5, seed=3)
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    M = 600
    rows = []
    for b in range(0, m.basis_size, m.basis_size // 24)[:24]:
        c = O.moments(m.csr(), s, b, [b, (b + 5) % m.basis_size, (b + 1) % m.basis_size,
                                      (b + 17) % m.basis_size], M)
        rows.extend(c)
    c = np.array(rows[:96]) * kpm.jackson(M)[None, :]
    return c, s


@pytest.mark.parametrize("table", ["device", "uploaded"])
@pytest.mark.parametrize("gtype", [X.RETARDED, X.ADVANCED, X.PRINCIPAL, X.NONPRINCIPAL])
def test_mfma_contraction_vs_oracle(engine, oracle_mod, chebyshev_coefficients, gtype, table):
    c, s = chebyshev_coefficients
    K, M = c.shape
    E, lo, hi = 1000, -0.99 * s, 0.99 * s
    T = oracle_mod.lookup_table(M, E, lo, hi, s)
    if table == "device":
        engine.lut_generate(M, E, lo, hi, s)
    else:
        engine.lut_upload(T)
    engine.coeffs_upload(c)
    got = engine.contract_mfma(gtype)
    ref = np.array([oracle_mod.gf_lookup(gtype, T, row) for row in c])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 3e-5, err
    engine.lut_destroy()


@pytest.mark.parametrize("gtype", [X.RETARDED, X.ADVANCED, X.PRINCIPAL, X.NONPRINCIPAL])
def test_direct_contraction_vs_oracle_bitwise(engine, oracle_mod, chebyshev_coefficients, gtype):
    """The direct evaluation without a table (Lib/src/Solver/
    ChebyshevExpander.cpp:371-432): factors from the host's libm, sums on
    the device in the reference's order, equal to the oracle's bit for bit
    for 40 vectors; then a second grid and a shorter moment count (the cached
    factors are replaced), and the first again."""
    c, s = chebyshev_coefficients
    c = c[:40]
    for (E, lo, hi, M) in ((700, -0.97 * s, 0.96 * s, c.shape[1]),
                           (333, -0.5 * s, 0.8 * s, 157),
                           (700, -0.97 * s, 0.96 * s, c.shape[1])):
        de = (hi - lo) / (E - 1)
        cm = np.ascontiguousarray(c[:, :M])
        got = engine.contract_direct(cm, gtype, E, lo, de, s)
        ref = np.array([oracle_mod.gf_direct(gtype, row, lo, hi, E, s) for row in cm])
        assert np.array_equal(got, ref), (E, M, np.abs(got - ref).max())


def test_batched_full_size_vs_oracle(engine, oracle_mod):
    """The 512x512 Nambu model at full size (1,048,576 states, seed 3), 3
    seeded sites, 128 moments: the batched FP32 path (gather and doubling;
    a real H kept in FP64) within 2e-5 of the oracle's moments (|mu| <= 1),
    and the exact FP64 blocks (tbtkx_moments_multi) equal to them bit for
    bit."""
    m = L.config_nambu()
    csr = m.csr()
    s = kpm.gershgorin_scale(csr, 1.05)
    M = 128
    rng = np.random.default_rng(3)
    sites = np.sort(rng.choice(m.basis_size, 3, replace=False))
    exact = np.array([oracle_mod.moments(csr, s, int(b), [int(b)], M)[0] for b in sites])
    engine.set_csr(csr, s)
    for flags in (0, X.DOUBLING):
        got = engine.moments_block(sites, M, flags)
        err = np.abs(got - exact).max()
        print(f"flags={flags}: max |mu - oracle| {err:.2e}", flush=True)
        assert err < 2e-5, (flags, err)
        assert engine.get_option(X.OPT_REAL) == 1
    multi = engine.moments_multi(sites, [[int(b)] for b in sites], M)
    for k in range(len(sites)):
        assert np.array_equal(multi[k][0], exact[k]), k


def test_trace_full_size_vs_oracle(engine, oracle_mod):
    """The 128^3 cubic lattice at full size (2,097,152 states, disorder
    U[-1, 1], seed 5), 2 random-phase vectors (seed 17), 64 moments: the
    stochastic trace in FP32 vectors (doubling, a real H in FP64) within
    3e-5 R N of the oracle's trace sums, and in FP64 within 1e-10 R N."""
    spec = L.config_cubic_spec(128)
    m = L.config_cubic(128)
    csr = m.csr()
    s = engine.lattice_gershgorin(spec) * 1.05
    N = m.basis_size
    R, seed, M = 2, 17, 64
    u = oracle_mod.uniform(seed, np.arange(R * N)).reshape(R, N)
    ref = oracle_mod.moments_trace(csr, s, np.exp(2j * np.pi * u), M)
    engine.set_lattice(spec, s)
    for flags, tol in ((X.DOUBLING, 3e-5), (X.VEC_FP64 | X.H_FP64, 1e-10)):
        got = engine.moments_trace(R, seed, M, flags)
        err = np.abs(got - ref).max() / (R * N)
        print(f"flags={flags}: max |mu - oracle| / (R N) {err:.2e}", flush=True)
        assert err < tol, (flags, err)


def test_uneven_full_size_vs_oracle(engine, oracle_mod):
    """An uneven batch at the configuration's largest model size: a
    1,000,000-state honeycomb ribbon (t = 2.7, 400 x 625 cells) beside a
    300 x 1000 square lattice, 2 sites each, 96 and 48 moments, doubling:
    within 2e-5 of the oracle's moments."""
    models = [L.honeycomb_ribbon(400, 625, t=2.7, disorder=0.5, seed=41),
              L.square(300, 1000, t=1.0, disorder=0.5, seed=42)]
    Ms = [96, 48]
    csrs = [m.csr() for m in models]
    scales = [m.gershgorin_bound() * 1.05 for m in models]
    rng = np.random.default_rng(43)
    sites = [np.sort(rng.choice(m.basis_size, 2, replace=False)) for m in models]
    engine.multi_set(csrs, scales)
    got = engine.multi_moments(sites, Ms, X.DOUBLING)
    assert models[0].basis_size == 1_000_000
    for m, csr, s, st, M, g in zip(models, csrs, scales, sites, Ms, got):
        exact = np.array([oracle_mod.moments(csr, s, int(b), [int(b)], M)[0] for b in st])
        err = np.abs(g - exact).max()
        print(f"{m.name} N={m.basis_size}: max |mu - oracle| {err:.2e}", flush=True)
        assert err < 2e-5, (m.name, err)


def test_mfma_full_shape_vs_oracle(engine, oracle_mod):
    """The contraction at the configuration's full moment and energy counts
    (M = 4096, E = 8192, s = 10), 8 coefficient vectors of single levels at
    seeded energies (Jackson-damped, random phases): the split-BF16 MFMA
    contraction within 3e-5 of max|G| of the oracle's lookup contraction,
    Retarded and NonPrincipal, with the device table and the oracle's table
    uploaded."""
    M, E, s = 4096, 8192, 10.0
    rng = np.random.default_rng(21)
    x = rng.uniform(-0.9, 0.9, 8)
    c = np.cos(np.outer(np.arccos(x), np.arange(M))) * kpm.jackson(M)[None, :]
    c = c * np.exp(1j * rng.uniform(0, 2 * np.pi, (8, 1)))
    lo, hi = -0.99 * s, 0.99 * s
    T = oracle_mod.lookup_table(M, E, lo, hi, s)
    for table in ("device", "uploaded"):
        if table == "device":
            engine.lut_generate(M, E, lo, hi, s)
        else:
            engine.lut_upload(T)
        engine.coeffs_upload(c)
        for gtype in (X.RETARDED, X.NONPRINCIPAL):
            got = engine.contract_mfma(gtype)
            ref = np.array([oracle_mod.gf_lookup(gtype, T, row) for row in c])
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"{table} type {gtype}: {err:.2e}", flush=True)
            assert err < 3e-5, (table, gtype, err)
        engine.lut_destroy()
