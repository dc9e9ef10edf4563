"""Uneven multi-model batch (tbtkx_multi_set / tbtkx_multi_moments) on the GPU.

Checks, per model of a small uneven batch:
- gather mode is bitwise equal to the single-model batched path
  (tbtkx_moments_block, same recurrence arithmetic; mu_n is read from the
  vectors directly): both keep a complex H as FP32 values with FP32 row sums
  and a real H as one FP64 value per ELL entry with FP64 row sums, per model
  (the "complex" batch mixes both forms); FP32 vectors in both;
- doubling mode agrees with the single-model doubling path to FP64 rounding of
  the dot-product sums (1e-12 relative);
- both agree with the FP64 oracle (oracle/kpm_oracle.c, kpm_moments) within
  the FP32-storage tolerance of the batched path (2e-5 absolute on moments of
  unit start vectors, |mu_n| <= 1);
- autotuned, fixed-tile and hipGraph-replayed launches agree.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


def small_batch():
    """Models of uneven size, width and moment count, including a complex H
    (square lattice with flux), a model smaller than one row group, widths 1
    and 64, M = 1 and M = 2."""
    rng = np.random.default_rng(11)
    specs = [
        (L.square(9, 7, disorder=0.5, seed=1), 3, 41),
        (L.honeycomb_ribbon(4, 6, disorder=0.5, seed=2), 1, 120),
        (L.square(6, 5, disorder=0.5, seed=3, flux=0.13), 7, 77),
        (L.honeycomb_ribbon(12, 9, disorder=0.5, seed=4), 64, 64),
        (L.square(3, 2, disorder=0.5, seed=5), 2, 2),
        (L.square(30, 40, disorder=0.5, seed=6), 33, 150),
        (L.square(5, 5, disorder=0.5, seed=7), 4, 1),
        (L.honeycomb_ribbon(20, 30, disorder=0.5, seed=8), 17, 93),
    ]
    out = []
    for model, B, M in specs:
        sites = np.sort(rng.choice(model.basis_size, B, replace=False))
        out.append((model, sites, M))
    return out


def real_batch():
    return [b for b in small_batch() if not b[0].params.get("flux")]


def upload(engine, batch):
    csrs = [m.csr() for m, _, _ in batch]
    scales = [m.gershgorin_bound() * 1.05 for m, _, _ in batch]
    engine.multi_set(csrs, scales)
    return csrs, scales


def single_model(engine, csr, s, sites, M, flags):
    """The single-model batched path on one model of the batch."""
    engine.set_csr(csr, s)
    return engine.moments_block(sites, M, flags)


@pytest.mark.parametrize("which", ["complex", "real"])
def test_multi_gather_bitwise_vs_block(engine, which):
    batch = small_batch() if which == "complex" else real_batch()
    csrs, scales = upload(engine, batch)
    got = engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], 0)
    for (m, sites, M), csr, s, g in zip(batch, csrs, scales, got):
        ref = single_model(engine, csr, s, sites, M, 0)
        assert g.shape == ref.shape
        assert np.array_equal(g, ref), (m.name, m.basis_size, np.abs(g - ref).max())


@pytest.mark.parametrize("which", ["complex", "real"])
def test_multi_doubling_vs_block_and_oracle(engine, oracle_mod, which):
    batch = small_batch() if which == "complex" else real_batch()
    csrs, scales = upload(engine, batch)
    got = engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], X.DOUBLING)
    for (m, sites, M), csr, s, g in zip(batch, csrs, scales, got):
        ref = single_model(engine, csr, s, sites, M, X.DOUBLING)
        assert np.abs(g - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), m.name
        exact = np.array([oracle_mod.moments(csr, s, b, [b], M)[0] for b in sites])
        assert np.abs(g - exact).max() < 2e-5, (m.name, np.abs(g - exact).max())


def test_multi_gather_vs_oracle(engine, oracle_mod):
    batch = small_batch()
    csrs, scales = upload(engine, batch)
    got = engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], 0)
    for (m, sites, M), csr, s, g in zip(batch, csrs, scales, got):
        exact = np.array([oracle_mod.moments(csr, s, b, [b], M)[0] for b in sites])
        assert np.abs(g - exact).max() < 2e-5, (m.name, np.abs(g - exact).max())
        # mu_0 of a unit start vector is exactly 1
        assert np.all(g[:, 0] == 1.0)


@pytest.mark.parametrize("flags", [0, X.DOUBLING])
def test_multi_launch_shapes_and_graph_agree(engine, flags):
    batch = small_batch()
    upload(engine, batch)
    sites = [b[1] for b in batch]
    Ms = [b[2] for b in batch]
    base = engine.multi_moments(sites, Ms, flags)
    results = []
    try:
        # fixed tiles on a persistent grid (share 1, 2) and one tile per
        # wavefront (share 0)
        for tile, share in ((2, 1), (6, 2), (16, 1), (2, 0), (8, 0), (32, 0)):
            engine.set_option(X.OPT_TILE_ROWS, tile)
            engine.set_option(X.OPT_GRID_SHARE, share)
            results.append(engine.multi_moments(sites, Ms, flags))
            assert engine.get_option(X.OPT_GRID_SHARE) == share
        engine.set_option(X.OPT_TILE_ROWS, 0)
        engine.set_option(X.OPT_GRID_SHARE, 0)
        engine.set_option(X.OPT_GRAPH, 1)
        for _ in range(3):
            results.append(engine.multi_moments(sites, Ms, flags))
    finally:
        engine.set_option(X.OPT_TILE_ROWS, 0)
        engine.set_option(X.OPT_GRID_SHARE, 0)
        engine.set_option(X.OPT_GRAPH, 0)
    for r in results:
        for a, b in zip(r, base):
            if flags == 0:
                assert np.array_equal(a, b)
            else:
                assert np.abs(a - b).max() <= 1e-12


def test_multi_graph_replays_with_eager_work_between(engine):
    batch = small_batch()
    csrs, scales = upload(engine, batch)
    sites = [b[1] for b in batch]
    Ms = [b[2] for b in batch]
    eager = engine.multi_moments(sites, Ms, X.DOUBLING)
    engine.set_option(X.OPT_GRAPH, 1)
    try:
        for _ in range(3):
            g = engine.multi_moments(sites, Ms, X.DOUBLING)
            for a, b in zip(g, eager):
                assert np.abs(a - b).max() <= 1e-12
            # unrelated eager work on the same handle between replays
            engine.set_csr(csrs[0], scales[0])
            engine.moments_block(sites[0], 9, 0)
    finally:
        engine.set_option(X.OPT_GRAPH, 0)


@pytest.mark.parametrize("which", ["complex", "real"])
def test_multi_mixed_ell_widths(engine, oracle_mod, which):
    """One batch with ELL widths 4 (honeycomb ribbon), 5 (square lattice), 6
    (Nambu s-wave) and 8 (cubic lattice, 7 entries per row): each model runs
    on the path compiled for its own width (two-column lanes, one-column
    pipelined, per-lane entries), bitwise equal in gather mode to the
    single-model batched path, and within the FP32 tolerance of the FP64
    oracle in doubling mode. 'complex' adds a square lattice with flux, which
    makes the whole batch complex (no packing)."""
    models = [L.honeycomb_ribbon(11, 7, disorder=0.5, seed=1),
              L.square(13, 9, disorder=0.5, seed=2),
              L.nambu_swave(7, 6, disorder=0.5, seed=3),
              L.cubic(6, 5, 4, disorder=1.0, seed=4)]
    widths = [64, 5, 27, 9]
    Ms = [40, 57, 33, 64]
    if which == "complex":
        models.append(L.square(6, 5, disorder=0.5, seed=5, flux=0.13))
        widths.append(3)
        Ms.append(21)
    rng = np.random.default_rng(8)
    batch = [(m, np.sort(rng.choice(m.basis_size, b, replace=False)), M)
             for m, b, M in zip(models, widths, Ms)]
    csrs, scales = upload(engine, batch)
    got = engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], 0)
    dbl = engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], X.DOUBLING)
    for (m, sites, M), csr, s, g, d in zip(batch, csrs, scales, got, dbl):
        ref = single_model(engine, csr, s, sites, M, 0)
        assert np.array_equal(g, ref), (m.name, np.abs(g - ref).max())
        exact = np.array([oracle_mod.moments(csr, s, b, [b], M)[0] for b in sites])
        assert np.abs(d - exact).max() < 2e-5, (m.name, np.abs(d - exact).max())


def test_multi_rejects_bad_arguments(engine):
    batch = small_batch()[:2]
    upload(engine, batch)
    with pytest.raises(X.EngineError, match="1..64 sites"):
        engine.multi_moments([np.arange(65), [0]], [10, 10], 0)
    with pytest.raises(X.EngineError, match="outside the basis"):
        engine.multi_moments([[0], [batch[1][0].basis_size]], [10, 10], 0)
    with pytest.raises(X.EngineError, match="larger than 0"):
        engine.multi_moments([[0], [0]], [10, 0], 0)
    with pytest.raises(X.EngineError, match="only TBTKX_DOUBLING"):
        engine.multi_moments([[0], [0]], [10, 10], X.H_FP64)
    dense = L.coo_to_csr(np.repeat(np.arange(10), 10), np.tile(np.arange(10), 10),
                         np.ones(100, np.complex128), 10)
    with pytest.raises(X.EngineError, match="more than 8"):
        engine.multi_set([dense], [20.0])
    with pytest.raises(X.EngineError, match="scale factor"):
        engine.multi_set([batch[0][0].csr()], [0.0])


def test_multi_rejects_vectors_past_4gib(engine):
    """The step addresses a model's vectors with 32-bit byte offsets, so a
    model whose stored vectors (rows x (columns + 1) x 8 B) reach 4 GiB is
    refused before anything runs: 8,410,000 states with a complex H (no
    packing) and 64 sites. The same model with 32 sites runs."""
    spec = L.square_spec(2900, 2900)
    spec.terms = [(d, to, frm, re, 0.25 if d[1] == 1 else (-0.25 if d[1] == -1 else im))
                  for d, to, frm, re, im in spec.terms]
    engine.multi_set_lattices([spec], [6.0])
    sites = np.arange(64, dtype=np.int64) * 131_000
    with pytest.raises(X.EngineError, match="4 GiB"):
        engine.multi_moments([sites], [4], X.DOUBLING)
    mu = engine.multi_moments([sites[:32]], [4], X.DOUBLING)
    assert mu[0].shape == (32, 4) and np.all(mu[0][:, 0] == 1.0)


@pytest.mark.parametrize("flags", [0, X.DOUBLING])
def test_multi_graph_chunks_agree_with_direct(engine, flags):
    """A run of more steps than one captured graph holds (1099 gather steps,
    550 with the doubling identities: 5 and 3 graphs of at most 256 steps,
    models finishing inside and across chunk boundaries) replays to the
    direct launches' moments, three times in a row."""
    rng = np.random.default_rng(23)
    specs = [(L.square(8, 9, disorder=0.5, seed=1), 3, 1100),
             (L.honeycomb_ribbon(5, 6, disorder=0.5, seed=2), 2, 513),
             (L.square(12, 7, disorder=0.5, seed=3), 5, 257),
             (L.square(4, 4, disorder=0.5, seed=4), 1, 40)]
    batch = [(m, np.sort(rng.choice(m.basis_size, B, replace=False)), M) for m, B, M in specs]
    upload(engine, batch)
    sites = [b[1] for b in batch]
    Ms = [b[2] for b in batch]
    direct = engine.multi_moments(sites, Ms, flags)
    engine.set_option(X.OPT_GRAPH, 1)
    try:
        for _ in range(3):
            got = engine.multi_moments(sites, Ms, flags)
            for a, b in zip(got, direct):
                if flags == 0:
                    assert np.array_equal(a, b)
                else:
                    assert np.abs(a - b).max() <= 1e-12
    finally:
        engine.set_option(X.OPT_GRAPH, 0)


def test_multi_reports_launch_shape(engine):
    """With TBTKX_OPT_TILE_ROWS at 0 the option reads back the autotuned row
    groups per tile of the last call, and the grid share it ran with (0: one
    tile per wavefront); a fixed tile runs with the set grid share."""
    batch = small_batch()
    upload(engine, batch)
    engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], X.DOUBLING)
    tile, share = engine.get_option(X.OPT_TILE_ROWS), engine.get_option(X.OPT_GRID_SHARE)
    assert (share, tile) in [(s, g) for s in (1, 2) for g in (2, 4, 8, 16)] + \
        [(0, g) for g in (8, 16, 32, 64)]
    try:
        engine.set_option(X.OPT_TILE_ROWS, 6)
        engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], X.DOUBLING)
        assert engine.get_option(X.OPT_TILE_ROWS) == 6
        assert engine.get_option(X.OPT_GRID_SHARE) == 0
        engine.set_option(X.OPT_GRID_SHARE, 2)
        engine.multi_moments([b[1] for b in batch], [b[2] for b in batch], X.DOUBLING)
        assert engine.get_option(X.OPT_GRID_SHARE) == 2
    finally:
        engine.set_option(X.OPT_TILE_ROWS, 0)
        engine.set_option(X.OPT_GRID_SHARE, 0)


@pytest.mark.parametrize("path", ["uneven", "block"])
def test_real_high_moment_count_ldos_error(engine, oracle_mod, path):
    """A real honeycomb ribbon (t = 2.7, not exact in FP32) at 6000 moments:
    the LDOS (Jackson damping, the FP64 contraction) of the uneven batch and
    of the single-model batched step is within 1e-5 of the oracle's relative
    to its maximum, in gather and doubling mode. Both keep a real H in FP64
    (FP32 rounding of H shifts the eigenvalues by ~1e-7 of the band, which
    at this resolution moves the LDOS by ~1e-5)."""
    m = L.honeycomb_ribbon(30, 40, t=2.7, disorder=0.5, seed=31)
    csr = m.csr()
    s = m.gershgorin_bound() * 1.05
    sites = np.array([7, m.basis_size // 2])
    M, E = 6000, 2000
    if path == "uneven":
        engine.multi_set([csr], [s])
    else:
        engine.set_csr(csr, s)
    for flags in (0, X.DOUBLING):
        if path == "uneven":
            got = engine.multi_moments([sites], [M], flags)[0]
        else:
            got = engine.moments_block(sites, M, flags)
        engine.lut_generate(M, E, -0.99 * s, 0.99 * s, s)
        engine.set_damping(X.DAMP_JACKSON)
        try:
            for k, b in enumerate(sites):
                exact = oracle_mod.moments(csr, s, int(b), [int(b)], M)[0]
                l_ref = kpm.ldos_from_gf(engine.contract(exact, X.NONPRINCIPAL))
                l_got = kpm.ldos_from_gf(engine.contract(got[k], X.NONPRINCIPAL))
                err = np.abs(l_got - l_ref).max() / np.abs(l_ref).max()
                print(f"{path} flags={flags} site={b}: LDOS error {err:.2e}", flush=True)
                assert err <= 1e-5, (path, flags, int(b), err)
        finally:
            engine.set_damping(X.DAMP_NONE)
            engine.lut_destroy()
