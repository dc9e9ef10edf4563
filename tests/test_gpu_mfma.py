"""GPU tests of the split-BF16 MFMA contraction against the FP64 contraction
on the same device table. Split-BF16 keeps about 16 significant bits per
operand (three-term products), so each product carries a relative error of
order 2^-16 ~ 1.5e-5; the tests bound the error of the contraction at
3e-5 of max|G| for coefficient vectors that decay like Chebyshev moments and
at 1e-4 for unstructured random ones, for all four types and shapes that
are not multiples of the 128 x 128 tile."""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import kpm
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["pipelined", "single-stage"])
def mfma_kernel(engine, request):
    engine.set_option(X.OPT_MFMA_PIPE, request.param)
    yield request.param
    engine.set_option(X.OPT_MFMA_PIPE, 1)


@pytest.mark.parametrize("gtype", [X.RETARDED, X.ADVANCED, X.PRINCIPAL, X.NONPRINCIPAL])
def test_mfma_random_coefficients(engine, mfma_kernel, gtype):
    K, M, E, s = 300, 500, 700, 8.0
    rng = np.random.default_rng(gtype)
    c = (rng.normal(size=(K, M)) + 1j * rng.normal(size=(K, M))) / np.sqrt(M)
    engine.lut_generate(M, E, -0.97 * s, 0.97 * s, s)
    ref = engine.contract(c, gtype)
    engine.coeffs_upload(c)
    got = engine.contract_mfma(gtype)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-4, err


@pytest.mark.parametrize("gtype", [X.RETARDED, X.ADVANCED, X.PRINCIPAL, X.NONPRINCIPAL])
def test_mfma_kernels_agree_bitwise(engine, gtype):
    """The single-stage and the pipelined kernels (128 x 128 tiles for the
    complex types, 128 x 256 tiles over two table planes for Principal and
    NonPrincipal) add the same BF16 products into FP32 in the same order per
    accumulator, so their outputs are identical; K and E are not multiples
    of either tile."""
    K, M, E, s = 260, 96, 390, 6.0
    rng = np.random.default_rng(9)
    c = (rng.normal(size=(K, M)) + 1j * rng.normal(size=(K, M))) / np.sqrt(M)
    engine.lut_generate(M, E, -0.95 * s, 0.95 * s, s)
    engine.coeffs_upload(c)
    out = {}
    for pipe in (1, 0):
        engine.set_option(X.OPT_MFMA_PIPE, pipe)
        out[pipe] = engine.contract_mfma(gtype)
    engine.set_option(X.OPT_MFMA_PIPE, 1)
    assert np.array_equal(out[0], out[1])


def test_mfma_ldos_from_recurrence(engine, mfma_kernel):
    m = L.nambu_swave(10, 9, disorder=0.5, seed=3)
    s = kpm.gershgorin_scale(m.csr(), 1.05)
    engine.set_csr(m.csr(), s)
    M, E = 600, 1000
    sites = np.arange(0, m.basis_size, 3)[:200]
    mu = engine.moments_block(sites, M, X.VEC_FP64 | X.H_FP64) * kpm.jackson(M)[None, :]
    engine.lut_generate(M, E, -0.99 * s, 0.99 * s, s)
    ref = engine.contract(mu, X.NONPRINCIPAL).imag / np.pi
    engine.coeffs_upload(mu)
    got = engine.contract_mfma(X.NONPRINCIPAL).imag / np.pi
    err = (np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
    assert err < 3e-5, err
    engine.lut_destroy()
