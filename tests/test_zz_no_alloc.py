"""tbtkx_moments_noalloc, the allocation-free variant of tbtkx_moments (fixed
per-GPU scratch, no device allocation during the call): called through the
C ABI, its moments equal tbtkx_moments' bit for bit on two calls in a row,
the second on a smaller input (a smaller model, fewer to-indices, fewer
moments), for a real H (real recurrence) and a complex H; an oversize call
is rejected before any device work.
"""
import numpy as np
import pytest

from tbtk_amd import engine as X
from tbtk_amd import lattice as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = X.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_noalloc_equals_moments_twice(eng, kind):
    if kind == "real":
        big = L.square_spin(60, 50, t=1.0, J=0.5, disorder=1.0, seed=2)
        small = L.square_spin(9, 7, t=1.0, J=0.5, disorder=1.0, seed=3)
    else:
        big = L.square(70, 60, t=1.0, disorder=0.5, seed=6, flux=0.05)
        small = L.square(8, 9, t=1.0, disorder=0.5, seed=7, flux=0.1)
    calls = [(big, 8.0, 700, [5, 17, 3000, 5, 1234, 0]), (small, 6.0, 90, [4, 4])]
    for model, s, M, to in calls:
        print(f"{kind}: N={model.basis_size} K={len(to)} M={M}", flush=True)
        eng.set_csr(model.csr(), s)
        frm = to[0]
        ref = eng.moments(frm, to, M)
        got = eng.moments_noalloc(frm, to, M)
        assert eng.get_option(X.OPT_REAL) == (1 if kind == "real" else 0)
        assert np.array_equal(got, ref)
        assert np.count_nonzero(got[0]) == 0   # repeated to-index: last position only


def test_noalloc_rejects_oversize(eng):
    m = L.square_spin(10, 10, seed=1)
    eng.set_csr(m.csr(), 8.0)
    # 4096 to-indices would exceed N = 200
    with pytest.raises(X.EngineError, match="to-indices"):
        eng.moments_noalloc(0, np.arange(201) % 200, 10)
    # K x M x 16 B of moments beyond the scratch
    with pytest.raises(X.EngineError, match="scratch"):
        eng.moments_noalloc(0, np.arange(100), 50000)
    # the scratch still works afterwards
    ref = eng.moments(3, [3, 4], 40)
    assert np.array_equal(eng.moments_noalloc(3, [3, 4], 40), ref)
