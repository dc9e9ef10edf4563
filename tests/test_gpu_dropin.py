"""GPU tests of the C++ drop-in (TBTK::Solver / PropertyExtractor::
ChebyshevExpander on the tbtkx engine).

tests/cpp/pe_outputs builds the fixture models with the public API (the
loops of the reference's README and tests) and writes coefficients, LDOS,
Green's functions of all four types (multi-to and pattern forms, with the
lookup table and the direct evaluation) and a sampleDOS; they must equal the
reference's outputs in tests/golden bit for bit, both with the reference's
default flags (both GPU flags false, as its own callers set them) and with
the GPU flags on. test_solver Gpu.* holds the reference's known-answer and
closed-form tests; the reference's documentation example, compiled as
written against the drop-in, must print the reference's text.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


@pytest.fixture(scope="module", params=["default_flags", "gpu_flags"])
def outputs(request, tmp_path_factory):
    d = tmp_path_factory.mktemp("pe_" + request.param)
    crand = d / "crand.bin"
    load_fixture("gf_types")["crand_coef"].tofile(crand)
    cmd = [os.path.join(BIN, "pe_outputs"), str(d), str(crand)]
    if request.param == "gpu_flags":
        cmd.append("gpu")
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return d


def _read(d, name, dtype):
    return np.fromfile(d / f"{name}.bin", dtype=dtype)


def test_cpp_gpu_cases():
    r = subprocess.run([os.path.join(BIN, "test_solver"), "Gpu."], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_coefficients_bitexact(outputs):
    fx = load_fixture("square30_spin")
    for b in ("0", "1e-06"):
        got = _read(outputs, f"square30_spin_coef_b{b}", np.complex128).reshape(5, 1000)
        assert np.array_equal(got, fx[f"coef_b{b}"]), b
    fx = load_fixture("periodic5")
    got = _read(outputs, "periodic5_coef_b0", np.complex128).reshape(3, 100)
    assert np.array_equal(got, fx["coef_b0"])


def test_ldos_and_gf_bitexact(outputs):
    fx = load_fixture("square30_spin_ldos")
    assert np.array_equal(_read(outputs, "square30_spin_ldos_lookup1", np.float64),
                          fx["ldos_lookup1"])
    assert np.array_equal(_read(outputs, "square30_spin_ldos_lookup0", np.float64),
                          fx["ldos_lookup0"])
    for t in ("Advanced", "Retarded", "Principal", "NonPrincipal"):
        got = _read(outputs, f"square30_spin_gf_{t}", np.complex128).reshape(2, 300)
        assert np.array_equal(got, fx[f"gf_{t}"]), t
    pat = _read(outputs, "square30_spin_gfpattern_NonPrincipal", np.complex128).reshape(2, 300)
    assert np.array_equal(pat, fx["gf_NonPrincipal"])


def test_pattern_gf_many_from_indices(outputs):
    """The pattern form of calculateGreensFunction with 900 from-indices
    (G((15,15,0), (x,y,0)), Retarded, M = 400) runs as one batched call,
    as exact FP64 blocks of start vectors, and equals the reference's 900
    sequential recurrences bit for bit (tests/golden/
    square30_spin_gfpattern900.npz: SHA-256 of all values, the first four
    from-indices in full)."""
    fx = load_fixture("square30_spin_gfpattern900")
    got = _read(outputs, "square30_spin_gfpattern900", np.complex128)
    assert got.size == 900 * 300
    first = got.reshape(900, 300)[:4]
    assert np.array_equal(first, fx["first"]), np.abs(first - fx["first"]).max()
    assert hashlib.sha256(got.astype("<c16").tobytes()).hexdigest() == str(fx["sha256"])


def test_gf_types_from_coefficients(outputs):
    fx = load_fixture("gf_types")
    for c in ("c123", "crand"):
        for t in ("Advanced", "Retarded", "Principal", "NonPrincipal"):
            ref1 = fx[f"{c}_{t}_lookup1"]
            assert np.array_equal(_read(outputs, f"gf_types_{c}_{t}_lookup1", np.complex128),
                                  ref1), (c, t)
            assert np.array_equal(_read(outputs, f"gf_types_{c}_{t}_lookup0", np.complex128),
                                  fx[f"{c}_{t}_lookup0"]), (c, t)


def test_ring_ldos_and_sampledos(outputs):
    fx = load_fixture("ring10_ldos")
    for f in ("0.9", "0.7"):
        assert np.array_equal(_read(outputs, f"ring10_ldos_{f}", np.float64),
                              fx[f"ldos_{f}"]), f
    fx = load_fixture("periodic5_sampledos")
    assert np.array_equal(_read(outputs, "periodic5_sampledos", np.float64), fx["dos"])


def test_basic_chebyshev_template(outputs):
    """The calls of Templates/BasicChebyshev/src/main.cpp:91-116 (both GPU
    flags false, lookup table, 100x100 with spin, 500 coefficients, LDOS
    along y = 50 summed over spin) give the reference's LDOS bit for bit
    (tests/golden/basic_chebyshev_template.npz: SHA-256 of all 100 x 1000
    values and three rows)."""
    fx = load_fixture("basic_chebyshev_template")
    got = _read(outputs, "basic_chebyshev_template_ldos", np.float64)
    assert got.size == 100 * 1000
    rows = got.reshape(100, 1000)[[0, 50, 99]]
    assert np.array_equal(rows, fx["rows"]), np.abs(rows - fx["rows"]).max()
    assert hashlib.sha256(got.astype("<f8").tobytes()).hexdigest() == str(fx["sha256"])


def test_documentation_example_as_written():
    """doc/examples/Solver/ChebyshevExpander.cpp, compiled unchanged against
    the drop-in (oracle/_ref/doc_example_dropin, built by oracle/
    Makefile.ref), prints the text the same source prints when compiled
    against the reference library (tests/golden/doc_example_chebyshev.txt)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "doc_example_dropin")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/doc_example_dropin not built (needs the reference sources)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(os.path.join(ROOT, "tests", "golden", "doc_example_chebyshev.txt")) as f:
        assert r.stdout == f.read()


def _oracle_dos_inputs(O):
    from tbtk_amd import lattice as L
    m = L.periodic_square(6, t=1.0, mu=0.3)
    s, M, lo, hi, E = 6.0, 200, -5.9, 5.9, 500
    return m.csr(), s, M, O.lookup_table(M, E, lo, hi, s), m.basis_size


def test_calculate_dos_exact_vs_oracle(outputs, oracle_mod):
    """PE::calculateDOS without stochastic vectors is the sum of the LDOS of
    every basis state (Tr delta(E - H)): the oracle's moments, Lorentz
    damping (broadening 1e-6), lookup table and NonPrincipal contraction
    (Lib/src/Solver/ChebyshevExpander.cpp:103-199, 201-270, 324-370), summed
    over the basis in order. The plain recurrence is bitwise the oracle's,
    the FP64 block recurrence (doubling identities) within 1e-12."""
    O = oracle_mod
    csr, s, M, T, N = _oracle_dos_inputs(O)
    ref = np.zeros(T.shape[1])
    for b in range(N):
        c = O.lorentz(O.moments(csr, s, b, [b], M), 1e-6)[0]
        ref += O.gf_lookup(O.NONPRINCIPAL, T, c).imag / np.pi
    plain = _read(outputs, "periodic6_dos_exact_plain", np.float64)
    assert np.array_equal(plain, ref)
    block = _read(outputs, "periodic6_dos_exact_block", np.float64)
    assert np.abs(block - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(ref.sum() * (11.8 / 499) - N) < 0.05 * N


def test_calculate_dos_stochastic_vs_oracle(outputs, oracle_mod):
    """With setStochasticVectors(R, seed) the DOS comes from the stochastic
    trace sum_r <r|T_n|r> / R over R random-phase vectors exp(2 pi i u),
    u = the splitmix64 uniform of (seed, r N + i): the oracle's trace
    recurrence on the same vectors, damped and contracted as above. FP64
    block within 1e-10 of max DOS, FP32 block within 3e-5."""
    O = oracle_mod
    csr, s, M, T, N = _oracle_dos_inputs(O)
    R, seed = 50, 7
    vecs = np.exp(2j * np.pi * O.uniform(seed, np.arange(R * N)).reshape(R, N))
    mu = O.moments_trace(csr, s, vecs, M) / R
    ref = O.gf_lookup(O.NONPRINCIPAL, T, O.lorentz(mu, 1e-6)[0]).imag / np.pi
    peak = np.abs(ref).max()
    got = _read(outputs, "periodic6_dos_stochastic_fp64", np.float64)
    assert np.abs(got - ref).max() <= 1e-10 * peak
    got = _read(outputs, "periodic6_dos_stochastic_fp32", np.float64)
    assert np.abs(got - ref).max() <= 3e-5 * peak


def test_thermodynamic_properties_bitexact(outputs):
    """Density, Magnetization, SpinPolarizedLDOS (Custom and Ranges formats)
    and expectation values at T = 300 K and T = 0 equal the reference's bit
    for bit (tests/golden/square30_spin_thermo.npz)."""
    fx = load_fixture("square30_spin_thermo")
    for T in ("T300", "T0"):
        for name, dt in (("density_custom", np.float64), ("density_ranges", np.float64),
                         ("mag_custom", np.complex128), ("mag_ranges", np.complex128),
                         ("expval_offdiag", np.complex128), ("expval_diag", np.complex128)):
            got = _read(outputs, f"thermo_{name}_{T}", dt)
            assert np.array_equal(got, fx[f"{name}_{T}"]), (name, T)
    for name in ("splodos_custom", "splodos_ranges"):
        got = _read(outputs, f"thermo_{name}", np.complex128)
        assert np.array_equal(got, fx[name]), name


JSON_NAMES = ["ldos_custom", "ldos_ranges", "gf_NonPrincipal", "density_custom_T300",
              "density_ranges_T300", "sampledos"]


@pytest.mark.parametrize("name", JSON_NAMES)
def test_json_serialization_matches_reference(outputs, name):
    """serialize(Serializable::Mode::JSON) of the properties computed here
    equals the reference's serialization of the same computation
    (tests/golden/json), except the reference's uninitialized
    "defaultValue" (see tests/test_json.py)."""
    import re
    norm = lambda t: re.sub(r'"defaultValue":"[^"]*"', '"defaultValue":"<default>"', t)
    got = (outputs / f"json_{name}.json").read_text()
    with open(os.path.join(ROOT, "tests", "golden", "json", name + ".json")) as f:
        ref = f.read()
    assert norm(got) == norm(ref)
