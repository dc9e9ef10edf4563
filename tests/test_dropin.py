"""CPU tests of the boundary: the C ABI library loads and exports every
function include/tbtkx.h declares, and the C++ drop-in classes keep the
reference's setters, defaults and asserts (tests/cpp/test_solver.cpp,
Api.* cases). No GPU work is done here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def declared_symbols():
    with open(os.path.join(ROOT, "include", "tbtkx.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(tbtkx_[a-z_]+)\s*\(", text)))


def test_header_declares_the_engine():
    names = declared_symbols()
    for must in ("tbtkx_create", "tbtkx_set_csr", "tbtkx_moments", "tbtkx_contract",
                 "tbtkx_moments_block", "tbtkx_moments_trace"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from tbtk_amd import engine
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert set(declared_symbols()) <= set(engine.SIGNATURES)


def test_host_library_links_engine():
    from tbtk_amd import build
    out = subprocess.run(["ldd", build.HOSTLIB], capture_output=True, text=True).stdout
    assert "libtbtkx.so" in out


def test_cpp_api_cases():
    exe = os.path.join(BIN, "test_solver")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/test_solver missing: run __graft_entry__.build()")
    r = subprocess.run([exe, "Api."], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
