"""The host layer (tbtk_amd/host: the drop-in Solver/PropertyExtractor
classes, the model layer, Range, IndexTree, JSON) built with AddressSanitizer
and UndefinedBehaviorSanitizer (host code only, g++), running the CPU cases
of the C++ tests (tests/cpp/test_solver.cpp Api.*: setters, defaults,
asserts and exit paths, basis order and CSR assembly, IndexTree order) and
the JSON round trips of the reference's serializations. Any sanitizer
report fails the test (halt_on_error, exit code 86).
"""
import concurrent.futures as cf
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tbtk_amd", "host")
LIB = os.path.join(ROOT, "tbtk_amd", "lib")
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp",
         "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HOST, "include")]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:exitcode=86:detect_leaks=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=86")


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB, "libtbtkx.so")):
        pytest.fail("tbtk_amd/lib/libtbtkx.so missing: run __graft_entry__.build()")
    d = tmp_path_factory.mktemp("asan")
    srcs = sorted(glob.glob(os.path.join(HOST, "src", "*.cpp")))
    objs = [str(d / (os.path.basename(s) + ".o")) for s in srcs]

    def cc(args):
        r = subprocess.run(["g++", *FLAGS, *args], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(cc, [["-c", s, "-o", o] for s, o in zip(srcs, objs)]))
    exes = {}
    for name in ("test_solver", "json_roundtrip"):
        exe = str(d / name)
        cc([os.path.join(ROOT, "tests", "cpp", name + ".cpp"), *objs, "-o", exe,
            "-L", LIB, "-ltbtkx", "-Wl,-rpath," + LIB])
        exes[name] = exe
    return exes


def test_api_cases_clean_under_asan_ubsan(sanitized):
    r = subprocess.run([sanitized["test_solver"], "Api."], capture_output=True, text=True,
                       env=ENV, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_json_roundtrips_clean_under_asan_ubsan(sanitized):
    from test_json import JDIR, KINDS
    for name, kind in sorted(KINDS.items()):
        r = subprocess.run([sanitized["json_roundtrip"], kind, os.path.join(JDIR, name + ".json")],
                           capture_output=True, text=True, env=ENV, timeout=120)
        assert r.returncode == 0, (name, r.stderr[-4000:])
        assert "runtime error" not in r.stderr, name
