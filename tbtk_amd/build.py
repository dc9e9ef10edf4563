"""Build the native libraries in-tree.

  tbtk_amd/lib/libtbtkx.so      HIP engine (gfx950) behind include/tbtkx.h
  tbtk_amd/lib/libtbtk_host.so  C++ drop-in classes (TBTK::Solver::
                                ChebyshevExpander, TBTK::PropertyExtractor::
                                ChebyshevExpander and the Model subset)

Run ``python -m tbtk_amd.build`` (or __graft_entry__.build()). Every call
recompiles; objects go to tbtk_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
ENGINE = os.path.join(LIB, "libtbtkx.so")
HOSTLIB = os.path.join(LIB, "libtbtk_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function",
    # No packed FP32 instructions (numerics).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-fopenmp",
              "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HOST, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _stale(src: str, obj: str, deps) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in [src, *deps])


def build_engine(jobs: int = 8, changed_only: bool = False) -> str:
    """Compile every .hip source of csrc/ with hipcc and every .cpp source
    with the host compiler (changed_only: only those newer than their object
    or than a shared header) and link libtbtkx.so."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]
    deps = glob.glob(os.path.join(CSRC, "*.hpp")) + [os.path.join(ROOT, "include", "tbtkx.h")]
    todo = [(s, o) for s, o in zip(srcs, objs) if not changed_only or _stale(s, o, deps)]
    # Host-only sources (host_math.cpp) go through the host C++ compiler,
    # not HIP mode (see host_math.hpp).
    csrcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    cobjs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in csrcs]
    ctodo = [(s, o) for s, o in zip(csrcs, cobjs) if not changed_only or _stale(s, o, deps)]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", s, "-o", o] for s, o in todo]
    cmds += [["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-c", s, "-o", o] for s, o in ctodo]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = ENGINE + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, *cobjs, "-lpthread",
          "-o", tmp])
    os.replace(tmp, ENGINE)
    return ENGINE


def build_host(jobs: int = 8) -> str | None:
    srcs = sorted(glob.glob(os.path.join(HOST, "src", "*.cpp")))
    if not srcs:
        return None
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, "host_" + os.path.basename(s) + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, [["g++", *HOST_FLAGS, "-c", s, "-o", o]
                           for s, o in zip(srcs, objs)]))
    tmp = HOSTLIB + ".tmp"
    _run(["g++", "-shared", "-fopenmp", *objs, "-o", tmp, "-L", LIB, "-ltbtkx",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOSTLIB)
    return HOSTLIB


def build_cpp_tests(jobs: int = 8) -> list:
    """The C++ test programs of tests/cpp (drop-in API tests), linked against
    the two libraries with an $ORIGIN-relative rpath."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    srcs = sorted(glob.glob(os.path.join(tdir, "*.cpp")))
    if not srcs:
        return []
    bdir = os.path.join(tdir, "bin")
    os.makedirs(bdir, exist_ok=True)
    rpath = "-Wl,-rpath,$ORIGIN/../../../tbtk_amd/lib"
    outs = [os.path.join(bdir, os.path.splitext(os.path.basename(s))[0]) for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, [["g++", *HOST_FLAGS, s, "-o", o, "-L", LIB, "-ltbtk_host",
                            "-ltbtkx", rpath] for s, o in zip(srcs, outs)]))
    return outs


def build_all(changed_only: bool = False) -> None:
    build_engine(changed_only=changed_only)
    build_host()
    build_cpp_tests()


if __name__ == "__main__":
    # --changed: recompile only the HIP sources changed since their object
    build_all(changed_only="--changed" in sys.argv[1:])
    print("built", ENGINE, HOSTLIB if os.path.exists(HOSTLIB) else "")
    sys.exit(0)
