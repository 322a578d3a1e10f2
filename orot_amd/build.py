"""Build liborot_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m orot_amd.build          # compile and link
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liborot_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from forming packed-FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

SOURCES = ["deflate_kernels.hip", "deflate_l6.hip", "inflate.hip", "engine.cpp", "deflate_api.cpp",
           "inflate_api.cpp", "server.cpp", "stream_api.cpp", "noalloc.hip", "inflate_host.cpp",
           "deflate_big.hip", "server_kernels.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


CPP_TESTS = [("tests/cpp/wrappers.cpp", "tests/cpp/wrappers")]


def build_cpp_tests(verbose: bool = False) -> None:
    """Test programs written against include/orot/deflate.h, linked with the
    library (rpath relative to the program, so the tree can move)."""
    for src, exe in CPP_TESTS:
        out = os.path.join(ROOT, exe)
        _run(["g++", "-std=c++20", "-O2", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
              os.path.join(ROOT, src), "-o", out + ".tmp", "-L" + HERE, "-lorot_amd",
              "-Wl,-rpath,$ORIGIN/../../orot_amd", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"], verbose)
        os.replace(out + ".tmp", out)


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    build_cpp_tests(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
