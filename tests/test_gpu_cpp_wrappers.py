"""orot's C++ wrappers used by name against this library
(tests/cpp/wrappers.cpp, built with the library): compress, decompress,
Compressor, Decompressor, ParallelCompressor, adler32, crc32 and Error, and
the round-trip matrix of orot's parallel test
(tests/unit/test_deflate_parallel.cpp:56-107) through them, on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_wrappers_on_gpu():
    exe = os.path.join(ROOT, "tests", "cpp", "wrappers")
    assert os.path.exists(exe), "run python -m orot_amd.build"
    print("running", exe, "--require-gpu", flush=True)
    r = subprocess.run([exe, "--require-gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "0 failure(s)" in r.stdout and "host side only" not in r.stdout
