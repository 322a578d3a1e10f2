"""GPU tests of the whole-buffer entry point deflate_compress (orot
src/api/deflate_api.cpp:100-136, src/compress/block_compressor.cpp:16-73):
at levels 0-3 the bytes are orot's own (golden fixtures written by orot's
deflate_compress: one block per input, an input with no match stored, gzip
XFL by level); at levels 4-12 the L6-class stream of the oracle; capacity
below the bound -> 0; every output inflates through zlib."""
import hashlib
import json
import os
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

pytestmark = pytest.mark.gpu

WBITS = {0: -15, 1: 15, 2: 31}
HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "orot_fixtures.json")) as _f:
    _WHOLE = json.load(_f)["whole"]


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


def _input(e):
    n, seed = e["n"], e["seed"]
    if e["kind"] == "mixed":
        q = n // 4
        return np.concatenate([corpus.generate(k, q if i < 3 else n - 3 * q, seed + i)
                               for i, k in enumerate(corpus.GENERATORS)])
    return corpus.generate(e["kind"], n, seed)


@pytest.mark.parametrize("e", _WHOLE, ids=lambda e: f'{e["kind"]}-{e["n"]}-L{e["level"]}-f{e["format"]}')
def test_whole_buffer_is_orot_bytes(amd, e):
    data = _input(e)
    got = amd.compress(data, e["level"], e["format"])
    if "hex" in e:
        assert got.hex() == e["hex"]
    assert len(got) == e["size"] and hashlib.sha256(got).hexdigest() == e["sha256"]


@pytest.mark.parametrize("level", [4, 6, 9, 12])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_whole_buffer_l6_class(amd, level, fmt):
    data = np.concatenate([corpus.generate(k, 300_000, 60 + level) for k in corpus.GENERATORS])
    got = amd.compress(data, level, fmt)
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()
    assert got == orc.whole_compress(data, level, fmt)
    if fmt == 2:
        assert got[8] == (2 if level >= 9 else 0)


def test_whole_buffer_sizes_in_turn(amd):
    """One context per (level, format) serves inputs of any size in turn: its
    chunk capacity grows and shrinks with them."""
    for n in (10, 3_000_000, 70_000, 20_000_001, 5):
        data = np.concatenate([corpus.generate(k, n // 4 + 1, n + i) for i, k in enumerate(corpus.GENERATORS)])
        got = amd.compress(data, 1, 1)
        assert got == orc.whole_compress(data, 1, 1), n


def test_whole_buffer_long_zero_run(amd):
    """64 MiB of zeros as one block: 258-byte matches all the way."""
    data = np.zeros(64 << 20, dtype=np.uint8)
    got = amd.compress(data, 1, 2)
    assert got == orc.whole_compress(data, 1, 2)
    assert zlib.decompress(got, 31) == data.tobytes()


def test_capacity_below_bound(amd):
    import ctypes

    data = corpus.generate("text", 10_000, 1)
    bound = amd.compress_bound(data.size, 1)
    out = np.empty(bound, dtype=np.uint8)
    L = amd.lib()
    assert L.deflate_compress(data.ctypes.data, data.size, out.ctypes.data, bound - 1, 1, 1) == 0
    assert L.deflate_compress(data.ctypes.data, data.size, out.ctypes.data, bound, 1, 1) > 0
    assert L.deflate_compress(ctypes.c_void_p(0), 0, out.ctypes.data, amd.compress_bound(0, 1), 1, 1) > 0
