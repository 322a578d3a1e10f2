"""GPU parity of the L0-L3 parallel path against the oracle (byte-exact),
zlib inflate of every output, and orot's error behaviour."""
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

pytestmark = pytest.mark.gpu

WBITS = {0: -15, 1: 15, 2: 31}


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


def _check(amd, data, level, fmt, block):
    # output sized like orot's parallel test: bound + 4096 (test_deflate_parallel.cpp:32)
    out = np.empty(amd.compress_bound(len(data), fmt) + 4096, dtype=np.uint8)
    with amd.ParallelCompressor(level, fmt, 0, block) as pc:
        n = pc.compress_into(np.ascontiguousarray(data, dtype=np.uint8), out)
    assert n > 0, amd.last_error()
    got = out[:n].tobytes()
    want = orc.parallel_compress(data, level, fmt, block)
    assert got == want, (len(got), len(want))
    assert zlib.decompress(got, WBITS[fmt]) == bytes(np.asarray(data, dtype=np.uint8))


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 100, 4096, 70000])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_small_l1(amd, kind, n, fmt):
    data = corpus.generate(kind, n, 1000 + n)
    _check(amd, data, 1, fmt, 0)


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("level", [0, 1, 2, 3])
@pytest.mark.parametrize("block", [0, 1024, 65536, 131072])
def test_levels_blocks(amd, kind, level, block):
    data = corpus.generate(kind, 300_000, 77)
    _check(amd, data, level, 2, block)


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("block", [131072 + 8, 200_000, 262144])
def test_l1_wide_chunks(amd, kind, block):
    """Chunks above 128 KiB take the wide kernel (level 1 only)."""
    data = corpus.generate(kind, 700_000, 91)
    _check(amd, data, 1, 0, block)


def _orot_fixtures():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "orot_fixtures.json")) as f:
        return json.load(f)


_FIX = _orot_fixtures()


def _check_fixture(amd, e):
    import hashlib

    data = corpus.generate(e["kind"], e["n"], e["seed"])
    out = np.empty(amd.compress_bound(len(data), e["format"]) + 4096, dtype=np.uint8)
    with amd.ParallelCompressor(e["level"], e["format"], 0, e["block"]) as pc:
        n = pc.compress_into(data, out)
    assert n == e["size"], (n, e)
    assert hashlib.sha256(out[:n].tobytes()).hexdigest() == e["sha256"], e


@pytest.mark.parametrize("e", _FIX["large"], ids=lambda e: f'{e["kind"]}-{e["n"]}-L{e["level"]}-f{e["format"]}-b{e["block"]}')
def test_orot_bytes_large(amd, e):
    """GPU output equals the bytes orot's own parallel compressor wrote for
    the same input (golden fixtures, gzip trailers with the IEEE CRC-32)."""
    _check_fixture(amd, e)


def test_orot_bytes_text16(amd):
    """16 MiB Zipf text, level 1, gzip, 128 KiB chunks: orot's bytes."""
    _check_fixture(amd, _FIX["text16_l1_gzip"])


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("block", [4096, 65536])
def test_many_and_few_chunks(amd, kind, block):
    """3 MiB at 4 KiB chunks (768 chunks: two-wave candidate scan, four
    chunks per CU) and at 64 KiB chunks (48 chunks: six-wave scan); both
    scans give the same bytes."""
    data = corpus.generate(kind, 3 << 20, 4242)
    _check(amd, data, 1, 2, block)
