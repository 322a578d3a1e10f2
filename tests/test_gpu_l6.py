"""GPU parity of the L6-class path (levels 4-12) against the oracle's
restatement of the same design (byte-exact), zlib inflate of every output,
and the compressed size against orot's own at the same level (golden
fixtures written by orot: at most 3% larger)."""
import json
import os
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
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.empty(amd.compress_bound(len(data), fmt) + 4096, dtype=np.uint8)
    with amd.ParallelCompressor(level, fmt, 0, block) as pc:
        n = pc.compress_into(data, out)
    assert n > 0, amd.last_error()
    got = out[:n].tobytes()
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()
    want = orc.gpu_parallel_compress(data, level, fmt, block)
    assert got == want, (len(got), len(want))


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("n", [0, 1, 5, 100, 5000, 70000])
def test_l6_small(amd, kind, n):
    _check(amd, corpus.generate(kind, n, 300 + n), 6, 0, 0)


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("level", [4, 6, 9])
@pytest.mark.parametrize("fmt", [0, 2])
def test_l6_levels(amd, kind, level, fmt):
    _check(amd, corpus.generate(kind, 1_300_000, 17), level, fmt, 0)


@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("level", [5, 7, 8, 10, 11, 12])
def test_l6_other_levels(amd, kind, level):
    # levels 10-12 walk up to 2048 candidates: a smaller input keeps the oracle quick
    n = 600_000 if level >= 10 else 1_300_000
    _check(amd, corpus.generate(kind, n, 19 + level), level, 1, 0)


def _fixtures():
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "orot_fixtures.json")) as f:
        return json.load(f)["levels_l6"]


@pytest.mark.parametrize("e", _fixtures(), ids=lambda e: f"{e['kind']}-{e['n']}-L{e['level']}")
def test_l6_size_within_3pct_of_orot(amd, e):
    """Every level 4-12 fixture input compressed on the GPU: at most 1.03x
    orot's compressed size at that level, zlib-valid, and the oracle's bytes."""
    data = corpus.generate(e["kind"], e["n"], e["seed"])
    out = np.empty(amd.compress_bound(len(data), 0) + 4096, dtype=np.uint8)
    with amd.ParallelCompressor(e["level"], 0, 0, 0) as pc:
        n = pc.compress_into(data, out)
    assert n > 0, amd.last_error()
    got = out[:n].tobytes()
    assert zlib.decompress(got, -15) == data.tobytes()
    assert n <= 1.03 * e["size"], (n, e["size"])
    assert got == orc.gpu_parallel_compress(data, e["level"], 0, 0, threads=8)


@pytest.mark.parametrize("kind", ["text", "records"])
@pytest.mark.parametrize("block", [65536, 200_000])
def test_l6_block_sizes(amd, kind, block):
    _check(amd, corpus.generate(kind, 700_000, 5), 6, 1, block)
