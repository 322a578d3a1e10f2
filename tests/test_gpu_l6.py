"""GPU parity of the L6-class path (levels 4-12) against the oracle's
restatement of the same design (byte-exact), zlib inflate of every output."""
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


@pytest.mark.parametrize("kind", ["text", "records"])
@pytest.mark.parametrize("block", [65536, 200_000])
def test_l6_block_sizes(amd, kind, block):
    _check(amd, corpus.generate(kind, 700_000, 5), 6, 1, block)
