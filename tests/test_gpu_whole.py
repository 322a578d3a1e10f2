"""GPU tests of the whole-buffer entry point deflate_compress: orot's rules
for it (capacity below the bound -> 0, gzip XFL by level) over the chunked
pipeline, every output inflating through zlib, and the compressed size next
to orot's own single-block output when the reference build is present."""
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


@pytest.mark.parametrize("level", [0, 1, 2, 6, 9])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_whole_buffer(amd, level, fmt):
    data = np.concatenate([corpus.generate(k, 300_000, 60 + level) for k in corpus.GENERATORS])
    got = amd.compress(data, level, fmt)
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()
    with amd.ParallelCompressor(level, fmt, 0, 0) as pc:
        par = pc.compress(data)
    if fmt == 2:
        assert got[8] == (2 if level >= 9 else (4 if level <= 2 else 0))
        got = got[:8] + par[8:9] + got[9:]
    assert got == par


def test_capacity_below_bound(amd):
    import ctypes

    data = corpus.generate("text", 10_000, 1)
    bound = amd.compress_bound(data.size, 1)
    out = np.empty(bound, dtype=np.uint8)
    L = amd.lib()
    assert L.deflate_compress(data.ctypes.data, data.size, out.ctypes.data, bound - 1, 1, 1) == 0
    assert L.deflate_compress(data.ctypes.data, data.size, out.ctypes.data, bound, 1, 1) > 0
    assert L.deflate_compress(ctypes.c_void_p(0), 0, out.ctypes.data, amd.compress_bound(0, 1), 1, 1) > 0


@pytest.mark.skipif(not orc.ref_available(), reason="orot reference build (oracle/_ref) not present")
@pytest.mark.parametrize("level", [1, 6])
def test_ratio_vs_orot_whole_buffer(amd, level):
    """orot's whole-buffer output is one LZ77 pass and one block; the chunked
    stream stays within 5% of its size on text and records."""
    for kind in ("text", "records"):
        data = corpus.generate(kind, 4 << 20, 77)
        got = amd.compress(data, level, 2)
        want = orc.ref_compress(data, level, 2)
        assert zlib.decompress(got, 31) == data.tobytes()
        assert len(got) <= 1.05 * len(want), (kind, len(got), len(want))
