"""Block sizes above the single-workgroup kernels' limits, which orot accepts
(src/parallel/parallel_compressor.cpp:22-28): levels 1-3 run the big-chunk
path (segment candidate scans with carry-in tables, per-position match
lengths, one workgroup per chunk for the parse) and give orot's bytes for
that block size (the oracle's, which tests/test_oracle.py pins to orot's own
build); levels 4-12 run blocks above 1 MiB as 1 MiB pieces with the 32 KiB
dictionary, the oracle's bytes for the same rule."""
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


def _mixed(n, seed):
    q = n // 4
    return np.concatenate([corpus.generate(k, q if i < 3 else n - 3 * q, seed + i)
                           for i, k in enumerate(corpus.GENERATORS)])


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("block", [131072 + 1, 262144 + 8, 1 << 20, 3_000_017, 16 << 20])
@pytest.mark.parametrize("fmt", [0, 2])
def test_big_blocks_equal_orot(amd, level, block, fmt):
    data = _mixed(7_000_003, level + block % 97)
    print("big block", level, block, fmt, flush=True)
    with amd.ParallelCompressor(level, fmt, 0, block) as pc:
        got = pc.compress(data)
    assert got == orc.parallel_compress(data, level, fmt, block, threads=8)
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()


@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_big_block_one_kind(amd, kind):
    data = corpus.generate(kind, 5_000_000, 99)
    with amd.ParallelCompressor(1, 1, 0, 4 << 20) as pc:
        got = pc.compress(data)
    assert got == orc.parallel_compress(data, 1, 1, 4 << 20, threads=8)


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 65535, 65536, 65537, 131073])
def test_big_block_small_inputs(amd, n):
    data = _mixed(n, n)
    for level in (1, 3):
        with amd.ParallelCompressor(level, 2, 0, 1 << 20) as pc:
            got = pc.compress(data)
        assert got == orc.parallel_compress(data, level, 2, 1 << 20), (n, level)


def test_big_block_batch(amd):
    bufs = [_mixed(1000 + 300_001 * i, i) for i in range(12)]
    with amd.BatchCompressor(2, 1, 600_000, 0) as bc:
        outs = bc.compress(bufs)
    for b, o in zip(bufs, outs):
        assert o == orc.parallel_compress(b, 2, 1, 600_000)


@pytest.mark.parametrize("level,block", [(4, 2 << 20), (6, 3_000_000), (9, 8 << 20), (12, 1 << 30)])
def test_l6_class_blocks_above_1mib(amd, level, block):
    data = _mixed(5_000_000, level)
    with amd.ParallelCompressor(level, 2, 0, block) as pc:
        got = pc.compress(data)
    assert zlib.decompress(got, 31) == data.tobytes()
    assert got == orc.gpu_parallel_compress(data, level, 2, block, threads=8)


def test_noalloc_big_block(amd):
    data = _mixed(3_000_000, 5)
    d_in = amd.DeviceBuffer.from_host(data)
    cap = amd.compress_bound(data.size, 1) + 4096
    d_out = amd.DeviceBuffer(cap)
    n = int(amd.lib().deflate_parallel_compress_device_noalloc(1, 1, 1 << 20, d_in.ptr, data.size, d_out.ptr, cap))
    assert n > 0, amd.last_error()
    assert d_out.download(n).tobytes() == orc.parallel_compress(data, 1, 1, 1 << 20)
