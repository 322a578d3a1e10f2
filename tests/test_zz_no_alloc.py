"""deflate_parallel_compress_device_noalloc against
deflate_parallel_compress_device: byte-identical output for the same level,
format and block size, over two calls in a row on the shared scratch array
(the second with a smaller input, so stale state of the first call would
show), plus the size limit and a zlib inflate of every output."""
import zlib

import numpy as np
import pytest

from orot_amd import corpus

pytestmark = pytest.mark.gpu

WBITS = {0: -15, 1: 15, 2: 31}


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


def _noalloc(amd, d_in, n, d_out, cap, level, fmt, block):
    return int(amd.lib().deflate_parallel_compress_device_noalloc(level, fmt, block, d_in.ptr, n, d_out.ptr, cap))


def _mixed(n, seed):
    q = n // 4
    parts = [corpus.generate(k, q if i < 3 else n - 3 * q, seed + i) for i, k in enumerate(corpus.GENERATORS)]
    return np.concatenate(parts)


@pytest.mark.parametrize("level,fmt,block", [(0, 1, 0), (1, 0, 0), (1, 2, 65536), (2, 1, 0), (6, 2, 0),
                                             (9, 0, 0), (6, 1, 262144)])
def test_noalloc_equals_context_path(amd, level, fmt, block):
    sizes = [3_000_000, 700_001]  # second call smaller than the first
    for call, n in enumerate(sizes):
        data = _mixed(n, 1000 * level + 10 * call + fmt)
        cap = amd.compress_bound(n, fmt) + 4096
        d_in = amd.DeviceBuffer.from_host(data)
        d_out = amd.DeviceBuffer(cap)
        d_ref = amd.DeviceBuffer(cap)
        print(f"noalloc level {level} fmt {fmt} block {block} call {call} n {n}", flush=True)
        got_n = _noalloc(amd, d_in, n, d_out, cap, level, fmt, block)
        assert got_n > 0, amd.last_error()
        with amd.ParallelCompressor(level, fmt, 0, block) as pc:
            want_n = pc.compress_device(d_in.ptr, n, d_ref.ptr, cap)
        assert want_n > 0, amd.last_error()
        got = d_out.download(got_n).tobytes()
        want = d_ref.download(want_n).tobytes()
        assert got == want, (call, got_n, want_n)
        assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()


def test_noalloc_empty_and_tiny(amd):
    for n in (0, 1, 7):
        data = corpus.generate("text", n, 3)
        cap = amd.compress_bound(n, 2) + 64
        d_in = amd.DeviceBuffer.from_host(data) if n else amd.DeviceBuffer(1)
        d_out = amd.DeviceBuffer(cap)
        got_n = _noalloc(amd, d_in, n, d_out, cap, 1, 2, 0)
        assert got_n > 0, amd.last_error()
        assert zlib.decompress(d_out.download(got_n).tobytes(), 31) == data.tobytes()


def test_noalloc_rejects_beyond_scratch(amd):
    n = 64 << 20  # working set far above the fixed scratch at level 6
    d_in = amd.DeviceBuffer(n)
    d_out = amd.DeviceBuffer(1 << 20)
    assert _noalloc(amd, d_in, n, d_out, 1 << 20, 6, 0, 0) == 0
    assert "scratch" in amd.last_error()


def test_noalloc_small_output_capacity(amd):
    data = corpus.generate("random", 200_000, 9)
    d_in = amd.DeviceBuffer.from_host(data)
    d_out = amd.DeviceBuffer(1000)
    assert _noalloc(amd, d_in, data.size, d_out, 1000, 1, 0, 0) == 0
