"""Streaming compressor (deflate_stream_new / deflate_stream_compress, orot
include/orot/deflate.h:92-108, src/api/stream_api.cpp:134-247,
src/compress/compressor.cpp:21-97) on the GPU.

Fed without flushes, a stream is byte for byte the parallel path's output for
the concatenated input at every level (pieces are cut at chunk boundaries and
the L6 class carries its 32 KiB dictionary across pieces), and at levels 0-3
that is the oracle's bytes.  Flushes keep every prefix decodable (zlib's
Z_SYNC_FLUSH marker) and the whole stream inflates to its input."""
import gzip
import zlib

import numpy as np
import pytest

from orot_amd import corpus

import oracle_ffi

pytestmark = pytest.mark.gpu

WBITS = {0: -15, 1: 15, 2: 31}


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


def mixed(n: int, seed: int) -> np.ndarray:
    parts = [corpus.generate(k, n // 4 + (1 if i < n % 4 else 0), seed + i) for i, k in enumerate(corpus.GENERATORS)]
    return np.concatenate(parts)[:n]


def feed(amd, level, fmt, data: np.ndarray, seed: int, out_chunk: int = 1 << 20) -> bytes:
    """Irregular writes with NO_FLUSH, then FINISH."""
    rng = np.random.default_rng(seed)
    out = []
    with amd.Compressor(level, fmt) as c:
        pos = 0
        while pos < data.size:
            k = int(min(data.size - pos, rng.integers(1, 3 << 20)))
            out.append(c.write(data[pos:pos + k], amd.Flush.NO_FLUSH, out_chunk))
            pos += k
        out.append(c.write(b"", amd.Flush.FINISH, out_chunk))
    return b"".join(out)


def inflate(s: bytes, fmt: int) -> bytes:
    d = zlib.decompressobj(WBITS[fmt])
    r = d.decompress(s)
    assert d.eof and not d.unused_data
    return r


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("level", [0, 1, 2, 3, 4, 6, 9, 12])
@pytest.mark.parametrize("n", [0, 1, 1000, 131072, 131073, 3_000_001])
def test_stream_equals_parallel(amd, level, fmt, n):
    data = mixed(n, 40 + n % 97)
    s = feed(amd, level, fmt, data, n + level)
    with amd.ParallelCompressor(level, fmt, 0, 0) as pc:
        p = pc.compress(data)
    assert s == p
    assert inflate(s, fmt) == data.tobytes()
    if level <= 3:
        assert s == oracle_ffi.parallel_compress(data, level, fmt)


@pytest.mark.skipif(not oracle_ffi.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_stream_equals_orot_stream(amd, kind):
    """Directly against orot's own streaming compressor, built from its
    sources, wherever its stream is valid: up to one 128 KiB block, fed
    before DEFLATE_FINISH (under 128 KiB) or with it (exactly 128 KiB)."""
    for n, one_call in ((1000, False), (65536, False), (131071, False), (131072, True)):
        data = corpus.generate(kind, n, 61)
        for level in (0, 1, 2, 3):
            for fmt in (0, 1, 2):
                assert feed(amd, level, fmt, data, 3) == oracle_ffi.ref_stream(data, level, fmt, one_call), \
                    (n, level, fmt)


@pytest.mark.parametrize("level,fmt", [(1, 2), (6, 1), (2, 0), (12, 2)])
def test_stream_over_several_pieces(amd, level, fmt):
    """Input of more than two 32 MiB pieces: the pieces' bit offsets, the
    dictionary carried across them (level 6) and the checksum combine."""
    data = mixed(70_000_003, 5)
    s = feed(amd, level, fmt, data, 11)
    with amd.ParallelCompressor(level, fmt, 0, 0) as pc:
        assert s == pc.compress(data)
    assert inflate(s, fmt) == data.tobytes()
    if fmt == 2:
        assert gzip.decompress(s) == data.tobytes()


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("level", [0, 1, 6])
def test_stream_flushes(amd, level, fmt):
    """SYNC and FULL flushes between writes: everything written before a flush
    decodes from the bytes out so far, which end with 00 00 FF FF, and the
    finished stream inflates to the input."""
    rng = np.random.default_rng(level * 3 + fmt)
    data = mixed(2_500_000, 9)
    out = b""
    d = zlib.decompressobj(WBITS[fmt])
    got = b""
    pos = 0
    with amd.Compressor(level, fmt) as c:
        while pos < data.size:
            k = int(min(data.size - pos, rng.integers(1, 400_000)))
            flush = [amd.Flush.NO_FLUSH, amd.Flush.SYNC_FLUSH, amd.Flush.FULL_FLUSH][int(rng.integers(0, 3))]
            piece = c.write(data[pos:pos + k], flush)
            out += piece
            pos += k
            got += d.decompress(piece)
            if flush != amd.Flush.NO_FLUSH:
                assert out.endswith(b"\x00\x00\xff\xff")
                assert got == data[:pos].tobytes()
        out += c.flush()
        assert c.flush() == b""  # nothing new since the last marker: no output
        out += c.finish()
    assert inflate(out, fmt) == data.tobytes()


@pytest.mark.parametrize("level", [1, 6, 9])
def test_full_flush_is_a_restart_point(amd, level):
    """After DEFLATE_FULL_FLUSH no match reaches back past the flush: the
    bytes after its marker inflate on their own (RFC 1951 raw, no
    dictionary); after a sync flush the L6 class may reach back."""
    data = mixed(3_000_000, 77)
    cut = 1_234_567
    with amd.Compressor(level, 0) as c:
        head = c.write(data[:cut], amd.Flush.FULL_FLUSH)
        tail = c.write(data[cut:], amd.Flush.NO_FLUSH) + c.finish()
    assert head.endswith(b"\x00\x00\xff\xff")
    assert inflate(tail, 0) == data[cut:].tobytes()
    assert inflate(head + tail, 0) == data.tobytes()


@pytest.mark.parametrize("level", [6, 9])
def test_full_flush_after_sync_flush_is_a_restart_point(amd, level):
    """A full flush straight after a sync flush (no input between them)
    writes no second marker but still ends the dictionary: the bytes written
    after it inflate on their own."""
    data = mixed(2_000_000, 78)
    cut = 900_001
    with amd.Compressor(level, 0) as c:
        head = c.write(data[:cut], amd.Flush.SYNC_FLUSH) + c.flush(True)
        tail = c.write(data[cut:], amd.Flush.NO_FLUSH) + c.finish()
    assert head.endswith(b"\x00\x00\xff\xff")
    assert inflate(tail, 0) == data[cut:].tobytes()
    assert inflate(head + tail, 0) == data.tobytes()


def test_stream_flush_only(amd):
    """A sync flush before any input, flushes with nothing pending, and an
    empty stream finished after a flush."""
    for fmt in (0, 1, 2):
        with amd.Compressor(1, fmt) as c:
            s = c.flush() + c.write(b"abc", amd.Flush.SYNC_FLUSH) + c.flush(True) + c.finish()
        assert inflate(s, fmt) == b"abc"
        with amd.Compressor(6, fmt) as c:
            s = c.flush() + c.finish()
        assert inflate(s, fmt) == b""


def test_stream_small_output_buffers(amd):
    """Output space of 7 bytes per call: NEED_OUTPUT until drained, the same
    stream as with a large buffer."""
    data = mixed(300_000, 21)
    for fmt in (1, 2):
        assert feed(amd, 1, fmt, data, 4, out_chunk=7) == feed(amd, 1, fmt, data, 4)


def test_stream_results_and_end(amd):
    """Result codes of one call: NEED_OUTPUT with no output space, OK once
    drained, STREAM_END after FINISH and on every later call."""
    data = mixed(50_000, 2)
    c = amd.Compressor(1, 1)
    big = np.empty(1 << 20, dtype=np.uint8)
    r, used, n = c.compress_call(data, np.empty(0, dtype=np.uint8), amd.Flush.NO_FLUSH)
    assert r == amd.Result.NEED_OUTPUT and n == 0  # the zlib header waits for space
    r, used, n = c.compress_call(data, big, amd.Flush.NO_FLUSH)
    assert r == amd.Result.OK and used == data.size and n == 2
    head = big[:2].tobytes()
    r, used, n = c.compress_call(b"", big, amd.Flush.FINISH)
    assert r == amd.Result.STREAM_END
    s = head + big[:n].tobytes()
    assert zlib.decompress(s) == data.tobytes()
    r, used, n = c.compress_call(b"x", big, amd.Flush.NO_FLUSH)
    assert r == amd.Result.STREAM_END and used == 0 and n == 0
    c.close()


def test_stream_param_errors(amd):
    import ctypes

    L = amd.lib()
    s = L.deflate_stream_new(1, 0)
    assert s
    buf = ctypes.create_string_buffer(64)
    ni, ai = ctypes.c_void_p(ctypes.addressof(buf)), ctypes.c_size_t(0)
    no, ao = ctypes.c_void_p(ctypes.addressof(buf)), ctypes.c_size_t(64)
    assert L.deflate_stream_compress(s, ctypes.byref(ni), ctypes.byref(ai), ctypes.byref(no), ctypes.byref(ao), 4) \
        == amd.Result.PARAM_ERROR
    assert L.deflate_stream_compress(s, None, ctypes.byref(ai), ctypes.byref(no), ctypes.byref(ao), 0) \
        == amd.Result.PARAM_ERROR
    L.deflate_stream_free(s)
