"""The host inflater behind orot's decompress API (deflate_decompress,
inflate_stream_new / inflate_stream_decompress and the Decompressor wrapper;
orot src/api/deflate_api.cpp:138-215, src/api/stream_api.cpp:253-380) against
zlib, the library routine that pins RFC 1950-1952 decoding: streams made by
zlib at every level and strategy (stored, fixed and dynamic blocks, flush
points, empty blocks), streams of the GPU compress path (the oracle's bytes,
which the GPU tests show equal to the HIP output) and the stored-after-Huffman
streams orot's own inflater mishandles inflate to their input; corrupt,
truncated and oversized cases return orot's result codes
(src/formats/gzip_wrapper.cpp:76-138, src/formats/zlib_wrapper.cpp).  Every
case runs whole and streamed (irregular input pieces, small output space).
Needs no GPU."""
import ctypes
import os
import subprocess
import sys
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

WBITS = {0: -15, 1: 15, 2: 31}
OK, STREAM_END, NEED_OUTPUT, DATA_ERROR, PARAM_ERROR = 0, 1, 3, -1, -5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


def zstream(data: bytes, level: int, fmt: int, strategy=zlib.Z_DEFAULT_STRATEGY, pieces=1, flush=zlib.Z_SYNC_FLUSH,
            zdict=None):
    kw = {"zdict": zdict} if zdict else {}
    co = zlib.compressobj(level, zlib.DEFLATED, WBITS[fmt], 9, strategy, **kw)
    out = b""
    step = max(1, len(data) // pieces)
    for i in range(0, len(data), step):
        out += co.compress(data[i:i + step])
        if pieces > 1:
            out += co.flush(flush)
    return out + co.flush()


def streamed(amd, s: bytes, fmt: int, seed: int = 0, max_in: int = 5000, out_chunk: int = 777):
    """inflate_stream_decompress over irregular input pieces and small output
    space: (last result, bytes out, input bytes left over)."""
    rng = np.random.default_rng(seed)
    out = []
    buf = np.empty(out_chunk, dtype=np.uint8)
    src = np.frombuffer(s, dtype=np.uint8)
    pos = 0
    r = OK
    with amd.Decompressor(fmt) as d:
        while True:
            k = int(min(len(src) - pos, rng.integers(1, max_in + 1)))
            r, used, n = d.call(src[pos:pos + k], buf)
            out.append(buf[:n].tobytes())
            pos += used
            while r == NEED_OUTPUT:
                r, used, n = d.call(src[pos:pos], buf)
                out.append(buf[:n].tobytes())
            if r != OK or pos >= len(src):
                break
    return r, b"".join(out), len(src) - pos


def inflate(amd, s: bytes, cap: int, fmt: int):
    """Whole-buffer deflate_decompress, and the same stream streamed: both
    must agree on success."""
    out = np.empty(cap, dtype=np.uint8)
    r, n = amd.decompress_into(s, out, fmt)
    got = out[:n].tobytes() if r == OK else b""
    if r == OK:
        rs, gs, _ = streamed(amd, s, fmt, seed=len(s))
        assert rs == STREAM_END and gs == got
    return r, got


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("level", [0, 1, 4, 6, 9])
@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_zlib_levels(amd, kind, level, fmt):
    data = corpus.generate(kind, 300_000, 11 + level).tobytes()
    r, got = inflate(amd, zstream(data, level, fmt), len(data), fmt)
    assert r == OK and got == data


@pytest.mark.parametrize("strategy", [zlib.Z_FILTERED, zlib.Z_HUFFMAN_ONLY, zlib.Z_RLE, zlib.Z_FIXED])
@pytest.mark.parametrize("kind", ["text", "records", "zeros"])
def test_zlib_strategies(amd, kind, strategy):
    data = corpus.generate(kind, 200_000, 5).tobytes()
    s = zstream(data, 6, 2, strategy)
    r, got = inflate(amd, s, len(data), 2)
    assert r == OK and got == data


@pytest.mark.parametrize("flush", [zlib.Z_SYNC_FLUSH, zlib.Z_FULL_FLUSH])
def test_flush_points(amd, flush):
    """Flushes end blocks mid-stream with empty stored blocks."""
    data = corpus.generate("text", 150_000, 8).tobytes()
    for fmt in (0, 1, 2):
        s = zstream(data, 6, fmt, pieces=37, flush=flush)
        r, got = inflate(amd, s, len(data), fmt)
        assert r == OK and got == data


@pytest.mark.parametrize("n", [0, 1, 2, 3, 100, 65535, 65536, 65537, 200_003])
def test_sizes(amd, n):
    for kind in ("random", "text"):
        data = corpus.generate(kind, n, n + 1).tobytes() if n else b""
        for level in (0, 1, 9):
            for fmt in (0, 1, 2):
                r, got = inflate(amd, zstream(data, level, fmt), max(n, 1), fmt)
                assert r == OK and got == data, (kind, level, fmt)


@pytest.mark.parametrize("level", [0, 1, 2, 4, 6, 9, 12])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_own_streams(amd, level, fmt):
    """Mixed content: stored chunks after Huffman chunks and back."""
    data = np.concatenate([corpus.generate(k, 400_000, 3) for k in corpus.GENERATORS + ("text",)])
    s = orc.gpu_parallel_compress(data, level, fmt, 0, threads=4)
    r, got = inflate(amd, s, data.size, fmt)
    assert r == OK and got == data.tobytes()


def test_python_decompress_grows(amd):
    data = corpus.generate("zeros", 3_000_000, 1).tobytes()
    assert amd.decompress(zlib.compress(data, 9), amd.Format.ZLIB) == data


def test_output_capacity(amd):
    data = corpus.generate("text", 100_000, 4).tobytes()
    for fmt in (0, 1, 2):
        s = zstream(data, 6, fmt)
        assert inflate(amd, s, len(data), fmt)[0] == OK
        assert inflate(amd, s, len(data) - 1, fmt)[0] == NEED_OUTPUT
        s0 = zstream(data, 0, fmt)  # stored blocks
        assert inflate(amd, s0, len(data) - 1, fmt)[0] == NEED_OUTPUT


def test_truncated_and_corrupt(amd):
    data = corpus.generate("text", 80_000, 6).tobytes()
    for fmt in (0, 1, 2):
        s = zstream(data, 6, fmt)
        for cut in (1, len(s) // 3, len(s) // 2, len(s) - 9):
            assert inflate(amd, s[:cut], len(data), fmt)[0] == DATA_ERROR, (fmt, cut)
    g = bytearray(zstream(data, 6, 2))
    bad_crc = bytes(g[:-8]) + bytes([g[-8] ^ 1]) + bytes(g[-7:])
    bad_isize = bytes(g[:-4]) + bytes([g[-4] ^ 1]) + bytes(g[-3:])
    bad_magic = b"\x1f\x8c" + bytes(g[2:])
    for s in (bad_crc, bad_isize, bad_magic):
        assert inflate(amd, s, len(data), 2)[0] == DATA_ERROR
    z = bytearray(zstream(data, 6, 1))
    bad_adler = bytes(z[:-1]) + bytes([z[-1] ^ 0x80])
    bad_check = bytes([z[0], z[1] ^ 1]) + bytes(z[2:])
    for s in (bad_adler, bad_check):
        assert inflate(amd, s, len(data), 1)[0] == DATA_ERROR
    # a preset dictionary: FDICT in zlib framing, distances before the start in raw
    zd = corpus.generate("text", 4000, 9).tobytes()
    assert inflate(amd, zstream(data, 6, 1, zdict=zd), len(data), 1)[0] == DATA_ERROR
    assert inflate(amd, zstream(data[:5000], 6, 0, zdict=data[5000:9000]), 5000, 0)[0] == DATA_ERROR
    # block type 3; stored LEN / NLEN mismatch; an incomplete dynamic header
    assert inflate(amd, b"\x07\x00\x00", 16, 0)[0] == DATA_ERROR
    assert inflate(amd, b"\x01\x05\x00\x00\x00hello", 16, 0)[0] == DATA_ERROR
    assert inflate(amd, b"\x05", 16, 0)[0] == DATA_ERROR


def test_garbage_never_escapes_buffers(amd):
    """Seeded random bytes as raw streams: every call ends with one of the
    documented codes, and an OK result agrees with zlib on the same bytes."""
    rng = np.random.default_rng(123)
    for i in range(200):
        s = rng.integers(0, 256, int(rng.integers(1, 400)), dtype=np.uint8).tobytes()
        r, got = inflate(amd, s, 1 << 16, 0)
        assert r in (OK, DATA_ERROR, NEED_OUTPUT), (i, r)
        if r == OK:
            d = zlib.decompressobj(-15)
            assert got == d.decompress(s, 1 << 16) and d.eof


def test_stored_after_huffman(amd):
    """The stream orot's own inflater rejects (a fixed block, then a stored
    block whose first bytes sit in the bit buffer) inflates here, whole and
    fed one byte at a time."""
    a = corpus.generate("text", 1000, 5).tobytes()
    b = corpus.generate("random", 50, 6).tobytes()
    s1, bits = orc.encode_block_forced(np.frombuffer(a, dtype=np.uint8).astype(np.uint32), a, 1, is_last=0)
    buf = bytearray(s1 + bytes(64))
    buf[bits >> 3] |= 1 << (bits & 7)
    pos = ((bits + 3) + 7) & ~7
    buf[pos >> 3:(pos >> 3) + 4] = (50).to_bytes(2, "little") + (0xFFCD).to_bytes(2, "little")
    buf[(pos >> 3) + 4:(pos >> 3) + 54] = b
    s = bytes(buf[:(pos >> 3) + 54])
    assert zlib.decompress(s, -15) == a + b
    assert inflate(amd, s, len(a) + len(b), 0) == (OK, a + b)
    assert streamed(amd, s, 0, max_in=1, out_chunk=3)[:2] == (STREAM_END, a + b)


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_stream_leaves_trailing_input(amd, fmt):
    """Bytes after the end of the stream stay with the caller; later calls
    keep returning DEFLATE_STREAM_END."""
    data = corpus.generate("text", 50_000, 12).tobytes()
    s = zstream(data, 6, fmt) + b"TRAILING"
    buf = np.empty(len(data) + 100, dtype=np.uint8)
    with amd.Decompressor(fmt) as d:
        r, used, n = d.call(s, buf)
        assert r == STREAM_END and used == len(s) - 8 and buf[:n].tobytes() == data
        assert d.call(b"more", buf)[0] == STREAM_END
    out = np.empty(len(data), dtype=np.uint8)
    assert amd.decompress_into(s, out, fmt) == (OK, len(data))


def test_stream_errors_and_arguments(amd):
    L = amd.lib()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    p, n = vp(0), sz(0)
    assert L.inflate_stream_decompress(None, ctypes.byref(p), ctypes.byref(n), ctypes.byref(p),
                                       ctypes.byref(n)) == PARAM_ERROR
    assert not L.inflate_stream_new(9)
    buf = np.empty(100, dtype=np.uint8)
    with amd.Decompressor(1) as d:
        assert d.call(b"\x78\x9d", buf)[0] == DATA_ERROR  # FCHECK
        assert d.call(b"\x78\x9c", buf)[0] == DATA_ERROR  # stays failed
    with amd.Decompressor(2) as d:
        assert d.call(b"\x1f\x8b\x08\x00", buf)[0] == OK  # header not complete yet
        assert d.call(b"\x00" * 6 + b"\x07", buf)[0] == DATA_ERROR  # block type 3


def test_gzip_optional_header_fields(amd):
    """FEXTRA, FNAME, FCOMMENT and FHCRC are skipped (RFC 1952 2.3.1)."""
    data = corpus.generate("records", 30_000, 2).tobytes()
    raw = zstream(data, 6, 0)
    crc = zlib.crc32(data).to_bytes(4, "little") + len(data).to_bytes(4, "little")
    hdr = bytes([0x1F, 0x8B, 8, 0x04 | 0x08 | 0x10 | 0x02, 0, 0, 0, 0, 0, 3])
    hdr += (5).to_bytes(2, "little") + b"ABCDE" + b"name.txt\0" + b"a comment\0"
    hdr += (zlib.crc32(hdr) & 0xFFFF).to_bytes(2, "little")  # FHCRC
    s = hdr + raw + crc
    assert zlib.decompress(s, 31) == data
    assert inflate(amd, s, len(data), 2) == (OK, data)
    assert streamed(amd, s, 2, max_in=3)[:2] == (STREAM_END, data)


def test_allocator_hook():
    """deflate_set_allocator: the streaming decompressor's memory goes through
    the caller's allocator, every block is freed through it, and a NULL
    allocator is ignored (run in a child process: the hook is process-wide)."""
    code = r"""
import ctypes, sys, zlib
import orot_amd
from orot_amd import Allocator, ALLOC_FN, FREE_FN
libc = ctypes.CDLL(None)
libc.aligned_alloc.restype = ctypes.c_void_p
libc.aligned_alloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
libc.free.argtypes = [ctypes.c_void_p]
live, calls = set(), [0, 0]
def al(opaque, size, align):
    p = libc.aligned_alloc(align, (size + align - 1) // align * align)
    live.add(p); calls[0] += 1
    return p
def fr(opaque, p):
    live.discard(p); calls[1] += 1
    libc.free(p)
a = Allocator(ALLOC_FN(al), FREE_FN(fr), None)
L = orot_amd.lib()
L.deflate_set_allocator(ctypes.byref(a))
L.deflate_set_allocator(None)
data = bytes(range(256)) * 400
s = zlib.compress(data, 6)
with orot_amd.Decompressor(1) as d:
    out, done = d.feed(s[:1000])
    out2, done = d.feed(s[1000:])
assert done and out + out2 == data
assert orot_amd.decompress(s, 1) == data
sys.exit(0 if calls[0] >= 2 and calls[0] == calls[1] and not live else 1)
"""
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env={**os.environ, "PYTHONPATH": ROOT})
    assert r.returncode == 0


def test_cpp_wrappers_host_side():
    """orot's C++ wrappers by name (tests/cpp/wrappers.cpp, built with the
    library): the host half runs anywhere."""
    exe = os.path.join(ROOT, "tests", "cpp", "wrappers")
    assert os.path.exists(exe), "run python -m orot_amd.build"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr + r.stdout
