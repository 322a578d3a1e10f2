"""GPU inflate (deflate_decompress, deflate_batch_decompress_device) against
zlib, the library routine that pins RFC 1950-1952 decoding: streams made by
zlib at every level and strategy (stored, fixed and dynamic blocks, flush
points, empty blocks) and streams made by this library's compressor inflate
to their input; corrupt, truncated and oversized cases return orot's result
codes (src/api/deflate_api.cpp:138-215, src/formats/gzip_wrapper.cpp:76-138,
src/formats/zlib_wrapper.cpp)."""
import zlib

import numpy as np
import pytest

from orot_amd import corpus

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="GPU inflate switched off: its kernel did not finish on its first GPU run "
                            "(a stored-block stream) and no cause in the code has been shown; the same "
                            "kernel source passes tests/test_inflate_emu.py on the CPU"),
]

WBITS = {0: -15, 1: 15, 2: 31}
OK, NEED_OUTPUT, DATA_ERROR = 0, 3, -1


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


def inflate(amd, s: bytes, cap: int, fmt: int):
    out = np.empty(cap, dtype=np.uint8)
    r, n = amd.decompress_into(s, out, fmt)
    return r, out[:n].tobytes() if r == OK else b""


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


@pytest.mark.parametrize("level", [0, 1, 2, 4, 6, 9])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_own_streams(amd, level, fmt):
    data = np.concatenate([corpus.generate(k, 400_000, 3) for k in corpus.GENERATORS])
    with amd.ParallelCompressor(level, fmt, 0, 0) as pc:
        s = pc.compress(data)
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


def test_batch(amd):
    """Many streams in one launch, each with its own result: valid streams of
    every kind and level, one corrupt stream, one with too small an output."""
    rng = np.random.default_rng(7)
    datas, streams = [], []
    for i in range(300):
        n = int(np.exp(rng.uniform(np.log(1), np.log(600_000))))
        d = corpus.generate(corpus.GENERATORS[i % 4], n, 1000 + i).tobytes()
        datas.append(d)
        streams.append(zstream(d, (0, 1, 6, 9)[i % 4], 2))
    streams[17] = streams[17][:-3]                      # truncated trailer
    caps = [max(len(d), 1) for d in datas]
    caps[23] = max(0, len(datas[23]) - 1)               # output one byte short
    in_offs = np.cumsum([0] + [len(s) for s in streams])
    out_offs = np.cumsum([0] + caps)
    d_in = amd.DeviceBuffer.from_host(np.frombuffer(b"".join(streams) or b"\0", dtype=np.uint8))
    d_out = amd.DeviceBuffer(int(out_offs[-1]) + 1)
    items = [(int(in_offs[i]), len(streams[i]), int(out_offs[i]), caps[i]) for i in range(len(streams))]
    sizes, res = amd.batch_decompress_device(d_in.ptr, d_out.ptr, items, amd.Format.GZIP)
    for i, d in enumerate(datas):
        if i == 17:
            assert res[i] == DATA_ERROR
        elif i == 23 and len(d) > 0:
            assert res[i] == NEED_OUTPUT
        else:
            assert res[i] == OK and sizes[i] == len(d), i
            if d:
                assert d_out.download(len(d), int(out_offs[i])).tobytes() == d, i
    d_in.free()
    d_out.free()
