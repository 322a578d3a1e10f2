"""GPU inflate (deflate_batch_decompress_device): many zlib-made streams
in one launch against their inputs, with per-stream result codes (orot
src/api/deflate_api.cpp:138-215).  The host-buffer decompress API is tested
in test_host_inflate.py."""
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
