"""GPU parity of the batch entry points (include/orot/deflate_batch.h): every
buffer of a batch compresses to the same bytes as the oracle -- which is
orot's bytes at levels 0-3 -- alone and in any mix, with and without graph
replay, and inflates through zlib."""
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


def _mix(seed, sizes):
    kinds = corpus.GENERATORS
    return [corpus.generate(kinds[(seed + i) % 4], int(n), 1000 * seed + i) for i, n in enumerate(sizes)]


class Arena:
    """Device input / output arenas for a list of host buffers; input offsets
    start at `skew` (an odd skew exercises unaligned chunk starts)."""

    def __init__(self, amd, bufs, fmt, skew=0, cap_extra=4096):
        self.amd = amd
        self.bufs = bufs
        self.in_off, self.out_off, self.caps = [], [], []
        ip, op = skew, 0
        for b in bufs:
            self.in_off.append(ip)
            ip += b.size + 13
            cap = amd.compress_bound(b.size, fmt) + 6 * (b.size // 1024 + 4) + cap_extra
            self.out_off.append(op)
            self.caps.append(cap)
            op += cap + 7
        self.d_in = amd.DeviceBuffer(max(ip, 1))
        for b, o in zip(bufs, self.in_off):
            if b.size:
                self.d_in.upload(b, o)
        self.d_out = amd.DeviceBuffer(max(op, 1))

    def items(self, caps=None):
        return self.amd.BatchCompressor.items(self.in_off, [b.size for b in self.bufs], self.out_off,
                                              caps or self.caps)

    def outputs(self, sizes):
        return [self.d_out.download(int(s), self.out_off[i]).tobytes() if s else b"" for i, s in enumerate(sizes)]

    def free(self):
        self.d_in.free()
        self.d_out.free()


def _want(buf, level, fmt, block=0):
    if level <= 3:
        return orc.parallel_compress(buf, level, fmt, block)
    return orc.gpu_parallel_compress(buf, level, fmt, block)


@pytest.mark.parametrize("level,fmt", [(1, 1), (1, 2), (2, 0), (0, 2), (6, 2), (4, 1)])
def test_batch_matches_oracle(amd, level, fmt):
    sizes = [0, 1, 5, 100, 4096, 70000, 131072, 131073, 300001, 1_000_000]
    bufs = _mix(level + 3 * fmt, sizes)
    ar = Arena(amd, bufs, fmt)
    with amd.BatchCompressor(level, fmt) as bc:
        got = ar.outputs(bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
    for b, g in zip(bufs, got):
        assert g, amd.last_error()
        assert zlib.decompress(g, WBITS[fmt]) == b.tobytes()
        assert g == _want(b, level, fmt), (b.size, len(g))
    ar.free()


def test_alone_equals_in_mix(amd):
    """Each result is the same alone, in a mix, and through the single-buffer
    entry point."""
    bufs = _mix(7, [250_000, 3, 1_500_000, 64, 777_777, 0, 131072 * 3])
    ar = Arena(amd, bufs, 1, skew=3)
    with amd.BatchCompressor(1, 1) as bc:
        mixed = ar.outputs(bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
        for i, b in enumerate(bufs):
            one = amd.BatchCompressor.items([ar.in_off[i]], [b.size], [ar.out_off[i]], [ar.caps[i]])
            s = bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, one)
            assert ar.d_out.download(int(s[0]), ar.out_off[i]).tobytes() == mixed[i]
    with amd.ParallelCompressor(1, 1, 0, 0) as pc:
        for b, m in zip(bufs, mixed):
            assert pc.compress(b) == m
    ar.free()


@pytest.mark.parametrize("level,fmt", [(1, 1), (6, 2)])
def test_graph_replays_match_eager(amd, level, fmt):
    """A captured graph replays for changing batches; every replay matches the
    eager result, with eager work and synchronisations between replays."""
    pools = [_mix(11, [70000, 5, 400_000, 1_000_000]), _mix(12, [1, 250_000]), _mix(13, [999_999, 4096, 0])]
    with amd.BatchCompressor(level, fmt, 0, amd.BATCH_GRAPH) as bg, amd.BatchCompressor(level, fmt) as be:
        big = max(sum(b.size for b in p) for p in pools)
        ar_all = [Arena(amd, p, fmt) for p in pools]
        for rep in range(3):
            for ar in ar_all:
                eager = ar.outputs(be.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
                graph = ar.outputs(bg.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
                assert graph == eager, (rep, big)
                for b, g in zip(ar.bufs, graph):
                    assert zlib.decompress(g, WBITS[fmt]) == b.tobytes()
        for ar in ar_all:
            ar.free()


def test_overflow_is_per_item(amd):
    bufs = _mix(21, [10_000, 200_000, 50_000])
    ar = Arena(amd, bufs, 2)
    caps = list(ar.caps)
    caps[1] = 100  # cannot hold 200 KB of anything
    with amd.BatchCompressor(1, 2) as bc:
        sizes = bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items(caps))
    assert sizes[1] == 0 and sizes[0] > 0 and sizes[2] > 0
    got = ar.outputs(sizes)
    for i in (0, 2):
        assert got[i] == _want(bufs[i], 1, 2)
    ar.free()


def test_host_batch_api(amd):
    bufs = _mix(31, [0, 17, 300_000, 65536])
    with amd.BatchCompressor(6, 1) as bc:
        got = bc.compress(bufs)
    for b, g in zip(bufs, got):
        assert g == _want(b, 6, 1)


@pytest.mark.parametrize("level", [1, 6])
def test_autotune_output_valid(amd, level):
    bufs = _mix(41, [2_000_000, 1_000_000, 3_000_000, 500_000])
    ar = Arena(amd, bufs, 1)
    with amd.BatchCompressor(level, 1, 0, amd.BATCH_AUTOTUNE) as bc:
        sizes = bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items())
        cfg = bc.config()
        got = ar.outputs(sizes)
        # the tuned configuration is a valid one and is used again by the next batch
        again = ar.outputs(bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
    assert cfg["block_size"] in (65536, 131072, 262144, 524288, 1048576)
    assert got == again
    for b, g in zip(bufs, got):
        assert zlib.decompress(g, 15) == b.tobytes()
    ar.free()


def test_set_config_matches_single_path(amd):
    """An explicit configuration (64 KiB chunks, 14-bit hash at level 1) gives
    the bytes of the single-buffer path with the same chunk size at level 2,
    whose fast path is the same 14-bit one."""
    bufs = _mix(51, [500_000, 64, 131072])
    ar = Arena(amd, bufs, 0)
    with amd.BatchCompressor(1, 0) as bc:
        bc.set_config(65536, hash_bits=14)
        got = ar.outputs(bc.compress_device(ar.d_in.ptr, ar.d_out.ptr, ar.items()))
    for b, g in zip(bufs, got):
        assert g == orc.parallel_compress(b, 2, 0, 65536)
    ar.free()
