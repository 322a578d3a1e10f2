"""Dictionary mode at levels 1-3 (deflate_parallel_set_dictionary, the
batch config's `dictionary`): each chunk after the first of its buffer
matches into the 32 KiB before it.  The HIP bytes equal the oracle's
restatement (whose token stream the CPU tests pin to a brute force of the
definition); with the mode off the bytes stay orot's."""
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
@pytest.mark.parametrize("block", [0, 65536, 300_000, 1 << 20])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_dictionary_mode_equals_oracle(amd, level, block, fmt):
    data = _mixed(3_000_001, level * 7 + fmt)
    with amd.ParallelCompressor(level, fmt, 0, block) as pc:
        pc.set_dictionary(True)
        got = pc.compress(data)
        pc.set_dictionary(False)
        plain = pc.compress(data)
    assert got == orc.l1_dict_compress(data, level, fmt, block, threads=8)
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()
    assert plain == orc.parallel_compress(data, level, fmt, block, threads=8)
    assert len(got) <= len(plain)


@pytest.mark.parametrize("n", [0, 1, 100, 65535, 65536, 65537, 100_000])
def test_dictionary_mode_small(amd, n):
    data = _mixed(n, n)
    with amd.ParallelCompressor(1, 2, 0, 32768) as pc:
        pc.set_dictionary(True)
        got = pc.compress(data)
    assert got == orc.l1_dict_compress(data, 1, 2, 32768)


def test_dictionary_mode_batch(amd):
    bufs = [_mixed(1000 + 97_001 * i, i) for i in range(16)]
    with amd.BatchCompressor(1, 1, 0, amd.BATCH_GRAPH) as bc:
        bc.set_config(65536, 12, 0, 0, 0, 1)
        assert bc.config()["dictionary"] == 1
        for _ in range(2):  # graph capture, then replay
            outs = bc.compress(bufs)
            for b, o in zip(bufs, outs):
                assert o == orc.l1_dict_compress(b, 1, 1, 65536)
