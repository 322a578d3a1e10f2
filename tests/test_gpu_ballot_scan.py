"""The ballot match-any variant of the candidate scans (L1-class candidate
kernel, L6-class predecessor kernel), which the engine takes where
same-address LDS exchanges of one wave instruction do not apply in lane
order.  OROT_AMD_BALLOT_SCAN=1 forces it for contexts created while it is
set; the bytes must equal the oracle's, as with the exchange variant."""
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

pytestmark = pytest.mark.gpu

WBITS = {0: -15, 1: 15, 2: 31}


@pytest.fixture()
def amd(monkeypatch):
    import orot_amd

    orot_amd.lib()
    monkeypatch.setenv("OROT_AMD_BALLOT_SCAN", "1")
    return orot_amd


def _mixed(n, seed):
    q = n // 4
    return np.concatenate([corpus.generate(k, q, seed + i) for i, k in enumerate(corpus.GENERATORS)])


# (level, block, input size): few chunks (six-wave L1 scan), many chunks
# (two-wave L1 scan), 14-bit L2/L3 tables, the L6 class with and without
# segmented predecessor tables
@pytest.mark.parametrize("level,block,n", [(1, 0, 3_000_000), (1, 16384, 8_000_000), (2, 0, 2_000_000),
                                           (3, 65536, 1_000_000), (6, 0, 3_000_000), (9, 262144, 2_000_000)])
@pytest.mark.parametrize("fmt", [0, 2])
def test_ballot_scan_equals_oracle(amd, level, block, n, fmt):
    data = _mixed(n, 10 * level + fmt)
    with amd.ParallelCompressor(level, fmt, 0, block) as pc:
        got = pc.compress(data)
    assert zlib.decompress(got, WBITS[fmt]) == data.tobytes()
    assert got == orc.gpu_parallel_compress(data, level, fmt, block, threads=8)


def test_ballot_scan_batch(amd):
    bufs = [corpus.generate(corpus.GENERATORS[i % 4], 1000 + 37_000 * i, i) for i in range(24)]
    with amd.BatchCompressor(1, 1, 0, 0) as bc:
        outs = bc.compress(bufs)
    for b, o in zip(bufs, outs):
        assert o == orc.gpu_parallel_compress(b, 1, 1)
