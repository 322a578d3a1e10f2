"""The benchmark configurations at their real sizes on the GPU, checked by
size-independent properties and orot's own output:
- the 256 MiB four-segment buffer at level 1, raw, 128 KiB chunks: SHA-256
  equal to orot's bytes for it (tests/golden/bench_digest.json, written by
  orot) and a zlib inflate;
- the 4096-buffer mix (4 KiB - 4 MiB, 2.53 GB) at level 1, zlib, as one
  graph-submitted batch: every buffer inflates, equals the same buffer
  compressed alone, and the total equals orot's compressed size for the same
  buffers when orot's build is present;
- 16 of the 256 level-6 buffers of 8 MiB (8 Zipf text, 8 records), gzip:
  the oracle's bytes, zlib with its CRC-32 trailer, and within 3% of orot's
  level-6 size."""
import hashlib
import json
import os
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def amd():
    import orot_amd

    orot_amd.lib()
    return orot_amd


@pytest.fixture(scope="module")
def bench():
    import sys

    sys.path.insert(0, ROOT)
    import bench as b

    return b


def _generate(bench, specs):
    """The buffers of `specs`, made by freshly spawned workers: this process
    already holds GPU state, which forked children must not inherit."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    workers = max(1, min(os.cpu_count() or 1, 16))
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(bench._gen, specs, chunksize=max(1, len(specs) // (4 * workers))))


def test_plain_256mib_is_orot_bytes(amd, bench):
    with open(os.path.join(ROOT, "tests", "golden", "bench_digest.json")) as f:
        golden = json.load(f)
    data = bench.workload()
    assert data.size == golden["input_bytes"]
    print("256 MiB plain path", flush=True)
    d_in = amd.DeviceBuffer.from_host(data)
    cap = amd.compress_bound(data.size, 0) + (1 << 20)
    d_out = amd.DeviceBuffer(cap)
    with amd.ParallelCompressor(1, 0, 0, 131072) as pc:
        n = pc.compress_device(d_in.ptr, data.size, d_out.ptr, cap)
    out = d_out.download(n).tobytes()
    assert n == golden["size"] and hashlib.sha256(out).hexdigest() == golden["sha256"]
    assert zlib.decompress(out, -15) == data.tobytes()


def test_batch_4096_mix(amd, bench):
    specs = bench.batch_specs()
    bufs = _generate(bench, specs)
    sizes = np.array([b.size for b in bufs], dtype=np.int64)
    in_off = np.zeros(len(bufs), dtype=np.int64)
    in_off[1:] = np.cumsum((sizes[:-1] + 255) // 256 * 256)
    caps = np.array([amd.compress_bound(int(k), 1) + 64 * (int(k) // 65535 + 4) + 4096 for k in sizes])
    out_off = np.zeros(len(bufs), dtype=np.int64)
    out_off[1:] = np.cumsum((caps[:-1] + 255) // 256 * 256)
    print("batch of", len(bufs), "buffers,", int(sizes.sum()), "bytes", flush=True)
    d_in = amd.DeviceBuffer(int(in_off[-1] + sizes[-1]) + 256)
    for b, o in zip(bufs, in_off):
        if b.size:
            d_in.upload(b, int(o))
    d_out = amd.DeviceBuffer(int(out_off[-1] + caps[-1]))
    items = amd.BatchCompressor.items(in_off, sizes, out_off, caps)
    with amd.BatchCompressor(1, 1, 0, amd.BATCH_GRAPH) as bc:
        got = bc.compress_device(d_in.ptr, d_out.ptr, items)
    assert np.all(got > 0)
    rng = np.random.default_rng(3)
    alone = set(int(i) for i in rng.choice(len(bufs), 64, replace=False))
    with amd.ParallelCompressor(1, 1, 0, 0) as pc:
        for i, b in enumerate(bufs):
            s = d_out.download(int(got[i]), int(out_off[i])).tobytes()
            assert zlib.decompress(s) == b.tobytes(), i
            if i in alone:
                assert s == pc.compress(b), i
    if orc.ref_available():
        total = 0
        for b in bufs:
            total += len(orc.ref_parallel_compress(b, 1, 1, 0, threads=8))
        assert int(got.sum()) == total


def test_l6_config_slice(amd, bench):
    specs = bench.l6_specs()[:16]
    bufs = _generate(bench, specs)
    print("L6 slice of", len(bufs), "x 8 MiB", flush=True)
    with amd.BatchCompressor(6, 2, 0, 0) as bc:
        outs = bc.compress(bufs)
    ours = orot = 0
    for (kind, n, seed), b, o in zip(specs, bufs, outs):
        assert zlib.decompress(o, 31) == b.tobytes()
        assert o == orc.gpu_parallel_compress(b, 6, 2, 0, threads=8)
        ours += len(o)
        if orc.ref_available():
            orot += len(orc.ref_parallel_compress(b, 6, 2, 0, threads=8))
    if orc.ref_available():
        assert ours <= 1.03 * orot, (ours, orot)
