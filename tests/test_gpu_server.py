"""GPU tests of the compression server (include/orot/deflate_server.h):
every result is the single-buffer output of that request alone, whatever it
was batched with; admission waits for HBM budget; requests that arrive while
batches run join later batches."""
import threading
import time
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


def _requests(seed, count, max_size=2_000_000):
    sizes = corpus.log_uniform_sizes(count, 1024, max_size, seed)
    out = []
    for i, n in enumerate(sizes):
        kind = corpus.GENERATORS[(seed + i) % 4]
        level = (1, 6, 2, 4)[i % 4]
        fmt = (2, 1, 0)[i % 3]
        out.append((corpus.generate(kind, int(n), seed * 7919 + i), level, fmt))
    return out


def _want(buf, level, fmt):
    if level <= 3:
        return orc.parallel_compress(buf, level, fmt)
    return orc.gpu_parallel_compress(buf, level, fmt)


def _drain(srv, expect, timeout_s=120.0):
    done = {}
    t0 = time.time()
    while len(done) < expect and time.time() - t0 < timeout_s:
        for c in srv.poll(200):
            done[c[0]] = c
    assert len(done) == expect, srv.stats()
    return done


def test_results_equal_alone(amd):
    reqs = _requests(3, 40)
    offs = np.cumsum([0] + [(b.size + 255) // 256 * 256 for b, _, _ in reqs])
    d_in = amd.DeviceBuffer(int(offs[-1]) + 256)
    for (b, _, _), o in zip(reqs, offs):
        if b.size:
            d_in.upload(b, int(o))
    with amd.Server(hbm_budget=1 << 30) as srv:
        ids = [srv.submit_device(d_in.ptr + int(o), b.size, lv, fmt) for (b, lv, fmt), o in zip(reqs, offs)]
        done = _drain(srv, len(ids))
        for rid, (b, lv, fmt) in zip(ids, reqs):
            c = done[rid]
            assert c[3] == 0 and c[5] <= c[6] <= c[7]
            got = srv.read(rid, c[2])
            assert zlib.decompress(got, WBITS[fmt]) == b.tobytes()
            assert got == _want(b, lv, fmt), (b.size, lv, fmt)
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(ids) and st["pool_bytes"] == 0
    d_in.free()


def test_host_submit(amd):
    reqs = _requests(5, 12, 500_000)
    with amd.Server(hbm_budget=256 << 20) as srv:
        ids = [srv.submit(b, lv, fmt) for b, lv, fmt in reqs]
        done = _drain(srv, len(ids))
        for rid, (b, lv, fmt) in zip(ids, reqs):
            assert srv.read(rid, done[rid][2]) == _want(b, lv, fmt)
            srv.release(rid)


def test_admission_waits_for_budget(amd):
    """A budget that holds only a few outputs at a time: the queue advances as
    the client releases completed requests; a request larger than the whole
    budget is refused at submission."""
    reqs = [(corpus.generate("random", 1_000_000, 90 + i), 1, 2) for i in range(12)]
    with amd.Server(hbm_budget=4 << 20, lanes_per_level=1) as srv:
        with pytest.raises(amd.Error):
            srv.submit(np.zeros(8 << 20, dtype=np.uint8), 1, 2)
        ids = [srv.submit(b, lv, fmt) for b, lv, fmt in reqs]
        got = {}
        t0 = time.time()
        while len(got) < len(ids) and time.time() - t0 < 120:
            for c in srv.poll(200):
                assert c[3] == 0
                got[c[0]] = srv.read(c[0], c[2])
                srv.release(c[0])
            assert srv.stats()["pool_bytes"] <= 4 << 20
        assert len(got) == len(ids)
        st = srv.stats()
        assert st["batches"] >= 3  # never more than what fits in 4 MiB at once
        for rid, (b, lv, fmt) in zip(ids, reqs):
            assert got[rid] == _want(b, lv, fmt)


def test_arrivals_during_batches(amd):
    """Requests submitted from another thread while batches run are batched
    as lanes free up; each result still equals the request alone."""
    reqs = _requests(9, 60, 1_000_000)
    offs = np.cumsum([0] + [(b.size + 255) // 256 * 256 for b, _, _ in reqs])
    d_in = amd.DeviceBuffer(int(offs[-1]) + 256)
    for (b, _, _), o in zip(reqs, offs):
        if b.size:
            d_in.upload(b, int(o))
    ids = [None] * len(reqs)
    with amd.Server(hbm_budget=1 << 30) as srv:
        def producer():
            for i, ((b, lv, fmt), o) in enumerate(zip(reqs, offs)):
                ids[i] = srv.submit_device(d_in.ptr + int(o), b.size, lv, fmt)
                time.sleep(0.002)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        done = _drain(srv, len(reqs))
        for rid, (b, lv, fmt) in zip(ids, reqs):
            assert srv.read(rid, done[rid][2]) == _want(b, lv, fmt)
            srv.release(rid)
        assert srv.stats()["batches"] > 2
    d_in.free()


def test_pool_wraps_and_reclaims_out_of_order(amd):
    """A pool a few requests wide: allocations wrap around its end, the
    client releases completed requests in a shuffled order and holds some
    back; outputs that are held at the same time never overlap, the pool
    figure never exceeds the budget, and it returns to 0 at the end."""
    rng = np.random.default_rng(17)
    budget = 3 << 20
    reqs = []
    for i in range(80):
        n = int(rng.integers(1024, 400_000))
        reqs.append((corpus.generate(corpus.GENERATORS[i % 4], n, 300 + i), (1, 2)[i % 2], 2))
    with amd.Server(hbm_budget=budget, lanes_per_level=2) as srv:
        ids = [srv.submit(b, lv, fmt) for b, lv, fmt in reqs]
        index = {rid: i for i, rid in enumerate(ids)}
        held = {}
        got = {}
        t0 = time.time()
        while len(got) < len(ids) and time.time() - t0 < 120:
            for c in srv.poll(100):
                assert c[3] == 0
                got[c[0]] = srv.read(c[0], c[2])
                held[c[0]] = (srv.output_ptr(c[0]), c[2])
            spans = sorted(held.values())
            for (p0, n0), (p1, _) in zip(spans, spans[1:]):
                assert p0 + n0 <= p1, "outputs held at once overlap"
            # release a shuffled part of what is held, keep the rest a while
            keys = list(held)
            rng.shuffle(keys)
            for rid in keys[: max(1, len(keys) * 2 // 3)]:
                srv.release(rid)
                del held[rid]
            assert srv.stats()["pool_bytes"] <= budget
        assert len(got) == len(ids), srv.stats()
        for rid in list(held):
            srv.release(rid)
        st = srv.stats()
        assert st["pool_bytes"] == 0 and st["completed"] == len(ids) and st["queued"] == 0
        for rid, i in index.items():
            b, lv, fmt = reqs[i]
            assert got[rid] == _want(b, lv, fmt), (i, b.size, lv)


def test_pool_accounting_is_on_the_device(amd):
    """Completed outputs hold pool bytes until they are released; the figure
    is the device allocator's (it includes the host input's copy)."""
    b = corpus.generate("text", 300_000, 5)
    with amd.Server(hbm_budget=64 << 20) as srv:
        rid = srv.submit(b, 1, 2)
        done = _drain(srv, 1)
        assert done[rid][3] == 0
        used = srv.stats()["pool_bytes"]
        assert used >= 2 * 300_000
        assert srv.read(rid, done[rid][2]) == _want(b, 1, 2)
        srv.release(rid)
        assert srv.stats()["pool_bytes"] == 0


def test_lane_capacity_bounds_a_request(amd):
    """A request with more blocks than a lane's working memory holds is
    refused at submission; one that fits runs."""
    with amd.Server(hbm_budget=256 << 20, lane_arena_bytes=8 << 20) as srv:
        with pytest.raises(amd.Error):
            srv.submit(np.zeros(64 << 20, dtype=np.uint8), 1, 0)
        b = corpus.generate("records", 200_000, 8)
        rid = srv.submit(b, 1, 0)
        done = _drain(srv, 1)
        assert srv.read(rid, done[rid][2]) == _want(b, 1, 0)
        srv.release(rid)
