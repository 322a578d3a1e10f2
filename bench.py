#!/usr/bin/env python3
"""Benchmark of the MI355X parallel DEFLATE compressor.

Default workload (one GPU): the batched path at level 1, zlib framing, on 4096
independent buffers with sizes log-uniform in [4 KiB, 4 MiB] (2.53 GB in
total) and content drawn uniformly from four generators -- all-zero bytes,
Zipf-word text, uniform random bytes and 64-byte records.  A step is one
graph-submitted batch of all 4096 buffers, resident in HBM, in the
configuration the autotuner keeps.  Every output is checked once (untimed) to
inflate back to its input through zlib, to equal that buffer compressed alone,
and its size is compared with orot's own output for the same buffer.  The
plain path (one deflate_parallel_compress_device call per buffer) is timed on
the same buffers, and so is the host-buffer path (deflate_batch_compress,
PCIe copies included).

Prints one JSON line with the metric, per-step latency percentiles, the
roofline of the dominant kernels (the L1 LZ stage), and orot's parallel CPU
path timed on this host on orot's own configuration (16 MiB Zipf text, L1,
gzip, 128 KiB chunks) with all host threads and with one.
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import re
import statistics
import sys
import time
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import orot_amd  # noqa: E402
from orot_amd import corpus  # noqa: E402

SEGMENT = 64 << 20
SEEDS = {"zeros": 0, "text": 0x5EED0001, "random": 0x5EED0002, "records": 0x5EED0003}
HBM_PEAK = 8.0e12
# rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the dominant kernel on the
# default workload, scaled by the same counters on a plain copy kernel.
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "l1_lz_stage_256MiB_pmc.txt")
BATCH_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "l1_lz_stage_batch_4096_pmc.txt")
L6_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "l6_kernels_256x8MiB_pmc.txt")
# the L1-class LZ stage at 12-bit hashes and chunks up to 128 KiB: candidate
# scan, then match / parse / emission
L1_STAGE = "l1_candidate_kernel + l1_barrier_kernel"


def measured_traffic(path: str = TRAFFIC_SUMMARY):
    """HBM bytes per launch from the committed counter summary (its line
    'traffic_bytes_per_launch: N'), or None when there is none."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def measured_issue(path: str = TRAFFIC_SUMMARY):
    """Issue-side ceilings of the dominant kernels from the committed counter
    summary: per kernel, the fraction of the VALU issue slots and of the LDS
    pipe its launch used (the summary's 'VALU x  LDS pipe y' lines, or its
    VALU_frac / LDS_frac columns), or None when there are none."""
    out = {}
    try:
        with open(path) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    cols = None
    for line in lines:
        m = re.match(r"\s+(\S+(?:, \S+)*)\s+t =\s+[\d.]+ us\s+VALU ([\d.]+)\s+LDS pipe ([\d.]+)", line)
        if m:
            out[m.group(1)] = {"valu": float(m.group(2)), "lds_pipe": float(m.group(3))}
            continue
        f = line.split()
        if f and f[0] == "kernel" and "VALU_frac" in f:
            cols = f
            continue
        if cols and len(f) == len(cols) and f[0].startswith("l6_match"):
            out[f[0]] = {"valu": float(f[cols.index("VALU_frac")]), "lds_pipe": float(f[cols.index("LDS_frac")])}
    if not out:
        return None
    return {"source": os.path.relpath(path, ROOT), "kernels": out}


def calibrate_copy(nbytes: int = 256 << 20, reps: int = 3) -> int:
    """A plain elementwise kernel that reads and writes nbytes each (bitwise
    NOT of a uint8 tensor): FETCH_SIZE / WRITE_SIZE of this kernel give the
    counter's scale against bytes known to move."""
    import torch

    a = torch.full((nbytes,), 0x5A, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    ok = bool((b[:4096] == 0xA5).all().item())
    print(json.dumps({"calibrate_copy_bytes": nbytes, "reps": reps, "ok": ok}), flush=True)
    return 0 if ok else 1


def workload() -> np.ndarray:
    parts = [corpus.generate(k, SEGMENT, SEEDS[k]) for k in ("zeros", "text", "random", "records")]
    return np.concatenate(parts)


def _gen(spec):
    kind, size, seed = spec
    return corpus.generate(kind, int(size), int(seed))


_BUFFERS = {}  # buffers generated ahead of time, by their specs


def make_buffers(specs):
    """Generate (kind, size, seed) buffers in a process pool.  Called before
    anything touches the GPU, so the forked workers hold no device state
    (buffers a later leg needs are generated up front by pregenerate)."""
    from concurrent.futures import ProcessPoolExecutor

    got = _BUFFERS.pop(tuple(specs), None)
    if got is not None:
        return got

    workers = max(1, min(cpu_threads(), 16))
    if workers == 1 or len(specs) < 4:
        return [_gen(s) for s in specs]
    with ProcessPoolExecutor(workers) as ex:
        return list(ex.map(_gen, specs, chunksize=max(1, len(specs) // (4 * workers))))


# Level-6 workload: 1 GiB Zipf-word text plus 1 GiB 64-byte records as
# 256 buffers of 8 MiB (even index text, odd index records), gzip.
L6_COUNT = 256
L6_SIZE = 8 << 20
L6_MASTER = 0x6C360000


def l6_specs():
    return [("text" if i % 2 == 0 else "records", L6_SIZE, corpus.buffer_seed(L6_MASTER, i))
            for i in range(L6_COUNT)]


def cpu_threads() -> int:
    n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit():
        n = min(n, int(share))
    return max(1, n)


class OrotCPU:
    """orot's parallel compressor on host cores: the reference build when
    present (kind "reference"), else the oracle's restatement (kind "port")."""

    def __init__(self, level: int, fmt: int, block: int, threads: int):
        ref_x86 = os.path.join(ROOT, "oracle", "_ref", "liborot_ref_x86.so")
        self.level, self.fmt, self.block, self.threads = level, fmt, block, threads
        self.ctx = None
        if os.path.exists(ref_x86):
            L = ctypes.CDLL(ref_x86)
            L.deflate_parallel_new.restype = ctypes.c_void_p
            L.deflate_parallel_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
            L.deflate_parallel_compress.restype = ctypes.c_size_t
            L.deflate_parallel_compress.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                    ctypes.c_void_p, ctypes.c_size_t]
            L.deflate_parallel_free.argtypes = [ctypes.c_void_p]
            self.L = L
            self.ctx = L.deflate_parallel_new(level, fmt, threads, block)
            self.kind = "reference"
        else:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_ffi  # noqa: E402

            self.O = oracle_ffi.oracle()
            self.kind = "port"

    def compress(self, data: np.ndarray, out: np.ndarray) -> int:
        if self.ctx is not None:
            return self.L.deflate_parallel_compress(self.ctx, data.ctypes.data, data.size, out.ctypes.data, out.size)
        return self.O.oracle_parallel_compress(self.level, self.fmt, self.threads, self.block, data.ctypes.data,
                                               data.size, out.ctypes.data, out.size)

    def close(self):
        if self.ctx is not None:
            self.L.deflate_parallel_free(self.ctx)
            self.ctx = None


def cpu_model() -> str:
    """The host CPU's model name as lscpu reports it."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return ""
    for line in out.splitlines():
        if line.startswith("Model name:"):
            return line.split(":", 1)[1].strip()
    return ""


def cpu_baseline(data: np.ndarray, level: int = 1, fmt: int = 0, chunk: int = 131072, sample: str = "",
                 budget_s: float = 15.0, threads: int = 0):
    """orot's parallel compressor on the host cores: median of up to 10 runs
    over the whole buffer after one warm-up."""
    threads = threads or cpu_threads()
    cpu = OrotCPU(level, fmt, chunk, threads)
    out = np.empty(data.size + data.size // 8 + 4096, dtype=np.uint8)
    cpu.compress(data, out)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 10 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        clen = cpu.compress(data, out)
        times.append(time.perf_counter() - t0)
    cpu.close()
    med = statistics.median(times)
    return {
        "value": round(data.size / med / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": cpu.kind,
        "sample": f"{sample}, median of {len(times)} runs after 1 warm-up ({clen} B out)",
    }


def cpu_config1() -> dict:
    """orot's parallel L1 on its own CPU configuration: one 16 MiB Zipf-word
    text buffer, gzip, 128 KiB chunks (128 chunks), with every host thread
    this process may use and with one thread."""
    data = corpus.generate("text", 16 << 20, 0xC0F1)
    sample = "orot parallel L1 gzip, 128 KiB chunks, one 16 MiB Zipf-word text buffer"
    rec = cpu_baseline(data, 1, 2, 131072, sample)
    one = cpu_baseline(data, 1, 2, 131072, sample, threads=1)
    rec["one_thread"] = {"value": one["value"], "unit": "GB/s", "cores": 1, "sample": one["sample"]}
    rec["cpu_model"] = cpu_model()
    rec["host_cpus"] = os.cpu_count()
    return rec


def cpu_multi(bufs, level: int, fmt: int, budget_s: float = 20.0):
    """orot's parallel compressor over a prefix of the buffers, one call per
    buffer, until budget_s has passed.  Returns (baseline record, per-buffer
    compressed sizes of the sample)."""
    threads = cpu_threads()
    cpu = OrotCPU(level, fmt, 0, threads)
    out = np.empty(max(b.size for b in bufs) * 9 // 8 + 4096, dtype=np.uint8)
    cpu.compress(bufs[0], out)  # warm-up
    sizes = []
    nbytes = 0
    t0 = time.perf_counter()
    while len(sizes) < len(bufs) and (time.perf_counter() - t0) < budget_s:
        b = bufs[len(sizes)]
        sizes.append(cpu.compress(b, out))
        nbytes += b.size
    dt = time.perf_counter() - t0
    cpu.close()
    rec = {
        "value": round(nbytes / dt / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": cpu.kind,
        "sample": f"orot parallel L{level} {['raw', 'zlib', 'gzip'][fmt]}, default chunks, the first "
                  f"{len(sizes)} of {len(bufs)} buffers ({nbytes >> 20} MiB), one call per buffer",
    }
    return rec, sizes


METRIC = "input GB/s at ratio within 3% of orot L1/L6 (bit-exact inflate); p50/p99 latency"


def pct(vals, q):
    v = sorted(vals)
    return v[min(len(v) - 1, int(len(v) * q))]


def base_record(args, value, ms_per_step, config):
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded zeros / Zipf text / random / records generators)",
        "config": config,
    }


def bench_single(args, data: np.ndarray, level: int, fmt: int, chunk: int, config: dict, golden: dict,
                 cpu_sample: str, traffic) -> dict:
    """One buffer resident in HBM; a step is one
    deflate_parallel_compress_device call over it."""
    n = data.size
    d_in = orot_amd.DeviceBuffer.from_host(data)
    cap = orot_amd.compress_bound(n, fmt) + (1 << 20)
    d_out = orot_amd.DeviceBuffer(cap)
    pc = orot_amd.ParallelCompressor(level, fmt, 0, chunk)
    pc.set_timing(True)

    # correctness (untimed): zlib inflate and orot's bytes
    clen = pc.compress_device(d_in.ptr, n, d_out.ptr, cap)
    if clen == 0:
        raise SystemExit(f"compress failed: {orot_amd.last_error()}")
    comp = d_out.download(clen).tobytes()
    inflated_ok = zlib.decompress(comp, {0: -15, 1: 15, 2: 31}[fmt]) == data.tobytes()
    digest = hashlib.sha256(comp).hexdigest()
    matches_orot = golden is not None and golden.get("sha256") == digest and golden.get("size") == clen
    del comp

    for _ in range(args.warmup):
        pc.compress_device(d_in.ptr, n, d_out.ptr, cap)
    orot_amd.lib().orot_gpu_synchronize()
    lat, lz = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        pc.compress_device(d_in.ptr, n, d_out.ptr, cap)
        lat.append(time.perf_counter() - s0)
        st = pc.stage_times()
        lz.append(st[1])
    orot_amd.lib().orot_gpu_synchronize()
    total = time.perf_counter() - t0
    lat_ms = [x * 1e3 for x in lat]
    t_kernel = statistics.mean(lz) * 1e-3
    b_alg = n + clen
    achieved = b_alg / t_kernel / 1e9

    rec = base_record(args, n * args.steps / total / 1e9, total / args.steps * 1e3, config)
    rec.update({
        "p50_ms": round(pct(lat_ms, 0.5), 3),
        "p99_ms": round(pct(lat_ms, 0.99), 3),
        "compressed_bytes": clen,
        "ratio": round(clen / n, 5),
        "inflate_ok": inflated_ok,
        "matches_orot_bytes": matches_orot,
        "stage_ms_last": [round(x, 4) for x in pc.stage_times()],
        "roofline": {
            "bound": "hbm",
            "kernel": L1_STAGE,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved * 1e9 / HBM_PEAK, 5),
            "traffic": traffic,
            "issue": measured_issue(),
        },
    })
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(data, level, fmt, chunk, cpu_sample)
    pc.close()
    d_in.free()
    d_out.free()
    return rec


def bench_l1(args) -> dict:
    """Default: one 256 MiB buffer (4 x 64 MiB segments), level 1, raw,
    128 KiB chunks."""
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "bench_digest.json")
    if os.path.exists(gpath):
        with open(gpath) as f:
            golden = json.load(f)
    data = workload()
    rec = bench_single(args, data, 1, 0, 131072, {
        "workload": "plain GPU path, L1 greedy, raw DEFLATE, 256 MiB = 4 x 64 MiB segments "
                    "(zeros, Zipf text, random, 64-byte records), 128 KiB chunks",
        "level": 1,
        "format": "raw",
        "chunk": 131072,
        "input_bytes": data.size,
    }, golden, f"orot parallel L1 raw, 128 KiB chunks, the full {data.size >> 20} MiB 4-segment buffer",
        measured_traffic())
    return rec


def bench_text16(args) -> dict:
    """orot's own CPU configuration: one 16 MiB Zipf-text buffer, level 1,
    gzip, 128 KiB chunks (orot's bytes for it are a golden fixture)."""
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "orot_fixtures.json")
    if os.path.exists(gpath):
        with open(gpath) as f:
            golden = json.load(f).get("text16_l1_gzip")
    data = corpus.generate("text", 16 << 20, 0xC0F1)
    return bench_single(args, data, 1, 2, 131072, {
        "workload": "orot's CPU configuration on the GPU: one 16 MiB Zipf-word text buffer, L1 greedy, gzip, "
                    "128 KiB chunks",
        "level": 1,
        "format": "gzip",
        "chunk": 131072,
        "input_bytes": data.size,
    }, golden, "orot parallel L1 gzip, 128 KiB chunks, the same 16 MiB text buffer", None)


# Streaming workload: the default 256 MiB buffer through deflate_stream_compress
# in 4 MiB writes, host memory to host memory, level 1, raw.
STREAM_WRITE = 4 << 20
OROT_STREAM_BLOCK = 131071  # largest input orot's stream writes as one valid final block


def cpu_stream_baseline(data: np.ndarray, budget_s: float = 15.0) -> dict:
    """orot's own streaming compressor on one core, over consecutive
    131071-byte streams of the buffer (its streams are valid only up to one
    128 KiB block), until budget_s has passed; the oracle's parallel path on
    the same pieces (the same bytes) when orot is not built here."""
    ref_x86 = os.path.join(ROOT, "oracle", "_ref", "liborot_ref_x86.so")
    out = np.empty(OROT_STREAM_BLOCK * 2 + 4096, dtype=np.uint8)
    P = ctypes.POINTER
    if os.path.exists(ref_x86):
        L = ctypes.CDLL(ref_x86)
        L.deflate_stream_new.restype = ctypes.c_void_p
        L.deflate_stream_new.argtypes = [ctypes.c_int, ctypes.c_int]
        L.deflate_stream_free.argtypes = [ctypes.c_void_p]
        L.deflate_stream_compress.restype = ctypes.c_int
        L.deflate_stream_compress.argtypes = [ctypes.c_void_p, P(ctypes.c_void_p), P(ctypes.c_size_t),
                                              P(ctypes.c_void_p), P(ctypes.c_size_t), ctypes.c_int]

        def one(piece):
            st = L.deflate_stream_new(1, 0)
            ni, ai = ctypes.c_void_p(piece.ctypes.data), ctypes.c_size_t(piece.size)
            no, ao = ctypes.c_void_p(out.ctypes.data), ctypes.c_size_t(out.size)
            L.deflate_stream_compress(st, ctypes.byref(ni), ctypes.byref(ai), ctypes.byref(no), ctypes.byref(ao), 3)
            L.deflate_stream_free(st)
            return out.size - ao.value
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi  # noqa: E402

        O = oracle_ffi.oracle()

        def one(piece):
            return O.oracle_parallel_compress(1, 0, 1, 131072, piece.ctypes.data, piece.size, out.ctypes.data,
                                              out.size)
        kind = "port"
    one(data[:OROT_STREAM_BLOCK])  # warm-up
    nbytes = clen = 0
    t0 = time.perf_counter()
    for p0 in range(0, data.size - OROT_STREAM_BLOCK + 1, OROT_STREAM_BLOCK):
        clen += one(data[p0:p0 + OROT_STREAM_BLOCK])
        nbytes += OROT_STREAM_BLOCK
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": round(nbytes / dt / 1e9, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": kind,
        "sample": f"orot deflate_stream_compress L1 raw, one stream per {OROT_STREAM_BLOCK} bytes of the 4-segment "
                  f"buffer, the first {nbytes >> 20} MiB ({clen} B out)",
    }


def bench_stream(args) -> dict:
    """A step is one whole stream: a fresh deflate_stream over the 256 MiB
    four-segment buffer, fed in 4 MiB writes from host memory, output into a
    host buffer, DEFLATE_FINISH at the end."""
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "bench_digest.json")
    if os.path.exists(gpath):
        with open(gpath) as f:
            golden = json.load(f)
    data = workload()
    n = data.size
    out = np.empty(orot_amd.compress_bound(n, 0) + (1 << 20), dtype=np.uint8)

    def one_stream() -> int:
        pos = 0
        with orot_amd.Compressor(1, orot_amd.Format.RAW) as c:
            for p0 in range(0, n, STREAM_WRITE):
                src = data[p0:p0 + STREAM_WRITE]
                used = 0
                while used < src.size:
                    r, u, w = c.compress_call(src[used:], out[pos:], orot_amd.Flush.NO_FLUSH)
                    used += u
                    pos += w
                    if r < 0:
                        raise SystemExit(f"stream failed: {orot_amd.last_error()}")
            while True:
                r, _, w = c.compress_call(b"", out[pos:], orot_amd.Flush.FINISH)
                pos += w
                if r == orot_amd.Result.STREAM_END:
                    return pos
                if r < 0:
                    raise SystemExit(f"stream failed: {orot_amd.last_error()}")

    clen = one_stream()
    comp = out[:clen].tobytes()
    inflated_ok = zlib.decompress(comp, -15) == data.tobytes()
    digest = hashlib.sha256(comp).hexdigest()
    matches_orot = golden is not None and golden.get("sha256") == digest and golden.get("size") == clen
    del comp
    for _ in range(args.warmup):
        one_stream()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        one_stream()
        lat.append(time.perf_counter() - s0)
    total = time.perf_counter() - t0
    lat_ms = [x * 1e3 for x in lat]
    rec = base_record(args, n * args.steps / total / 1e9, total / args.steps * 1e3, {
        "workload": "streaming compressor: the 256 MiB 4-segment buffer (zeros, Zipf text, random, 64-byte "
                    "records) through deflate_stream_compress in 4 MiB writes, host memory to host memory, "
                    "L1 greedy, raw DEFLATE, 32 MiB GPU pieces of 128 KiB chunks",
        "level": 1,
        "format": "raw",
        "chunk": 131072,
        "input_bytes": n,
    })
    rec.update({
        "p50_ms": round(pct(lat_ms, 0.5), 3),
        "p99_ms": round(pct(lat_ms, 0.99), 3),
        "compressed_bytes": clen,
        "ratio": round(clen / n, 5),
        "inflate_ok": inflated_ok,
        "matches_orot_bytes": matches_orot,
        # the kernels are the default workload's; a stream is bound by the
        # host copies and PCIe around them
        "roofline": None,
    })
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_stream_baseline(data)
    return rec


# Batch workload: 4096 buffers, sizes log-uniform in [4 KiB, 4 MiB], content
# type uniform over the four generators, level 1, zlib.
BATCH_COUNT = 4096
BATCH_LO = 4 << 10
BATCH_HI = 4 << 20
BATCH_MASTER = 0xB47C0003


def batch_specs():
    sizes = corpus.log_uniform_sizes(BATCH_COUNT, BATCH_LO, BATCH_HI, BATCH_MASTER)
    kinds = corpus.stream(BATCH_MASTER ^ 0x4B1D, BATCH_COUNT) % np.uint64(4)
    return [(corpus.GENERATORS[int(k)], int(n), corpus.buffer_seed(BATCH_MASTER, i))
            for i, (k, n) in enumerate(zip(kinds, sizes))]


def host_batch_seconds(bufs, level: int, fmt: int, reps: int = 2) -> float:
    """Best of `reps` deflate_batch_compress calls over the host buffers
    (H2D copies, the batch, D2H copies of the outputs)."""
    n = len(bufs)
    outs = [np.empty(orot_amd.compress_bound(b.size, 2) + 64 * (b.size // 65535 + 4), dtype=np.uint8) for b in bufs]
    in_p = (ctypes.c_void_p * n)(*[b.ctypes.data for b in bufs])
    out_p = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    in_s = (ctypes.c_size_t * n)(*[b.size for b in bufs])
    out_c = (ctypes.c_size_t * n)(*[o.size for o in outs])
    got = (ctypes.c_size_t * n)()
    best = None
    with orot_amd.BatchCompressor(level, fmt, 0, 0) as hb:
        for _ in range(reps + 1):  # the first call grows the arenas
            t0 = time.perf_counter()
            if orot_amd.lib().deflate_batch_compress(hb._ctx, in_p, in_s, out_p, out_c, got, n) != 0:
                raise SystemExit(f"deflate_batch_compress: {orot_amd.last_error()}")
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
    return best


def batch_only(args, level: int, fmt: int, d_in, d_out, items, n: int) -> dict:
    """Profiling run: only the graph-submitted batch in the default
    configuration (no plain path, autotuner or CPU leg), warm-up then steps."""
    with orot_amd.BatchCompressor(level, fmt, 0, orot_amd.BATCH_GRAPH) as bc:
        for _ in range(args.warmup):
            bc.compress_device(d_in.ptr, d_out.ptr, items)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out_sizes = bc.compress_device(d_in.ptr, d_out.ptr, items)
        total = time.perf_counter() - t0
    rec = base_record(args, n * args.steps / total / 1e9, total / args.steps * 1e3,
                      {"workload": "batch only (profiling)", "level": level, "buffers": len(items),
                       "input_bytes": n})
    rec["compressed_bytes"] = int(np.sum(out_sizes))
    d_in.free()
    d_out.free()
    return rec


def bench_multi(args, specs, level: int, fmt: int, workload: str, cpu_kinds, traffic_summary=None) -> dict:
    """Many buffers: the plain path (one deflate_parallel_compress_device call
    per buffer, orot's default chunks) and the batch path (one submission of
    every buffer: graph replay, autotuned configuration).  A step is one
    submission of the whole batch."""
    bufs = make_buffers(specs)
    sizes = np.array([b.size for b in bufs], dtype=np.int64)
    n = int(sizes.sum())
    in_off = np.zeros(len(bufs), dtype=np.int64)
    in_off[1:] = np.cumsum((sizes[:-1] + 255) // 256 * 256)
    caps = np.array([orot_amd.compress_bound(int(k), fmt) + 64 * (int(k) // 65535 + 4) + 4096 for k in sizes],
                    dtype=np.int64)
    out_off = np.zeros(len(bufs), dtype=np.int64)
    out_off[1:] = np.cumsum((caps[:-1] + 255) // 256 * 256)
    d_in = orot_amd.DeviceBuffer(int(in_off[-1] + sizes[-1]) + 256)
    for b, o in zip(bufs, in_off):
        if b.size:
            d_in.upload(b, int(o))
    d_out = orot_amd.DeviceBuffer(int(out_off[-1] + caps[-1]))
    items = orot_amd.BatchCompressor.items(in_off, sizes, out_off, caps)

    def digests(out_sizes):
        return [hashlib.sha256(d_out.download(int(k), int(o)).tobytes()).digest() for k, o in zip(out_sizes, out_off)]

    if args.batch_only:
        return batch_only(args, level, fmt, d_in, d_out, items, n)

    # plain path: one call per buffer, orot's default configuration
    pc = orot_amd.ParallelCompressor(level, fmt, 0, 0)

    def plain_pass():
        out = []
        for i in range(len(bufs)):
            k = pc.compress_device(d_in.ptr + int(in_off[i]), int(sizes[i]), d_out.ptr + int(out_off[i]),
                                   int(caps[i]))
            if k == 0:
                raise SystemExit(f"compress failed on buffer {i}: {orot_amd.last_error()}")
            out.append(k)
        return out

    plain_sizes = plain_pass()
    ok = True
    wb = {0: -15, 1: 15, 2: 31}[fmt]
    for i, b in enumerate(bufs):
        ok = ok and zlib.decompress(d_out.download(plain_sizes[i], int(out_off[i])).tobytes(), wb) == b.tobytes()
    plain_dig = digests(plain_sizes)
    plain_runs = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(plain_runs):
        plain_pass()
    plain_t = (time.perf_counter() - t0) / plain_runs
    pc.close()

    # batch path, default configuration: every buffer's bytes equal the plain path's
    with orot_amd.BatchCompressor(level, fmt, 0, 0) as bd:
        mix_equal = digests(bd.compress_device(d_in.ptr, d_out.ptr, items)) == plain_dig

    # batch path, autotuned, graph-submitted
    bc = orot_amd.BatchCompressor(level, fmt, 0, orot_amd.BATCH_GRAPH | orot_amd.BATCH_AUTOTUNE)
    t_tune = time.perf_counter()
    out_sizes = bc.compress_device(d_in.ptr, d_out.ptr, items)  # tunes on a prefix, then runs
    t_tune = time.perf_counter() - t_tune
    tuned = bc.config()
    for i, b in enumerate(bufs):
        if out_sizes[i] == 0:
            raise SystemExit(f"batch buffer {i} did not fit")
        ok = ok and zlib.decompress(d_out.download(int(out_sizes[i]), int(out_off[i])).tobytes(), wb) == b.tobytes()
    for _ in range(args.warmup):
        bc.compress_device(d_in.ptr, d_out.ptr, items)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        bc.compress_device(d_in.ptr, d_out.ptr, items)
        lat.append((time.perf_counter() - s0) * 1e3)
    total = time.perf_counter() - t0
    bc.close()
    clen = int(np.sum(out_sizes))
    plain_clen = int(sum(plain_sizes))

    # host buffers in and out: deflate_batch_compress, PCIe copies included
    host_t = host_batch_seconds(bufs, level, fmt, reps=2)

    # stage times of the tuned configuration (direct launches, HIP events)
    with orot_amd.BatchCompressor(level, fmt, 0, 0) as bt:
        bt.set_config(tuned["block_size"], tuned["hash_bits"], tuned["max_chain"], tuned["nice_len"],
                      tuned["lazy_depth"])
        bt.set_timing(True)
        bt.compress_device(d_in.ptr, d_out.ptr, items)
        stages = bt.stage_times()
    front = stages[1] * 1e-3
    b_alg = n + clen
    kernel = L1_STAGE if level <= 3 else "l6 front (pred + match + decide + parse + block analysis)"
    if level <= 3 and tuned["block_size"] > 131072:
        kernel = "l1_lz77_kernel"
    if level <= 3 and tuned["hash_bits"] == 14:
        kernel = "l1_chunk_kernel"

    rec = base_record(args, n * args.steps / total / 1e9, total / args.steps * 1e3, {
        "workload": workload,
        "level": level,
        "format": ["raw", "zlib", "gzip"][fmt],
        "buffers": len(bufs),
        "input_bytes": n,
        "tuned": tuned,
    })
    rec.update({
        "p50_ms": round(pct(lat, 0.5), 3),
        "p99_ms": round(pct(lat, 0.99), 3),
        "compressed_bytes": clen,
        "ratio": round(clen / n, 5),
        "inflate_ok": ok,
        "plain_path_gbps": round(n / plain_t / 1e9, 3),
        "speedup_vs_plain": round((n * args.steps / total) / (n / plain_t), 2),
        "alone_equals_in_mix": mix_equal,
        "size_vs_default_config": round(clen / plain_clen, 5),
        "autotune_s": round(t_tune, 3),
        "host_path_gbps": round(n / host_t / 1e9, 3),
        "stage_ms_tuned": [round(x, 4) for x in stages],
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(b_alg / front / 1e9, 2),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(b_alg / front / HBM_PEAK, 5),
            "traffic": measured_traffic(traffic_summary) if traffic_summary else None,
            "issue": measured_issue(traffic_summary) if traffic_summary else None,
        },
    })
    if not args.no_cpu:
        base, orot_sizes = cpu_multi(bufs, level, fmt)
        rec["cpu_baseline"] = cpu_config1() if level <= 3 else base
        rec["cpu_batch"] = base
        m = len(orot_sizes)
        by_kind = {}
        for i in range(m):
            k = cpu_kinds(i)
            g, o = by_kind.get(k, (0, 0))
            by_kind[k] = (g + int(out_sizes[i]), o + orot_sizes[i])
        rec["ratio_vs_orot"] = {
            "buffers": m,
            "gpu_over_orot": round(int(np.sum(out_sizes[:m])) / max(1, sum(orot_sizes)), 5),
            "per_kind": {k: round(g / max(1, o), 5) for k, (g, o) in by_kind.items()},
        }
    d_in.free()
    d_out.free()
    return rec


def bench_l6(args) -> dict:
    """Level 6, gzip: 256 x 8 MiB buffers (text / records)."""
    specs = l6_specs()
    return bench_multi(args, specs, 6, 2,
                       "L6-class lazy parse, dynamic Huffman blocks (15-bit codes, <= 16384 symbols), gzip, "
                       "256 x 8 MiB buffers (128 Zipf text + 128 64-byte records), batched",
                       lambda i: specs[i][0], L6_TRAFFIC_SUMMARY)


def bench_batch(args) -> dict:
    """Level 1, zlib: 4096 buffers of 4 KiB - 4 MiB, mixed content."""
    specs = batch_specs()
    rec = bench_multi(args, specs, 1, 1,
                      "batched L1-class greedy, zlib, 4096 buffers log-uniform in [4 KiB, 4 MiB], content "
                      "uniform over zeros / Zipf text / random / records, one graph-submitted batch, autotuned",
                      lambda i: specs[i][0], BATCH_TRAFFIC_SUMMARY)
    if not args.batch_only and not args.no_plain256:
        # the plain-path line of the 256 MiB four-segment buffer (L1, raw), no CPU leg
        sub = argparse.Namespace(**vars(args))
        sub.no_cpu = True
        sub.steps = min(args.steps, 10)
        p = bench_l1(sub)
        rec["plain_256MiB"] = {k: p[k] for k in ("value", "ms_per_step", "p50_ms", "p99_ms", "compressed_bytes",
                                                 "inflate_ok", "matches_orot_bytes", "roofline")}
    if not args.batch_only and not args.no_extras:
        # the level-6 batch (orot's ratio at L6 from its host path over a sample)
        sub = argparse.Namespace(**vars(args))
        sub.steps = min(args.steps, 5)
        sub.warmup = min(args.warmup, 1)
        l6 = bench_l6(sub)
        rec["l6_batch"] = {"workload": l6["config"]["workload"], "steps": sub.steps}
        rec["l6_batch"].update({k: l6[k] for k in (
            "value", "unit", "ms_per_step", "p50_ms", "p99_ms", "compressed_bytes", "ratio", "inflate_ok",
            "plain_path_gbps", "speedup_vs_plain", "alone_equals_in_mix", "roofline") if k in l6})
        if "ratio_vs_orot" in l6:
            rec["l6_batch"]["ratio_vs_orot"] = l6["ratio_vs_orot"]
            rec["l6_batch"]["cpu_batch"] = l6.get("cpu_batch")
        # a shorter run of the serving stream
        sub = argparse.Namespace(**vars(args))
        sub.duration = EXTRA_SERVE_S
        sub.no_cpu = True
        sv = bench_serve(sub)
        rec["serving"] = {"workload": sv["config"]["workload"]}
        rec["serving"].update({k: sv[k] for k in (
            "value", "unit", "offered_gbps", "completed", "completed_ok", "p50_ms", "p99_ms", "p999_ms",
            "latency_by_level_ms", "queue_p50_ms", "queue_p99_ms", "inflate_ok", "outputs_checked", "server")})
    return rec


# Serving workload: Poisson arrivals, sizes log-uniform in [1 KiB, 8 MiB],
# content uniform over the four generators, level uniform over {1, 6}, gzip,
# 32 GiB HBM budget.
SERVE_RATE = 2000.0
EXTRA_SERVE_S = 20.0  # the serving run inside the default line
SERVE_LO = 1 << 10
SERVE_HI = 8 << 20
SERVE_MASTER = 0x5E7E0004


def serve_requests(duration: float):
    count = int(SERVE_RATE * duration)
    r = corpus.stream(SERVE_MASTER, 4 * count)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    gaps = -np.log(1.0 - u[:count]) / SERVE_RATE
    arrivals = np.cumsum(gaps)
    sizes = corpus.log_uniform_sizes(count, SERVE_LO, SERVE_HI, SERVE_MASTER ^ 0x51)
    kinds = (r[count:2 * count] % np.uint64(4)).astype(np.int64)
    levels = np.where((r[2 * count:3 * count] & np.uint64(1)) == 0, 1, 6)
    offs = (u[3 * count:] * (SEGMENT - sizes)).astype(np.int64)
    return arrivals, sizes, kinds, levels, offs


def bench_serve(args) -> dict:
    """Open-loop serving: a producer thread submits each request at its
    arrival time (inputs are slices of the device-resident four-segment
    corpus); the main thread polls completions, checks a sample of outputs
    and releases every output."""
    import threading

    data = workload()
    d_in = orot_amd.DeviceBuffer.from_host(data)
    arrivals, sizes, kinds, levels, offs = serve_requests(args.duration)
    count = len(sizes)
    srv = orot_amd.Server(hbm_budget=32 << 30)
    ids = np.zeros(count, dtype=np.uint64)
    index = {}
    t_start = time.perf_counter()

    def producer():
        for i in range(count):
            wait = arrivals[i] - (time.perf_counter() - t_start)
            if wait > 0.0005:
                time.sleep(wait - 0.0003)
            while time.perf_counter() - t_start < arrivals[i]:
                pass
            src = int(kinds[i]) * SEGMENT + int(offs[i])
            rid = srv.submit_device(d_in.ptr + src, int(sizes[i]), int(levels[i]), orot_amd.Format.GZIP)
            ids[i] = rid
            index[rid] = i

    th = threading.Thread(target=producer)
    th.start()
    lat, queue_ms, done_bytes = [], [], 0
    lat_by_level = {1: [], 6: []}
    checked, ok = 0, True
    first_ns, last_ns = None, 0
    got = 0
    deadline = time.perf_counter() + args.duration + 120
    while got < count and time.perf_counter() < deadline:
        for c in srv.poll(100):
            rid, n_in, n_out, status, lv, t_sub, t_launch, t_done = c
            got += 1
            if status != 0:
                ok = False
                continue
            lat.append((t_done - t_sub) / 1e6)
            lat_by_level.setdefault(lv, []).append(lat[-1])
            queue_ms.append((t_launch - t_sub) / 1e6)
            done_bytes += n_in
            first_ns = t_sub if first_ns is None else min(first_ns, t_sub)
            last_ns = max(last_ns, t_done)
            while rid not in index:  # the producer records the id right after submit returns
                time.sleep(0)
            i = index[rid]
            if i % 64 == 0:
                comp = srv.read(rid, n_out)
                src = int(kinds[i]) * SEGMENT + int(offs[i])
                ok = ok and zlib.decompress(comp, 31) == data[src:src + int(sizes[i])].tobytes()
                checked += 1
            srv.release(rid)
    th.join()
    stats = srv.stats()
    srv.close()
    span = (last_ns - (first_ns or 0)) / 1e9
    offered = float(np.sum(sizes)) / float(arrivals[-1])
    rec = base_record(args, done_bytes / span / 1e9 if span > 0 else 0.0, span / max(1, got) * 1e3, {
        "workload": f"serving: Poisson arrivals at {SERVE_RATE:.0f} req/s for {args.duration:g} s, sizes "
                    "log-uniform in [1 KiB, 8 MiB], content uniform over zeros / Zipf text / random / records, "
                    "levels uniform over {1, 6}, gzip, 32 GiB HBM budget, continuous batching",
        "requests": count,
        "format": "gzip",
    })
    rec.pop("steps")
    rec.pop("warmup")
    rec["ms_per_step"] = None
    lat_s = sorted(lat)
    rec.update({
        "completed": got,
        "completed_ok": len(lat),
        "offered_gbps": round(offered / 1e9, 3),
        "p50_ms": round(pct(lat_s, 0.5), 3),
        "p99_ms": round(pct(lat_s, 0.99), 3),
        "p999_ms": round(pct(lat_s, 0.999), 3),
        "latency_by_level_ms": {str(k): {"p50": round(pct(sorted(v), 0.5), 3), "p99": round(pct(sorted(v), 0.99), 3)}
                                for k, v in lat_by_level.items() if v},
        "queue_p50_ms": round(pct(sorted(queue_ms), 0.5), 3),
        "queue_p99_ms": round(pct(sorted(queue_ms), 0.99), 3),
        "inflate_ok": ok,
        "outputs_checked": checked,
        "server": stats,
    })
    if not args.no_cpu:
        # orot on the host cores over a sample of the same request mix
        threads = cpu_threads()
        cpus = {1: OrotCPU(1, 2, 0, threads), 6: OrotCPU(6, 2, 0, threads)}
        out = np.empty(int(SERVE_HI) * 9 // 8 + 4096, dtype=np.uint8)
        nb, t0, k = 0, time.perf_counter(), 0
        while k < count and time.perf_counter() - t0 < 20.0:
            src = int(kinds[k]) * SEGMENT + int(offs[k])
            cpus[int(levels[k])].compress(data[src:src + int(sizes[k])], out)
            nb += int(sizes[k])
            k += 1
        dt = time.perf_counter() - t0
        for c in cpus.values():
            c.close()
        rec["cpu_baseline"] = {
            "value": round(nb / dt / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": cpus[1].kind,
            "sample": f"orot parallel L1/L6 gzip, one call per request, the first {k} requests of the same "
                      f"stream back to back ({nb >> 20} MiB)",
        }
    d_in.free()
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--workload", choices=("l1", "text16", "l6", "batch", "serve", "stream"), default="batch",
                    help="batch: 4096 mixed buffers of 4 KiB - 4 MiB at level 1, zlib, batched (default); "
                         "l1: 256 MiB four-segment buffer at level 1, raw, one plain call; "
                         "text16: 16 MiB Zipf text at level 1, gzip (orot's own CPU configuration); "
                         "l6: 256 x 8 MiB text/records buffers at level 6, gzip, batched; "
                         "serve: Poisson request stream through the continuous-batching server; "
                         "stream: the l1 buffer through the streaming compressor, host to host, 4 MiB writes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch / l6: time only the graph-submitted batch (for profilers)")
    ap.add_argument("--no-extras", action="store_true",
                    help="batch: skip the level-6 batch and the serving run added to the default line")
    ap.add_argument("--no-plain256", action="store_true",
                    help="batch: skip the extra plain-path line on the 256 MiB four-segment buffer")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to scale HBM counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1). Across GPUs a batch would split into "
              "independent sub-batches, one per GPU, with no exchange between them.", file=sys.stderr)
        return 2
    if args.workload == "batch" and not args.batch_only and not args.no_extras:
        specs = l6_specs()
        _BUFFERS[tuple(specs)] = make_buffers(specs)
    rec = {"l1": bench_l1, "text16": bench_text16, "l6": bench_l6, "batch": bench_batch,
           "serve": bench_serve, "stream": bench_stream}[args.workload](args)
    try:
        rec["device"] = orot_amd.device_info()[0]
    except Exception:
        pass
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
