#!/usr/bin/env python3
"""Write the golden fixtures of this directory from orot itself.

Runs only where the orot tree is present and oracle/_ref/liborot_ref.so has
been built from it (`make -C oracle ref`): that library is orot's own
parallel compressor and LZ77 stage compiled from its sources, scalar kernels
(so gzip trailers carry the IEEE CRC-32).  The fixtures are data: inputs are
named by (generator, size, seed) of orot_amd.corpus, outputs are orot's bytes
(hex, small inputs), or their size and SHA-256 (large inputs), or the SHA-256
of orot's LZ77 token array (u32 little-endian, orot's Token layout).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import oracle_ffi as orc  # noqa: E402
from orot_amd import corpus  # noqa: E402

SMALL_SIZES = [0, 1, 2, 3, 4, 5, 7, 64, 300, 1000, 4096]
TOKEN_SIZES = [1000, 65536, 300000]
WBITS = {0: -15, 1: 15, 2: 31}


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def main() -> int:
    if not orc.ref_available():
        print("oracle/_ref/liborot_ref.so not built: run `make -C oracle ref` where the orot tree exists")
        return 1
    out = {"small": [], "large": [], "tokens": [], "levels_l6": []}
    # 1. exact payloads of small inputs, levels 0-3, every format, a few block sizes
    for kind in corpus.GENERATORS:
        for n in SMALL_SIZES:
            seed = 4242 + n
            data = corpus.generate(kind, n, seed)
            for level in (0, 1, 2, 3):
                for fmt in (0, 1, 2):
                    for block in (0, 64, 1024):
                        comp = orc.ref_parallel_compress(data, level, fmt, block)
                        out["small"].append({"kind": kind, "n": n, "seed": seed, "level": level, "format": fmt,
                                             "block": block, "hex": comp.hex(),
                                             "zlib_ok": zlib.decompress(comp, WBITS[fmt]) == data.tobytes()})
    # 2. size + digest of larger inputs (levels 0-3)
    for kind in corpus.GENERATORS:
        for n in (300_000, 2_000_000):
            seed = 777 + n
            data = corpus.generate(kind, n, seed)
            for level in (0, 1, 2, 3):
                for fmt, block in ((0, 0), (2, 0), (1, 65536), (2, 262144)):
                    if level >= 2 and block > 131072:
                        continue
                    comp = orc.ref_parallel_compress(data, level, fmt, block, threads=8)
                    out["large"].append({"kind": kind, "n": n, "seed": seed, "level": level, "format": fmt,
                                         "block": block, "size": len(comp), "sha256": sha(comp)})
    # the CPU reference configuration: 16 MiB Zipf text, L1, gzip, 128 KiB chunks
    text16 = corpus.generate("text", 16 << 20, 0xC0F1)
    comp = orc.ref_parallel_compress(text16, 1, 2, 131072, threads=8)
    out["text16_l1_gzip"] = {"kind": "text", "n": 16 << 20, "seed": 0xC0F1, "level": 1, "format": 2,
                             "block": 131072, "size": len(comp), "sha256": sha(comp),
                             "zlib_ok": zlib.decompress(comp, 31) == text16.tobytes()}
    # 3. LZ77 token streams, levels 1-9 (one chunk, fresh tables)
    for kind in corpus.GENERATORS:
        for n in TOKEN_SIZES:
            seed = 99 + n
            data = corpus.generate(kind, n, seed)
            for level in range(1, 10):
                t = orc.ref_lz77(data, level)
                out["tokens"].append({"kind": kind, "n": n, "seed": seed, "level": level, "count": int(t.size),
                                      "sha256": sha(t.astype("<u4").tobytes())})
    # 4. orot's sizes at levels 4-12 (ratio reference of the L6-class path) and
    #    whether zlib accepts orot's own output there
    for kind in corpus.GENERATORS:
        for n in (300_000, 4_000_000):
            seed = 555 + n
            data = corpus.generate(kind, n, seed)
            for level in range(4, 13):
                comp = orc.ref_parallel_compress(data, level, 0, 0, threads=8)
                try:
                    ok = zlib.decompress(comp, -15) == data.tobytes()
                except zlib.error:
                    ok = False
                out["levels_l6"].append({"kind": kind, "n": n, "seed": seed, "level": level, "format": 0,
                                         "block": 0, "size": len(comp), "zlib_ok": ok})
    # 5. orot's whole-buffer deflate_compress (one block per input) at levels 0-3
    out["whole"] = []
    for kind in corpus.GENERATORS + ("mixed",):
        for n in (0, 1, 5, 4096, 300_000, 2_000_000):
            seed = 888 + n
            if kind == "mixed":
                q = n // 4
                data = np.concatenate([corpus.generate(k, q if i < 3 else n - 3 * q, seed + i)
                                       for i, k in enumerate(corpus.GENERATORS)])
            else:
                data = corpus.generate(kind, n, seed)
            for level in (0, 1, 2, 3):
                for fmt in (0, 1, 2):
                    comp = orc.ref_compress(data, level, fmt)
                    e = {"kind": kind, "n": n, "seed": seed, "level": level, "format": fmt, "size": len(comp),
                         "sha256": sha(comp)}
                    if n <= 4096:
                        e["hex"] = comp.hex()
                    out["whole"].append(e)
    with open(os.path.join(HERE, "orot_fixtures.json"), "w") as f:
        json.dump(out, f, indent=0)
    # 6. the benchmark workload (256 MiB, four 64 MiB segments, L1 raw, 128 KiB)
    import bench  # noqa: E402

    data = bench.workload()
    comp = orc.ref_parallel_compress(data, 1, 0, 131072, threads=8)
    with open(os.path.join(HERE, "bench_digest.json"), "w") as f:
        json.dump({"size": len(comp), "sha256": sha(comp), "level": 1, "format": 0, "block": 131072,
                   "input_bytes": int(data.size)}, f)
    print("wrote", os.path.join(HERE, "orot_fixtures.json"), "and bench_digest.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
