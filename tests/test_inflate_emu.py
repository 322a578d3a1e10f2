"""CPU check of the inflate kernel's source (orot_amd/csrc/inflate.hip)
against zlib, the library routine that pins RFC 1950-1952 decoding.  The
unchanged device code is compiled for the host with tests/inflate_emu, at
-O3 with AddressSanitizer and UndefinedBehaviorSanitizer, and run in its
launched shape: workgroups of four waves sharing the stream queue, each wave
64 host threads whose ballots, shuffles and wave fences meet at a barrier of
the wave (the lockstep of a wave's lanes), with the LDS and the output
buffers filled with a non-zero pattern first.  It decodes zlib-made streams
of every level, strategy and framing, this library's own streams, batches,
and truncated, corrupt and oversized cases, against zlib's output and orot's
result codes (src/formats/gzip_wrapper.cpp:76-138, zlib_wrapper.cpp).  This
checks the algorithm as written, not the GPU code generation."""
import os
import shutil
import subprocess
import zlib

import pytest

from orot_amd import corpus

HERE = os.path.dirname(os.path.abspath(__file__))
CLANG = "/opt/rocm/llvm/bin/clang++"
WBITS = {0: -15, 1: 15, 2: 31}
OK, NEED_OUTPUT, DATA_ERROR = 0, 3, -1


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("clang++ of the ROCm toolchain is not installed")
    d = tmp_path_factory.mktemp("inflate_emu")
    exe = str(d / "emu")
    subprocess.run([CLANG, "-std=c++20", "-O3", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", os.path.join(HERE, "inflate_emu", "shim"),
                    "-I", os.path.join(HERE, "..", "include"), "-o", exe,
                    os.path.join(HERE, "inflate_emu", "emu.cpp"), "-lpthread"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

    def run_batch(jobs, fmt: int):
        """jobs: (stream, capacity) pairs, one launch; [(status, output)]."""
        args = [exe, str(fmt), str(len(jobs))]
        for i, (stream, cap) in enumerate(jobs):
            src = d / f"in{i}.bin"
            src.write_bytes(stream)
            args += [str(src), str(cap), str(d / f"out{i}.bin")]
        r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
        got = []
        for i in range(len(jobs)):
            st = int(r.stdout.split(f"item {i}: status")[1].split()[0])
            got.append((st, (d / f"out{i}.bin").read_bytes() if st == OK else b""))
        return got

    def run(stream: bytes, fmt: int, cap: int):
        return run_batch([(stream, cap)], fmt)[0]

    run.batch = run_batch
    yield run
    shutil.rmtree(d, ignore_errors=True)


def zstream(data: bytes, level: int, fmt: int, strategy=zlib.Z_DEFAULT_STRATEGY):
    co = zlib.compressobj(level, zlib.DEFLATED, WBITS[fmt], 9, strategy)
    return co.compress(data) + co.flush()


@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_levels_and_framings(emu, kind):
    data = corpus.generate(kind, 60_000, 3).tobytes()
    for level in (0, 1, 6, 9):
        for fmt in (0, 1, 2):
            st, got = emu(zstream(data, level, fmt), fmt, len(data))
            assert st == OK and got == data, (level, fmt)


def test_strategies_and_sizes(emu):
    text = corpus.generate("text", 50_000, 5).tobytes()
    for strategy in (zlib.Z_FIXED, zlib.Z_HUFFMAN_ONLY, zlib.Z_RLE, zlib.Z_FILTERED):
        st, got = emu(zstream(text, 6, 0, strategy), 0, len(text))
        assert st == OK and got == text, strategy
    for n in (0, 1, 3, 65535, 65536, 65537):
        data = corpus.generate("random", n, n + 1).tobytes() if n else b""
        for level in (0, 9):
            st, got = emu(zstream(data, level, 2), 2, max(n, 1))
            assert st == OK and got == data, (n, level)


def test_own_streams(emu):
    """Streams of this library's compressor (fixed, stored and dynamic blocks
    at chunk boundaries that are not byte aligned), as the oracle writes them."""
    import oracle_ffi as orc

    data = b"".join(corpus.generate(k, 40_000, 9).tobytes() for k in corpus.GENERATORS)
    import numpy as np

    arr = np.frombuffer(data, dtype=np.uint8)
    for level, fmt in ((1, 0), (1, 2), (0, 1), (6, 2)):
        s = orc.parallel_compress(arr, level, fmt) if level <= 3 else orc.gpu_parallel_compress(arr, level, fmt)
        st, got = emu(s, fmt, len(data))
        assert st == OK and got == data, (level, fmt)


def test_errors(emu):
    data = corpus.generate("text", 40_000, 6).tobytes()
    for fmt in (0, 1, 2):
        s = zstream(data, 6, fmt)
        for cut in (1, len(s) // 3, len(s) - 9):
            assert emu(s[:cut], fmt, len(data))[0] == DATA_ERROR, (fmt, cut)
        assert emu(s, fmt, len(data) - 1)[0] == NEED_OUTPUT
        assert emu(zstream(data, 0, fmt), fmt, len(data) - 1)[0] == NEED_OUTPUT
    g = zstream(data, 6, 2)
    assert emu(g[:-8] + bytes([g[-8] ^ 1]) + g[-7:], 2, len(data))[0] == DATA_ERROR  # CRC
    assert emu(g[:-4] + bytes([g[-4] ^ 1]) + g[-3:], 2, len(data))[0] == DATA_ERROR  # ISIZE
    z = zstream(data, 6, 1)
    assert emu(z[:-1] + bytes([z[-1] ^ 0x80]), 1, len(data))[0] == DATA_ERROR  # Adler-32
    assert emu(b"\x07\x00\x00", 0, 16)[0] == DATA_ERROR  # block type 3
    assert emu(b"\x01\x05\x00\x00\x00hello", 0, 16)[0] == DATA_ERROR  # LEN / NLEN
    co = zlib.compressobj(6, zlib.DEFLATED, -15, 9, zlib.Z_DEFAULT_STRATEGY, zdict=data[:4000])
    s = co.compress(data[4000:9000]) + co.flush()
    assert emu(s, 0, 5000)[0] == DATA_ERROR  # distance before the start


def test_batch_launch_shape(emu):
    """Eleven streams in one launch: three workgroups' worth, taken from the
    one queue by four waves at once, longest first; every stream's result in
    its own slot."""
    jobs, want = [], []
    kinds = list(corpus.GENERATORS) * 3
    for i, kind in enumerate(kinds[:11]):
        data = corpus.generate(kind, 5_000 + 7_000 * i, 40 + i).tobytes()
        jobs.append((zstream(data, (0, 1, 6, 9)[i % 4], 1), len(data)))
        want.append(data)
    got = emu.batch(jobs, 1)
    for i, (st, out) in enumerate(got):
        assert st == OK and out == want[i], i


def test_long_zero_length_runs(emu):
    """Dynamic headers whose code-length runs (symbol 18) are longer than the
    64 lanes that write them, over LDS that starts non-zero: a block whose
    literals are 16 byte values leaves runs of 138 zero lengths."""
    rng = corpus.generate("random", 30_000, 77)
    data = bytes(b & 0x0F | 0x40 for b in rng.tobytes())
    for fmt in (0, 2):
        s = zstream(data, 9, fmt)
        st, got = emu(s, fmt, len(data))
        assert st == OK and got == data, fmt
