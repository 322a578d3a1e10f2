"""ctypes bindings of the oracle (oracle/liboracle.so) and, when it was built
in this checkout, of orot itself (oracle/_ref/liborot_ref.so).  Test
infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "liborot_ref.so")

_vp, _sz = ctypes.c_void_p, ctypes.c_size_t
_o = None
_r = None


def oracle() -> ctypes.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        L.oracle_crc32.restype = ctypes.c_uint32
        L.oracle_crc32.argtypes = [ctypes.c_uint32, _vp, _sz]
        L.oracle_adler32.restype = ctypes.c_uint32
        L.oracle_adler32.argtypes = [ctypes.c_uint32, _vp, _sz]
        L.oracle_crc32_combine.restype = ctypes.c_uint32
        L.oracle_crc32_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.oracle_adler32_combine.restype = ctypes.c_uint32
        L.oracle_adler32_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.oracle_compress_bound.restype = _sz
        L.oracle_compress_bound.argtypes = [_sz, ctypes.c_int]
        L.oracle_lz77.restype = _sz
        L.oracle_lz77.argtypes = [_vp, _sz, ctypes.c_int, _vp]
        L.oracle_l1_candidates.restype = None
        L.oracle_l1_candidates.argtypes = [_vp, _sz, ctypes.c_int, _vp]
        L.oracle_parallel_compress.restype = _sz
        L.oracle_parallel_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz, _vp, _sz, _vp, _sz]
        L.oracle_gpu_parallel_compress.restype = _sz
        L.oracle_gpu_parallel_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz, _vp, _sz, _vp, _sz]
        L.oracle_l6_lz77.restype = _sz
        L.oracle_l6_lz77.argtypes = [_vp, _sz, ctypes.c_int, _vp]
        L.oracle_l6_match.restype = None
        L.oracle_l6_match.argtypes = [_vp, _sz, ctypes.c_int, _vp, _vp]
        L.oracle_l6_match_h.restype = None
        L.oracle_l6_match_h.argtypes = [_vp, _sz, _sz, ctypes.c_int, _vp, _vp]
        L.oracle_l6_lz77_h.restype = _sz
        L.oracle_l6_lz77_h.argtypes = [_vp, _sz, _sz, ctypes.c_int, _vp]
        L.oracle_l1_lz77_h.restype = _sz
        L.oracle_l1_lz77_h.argtypes = [_vp, _sz, _sz, ctypes.c_int, _vp]
        L.oracle_l1_dict_compress.restype = _sz
        L.oracle_l1_dict_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz, _vp, _sz, _vp, _sz]
        L.oracle_whole_compress.restype = _sz
        L.oracle_whole_compress.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _sz, _vp, _sz]
        L.oracle_encode_blocks_dyn.restype = None
        L.oracle_encode_blocks_dyn.argtypes = [_vp, _sz, _vp, _sz, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.oracle_encode_block_forced.restype = None
        L.oracle_encode_block_forced.argtypes = [_vp, _sz, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.oracle_encode_block_fixed.restype = ctypes.c_int
        L.oracle_encode_block_fixed.argtypes = [_vp, _sz, _vp, _sz, ctypes.c_int, _vp, _vp]
        L.oracle_limited_lengths.restype = ctypes.c_int
        L.oracle_limited_lengths.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
        _o = L
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    global _r
    if _r is None:
        L = ctypes.CDLL(REF_PATH)
        L.deflate_parallel_new.restype = _vp
        L.deflate_parallel_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz]
        L.deflate_parallel_compress.restype = _sz
        L.deflate_parallel_compress.argtypes = [_vp, _vp, _sz, _vp, _sz]
        L.deflate_parallel_free.restype = None
        L.deflate_parallel_free.argtypes = [_vp]
        L.ref_lz77.restype = _sz
        L.ref_lz77.argtypes = [_vp, _sz, ctypes.c_int, _vp]
        _r = L
    return _r


def _u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def crc32(data, init=0) -> int:
    a = _u8(data)
    return int(oracle().oracle_crc32(init, a.ctypes.data, a.size))


def adler32(data, init=1) -> int:
    a = _u8(data)
    return int(oracle().oracle_adler32(init, a.ctypes.data, a.size))


def lz77(data, level: int) -> np.ndarray:
    a = _u8(data)
    t = np.zeros(max(1, a.size), dtype=np.uint32)
    n = oracle().oracle_lz77(a.ctypes.data, a.size, level, t.ctypes.data)
    return t[:n]


def l1_candidates(data, hash_bits: int = 12) -> np.ndarray:
    a = _u8(data)
    d = np.zeros(max(1, a.size), dtype=np.uint16)
    oracle().oracle_l1_candidates(a.ctypes.data, a.size, hash_bits, d.ctypes.data)
    return d[: a.size]


def parallel_compress(data, level: int, fmt: int, block_size: int = 0, threads: int = 1) -> bytes:
    a = _u8(data)
    cap = a.size + a.size // 8 + 4096
    out = np.zeros(cap, dtype=np.uint8)
    n = oracle().oracle_parallel_compress(level, fmt, threads, block_size, a.ctypes.data, a.size, out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("oracle_parallel_compress failed")
    return out[:n].tobytes()


def gpu_parallel_compress(data, level: int, fmt: int, block_size: int = 0, threads: int = 1) -> bytes:
    """What the HIP path computes (orot's bytes at levels 0-3, the L6-class
    design at levels 4-12)."""
    a = _u8(data)
    cap = a.size + a.size // 8 + 4096
    out = np.zeros(cap, dtype=np.uint8)
    n = oracle().oracle_gpu_parallel_compress(level, fmt, threads, block_size, a.ctypes.data, a.size,
                                              out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("oracle_gpu_parallel_compress failed")
    return out[:n].tobytes()


def l1_lz77_h(buf, halo: int, level: int) -> np.ndarray:
    a = _u8(buf)
    n = a.size - halo
    t = np.zeros(max(1, n), dtype=np.uint32)
    k = oracle().oracle_l1_lz77_h(a.ctypes.data, halo, n, level, t.ctypes.data)
    return t[:k]


def l1_dict_compress(data, level: int, fmt: int, block_size: int = 0, threads: int = 1) -> bytes:
    """Levels 1-3 in dictionary mode (each chunk sees the 32 KiB before it)."""
    a = _u8(data)
    cap = a.size + a.size // 8 + 4096
    out = np.zeros(cap, dtype=np.uint8)
    n = oracle().oracle_l1_dict_compress(level, fmt, threads, block_size, a.ctypes.data, a.size, out.ctypes.data,
                                         cap)
    if n == 0:
        raise RuntimeError("oracle_l1_dict_compress failed")
    return out[:n].tobytes()


def whole_compress(data, level: int, fmt: int) -> bytes:
    """orot's whole-buffer deflate_compress as the HIP path computes it
    (orot's bytes at levels 0-3)."""
    a = _u8(data)
    cap = a.size + (a.size >> 12) + (a.size >> 14) + (a.size >> 25) + 13 + 18
    out = np.zeros(cap, dtype=np.uint8)
    n = oracle().oracle_whole_compress(level, fmt, a.ctypes.data, a.size, out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("oracle_whole_compress failed")
    return out[:n].tobytes()


def l6_lz77(data, level: int) -> np.ndarray:
    a = _u8(data)
    t = np.zeros(max(1, a.size), dtype=np.uint32)
    n = oracle().oracle_l6_lz77(a.ctypes.data, a.size, level, t.ctypes.data)
    return t[:n]


def l6_match(data, level: int):
    """M(x) of the L6 class for every position: (lengths, distances)."""
    a = _u8(data)
    ml = np.zeros(max(1, a.size), dtype=np.uint32)
    md = np.zeros(max(1, a.size), dtype=np.uint32)
    oracle().oracle_l6_match(a.ctypes.data, a.size, level, ml.ctypes.data, md.ctypes.data)
    return ml[:a.size], md[:a.size]


def l6_match_h(buf, halo: int, level: int):
    """M(x) for the len(buf) - halo positions after a dictionary of `halo`
    bytes at the start of buf."""
    a = _u8(buf)
    n = a.size - halo
    ml = np.zeros(max(1, n), dtype=np.uint32)
    md = np.zeros(max(1, n), dtype=np.uint32)
    oracle().oracle_l6_match_h(a.ctypes.data, halo, n, level, ml.ctypes.data, md.ctypes.data)
    return ml[:n], md[:n]


def l6_lz77_h(buf, halo: int, level: int) -> np.ndarray:
    a = _u8(buf)
    n = a.size - halo
    t = np.zeros(max(1, n), dtype=np.uint32)
    k = oracle().oracle_l6_lz77_h(a.ctypes.data, halo, n, level, t.ctypes.data)
    return t[:k]


def encode_blocks_dyn(tokens, raw, is_last: int = 1, max_syms: int = 16384):
    """The L6-class block encoder from a byte-aligned start: (bytes, bits)."""
    t = np.ascontiguousarray(tokens, dtype=np.uint32)
    r = _u8(raw)
    dst = np.zeros(r.size + 8 * t.size // 7 + 4096, dtype=np.uint8)
    bp = np.zeros(1, dtype=np.uint64)
    oracle().oracle_encode_blocks_dyn(t.ctypes.data, t.size, r.ctypes.data, r.size, is_last, max_syms,
                                      dst.ctypes.data, bp.ctypes.data)
    bits = int(bp[0])
    return dst[:(bits + 7) // 8].tobytes(), bits


def encode_block_forced(tokens, raw, btype: int, is_last: int = 1):
    """One block of type btype (0 stored, 1 fixed, 2 dynamic) from a
    byte-aligned start: (bytes, bits)."""
    t = np.ascontiguousarray(tokens, dtype=np.uint32)
    r = _u8(raw)
    dst = np.zeros(r.size + 8 * t.size // 7 + 4096, dtype=np.uint8)
    bp = np.zeros(1, dtype=np.uint64)
    oracle().oracle_encode_block_forced(t.ctypes.data, t.size, r.ctypes.data, is_last, btype, dst.ctypes.data,
                                        bp.ctypes.data)
    bits = int(bp[0])
    return dst[:(bits + 7) // 8].tobytes(), bits


def limited_lengths(freq, max_bits: int) -> np.ndarray:
    f = np.ascontiguousarray(freq, dtype=np.uint32)
    lens = np.zeros(f.size, dtype=np.uint8)
    oracle().oracle_limited_lengths(f.ctypes.data, f.size, max_bits, lens.ctypes.data)
    return lens


def ref_parallel_compress(data, level: int, fmt: int, block_size: int = 0, threads: int = 1) -> bytes:
    a = _u8(data)
    L = ref()
    ctx = L.deflate_parallel_new(level, fmt, threads, block_size)
    cap = a.size + a.size // 8 + 4096
    out = np.zeros(cap, dtype=np.uint8)
    n = L.deflate_parallel_compress(ctx, a.ctypes.data, a.size, out.ctypes.data, cap)
    L.deflate_parallel_free(ctx)
    return out[:n].tobytes()


def ref_lz77(data, level: int) -> np.ndarray:
    a = _u8(data)
    t = np.zeros(max(1, a.size), dtype=np.uint32)
    n = ref().ref_lz77(a.ctypes.data, a.size, level, t.ctypes.data)
    return t[:n]


def ref_decompress(comp: bytes, fmt: int, cap: int):
    """orot's own deflate_decompress (src/api/deflate_api.cpp:138-215 with
    its built-in inflater, src/decompress/decompressor.cpp): (result, bytes)."""
    L = ref()
    L.deflate_decompress.restype = ctypes.c_int
    L.deflate_decompress.argtypes = [_vp, _sz, _vp, _sz, ctypes.POINTER(_sz), ctypes.c_int]
    src = np.frombuffer(comp, dtype=np.uint8) if comp else np.zeros(1, dtype=np.uint8)
    out = np.zeros(max(1, cap), dtype=np.uint8)
    got = _sz(0)
    r = L.deflate_decompress(src.ctypes.data, len(comp), out.ctypes.data, cap, ctypes.byref(got), fmt)
    return r, out[:got.value].tobytes()


def _ref_decompress_child(conn, comp, fmt, cap):
    conn.send(ref_decompress(comp, fmt, cap))
    conn.close()


def ref_decompress_guarded(comp: bytes, fmt: int, cap: int, timeout: float = 60.0):
    """ref_decompress in a child process: orot's inflater can loop or crash
    on valid streams that hold a stored block after a Huffman block, so a
    test must not call it in its own process.  Returns (result, bytes), or
    None when the child did not answer within `timeout` seconds."""
    import multiprocessing as mp

    ctx = mp.get_context("fork")
    a, b = ctx.Pipe(duplex=False)
    p = ctx.Process(target=_ref_decompress_child, args=(b, comp, fmt, cap))
    p.start()
    b.close()
    got = a.recv() if a.poll(timeout) else None
    p.join(1.0)
    if p.is_alive():
        p.kill()
        p.join()
    return got


def ref_stream(data, level: int, fmt: int, one_call: bool = False) -> bytes:
    """orot's own streaming compressor (src/api/stream_api.cpp:134-247): all
    input in one DEFLATE_NO_FLUSH call, then DEFLATE_FINISH (or the input with
    DEFLATE_FINISH in one call)."""
    a = _u8(data)
    L = ref()
    L.deflate_stream_new.restype = _vp
    L.deflate_stream_new.argtypes = [ctypes.c_int, ctypes.c_int]
    L.deflate_stream_free.restype = None
    L.deflate_stream_free.argtypes = [_vp]
    P = ctypes.POINTER
    L.deflate_stream_compress.restype = ctypes.c_int
    L.deflate_stream_compress.argtypes = [_vp, P(_vp), P(_sz), P(_vp), P(_sz), ctypes.c_int]
    s = L.deflate_stream_new(level, fmt)
    out = np.zeros(a.size + a.size // 8 + 4096, dtype=np.uint8)
    ni, ai = _vp(a.ctypes.data), _sz(a.size)
    no, ao = _vp(out.ctypes.data), _sz(out.size)
    try:
        for flush in ((3,) if one_call else (0, 3)):
            r = L.deflate_stream_compress(s, ctypes.byref(ni), ctypes.byref(ai), ctypes.byref(no), ctypes.byref(ao),
                                          flush)
            if r < 0 or r == 3:
                raise RuntimeError(f"orot deflate_stream_compress: {r}")
    finally:
        L.deflate_stream_free(s)
    return out[:out.size - ao.value].tobytes()


def ref_compress(data, level: int, fmt: int) -> bytes:
    """orot's own whole-buffer deflate_compress (src/api/deflate_api.cpp:100-136)."""
    a = _u8(data)
    L = ref()
    L.deflate_compress.restype = _sz
    L.deflate_compress.argtypes = [_vp, _sz, _vp, _sz, ctypes.c_int, ctypes.c_int]
    cap = a.size + (a.size >> 12) + (a.size >> 14) + (a.size >> 25) + 13 + 18
    out = np.zeros(cap, dtype=np.uint8)
    n = L.deflate_compress(a.ctypes.data, a.size, out.ctypes.data, cap, level, fmt)
    return out[:n].tobytes()
