"""MI355X (gfx950) DEFLATE compressor with orot's parallel compress API.

Python mirror of orot's C / C++ interface (orot include/orot/deflate.h):
``compress_bound``, ``ParallelCompressor``, the streaming ``Compressor``,
``crc32``, ``adler32``, the ``Format`` / ``Level`` / ``Flush`` constants and
the ``Error`` exception, all served by
liborot_amd.so through its C ABI (include/orot/deflate.h,
include/orot/deflate_gpu.h), plus ``BatchCompressor`` over the batch
entry points of include/orot/deflate_batch.h.  The library must be built
(``python -m orot_amd.build``); importing this package without it raises.
"""
from __future__ import annotations

import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborot_amd.so")

# Every C symbol declared in include/orot/*.h.
EXPORTED_SYMBOLS = (
    "deflate_compress_bound",
    "deflate_compress",
    "deflate_decompress",
    "deflate_parallel_new",
    "deflate_parallel_free",
    "deflate_parallel_compress",
    "deflate_adler32",
    "deflate_crc32",
    "deflate_parallel_compress_device",
    "deflate_parallel_compress_device_noalloc",
    "deflate_parallel_set_dictionary",
    "deflate_parallel_set_timing",
    "deflate_parallel_stage_times",
    "deflate_decompress_device",
    "deflate_batch_decompress_device",
    "orot_gpu_malloc",
    "orot_gpu_free",
    "orot_gpu_memcpy",
    "orot_gpu_memset",
    "orot_gpu_synchronize",
    "orot_gpu_device_info",
    "orot_gpu_last_error",
    "deflate_batch_new",
    "deflate_batch_free",
    "deflate_batch_compress_device",
    "deflate_batch_compress",
    "deflate_batch_get_config",
    "deflate_batch_set_config",
    "deflate_batch_autotune",
    "deflate_batch_set_timing",
    "deflate_batch_stage_times",
    "deflate_stream_new",
    "deflate_stream_free",
    "deflate_stream_compress",
    "inflate_stream_new",
    "inflate_stream_free",
    "inflate_stream_decompress",
    "deflate_set_allocator",
    "deflate_server_new",
    "deflate_server_free",
    "deflate_server_submit_device",
    "deflate_server_submit",
    "deflate_server_poll",
    "deflate_server_output",
    "deflate_server_read",
    "deflate_server_release",
    "deflate_server_stats",
)

BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2


class BatchItem(ctypes.Structure):
    """deflate_batch_item (include/orot/deflate_batch.h)."""

    _fields_ = [("in_offset", ctypes.c_uint64), ("in_size", ctypes.c_uint64),
                ("out_offset", ctypes.c_uint64), ("out_capacity", ctypes.c_uint64)]


class InflateItem(ctypes.Structure):
    """deflate_inflate_item (include/orot/deflate_gpu.h)."""

    _fields_ = [("in_offset", ctypes.c_uint64), ("in_size", ctypes.c_uint64),
                ("out_offset", ctypes.c_uint64), ("out_capacity", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    """deflate_server_config (include/orot/deflate_server.h)."""

    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_batch_bytes", ctypes.c_size_t),
                ("max_batch_requests", ctypes.c_int), ("lanes_per_level", ctypes.c_int),
                ("lane_arena_bytes", ctypes.c_size_t)]


class Completion(ctypes.Structure):
    """deflate_server_completion (include/orot/deflate_server.h)."""

    _fields_ = [("id", ctypes.c_uint64), ("in_size", ctypes.c_uint64), ("out_size", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("level", ctypes.c_int32), ("submit_ns", ctypes.c_uint64),
                ("launch_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64)]


class BatchConfig(ctypes.Structure):
    """deflate_batch_config (include/orot/deflate_batch.h)."""

    _fields_ = [("block_size", ctypes.c_size_t), ("hash_bits", ctypes.c_int), ("max_chain", ctypes.c_int),
                ("nice_len", ctypes.c_int), ("lazy_depth", ctypes.c_int), ("dictionary", ctypes.c_int)]


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)


class Allocator(ctypes.Structure):
    """deflate_allocator (orot include/orot/deflate.h:49-54)."""

    _fields_ = [("alloc", ALLOC_FN), ("free", FREE_FN), ("opaque", ctypes.c_void_p)]


class Format(enum.IntEnum):
    """deflate_format (orot include/orot/deflate.h:28-32)."""

    RAW = 0
    ZLIB = 1
    GZIP = 2


class Level:
    """Level constants (orot include/orot/deflate.h:35-39)."""

    STORE = 0
    FAST = 1
    DEFAULT = 6
    BETTER = 9
    MAX = 12


class Result(enum.IntEnum):
    """deflate_result (orot include/orot/deflate.h:15-25)."""

    OK = 0
    STREAM_END = 1
    NEED_INPUT = 2
    NEED_OUTPUT = 3
    DATA_ERROR = -1
    MEM_ERROR = -2
    BUF_ERROR = -3
    VERSION_ERROR = -4
    PARAM_ERROR = -5


class Error(RuntimeError):
    """orot::deflate::Error (orot include/orot/deflate.h:204-211)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load liborot_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m orot_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
    L.deflate_compress_bound.restype = sz
    L.deflate_compress_bound.argtypes = [sz, ctypes.c_int]
    L.deflate_compress.restype = sz
    L.deflate_compress.argtypes = [vp, sz, vp, sz, ctypes.c_int, ctypes.c_int]
    L.deflate_parallel_new.restype = vp
    L.deflate_parallel_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, sz]
    L.deflate_parallel_free.restype = None
    L.deflate_parallel_free.argtypes = [vp]
    L.deflate_parallel_compress.restype = sz
    L.deflate_parallel_compress.argtypes = [vp, vp, sz, vp, sz]
    L.deflate_parallel_compress_device.restype = sz
    L.deflate_parallel_compress_device.argtypes = [vp, vp, sz, vp, sz]
    L.deflate_parallel_set_dictionary.restype = ctypes.c_int
    L.deflate_parallel_set_dictionary.argtypes = [vp, ctypes.c_int]
    L.deflate_parallel_compress_device_noalloc.restype = sz
    L.deflate_parallel_compress_device_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, sz, vp, sz, vp, sz]
    L.deflate_parallel_set_timing.restype = None
    L.deflate_parallel_set_timing.argtypes = [vp, ctypes.c_int]
    L.deflate_parallel_stage_times.restype = ctypes.c_int
    L.deflate_parallel_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    L.deflate_decompress.restype = ctypes.c_int
    L.deflate_decompress.argtypes = [vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.c_int]
    L.deflate_decompress_device.restype = ctypes.c_int
    L.deflate_decompress_device.argtypes = [vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.c_int]
    L.deflate_batch_decompress_device.restype = ctypes.c_int
    L.deflate_batch_decompress_device.argtypes = [vp, vp, ctypes.POINTER(InflateItem), sz, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32)]
    L.deflate_crc32.restype = u32
    L.deflate_stream_new.restype = vp
    L.deflate_stream_new.argtypes = [ctypes.c_int, ctypes.c_int]
    L.deflate_stream_free.restype = None
    L.deflate_stream_free.argtypes = [vp]
    L.deflate_stream_compress.restype = ctypes.c_int
    L.deflate_stream_compress.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(sz), ctypes.POINTER(vp),
                                          ctypes.POINTER(sz), ctypes.c_int]
    L.inflate_stream_new.restype = vp
    L.inflate_stream_new.argtypes = [ctypes.c_int]
    L.inflate_stream_free.restype = None
    L.inflate_stream_free.argtypes = [vp]
    L.inflate_stream_decompress.restype = ctypes.c_int
    L.inflate_stream_decompress.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(sz), ctypes.POINTER(vp),
                                            ctypes.POINTER(sz)]
    L.deflate_set_allocator.restype = None
    L.deflate_set_allocator.argtypes = [ctypes.POINTER(Allocator)]
    L.deflate_crc32.argtypes = [u32, vp, sz]
    L.deflate_adler32.restype = u32
    L.deflate_adler32.argtypes = [u32, vp, sz]
    L.orot_gpu_malloc.restype = vp
    L.orot_gpu_malloc.argtypes = [sz]
    L.orot_gpu_free.restype = ctypes.c_int
    L.orot_gpu_free.argtypes = [vp]
    L.orot_gpu_memcpy.restype = ctypes.c_int
    L.orot_gpu_memcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    L.orot_gpu_memset.restype = ctypes.c_int
    L.orot_gpu_memset.argtypes = [vp, ctypes.c_int, sz]
    L.orot_gpu_synchronize.restype = ctypes.c_int
    L.orot_gpu_synchronize.argtypes = []
    L.orot_gpu_device_info.restype = ctypes.c_int
    L.orot_gpu_device_info.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_int)]
    L.orot_gpu_last_error.restype = ctypes.c_char_p
    L.orot_gpu_last_error.argtypes = []
    items_p = ctypes.POINTER(BatchItem)
    L.deflate_batch_new.restype = vp
    L.deflate_batch_new.argtypes = [ctypes.c_int, ctypes.c_int, sz, ctypes.c_uint]
    L.deflate_batch_free.restype = None
    L.deflate_batch_free.argtypes = [vp]
    L.deflate_batch_compress_device.restype = ctypes.c_int
    L.deflate_batch_compress_device.argtypes = [vp, vp, vp, items_p, sz, ctypes.POINTER(ctypes.c_uint64)]
    L.deflate_batch_compress.restype = ctypes.c_int
    L.deflate_batch_compress.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(sz), ctypes.POINTER(vp),
                                         ctypes.POINTER(sz), ctypes.POINTER(sz), sz]
    L.deflate_batch_get_config.restype = ctypes.c_int
    L.deflate_batch_get_config.argtypes = [vp, ctypes.POINTER(BatchConfig)]
    L.deflate_batch_set_config.restype = ctypes.c_int
    L.deflate_batch_set_config.argtypes = [vp, ctypes.POINTER(BatchConfig)]
    L.deflate_batch_autotune.restype = ctypes.c_int
    L.deflate_batch_autotune.argtypes = [vp, vp, vp, items_p, sz]
    L.deflate_batch_set_timing.restype = None
    L.deflate_batch_set_timing.argtypes = [vp, ctypes.c_int]
    L.deflate_batch_stage_times.restype = ctypes.c_int
    L.deflate_batch_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    u64 = ctypes.c_uint64
    L.deflate_server_new.restype = vp
    L.deflate_server_new.argtypes = [ctypes.POINTER(ServerConfig)]
    L.deflate_server_free.restype = None
    L.deflate_server_free.argtypes = [vp]
    L.deflate_server_submit_device.restype = u64
    L.deflate_server_submit_device.argtypes = [vp, vp, sz, ctypes.c_int, ctypes.c_int]
    L.deflate_server_submit.restype = u64
    L.deflate_server_submit.argtypes = [vp, vp, sz, ctypes.c_int, ctypes.c_int]
    L.deflate_server_poll.restype = ctypes.c_int
    L.deflate_server_poll.argtypes = [vp, ctypes.POINTER(Completion), ctypes.c_int, ctypes.c_int]
    L.deflate_server_output.restype = vp
    L.deflate_server_output.argtypes = [vp, u64]
    L.deflate_server_read.restype = sz
    L.deflate_server_read.argtypes = [vp, u64, vp, sz]
    L.deflate_server_release.restype = ctypes.c_int
    L.deflate_server_release.argtypes = [vp, u64]
    L.deflate_server_stats.restype = ctypes.c_int
    L.deflate_server_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.c_int]
    _lib = L
    return L


def last_error() -> str:
    return lib().orot_gpu_last_error().decode(errors="replace")


def _as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def compress_bound(n: int, fmt: int = Format.GZIP) -> int:
    """deflate_compress_bound (orot include/orot/deflate.h:66)."""
    return int(lib().deflate_compress_bound(n, int(fmt)))


def compress(data, level: int = Level.DEFAULT, fmt: int = Format.GZIP) -> bytes:
    """deflate_compress (orot include/orot/deflate.h:69-77): whole buffer,
    output capacity from compress_bound."""
    a = _as_u8(data)
    out = np.empty(compress_bound(a.size, fmt), dtype=np.uint8)
    n = lib().deflate_compress(a.ctypes.data, a.size, out.ctypes.data, out.size, int(level), int(fmt))
    if n == 0:
        raise Error(Result.BUF_ERROR, f"deflate_compress: {last_error()}")
    return out[:n].tobytes()


def decompress_into(data, out: np.ndarray, fmt: int = Format.GZIP) -> tuple[int, int]:
    """deflate_decompress (orot include/orot/deflate.h:79-89) into a
    preallocated uint8 array: returns (result code, decompressed size)."""
    a = _as_u8(data)
    got = ctypes.c_size_t(0)
    r = lib().deflate_decompress(a.ctypes.data if a.size else None, a.size, out.ctypes.data if out.size else None,
                                 out.size, ctypes.byref(got), int(fmt))
    return int(r), int(got.value)


def decompress(data, fmt: int = Format.GZIP, max_output: int = 256 << 20) -> bytes:
    """orot::deflate::decompress (orot include/orot/deflate.h:238-256), with
    the same max_output guard; unlike orot's single max_output-sized call, the
    buffer starts small and doubles while the result is NEED_OUTPUT."""
    a = _as_u8(data)
    cap = min(max_output, max(4 * a.size, 64 << 10))
    while True:
        out = np.empty(cap, dtype=np.uint8)
        r, n = decompress_into(a, out, fmt)
        if r == Result.OK:
            return out[:n].tobytes()
        if r == Result.NEED_OUTPUT and cap < max_output:
            cap = min(max_output, cap * 2)
            continue
        raise Error(r, f"deflate::decompress failed ({last_error()})")


class Decompressor:
    """orot::deflate::Decompressor (orot include/orot/deflate.h:308-342) over
    inflate_stream_new / inflate_stream_decompress."""

    def __init__(self, fmt: int = Format.ZLIB):
        self._s = lib().inflate_stream_new(int(fmt))
        if not self._s:
            raise Error(Result.MEM_ERROR, "Decompressor: allocation failed")

    def call(self, data, out: np.ndarray):
        """One inflate_stream_decompress call: (result, input bytes used,
        output bytes written)."""
        a = _as_u8(data)
        ni = ctypes.c_void_p(a.ctypes.data if a.size else None)
        ai = ctypes.c_size_t(a.size)
        no = ctypes.c_void_p(out.ctypes.data if out.size else None)
        ao = ctypes.c_size_t(out.size)
        r = lib().inflate_stream_decompress(self._s, ctypes.byref(ni), ctypes.byref(ai), ctypes.byref(no),
                                            ctypes.byref(ao))
        return int(r), a.size - ai.value, out.size - ao.value

    def feed(self, data, out_chunk: int = 1 << 20) -> tuple[bytes, bool]:
        """All of `data` through the stream, draining NEED_OUTPUT: (bytes
        out, stream ended).  Raises Error on a failure code."""
        a = _as_u8(data)
        got, pos = [], 0
        buf = np.empty(out_chunk, dtype=np.uint8)
        while True:
            r, used, n = self.call(a[pos:], buf)
            pos += used
            got.append(buf[:n].tobytes())
            if r < 0:
                raise Error(r, f"Decompressor::feed failed ({r})")
            if r == Result.STREAM_END:
                return b"".join(got), True
            if r == Result.OK:
                return b"".join(got), False

    def close(self) -> None:
        if self._s:
            lib().inflate_stream_free(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_decompress_device(d_in: int, d_out: int, items, fmt: int = Format.GZIP):
    """deflate_batch_decompress_device: items are (in_offset, in_size,
    out_offset, out_capacity); returns (sizes, result codes) as numpy arrays."""
    n = len(items)
    arr = (InflateItem * max(n, 1))(*[InflateItem(*map(int, it)) for it in items])
    sizes = (ctypes.c_uint64 * max(n, 1))()
    res = (ctypes.c_int32 * max(n, 1))()
    if lib().deflate_batch_decompress_device(d_in, d_out, arr, n, int(fmt), sizes, res) != 0:
        raise Error(Result.MEM_ERROR, f"deflate_batch_decompress_device: {last_error()}")
    return np.ctypeslib.as_array(sizes)[:n].copy(), np.ctypeslib.as_array(res)[:n].copy()


class ParallelCompressor:
    """orot::deflate::ParallelCompressor (orot include/orot/deflate.h:344-377)."""

    def __init__(self, level: int = Level.DEFAULT, fmt: int = Format.GZIP, threads: int = 0, block_size: int = 0):
        self._ctx = lib().deflate_parallel_new(int(level), int(fmt), int(threads), int(block_size))
        if not self._ctx:
            raise Error(Result.MEM_ERROR, f"ParallelCompressor: allocation failed ({last_error()})")
        self.level, self.format, self.block_size = int(level), int(fmt), int(block_size)

    def compress(self, data) -> bytes:
        """Compress host bytes; the output is sized with the gzip bound, as orot does."""
        src = _as_u8(data)
        out = np.empty(compress_bound(src.size, Format.GZIP), dtype=np.uint8)
        n = self.compress_into(src, out)
        if n == 0:
            raise Error(Result.BUF_ERROR, f"ParallelCompressor::compress: {last_error()}")
        return out[:n].tobytes()

    def compress_into(self, src: np.ndarray, out: np.ndarray) -> int:
        """deflate_parallel_compress on host arrays; returns bytes written (0 on failure)."""
        return int(lib().deflate_parallel_compress(self._ctx, src.ctypes.data, src.size, out.ctypes.data, out.size))

    def compress_device(self, d_in: int, n: int, d_out: int, cap: int) -> int:
        """deflate_parallel_compress_device on device pointers."""
        return int(lib().deflate_parallel_compress_device(self._ctx, d_in, n, d_out, cap))

    def set_dictionary(self, on: bool = True) -> None:
        """deflate_parallel_set_dictionary: levels 1-3 match into the 32 KiB
        before each chunk (not orot's bytes)."""
        if lib().deflate_parallel_set_dictionary(self._ctx, 1 if on else 0) != 0:
            raise Error(Result.PARAM_ERROR, f"deflate_parallel_set_dictionary: {last_error()}")

    def set_timing(self, on: bool = True) -> None:
        lib().deflate_parallel_set_timing(self._ctx, 1 if on else 0)

    def stage_times(self):
        """Per-stage device ms of the last call: checksum, lz77+emit, scan, assemble."""
        buf = (ctypes.c_float * 8)()
        k = lib().deflate_parallel_stage_times(self._ctx, buf, 8)
        return [float(buf[i]) for i in range(k)]

    def close(self) -> None:
        if self._ctx:
            lib().deflate_parallel_free(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Flush(enum.IntEnum):
    """deflate_flush (orot include/orot/deflate.h:42-47)."""
    NO_FLUSH = 0
    SYNC_FLUSH = 1
    FULL_FLUSH = 2
    FINISH = 3


class Compressor:
    """orot::deflate::Compressor (orot include/orot/deflate.h:258-304) over
    deflate_stream_new / deflate_stream_compress: input collects on the host
    and compresses on the GPU in pieces of whole chunks."""

    def __init__(self, level: int = Level.DEFAULT, fmt: int = Format.ZLIB):
        self._s = lib().deflate_stream_new(int(level), int(fmt))
        if not self._s:
            raise Error(Result.MEM_ERROR, f"Compressor: allocation failed ({last_error()})")

    def compress_call(self, data, out: np.ndarray, flush: int = Flush.NO_FLUSH):
        """One deflate_stream_compress call: returns (result, input bytes
        consumed, output bytes written)."""
        src = _as_u8(data)
        ni = ctypes.c_void_p(src.ctypes.data if src.size else None)
        ai = ctypes.c_size_t(src.size)
        no = ctypes.c_void_p(out.ctypes.data if out.size else None)
        ao = ctypes.c_size_t(out.size)
        r = lib().deflate_stream_compress(self._s, ctypes.byref(ni), ctypes.byref(ai), ctypes.byref(no),
                                          ctypes.byref(ao), int(flush))
        return int(r), src.size - ai.value, out.size - ao.value

    def write(self, data, flush: int = Flush.NO_FLUSH, out_chunk: int = 1 << 20) -> bytes:
        """Feeds all of `data` with `flush`, calling again while the result is
        NEED_OUTPUT; returns every output byte produced."""
        src = _as_u8(data)
        parts, pos = [], 0
        out = np.empty(out_chunk, dtype=np.uint8)
        while True:
            r, used, n = self.compress_call(src[pos:], out, flush)
            pos += used
            parts.append(out[:n].tobytes())
            if r < 0:
                raise Error(r, f"deflate_stream_compress failed ({last_error()})")
            if r == Result.STREAM_END or (r == Result.OK and pos == src.size):
                return b"".join(parts)

    def feed(self, data) -> bytes:
        return self.write(data, Flush.NO_FLUSH)

    def flush(self, full: bool = False) -> bytes:
        return self.write(b"", Flush.FULL_FLUSH if full else Flush.SYNC_FLUSH)

    def finish(self) -> bytes:
        return self.write(b"", Flush.FINISH)

    def close(self) -> None:
        if self._s:
            lib().deflate_stream_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def crc32(data, init: int = 0) -> int:
    """deflate_crc32 (orot include/orot/deflate.h:150)."""
    src = _as_u8(data)
    return int(lib().deflate_crc32(init, src.ctypes.data, src.size))


def adler32(data, init: int = 1) -> int:
    """deflate_adler32 (orot include/orot/deflate.h:149)."""
    src = _as_u8(data)
    return int(lib().deflate_adler32(init, src.ctypes.data, src.size))


class DeviceBuffer:
    """A block of GPU memory owned through the library's C ABI."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().orot_gpu_malloc(max(1, self.nbytes))
        if not self.ptr:
            raise Error(Result.MEM_ERROR, f"orot_gpu_malloc({nbytes}): {last_error()}")

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
        buf = cls(a.size)
        buf.upload(a)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
        if offset < 0 or offset + a.size > self.nbytes:
            raise ValueError(f"upload of {a.size} B at {offset} exceeds the {self.nbytes} B buffer")
        if a.size and lib().orot_gpu_memcpy(self.ptr + offset, a.ctypes.data, a.size, 0) != 0:
            raise Error(Result.MEM_ERROR, last_error())

    def download(self, n: int | None = None, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if n is None else int(n)
        if offset < 0 or n < 0 or offset + n > self.nbytes:
            raise ValueError(f"download of {n} B at {offset} exceeds the {self.nbytes} B buffer")
        out = np.empty(n, dtype=np.uint8)
        if n and lib().orot_gpu_memcpy(out.ctypes.data, self.ptr + offset, n, 1) != 0:
            raise Error(Result.MEM_ERROR, last_error())
        return out

    def fill(self, value: int) -> None:
        if lib().orot_gpu_memset(self.ptr, value, self.nbytes) != 0:
            raise Error(Result.MEM_ERROR, last_error())

    def free(self) -> None:
        if self.ptr:
            lib().orot_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int(0)
    if lib().orot_gpu_device_info(name, 256, ctypes.byref(cus)) != 0:
        raise Error(Result.MEM_ERROR, last_error())
    return name.value.decode(), cus.value


class BatchCompressor:
    """Many independent buffers per submission (include/orot/deflate_batch.h).
    Each output equals deflate_parallel_compress of that buffer alone with the
    same level, format and block size."""

    def __init__(self, level: int = Level.FAST, fmt: int = Format.ZLIB, block_size: int = 0, flags: int = 0):
        self._ctx = lib().deflate_batch_new(int(level), int(fmt), int(block_size), int(flags))
        if not self._ctx:
            raise Error(Result.MEM_ERROR, f"BatchCompressor: {last_error()}")
        self.level, self.format = int(level), int(fmt)

    @staticmethod
    def items(in_offsets, in_sizes, out_offsets, out_caps):
        n = len(in_sizes)
        arr = (BatchItem * n)()
        for i in range(n):
            arr[i].in_offset = int(in_offsets[i])
            arr[i].in_size = int(in_sizes[i])
            arr[i].out_offset = int(out_offsets[i])
            arr[i].out_capacity = int(out_caps[i])
        return arr

    def compress_device(self, d_in: int, d_out: int, items) -> np.ndarray:
        """Device-resident batch; returns the framed sizes (0: did not fit)."""
        n = len(items)
        sizes = np.zeros(max(n, 1), dtype=np.uint64)
        rc = lib().deflate_batch_compress_device(self._ctx, d_in, d_out, items, n,
                                                 sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc != 0:
            raise Error(Result.MEM_ERROR, f"deflate_batch_compress_device: {last_error()}")
        return sizes[:n]

    def compress(self, buffers) -> list:
        """Host buffers in, compressed bytes out (one entry per buffer)."""
        arrs = [_as_u8(b) for b in buffers]
        n = len(arrs)
        outs = [np.empty(compress_bound(a.size, Format.GZIP) + 64 * (a.size // 65535 + 4), dtype=np.uint8)
                for a in arrs]
        in_p = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in arrs])
        out_p = (ctypes.c_void_p * max(n, 1))(*[o.ctypes.data for o in outs])
        in_s = (ctypes.c_size_t * max(n, 1))(*[a.size for a in arrs])
        out_c = (ctypes.c_size_t * max(n, 1))(*[o.size for o in outs])
        got = (ctypes.c_size_t * max(n, 1))()
        rc = lib().deflate_batch_compress(self._ctx, in_p, in_s, out_p, out_c, got, n)
        if rc != 0:
            raise Error(Result.MEM_ERROR, f"deflate_batch_compress: {last_error()}")
        return [outs[i][:got[i]].tobytes() for i in range(n)]

    def config(self) -> dict:
        c = BatchConfig()
        if lib().deflate_batch_get_config(self._ctx, ctypes.byref(c)) != 0:
            raise Error(Result.PARAM_ERROR, last_error())
        return {"block_size": c.block_size, "hash_bits": c.hash_bits, "max_chain": c.max_chain,
                "nice_len": c.nice_len, "lazy_depth": c.lazy_depth, "dictionary": c.dictionary}

    def set_config(self, block_size: int, hash_bits: int = 0, max_chain: int = 0, nice_len: int = 0,
                   lazy_depth: int = 0, dictionary: int = 0) -> None:
        c = BatchConfig(block_size, hash_bits, max_chain, nice_len, lazy_depth, dictionary)
        if lib().deflate_batch_set_config(self._ctx, ctypes.byref(c)) != 0:
            raise Error(Result.PARAM_ERROR, f"deflate_batch_set_config: {last_error()}")

    def autotune(self, d_in: int, d_out: int, items) -> dict:
        if lib().deflate_batch_autotune(self._ctx, d_in, d_out, items, len(items)) != 0:
            raise Error(Result.MEM_ERROR, f"deflate_batch_autotune: {last_error()}")
        return self.config()

    def set_timing(self, on: bool) -> None:
        lib().deflate_batch_set_timing(self._ctx, 1 if on else 0)

    def stage_times(self) -> list:
        buf = (ctypes.c_float * 8)()
        k = lib().deflate_batch_stage_times(self._ctx, buf, 8)
        return [buf[i] for i in range(k)]

    def close(self) -> None:
        if self._ctx:
            lib().deflate_batch_free(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching compression server (include/orot/deflate_server.h)."""

    def __init__(self, hbm_budget: int = 0, max_batch_bytes: int = 0, max_batch_requests: int = 0,
                 lanes_per_level: int = 0, lane_arena_bytes: int = 0):
        cfg = ServerConfig(hbm_budget, max_batch_bytes, max_batch_requests, lanes_per_level, lane_arena_bytes)
        self._srv = lib().deflate_server_new(ctypes.byref(cfg))
        if not self._srv:
            raise Error(Result.MEM_ERROR, f"Server: {last_error()}")
        self._comp = (Completion * 1024)()

    def submit_device(self, d_in: int, size: int, level: int, fmt: int = Format.GZIP) -> int:
        rid = lib().deflate_server_submit_device(self._srv, d_in, int(size), int(level), int(fmt))
        if rid == 0:
            raise Error(Result.BUF_ERROR, f"deflate_server_submit_device: {last_error()}")
        return rid

    def submit(self, data, level: int, fmt: int = Format.GZIP) -> int:
        """Host input; the array must stay alive until the request completes."""
        a = _as_u8(data)
        rid = lib().deflate_server_submit(self._srv, a.ctypes.data, a.size, int(level), int(fmt))
        if rid == 0:
            raise Error(Result.BUF_ERROR, f"deflate_server_submit: {last_error()}")
        return rid

    def poll(self, timeout_ms: int = 0) -> list:
        k = lib().deflate_server_poll(self._srv, self._comp, len(self._comp), int(timeout_ms))
        return [(c.id, c.in_size, c.out_size, c.status, c.level, c.submit_ns, c.launch_ns, c.done_ns)
                for c in self._comp[:k]]

    def output_ptr(self, rid: int) -> int:
        return lib().deflate_server_output(self._srv, rid) or 0

    def read(self, rid: int, size: int) -> bytes:
        out = np.empty(max(int(size), 1), dtype=np.uint8)
        n = lib().deflate_server_read(self._srv, rid, out.ctypes.data, out.size)
        if n != size:
            raise Error(Result.BUF_ERROR, f"deflate_server_read: {last_error()}")
        return out[:n].tobytes()

    def release(self, rid: int) -> None:
        if lib().deflate_server_release(self._srv, rid) != 0:
            raise Error(Result.PARAM_ERROR, "deflate_server_release: unknown or unfinished request")

    def stats(self) -> dict:
        v = (ctypes.c_uint64 * 7)()
        lib().deflate_server_stats(self._srv, v, 7)
        keys = ("submitted", "completed", "queued", "in_flight", "pool_bytes", "batches", "max_queue")
        return dict(zip(keys, (int(x) for x in v)))

    def close(self) -> None:
        if self._srv:
            lib().deflate_server_free(self._srv)
            self._srv = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
