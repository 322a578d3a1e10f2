"""MI355X (gfx950) DEFLATE compressor with orot's parallel compress API.

Python mirror of orot's C / C++ interface (orot include/orot/deflate.h):
``compress_bound``, ``ParallelCompressor``, ``crc32``, ``adler32``, the
``Format`` / ``Level`` constants and the ``Error`` exception, all served by
liborot_amd.so through its C ABI (include/orot/deflate.h,
include/orot/deflate_gpu.h).  The library must be built
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
    "deflate_parallel_new",
    "deflate_parallel_free",
    "deflate_parallel_compress",
    "deflate_adler32",
    "deflate_crc32",
    "deflate_parallel_compress_device",
    "deflate_parallel_set_timing",
    "deflate_parallel_stage_times",
    "orot_gpu_malloc",
    "orot_gpu_free",
    "orot_gpu_memcpy",
    "orot_gpu_memset",
    "orot_gpu_synchronize",
    "orot_gpu_device_info",
    "orot_gpu_last_error",
)


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
    L.deflate_parallel_new.restype = vp
    L.deflate_parallel_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, sz]
    L.deflate_parallel_free.restype = None
    L.deflate_parallel_free.argtypes = [vp]
    L.deflate_parallel_compress.restype = sz
    L.deflate_parallel_compress.argtypes = [vp, vp, sz, vp, sz]
    L.deflate_parallel_compress_device.restype = sz
    L.deflate_parallel_compress_device.argtypes = [vp, vp, sz, vp, sz]
    L.deflate_parallel_set_timing.restype = None
    L.deflate_parallel_set_timing.argtypes = [vp, ctypes.c_int]
    L.deflate_parallel_stage_times.restype = ctypes.c_int
    L.deflate_parallel_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    L.deflate_crc32.restype = u32
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
