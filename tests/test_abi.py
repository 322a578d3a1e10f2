"""The C ABI: liborot_amd.so loads without a GPU and exports every symbol the
headers under include/ declare; host-only entry points behave as orot's."""
import ctypes
import os
import re

import pytest

import orot_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include", "orot")
    for hdr in sorted(os.listdir(inc)):
        if not hdr.endswith(".h"):
            continue
        with open(os.path.join(inc, hdr)) as f:
            text = f.read()
        c_part = text.split("#ifdef __cplusplus\n#include")[0]
        for m in re.finditer(r"^\s*[A-Za-z_][\w \*]*?\b(\w+)\s*\(", c_part, re.M):
            name = m.group(1)
            if name.startswith(("deflate_", "inflate_", "orot_gpu_")):
                names.add(name)
    return names


def test_library_built():
    assert os.path.exists(orot_amd.LIB_PATH), "run python -m orot_amd.build"


def test_every_declared_symbol_exported():
    lib = ctypes.CDLL(orot_amd.LIB_PATH)
    declared = _declared()
    assert declared == set(orot_amd.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("n", [0, 1, 4096, 65535, 1 << 20, (1 << 33) + 5])
def test_bound_matches_orot(n):
    raw = n + (n >> 12) + (n >> 14) + (n >> 25) + 13
    assert orot_amd.compress_bound(n, 0) == raw
    assert orot_amd.compress_bound(n, 1) == raw + 6
    assert orot_amd.compress_bound(n, 2) == raw + 18


def test_no_gpu_means_no_context():
    """Without a GPU the parallel context cannot be created: there is no CPU
    compression path behind the ABI."""
    import subprocess
    import sys

    code = (
        "import orot_amd,sys\n"
        "L=orot_amd.lib()\n"
        "ctx=L.deflate_parallel_new(1,0,0,0)\n"
        "bctx=L.deflate_batch_new(1,1,0,0)\n"
        "cfg=orot_amd.ServerConfig(1<<20,0,0,0)\n"
        "srv=L.deflate_server_new(cfg)\n"
        "nogpu=L.orot_gpu_synchronize()!=0\n"
        "sys.exit(0 if (ctx is None) == nogpu and (bctx is None) == nogpu and (srv is None) == nogpu else 1)\n"
    )
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env={**os.environ, "PYTHONPATH": ROOT})
    assert r.returncode == 0


def test_headers_compile_and_link(tmp_path):
    """Every header compiles as C99, and with its C++ wrappers as C++20; a
    program using orot::deflate::Compressor and ParallelCompressor links
    against the library.  Without a GPU the constructors throw
    orot::deflate::Error with DEFLATE_MEM_ERROR; with one, a stream round
    trip runs."""
    import shutil
    import subprocess

    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("no host compiler")
    inc = os.path.join(ROOT, "include")
    hdrs = sorted(h for h in os.listdir(os.path.join(inc, "orot")) if h.endswith(".h"))
    c_src = tmp_path / "all.c"
    c_src.write_text("".join(f"#include <orot/{h}>\n" for h in hdrs) + "int main(void) { return 0; }\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{inc}", str(c_src)], check=True)
    cpp = tmp_path / "use.cpp"
    cpp.write_text("".join(f"#include <orot/{h}>\n" for h in hdrs) + r"""
#include <cstdio>
#include <vector>
int main() {
    try {
        orot::deflate::Compressor c(1, orot::deflate::Format::Gzip);
        std::vector<uint8_t> in(100000, 'a'), out(200000);
        size_t n = c.feed(in, out);
        n += c.finish(std::span<uint8_t>(out).subspan(n));
        return n > 18 ? 0 : 1;
    } catch (const orot::deflate::Error& e) {
        if (e.code() != DEFLATE_MEM_ERROR) return 2;
    }
    try {
        orot::deflate::ParallelCompressor pc(1);
        return 0;
    } catch (const orot::deflate::Error& e) {
        return e.code() == DEFLATE_MEM_ERROR ? 0 : 3;
    }
}
""")
    exe = tmp_path / "use"
    libdir = os.path.join(ROOT, "orot_amd")
    subprocess.run(["g++", "-std=c++20", "-Wall", "-Werror", f"-I{inc}", str(cpp), "-o", str(exe), f"-L{libdir}",
                    "-lorot_amd", "-L/opt/rocm/lib", f"-Wl,-rpath,{libdir}", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_stream_arguments_and_no_gpu():
    """deflate_stream_compress rejects a NULL stream with DEFLATE_PARAM_ERROR,
    and deflate_stream_new returns NULL where no GPU is usable: the stream
    compresses on the GPU only.  Runs where torch sees no GPU."""
    import subprocess
    import sys

    code = (
        "import ctypes,orot_amd,sys,torch\n"
        "L=orot_amd.lib()\n"
        "if torch.cuda.is_available(): sys.exit(0)\n"
        "p=ctypes.c_void_p(0); n=ctypes.c_size_t(0)\n"
        "r=L.deflate_stream_compress(None,ctypes.byref(p),ctypes.byref(n),ctypes.byref(p),ctypes.byref(n),0)\n"
        "s=L.deflate_stream_new(1,1)\n"
        "sys.exit(0 if r==-5 and not s else 1)\n"
    )
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env={**os.environ, "PYTHONPATH": ROOT})
    assert r.returncode == 0


def test_decompress_arguments_and_host_inflate():
    """deflate_decompress rejects bad arguments with DEFLATE_PARAM_ERROR and
    inflates host buffers on the host, with no GPU needed; the GPU inflate
    entry points (device and batch) stay switched off."""
    import subprocess
    import sys

    code = (
        "import ctypes,orot_amd,sys,zlib\n"
        "L=orot_amd.lib()\n"
        "got=ctypes.c_size_t(0)\n"
        "buf=ctypes.create_string_buffer(64)\n"
        "bad=[L.deflate_decompress(None,5,buf,64,ctypes.byref(got),0),\n"
        "     L.deflate_decompress(buf,5,None,64,ctypes.byref(got),0),\n"
        "     L.deflate_decompress(buf,5,buf,64,ctypes.byref(got),7)]\n"
        "s=zlib.compress(b'abc')\n"
        "r=L.deflate_decompress(s,len(s),buf,64,ctypes.byref(got),1)\n"
        "dev=L.deflate_decompress_device(s,len(s),buf,64,ctypes.byref(got),1)\n"
        "ok=all(b==-5 for b in bad) and r==0 and buf.raw[:3]==b'abc' and dev==-2\n"
        "sys.exit(0 if ok else 1)\n"
    )
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env={**os.environ, "PYTHONPATH": ROOT})
    assert r.returncode == 0
