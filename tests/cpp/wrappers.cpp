// wrappers.cpp -- orot's C++ wrappers (include/orot/deflate.h:188-389 in orot)
// used by name against this library's header: compress, decompress,
// Compressor, Decompressor, ParallelCompressor, adler32, crc32, Error, Format,
// Level.  Test program, built by orot_amd.build next to the library.
//
// With a GPU it runs the round-trip matrix of orot's parallel test
// (tests/unit/test_deflate_parallel.cpp:56-107: {threads, block} of {1, 64 KiB},
// {2, 64 KiB}, {4, 128 KiB}, {0, 256 KiB} x raw / zlib / gzip x repeated text
// at L6, 256 KiB of zeros at L1 and 128 KiB of xorshift bytes at L6) through
// ParallelCompressor and decompress, then the other wrappers.  Without one,
// only the host side runs (Decompressor, decompress, and the compressors'
// constructors throwing Error with DEFLATE_MEM_ERROR); with --require-gpu
// that case fails.
//
// Exit status 0 when every check passes.
#include <orot/deflate.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace dz = orot::deflate;

static int failures = 0;
static void check(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

static std::vector<uint8_t> text_data() {
    std::string t;
    for (int i = 0; i < 1000; ++i) t += "GPU parallel deflate round trip: a sentence repeated to give the matcher work. ";
    return std::vector<uint8_t>(t.begin(), t.end());
}

static std::vector<uint8_t> xorshift_bytes(size_t n) {
    std::vector<uint8_t> v(n);
    uint32_t s = 0x9E3779B9u;
    for (auto& b : v) {
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        b = (uint8_t)s;
    }
    return v;
}

// A raw DEFLATE stream written by hand: one fixed-Huffman block holding the
// literals "abc" (not final), then a final stored block "xyz".
static std::vector<uint8_t> fixed_then_stored() {
    std::vector<uint8_t> out;
    uint64_t acc = 0;
    int n = 0;
    auto put = [&](uint32_t v, int k) {
        acc |= (uint64_t)v << n;
        n += k;
        while (n >= 8) {
            out.push_back((uint8_t)acc);
            acc >>= 8;
            n -= 8;
        }
    };
    auto put_code = [&](uint32_t code, int len) {  // Huffman codes go MSB first
        for (int i = len - 1; i >= 0; --i) put((code >> i) & 1u, 1);
    };
    put(0, 1);  // not final
    put(1, 2);  // fixed
    for (char c : std::string("abc")) put_code(0x30u + (uint8_t)c, 8);  // literals 0-143: 00110000 + v
    put_code(0, 7);  // end of block: 256 -> 0000000
    put(1, 1);       // final
    put(0, 2);       // stored
    if (n) put(0, 8 - n);
    put(3, 16);
    put(0xFFFCu, 16);
    for (char c : std::string("xyz")) put((uint8_t)c, 8);
    return out;
}

static bool host_side() {
    // Decompressor over a stream with a stored block after a Huffman block, fed a byte at a time
    const std::vector<uint8_t> raw = fixed_then_stored();
    std::vector<uint8_t> got;
    {
        dz::Decompressor d(dz::Format::Raw);
        bool done = false;
        uint8_t buf[2];
        for (size_t i = 0; i < raw.size() && !done; ++i) {
            size_t k = d.feed(std::span<const uint8_t>(&raw[i], 1), buf, done);
            got.insert(got.end(), buf, buf + k);
            while (!done && k == sizeof(buf)) {  // drain output held back
                k = d.feed({}, buf, done);
                got.insert(got.end(), buf, buf + k);
            }
        }
        check(done, "Decompressor: stream end");
    }
    check(std::string(got.begin(), got.end()) == "abcxyz", "Decompressor: fixed then stored block");
    check(dz::decompress(raw, dz::Format::Raw, 64) == std::vector<uint8_t>({'a', 'b', 'c', 'x', 'y', 'z'}),
          "decompress: fixed then stored block");
    try {
        const std::vector<uint8_t> bad = {0x78, 0x9C, 0xFF, 0xFF};
        (void)dz::decompress(bad, dz::Format::Zlib, 64);
        check(false, "decompress: corrupt stream accepted");
    } catch (const dz::Error& e) {
        check(e.code() == DEFLATE_DATA_ERROR, "decompress: corrupt stream -> DEFLATE_DATA_ERROR");
    }
    // a usable GPU decides the rest
    try {
        dz::ParallelCompressor pc(dz::Level::Fast);
        return true;
    } catch (const dz::Error& e) {
        check(e.code() == DEFLATE_MEM_ERROR, "ParallelCompressor without a GPU -> DEFLATE_MEM_ERROR");
    }
    try {
        dz::Compressor c;
        check(false, "Compressor without a GPU constructed");
    } catch (const dz::Error& e) {
        check(e.code() == DEFLATE_MEM_ERROR, "Compressor without a GPU -> DEFLATE_MEM_ERROR");
    }
    return false;
}

static void gpu_side() {
    // checksums (computed by GPU kernels in this library): orot's KATs
    const std::string kat = "123456789";
    const auto* k = reinterpret_cast<const uint8_t*>(kat.data());
    check(dz::crc32(std::span<const uint8_t>(k, kat.size())) == 0xCBF43926u, "crc32 KAT");
    const std::string wiki = "Wikipedia";
    check(dz::adler32(std::span<const uint8_t>(reinterpret_cast<const uint8_t*>(wiki.data()), wiki.size())) ==
              0x11E60398u,
          "adler32 KAT");
    const std::vector<uint8_t> text = text_data();
    const std::vector<uint8_t> zeros(256 * 1024, 0);
    const std::vector<uint8_t> rnd = xorshift_bytes(128 * 1024);
    const struct {
        int threads;
        size_t block;
    } cfgs[] = {{1, 64 * 1024}, {2, 64 * 1024}, {4, 128 * 1024}, {0, 256 * 1024}};
    const dz::Format fmts[] = {dz::Format::Raw, dz::Format::Zlib, dz::Format::Gzip};
    const struct {
        const std::vector<uint8_t>* data;
        int level;
        const char* name;
    } sets[] = {{&text, dz::Level::Default, "text/L6"}, {&zeros, dz::Level::Fast, "zeros/L1"},
                {&rnd, dz::Level::Default, "random/L6"}};
    int cases = 0;
    for (const auto& c : cfgs)
        for (dz::Format f : fmts)
            for (const auto& s : sets) {
                dz::ParallelCompressor pc(s.level, f, c.threads, c.block);
                const std::vector<uint8_t> comp = pc.compress(*s.data);
                const std::vector<uint8_t> back = dz::decompress(comp, f, s.data->size() + 64);
                check(back == *s.data, std::string("parallel round trip ") + s.name + " format " +
                                           std::to_string((int)f) + " block " + std::to_string(c.block));
                ++cases;
            }
    check(cases == 36, "matrix size");
    // whole-buffer compress at every level and format
    for (int level = dz::Level::Store; level <= dz::Level::Max; ++level)
        for (dz::Format f : fmts) {
            const std::vector<uint8_t> comp = dz::compress(text, level, f);
            check(dz::decompress(comp, f, text.size()) == text, "compress / decompress level " + std::to_string(level));
        }
    // streaming compressor into the streaming decompressor
    for (dz::Format f : fmts) {
        dz::Compressor c(dz::Level::Default, f);
        std::vector<uint8_t> comp(deflate_compress_bound(text.size(), DEFLATE_FORMAT_GZIP) + 64);
        size_t n = c.feed(std::span<const uint8_t>(text.data(), 5000), comp);
        n += c.feed(std::span<const uint8_t>(text.data() + 5000, text.size() - 5000),
                    std::span<uint8_t>(comp).subspan(n));
        n += c.finish(std::span<uint8_t>(comp).subspan(n));
        dz::Decompressor d(f);
        std::vector<uint8_t> out(text.size() + 16);
        bool done = false;
        const size_t m = d.feed(std::span<const uint8_t>(comp.data(), n), out, done);
        check(done && m == text.size() && std::memcmp(out.data(), text.data(), m) == 0,
              "Compressor -> Decompressor format " + std::to_string((int)f));
    }
}

int main(int argc, char** argv) {
    const bool require_gpu = argc > 1 && std::strcmp(argv[1], "--require-gpu") == 0;
    const bool gpu = host_side();
    if (gpu) {
        try {
            gpu_side();
        } catch (const dz::Error& e) {
            check(false, std::string("unexpected Error: ") + e.what());
        }
    } else if (require_gpu) {
        check(false, "no usable GPU");
    }
    std::printf("%s: %d failure(s)%s\n", argv[0], failures, gpu ? "" : " (host side only)");
    return failures == 0 ? 0 : 1;
}
