// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Compiled together with orot's own sources (from the orot tree, never copied
// here) into oracle/_ref/liborot_ref*.so by oracle/Makefile.  It exposes the
// token stream of orot's LZ77 stage (src/deflate/lz77.cpp:352-499, fresh
// state per chunk as in src/parallel/parallel_compressor.cpp:65-75) through a
// C symbol so that tests/golden/make_golden.py can pin the oracle's
// restatement token by token.  orot's public C API (deflate_parallel_*,
// deflate_compress, deflate_crc32, ...) is exported by the orot sources
// themselves.
#include <cstddef>
#include <cstdint>
#include <memory>

#include "compress/level_config.hpp"
#include "deflate/lz77.hpp"

extern "C" size_t ref_lz77(const uint8_t* src, size_t n, int level, uint32_t* tokens) {
    using namespace orot::deflate;
    const CompressConfig cfg = compress_config_for_level(level);
    std::unique_ptr<LZ77State> st(new LZ77State{});
    st->reset(cfg.lz77.hash_bits, cfg.lz77.bt4);
    static_assert(sizeof(Token) == sizeof(uint32_t), "token layout");
    return lz77_compress(src, n, reinterpret_cast<Token*>(tokens), *st, cfg.lz77);
}
