"""Pins the oracle (oracle/orot_oracle.c) before it is trusted as the checker.

Each test names what fixes the expected value: orot's own KATs and table
facts, orot's bytes (golden fixtures written by orot, tests/golden/), zlib,
closed forms, or brute force.  CPU only.
"""
import hashlib
import itertools
import json
import os
import zlib

import numpy as np
import pytest

import oracle_ffi as orc
from orot_amd import corpus

HERE = os.path.dirname(os.path.abspath(__file__))
WBITS = {0: -15, 1: 15, 2: 31}


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(HERE, "golden", "orot_fixtures.json")) as f:
        return json.load(f)


# ---- checksums: orot KATs (tests/unit/test_deflate_formats.cpp:74-88) and zlib
def test_crc32_kat():
    assert orc.crc32(b"123456789") == 0xCBF43926


def test_adler32_kat():
    assert orc.adler32(b"Wikipedia") == 0x11E60398


def test_checksums_empty_is_initial():  # test_deflate_simd.cpp: empty = initial value
    assert orc.crc32(b"", 0) == 0
    assert orc.crc32(b"", 0x12345678) == 0x12345678
    assert orc.adler32(b"", 1) == 1


@pytest.mark.parametrize("n", [1, 7, 8, 9, 5551, 5552, 5553, 100_000, 3_000_001])
def test_checksums_vs_zlib(n):
    d = corpus.random_bytes(n, n)
    assert orc.crc32(d) == zlib.crc32(d.tobytes())
    assert orc.adler32(d) == zlib.adler32(d.tobytes())
    assert orc.crc32(d, 0xDEADBEEF) == zlib.crc32(d.tobytes(), 0xDEADBEEF)
    assert orc.adler32(d, 0x00450011) == zlib.adler32(d.tobytes(), 0x00450011)


def test_checksums_long_runs_vs_zlib():  # all-0xFF stresses Adler's NMAX reduction
    d = np.full(1_000_003, 0xFF, dtype=np.uint8)
    assert orc.adler32(d) == zlib.adler32(d.tobytes())
    assert orc.crc32(d) == zlib.crc32(d.tobytes())


@pytest.mark.parametrize("split", [0, 1, 4, 1000, 65536, 99_999])
def test_combine_vs_zlib(split):
    d = corpus.random_bytes(100_000, 5).tobytes()
    a, b = d[:split], d[split:]
    O = orc.oracle()
    assert O.oracle_crc32_combine(zlib.crc32(a), zlib.crc32(b), len(b)) == zlib.crc32(d)
    assert O.oracle_adler32_combine(zlib.adler32(a), zlib.adler32(b), len(b)) == zlib.adler32(d)


# ---- bound (src/formats/raw_deflate.cpp:9-12, src/api/deflate_api.cpp:81-98)
@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 16384, 1 << 25, (1 << 32) + 7])
def test_bound_closed_form(n):
    raw = n + (n >> 12) + (n >> 14) + (n >> 25) + 13
    O = orc.oracle()
    assert O.oracle_compress_bound(n, 0) == raw
    assert O.oracle_compress_bound(n, 1) == raw + 6
    assert O.oracle_compress_bound(n, 2) == raw + 18


# ---- closed-form outputs derivable without orot (empty input, one byte)
def test_empty_and_one_byte_closed_forms():
    e = np.zeros(0, dtype=np.uint8)
    assert orc.parallel_compress(e, 1, 0) == bytes.fromhex("0300")
    assert orc.parallel_compress(e, 6, 0) == bytes.fromhex("0300")
    assert orc.parallel_compress(e, 1, 1) == bytes.fromhex("789c030000000001")
    assert orc.parallel_compress(e, 1, 2) == bytes.fromhex("1f8b08000000000000ff0300000000000000000000")[:20]
    assert orc.parallel_compress(e, 0, 0) == bytes.fromhex("010000ffff")
    a = np.frombuffer(b"a", dtype=np.uint8)
    assert orc.parallel_compress(a, 1, 0) == bytes.fromhex("4b0400")
    assert orc.parallel_compress(a, 6, 0) == bytes.fromhex("4b0400")
    assert orc.parallel_compress(a, 1, 1) == bytes.fromhex("789c4b0400006200 62".replace(" ", ""))


# ---- RFC 1951 tables, through the fixed-Huffman encoding of single symbols
# (orot test_deflate_huffman.cpp:21-76: fixed lens 8/9/7/8, LENGTH_BASE,
# DIST_BASE, length_to_code(3)=0, (258)=28).  A stream holding one match of
# length L at distance d inflates through zlib only if base / extra tables
# are right.
@pytest.mark.parametrize("L", [3, 4, 10, 11, 12, 18, 19, 34, 35, 66, 67, 130, 131, 257, 258])
@pytest.mark.parametrize("d", [1, 2, 4, 5, 8, 9, 24577, 32768])
def test_match_encoding_via_zlib(L, d):
    # d distinct literal bytes prefix so a match at distance d is legal, then L bytes copying
    rng = np.random.default_rng(L * 131 + d)
    pre = rng.integers(0, 256, size=d, dtype=np.uint8)
    body = np.concatenate([pre, np.resize(pre, L) if d >= L else np.resize(pre, L)])
    toks = [int(x) for x in pre] + [0x80000000 | (L << 16) | d]
    t = np.array(toks, dtype=np.uint32)
    dst = np.zeros(len(toks) * 4 + 64, dtype=np.uint8)
    bp = np.zeros(1, dtype=np.uint64)
    O = orc.oracle()
    O.oracle_encode_block_fixed(t.ctypes.data, t.size, body.ctypes.data, 10**9, 1, dst.ctypes.data, bp.ctypes.data)
    stream = dst[: (int(bp[0]) + 7) // 8].tobytes()
    want = bytearray(pre.tobytes())
    for i in range(L):
        want.append(want[len(want) - d])
    assert zlib.decompress(stream, -15) == bytes(want)


# ---- L1 candidate rule against a brute-force restatement of orot's fast path
def _brute_candidates(buf, bits):
    n = len(buf)
    out = [0] * n
    hist = {}
    for p in range(n):
        if p + 4 > n:
            continue
        v = int.from_bytes(buf[p:p + 4], "little")
        h = ((v * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - bits)
        c = hist.get(h, 0)
        hist[h] = p
        if c == 0:
            continue
        d = (p - c) & 0x7FFF
        if d and buf[c] == buf[p]:
            out[p] = d
    return out


@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_l1_candidates_bruteforce(kind):
    d = corpus.generate(kind, 40_000, 3)
    assert list(orc.l1_candidates(d, 12)) == _brute_candidates(d.tobytes(), 12)
    assert list(orc.l1_candidates(d, 14)) == _brute_candidates(d.tobytes(), 14)


def _greedy_from_candidates(buf, cand):
    """orot's fast-path parse (lz77.cpp:378-406) restated over candidates."""
    n = len(buf)
    toks = []
    p = 0
    while p < n:
        l = 0
        if p + 4 <= n and cand[p]:
            m = min(258, n - p)
            d = cand[p]
            while l < m and buf[p + l] == buf[p - d + l]:
                l += 1
        if l >= 3:
            toks.append(0x80000000 | (l << 16) | cand[p])
            p += l
        else:
            toks.append(buf[p])
            p += 1
    return toks


@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_l1_parse_is_greedy_over_candidates(kind):
    d = corpus.generate(kind, 30_000, 11)
    assert list(orc.lz77(d, 1)) == _greedy_from_candidates(d.tobytes(), [int(x) for x in orc.l1_candidates(d, 12)])


# ---- orot's own bytes and token streams (golden fixtures written by orot)
def test_golden_small_payloads(fixtures):
    bad = []
    for e in fixtures["small"]:
        data = corpus.generate(e["kind"], e["n"], e["seed"])
        got = orc.parallel_compress(data, e["level"], e["format"], e["block"])
        if got.hex() != e["hex"]:
            bad.append(e)
    assert not bad, bad[:3]


def test_golden_large_digests(fixtures):
    for e in fixtures["large"]:
        data = corpus.generate(e["kind"], e["n"], e["seed"])
        got = orc.parallel_compress(data, e["level"], e["format"], e["block"], threads=4)
        assert len(got) == e["size"] and hashlib.sha256(got).hexdigest() == e["sha256"], e


def test_golden_text16_l1_gzip(fixtures):
    e = fixtures["text16_l1_gzip"]
    data = corpus.generate(e["kind"], e["n"], e["seed"])
    got = orc.parallel_compress(data, e["level"], e["format"], e["block"], threads=8)
    assert len(got) == e["size"] and hashlib.sha256(got).hexdigest() == e["sha256"]


def test_golden_tokens(fixtures):
    for e in fixtures["tokens"]:
        data = corpus.generate(e["kind"], e["n"], e["seed"])
        t = orc.lz77(data, e["level"])
        assert t.size == e["count"] and hashlib.sha256(t.astype("<u4").tobytes()).hexdigest() == e["sha256"], e


def test_tokens_reconstruct_input():
    for kind in corpus.GENERATORS:
        data = corpus.generate(kind, 50_000, 8)
        for level in (1, 2, 3, 4, 6, 9):
            out = bytearray()
            for t in orc.lz77(data, level):
                t = int(t)
                if t & 0x80000000:
                    L, d = (t >> 16) & 0x1FF, t & 0xFFFF
                    assert 3 <= L <= 258 and 1 <= d <= 32768
                    for _ in range(L):
                        out.append(out[len(out) - d])
                else:
                    out.append(t & 0xFF)
            assert bytes(out) == data.tobytes(), (kind, level)


def test_l6_class_ratio_vs_orot(fixtures):
    """What the HIP path computes at levels 4-12 (oracle_gpu_parallel_compress:
    the L6-class matcher, lazy parse and block encoder, with the cross-chunk
    dictionary) stays within 3% of orot's own size at the same level on
    every fixture input, and is always zlib-valid, also where orot's own
    output is not (orot's levels 10-12 and its large level 4-9 records)."""
    for e in fixtures["levels_l6"]:
        data = corpus.generate(e["kind"], e["n"], e["seed"])
        got = orc.gpu_parallel_compress(data, e["level"], 0, 0, threads=4)
        assert zlib.decompress(got, -15) == data.tobytes()
        assert len(got) <= 1.03 * e["size"], (e, len(got))


def test_l6_class_ratio_covers_every_level(fixtures):
    assert sorted({e["level"] for e in fixtures["levels_l6"]}) == list(range(4, 13))


# ---- optimal length-limited Huffman lengths: brute force on tiny alphabets
def _kraft(lens):
    return sum(2.0 ** -int(l) for l in lens if l)


def _best_cost(freq, max_bits):
    act = [f for f in freq if f]
    best = None
    for ls in itertools.product(range(1, max_bits + 1), repeat=len(act)):
        if sum(2.0 ** -l for l in ls) <= 1.0 + 1e-12:
            c = sum(f * l for f, l in zip(act, ls))
            best = c if best is None or c < best else best
    return best


@pytest.mark.parametrize("seed", range(40))
def test_limited_lengths_optimal(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(2, 7))
    freq = np.zeros(k + 2, dtype=np.uint32)
    freq[rng.choice(k + 2, size=k, replace=False)] = rng.integers(1, 1000, size=k) ** int(rng.integers(1, 3))
    max_bits = int(rng.integers(3, 6))
    lens = orc.limited_lengths(freq, max_bits)
    assert all((l == 0) == (f == 0) for l, f in zip(lens, freq))
    assert max(lens) <= max_bits
    assert abs(_kraft(lens) - 1.0) < 1e-12
    assert sum(int(f) * int(l) for f, l in zip(freq, lens)) == _best_cost(list(freq), max_bits)


def test_limited_lengths_large_alphabet_bounds():
    rng = np.random.default_rng(1)
    for _ in range(50):
        freq = (rng.zipf(1.2, size=288) % 100000).astype(np.uint32)
        lens = orc.limited_lengths(freq, 15)
        assert max(lens) <= 15 and abs(_kraft(lens) - 1.0) < 1e-9
        lens7 = orc.limited_lengths(freq[:19], 7)
        assert max(lens7) <= 7 and abs(_kraft(lens7) - 1.0) < 1e-9


def test_single_symbol_gets_length_one():
    f = np.zeros(32, dtype=np.uint32)
    f[5] = 9
    assert list(orc.limited_lengths(f, 15)) == [0] * 5 + [1] + [0] * 26


# ---- restatement == orot, compared directly where orot was built here
@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_oracle_equals_orot_build(kind):
    data = corpus.generate(kind, 200_000, 21)
    for level in range(1, 10):
        assert np.array_equal(orc.lz77(data, level), orc.ref_lz77(data, level)), level
    for level in (0, 1, 2, 3):
        for fmt in (0, 1, 2):
            assert orc.parallel_compress(data, level, fmt) == orc.ref_parallel_compress(data, level, fmt)


# ---- orot's streaming compressor writes its parallel path's bytes wherever its
# stream is valid: the whole input in one 128 KiB block that is flushed as the
# final one, with no flush before (src/compress/compressor.cpp:21-97: a full
# block is written as soon as it fills, without BFINAL, so 128 KiB fed before
# DEFLATE_FINISH never gets a final block).  This is the definition the GPU
# stream keeps.
@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_orot_stream_is_parallel_bytes(kind):
    for n, one_call in ((1000, False), (4096, False), (65536, False), (131071, False), (131072, True)):
        data = corpus.generate(kind, n, 33)
        for level in (0, 1, 2, 3):
            for fmt in (0, 1, 2):
                got = orc.ref_stream(data, level, fmt, one_call)
                assert got == orc.parallel_compress(data, level, fmt), (n, level, fmt)


# ---- L6-class matcher against a brute-force restatement of its definition:
# the chain of x is every earlier position with the same 12-bit hash (whole
# 4-byte window in the buffer), nearest first; at most max_chain of them are
# examined, the walk ends at the first one more than 32 KiB back, a candidate
# counts only when its first 4 bytes equal x's, a longer match replaces the
# best (ties keep the nearer), and the walk stops at nice_len.
_L6_LEVELS = {4: (16, 128), 6: (32, 258), 9: (256, 258), 12: (2048, 258)}
_L6_LAZY = {4: 1, 5: 1, 6: 1, 7: 2, 8: 2, 9: 3, 10: 3, 11: 3, 12: 3}


def _h12(b: bytes, p: int) -> int:
    return ((int.from_bytes(b[p:p + 4], "little") * 0x9E3779B1) & 0xFFFFFFFF) >> 20


def _brute_l6_match(b: bytes, level: int):
    max_chain, nice = _L6_LEVELS[level]
    n = len(b)
    by_hash = {}
    ml, md = [0] * n, [0] * n
    for x in range(n):
        if x + 4 > n:
            continue
        h = _h12(b, x)
        chain = by_hash.get(h, [])
        best = bd = 0
        cap = min(258, n - x)
        for k, q in enumerate(reversed(chain)):
            if k >= max_chain or x - q > 32768:
                break
            if b[q:q + 4] != b[x:x + 4]:
                continue
            l = 0
            while l < cap and b[q + l] == b[x + l]:
                l += 1
            if l > best:
                best, bd = l, x - q
                if l >= nice:
                    break
        if best >= 3:
            ml[x], md[x] = best, bd
        by_hash.setdefault(h, []).append(x)
    return ml, md


@pytest.mark.parametrize("level", sorted(_L6_LEVELS))
@pytest.mark.parametrize("kind", ["text", "records", "zeros", "random"])
def test_l6_match_bruteforce(level, kind):
    d = corpus.generate(kind, 6000, 41 + level).tobytes()
    # a repeated piece far back exercises the 32 KiB window
    d = d + corpus.generate("random", 30000, 5).tobytes() + d[:3000]
    ml, md = orc.l6_match(d, level)
    bml, bmd = _brute_l6_match(d, level)
    assert list(ml) == bml and list(md) == bmd


@pytest.mark.parametrize("level", [4, 6, 9])
def test_l6_tokens_rebuild_input(level):
    """Invariant of the token stream: literals and (length, distance) copies
    rebuild the input, every distance is within 32 KiB and every length is
    in [3, 258]."""
    d = np.concatenate([corpus.generate(k, 20000, 8) for k in corpus.GENERATORS]).tobytes()
    out = bytearray()
    for t in orc.l6_lz77(d, level):
        t = int(t)
        if t & 0x80000000:
            L, dist = (t >> 16) & 0x1FF, t & 0xFFFF
            assert 3 <= L <= 258 and 1 <= dist <= 32768 and dist <= len(out)
            for _ in range(L):
                out.append(out[-dist])
        else:
            out.append(t & 0xFF)
    assert bytes(out) == d


def test_l6_match_window_edge():
    """A copy exactly 32768 bytes back is found; one 32769 back is not (the
    window of RFC 1951 distances), and the brute force agrees everywhere."""
    rng = np.random.default_rng(77)
    a = rng.integers(0, 256, 200, dtype=np.uint8).tobytes()
    fill = rng.integers(0, 256, 32768 - 200, dtype=np.uint8).tobytes()
    near = a + fill + a                     # second copy at distance 32768
    far = a + fill + b"\x00" + a            # second copy at distance 32769
    for level in (4, 6):
        ml, md = orc.l6_match(near, level)
        assert ml[32768] >= 128 and md[32768] == 32768
        ml2, md2 = orc.l6_match(far, level)
        assert md2[32769] != 32769
        assert (list(ml), list(md)) == _brute_l6_match(near, level)


@pytest.mark.parametrize("level", [4, 9, 12])
@pytest.mark.parametrize("halo", [1, 5000, 32768])
def test_l6_match_bruteforce_with_dictionary(level, halo):
    """With a dictionary of `halo` bytes before the chunk, M(x) of the
    chunk's positions is the brute force over the whole buffer: dictionary
    positions are in the chains and may be match sources."""
    pre = np.concatenate([corpus.generate("text", 3000, 7), corpus.generate("records", 2000, 8),
                          corpus.generate("random", 40000, 9)]).tobytes()
    dic = pre[len(pre) - halo:]
    body = (pre[:2500] + corpus.generate("text", 3000, 10).tobytes() + dic[-1500:])
    buf = dic + body
    ml, md = orc.l6_match_h(buf, halo, level)
    bml, bmd = _brute_l6_match(buf, level)
    assert list(ml) == bml[halo:] and list(md) == bmd[halo:]


def _l6_lazy_restated(s: bytes, ml, md, depth: int):
    """orot's lazy parse (src/deflate/lz77.cpp:442-494; the restatement of
    SURVEY appendix A.5) with the L6 class's M(p) as the main probe and
    M(p + 1) as the lazy probe: a literal is emitted while the probe beats
    the current match by more than one byte, at most `depth` times."""
    n = len(s)
    toks = []
    pos = 0
    while pos < n:
        if pos + 3 > n or ml[pos] < 3:
            toks.append(s[pos])
            pos += 1
            continue
        b, d = int(ml[pos]), int(md[pos])
        k = depth
        while k > 0 and pos + b + 1 < n and pos + 4 <= n and int(ml[pos + 1]) > b + 1:
            toks.append(s[pos])
            pos += 1
            b, d = int(ml[pos]), int(md[pos])
            k -= 1
        toks.append(0x80000000 | (b << 16) | d)
        pos += b
    return toks


@pytest.mark.parametrize("level", [4, 6, 9, 12])
@pytest.mark.parametrize("kind", corpus.GENERATORS + ("mixed",))
@pytest.mark.parametrize("halo", [0, 777, 32768])
def test_l6_lazy_parse_restated(level, kind, halo):
    """The oracle's L6-class token stream is the lazy parse restated over its
    M(x) (pinned by the brute force above), with and without a dictionary."""
    if kind == "mixed":
        body = np.concatenate([corpus.generate(k, 9000, 31 + level) for k in corpus.GENERATORS]).tobytes()
    else:
        body = corpus.generate(kind, 36000, 31 + level).tobytes()
    dic = corpus.generate("text", halo, 5).tobytes() if halo else b""
    buf = dic + body
    ml, md = orc.l6_match_h(buf, halo, level)
    got = [int(t) for t in orc.l6_lz77_h(buf, halo, level)]
    assert got == _l6_lazy_restated(body, ml, md, _L6_LAZY[level])


def test_l6_lazy_parse_defers_literals():
    """The restated rule is not vacuous on these inputs: at levels 7-12 the
    parse defers more than one literal in a row somewhere."""
    body = corpus.generate("text", 36000, 38).tobytes()
    ml, md = orc.l6_match_h(body, 0, 9)
    toks = _l6_lazy_restated(body, ml, md, 3)
    greedy = _l6_lazy_restated(body, ml, md, 0)
    assert toks != greedy


def _blocks(tokens, raw, max_syms=16384):
    """Token blocks of at most max_syms tokens and their raw spans."""
    out = []
    r = 0
    for t0 in range(0, max(1, len(tokens)), max_syms):
        tb = tokens[t0:t0 + max_syms]
        span = sum(((int(t) >> 16) & 0x1FF) if int(t) & 0x80000000 else 1 for t in tb)
        out.append((np.asarray(tb, dtype=np.uint32), raw[r:r + span]))
        r += span
    return out


@pytest.mark.parametrize("level", [4, 6, 9])
@pytest.mark.parametrize("kind", corpus.GENERATORS + ("mixed", "tiny"))
def test_l6_block_choice_is_smallest(level, kind):
    """Each L6-class block is emitted as stored, fixed and dynamic; every
    forced emission inflates, and the encoder picks the smallest of the three
    emitted sizes (from a byte-aligned start), ties going stored, then fixed,
    then dynamic."""
    if kind == "mixed":
        raw = np.concatenate([corpus.generate(k, 30000, 7 + level) for k in corpus.GENERATORS]).tobytes()
    elif kind == "tiny":
        raw = b"abcabcabcab"
    else:
        raw = corpus.generate(kind, 120000, 7 + level).tobytes()
    toks = orc.l6_lz77(raw, level)
    seen = set()
    done = b""
    for tb, rb in _blocks(list(toks), raw):
        sizes = []
        for btype in (0, 1, 2):
            b, bits = orc.encode_block_forced(tb, rb, btype)
            # matches may reach into the blocks before: they are the dictionary
            z = zlib.decompressobj(-15, zdict=done[-32768:]) if done else zlib.decompressobj(-15)
            assert z.decompress(b) == rb, btype
            assert (b[0] >> 1) & 3 == btype
            sizes.append(bits)
        chosen, bits = orc.encode_blocks_dyn(tb, rb, 1, 1 << 30)
        want = min(range(3), key=lambda t: (sizes[t], t))
        assert (chosen[0] >> 1) & 3 == want and bits == sizes[want], (sizes, bits)
        seen.add(want)
        done += rb
    if kind == "random":
        assert seen == {0}
    if kind == "text":
        assert 2 in seen


# ---- orot's own inflater accepts what the HIP path writes (every output
# inflates through orot's decompressor and through zlib), within the range
# where orot's inflater itself is sound.
def _stored_after_fixed(n_lit: int, n_stored: int) -> tuple[bytes, bytes]:
    """A non-final fixed-Huffman block of n_lit literals followed by a final
    stored block of n_stored bytes: a valid RFC 1951 stream."""
    a = corpus.generate("text", n_lit, 5).tobytes()
    b = corpus.generate("random", n_stored, 6).tobytes()
    s1, bits = orc.encode_block_forced(np.frombuffer(a, dtype=np.uint8).astype(np.uint32), a, 1, is_last=0)
    buf = bytearray(s1 + bytes(n_stored + 8))
    pos = bits + 3                      # BFINAL = 1, BTYPE = 00
    buf[bits >> 3] |= 1 << (bits & 7)
    pos = (pos + 7) & ~7
    buf[pos >> 3:(pos >> 3) + 4] = n_stored.to_bytes(2, "little") + ((~n_stored) & 0xFFFF).to_bytes(2, "little")
    pos += 32
    buf[pos >> 3:(pos >> 3) + n_stored] = b
    return bytes(buf[:(pos >> 3) + n_stored]), a + b


@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("n_lit", [10, 1000])
def test_orot_inflater_rejects_stored_after_huffman(n_lit):
    """orot's inflater loses the payload bytes its bit buffer already holds
    when a stored block follows a Huffman block (src/decompress/
    decompressor.cpp:253-300 copies the payload from next_in only).  zlib
    accepts the stream; orot's deflate_decompress reports DEFLATE_DATA_ERROR
    (on longer inputs it also loops or crashes).  orot's own parallel
    compressor writes such streams: a Zipf-text chunk followed by a
    random-bytes chunk at level 1.  Outputs of this kind are checked through
    zlib only."""
    comp, want = _stored_after_fixed(n_lit, 50)
    assert zlib.decompress(comp, -15) == want
    got = orc.ref_decompress_guarded(comp, 0, len(want) + 64)
    assert got is not None and got[0] == -1


@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("level", [0, 1, 3, 4, 6, 9, 12])
@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("kind", corpus.GENERATORS)
def test_orot_inflates_gpu_path_output(level, fmt, kind):
    """One content kind per buffer: no stored block follows a Huffman block
    (random bytes are stored throughout, the other kinds never stored)."""
    data = corpus.generate(kind, 300_000, 61 + level)
    comp = orc.gpu_parallel_compress(data, level, fmt, 0, threads=4)
    assert zlib.decompress(comp, WBITS[fmt]) == data.tobytes()
    got = orc.ref_decompress_guarded(comp, fmt, data.size + 1024)
    assert got is not None and got[0] == 0 and got[1] == data.tobytes(), got and got[0]


@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("kind", ["text", "records", "zeros"])
def test_orot_inflates_gpu_path_multichunk(kind):
    """Several chunks per buffer, with dictionary references across chunk
    boundaries at levels 4-12 and small blocks."""
    data = corpus.generate(kind, 1_500_000, 62)
    for level, block in ((1, 65536), (6, 0), (6, 65536), (9, 200_000), (12, 0)):
        comp = orc.gpu_parallel_compress(data, level, 2, block, threads=4)
        assert zlib.decompress(comp, 31) == data.tobytes()
        got = orc.ref_decompress_guarded(comp, 2, data.size)
        assert got is not None and got[0] == 0 and got[1] == data.tobytes(), (level, block, got and got[0])


# ---- block sizes beyond orot's defaults and the whole-buffer entry point,
# against orot's own build and orot's golden whole-buffer bytes
@pytest.mark.skipif(not orc.ref_available(), reason="orot not built in this checkout (oracle/_ref)")
@pytest.mark.parametrize("block", [131073, 600_000, 1 << 20, 3_000_017])
def test_oracle_big_blocks_equal_orot_build(block):
    data = np.concatenate([corpus.generate(k, 1_000_000, 13 + i) for i, k in enumerate(corpus.GENERATORS)])
    for level in (1, 2, 3):
        for fmt in (0, 2):
            assert orc.parallel_compress(data, level, fmt, block, threads=4) == \
                orc.ref_parallel_compress(data, level, fmt, block, threads=4), (level, fmt)


def test_golden_whole_buffer(fixtures):
    for e in fixtures["whole"]:
        n, seed = e["n"], e["seed"]
        if e["kind"] == "mixed":
            q = n // 4
            data = np.concatenate([corpus.generate(k, q if i < 3 else n - 3 * q, seed + i)
                                   for i, k in enumerate(corpus.GENERATORS)])
        else:
            data = corpus.generate(e["kind"], n, seed)
        got = orc.whole_compress(data, e["level"], e["format"])
        if "hex" in e:
            assert got.hex() == e["hex"], e
        assert len(got) == e["size"] and hashlib.sha256(got).hexdigest() == e["sha256"], e


def test_whole_buffer_closed_forms():
    """orot stores an input with no match (block_compressor.cpp:60-61): the
    empty input is one empty stored block, "a" a one-byte stored block."""
    assert orc.whole_compress(b"", 1, 0) == bytes.fromhex("010000ffff")
    assert orc.whole_compress(b"a", 1, 0) == bytes.fromhex("010100feff61")
    assert orc.whole_compress(b"a", 6, 0) == bytes.fromhex("4b0400")  # L6 class: the chunked pipeline


# ---- the L1-class dictionary mode (the HIP path's own, off by default): its
# token stream against a brute force of its definition (every dictionary and
# chunk position before p is in the table; orot's fast-path rules otherwise)
def _greedy_from(buf, cand, start):
    n = len(buf)
    toks = []
    p = start
    while p < n:
        l = 0
        if p + 4 <= n and cand[p]:
            m = min(258, n - p)
            d = cand[p]
            while l < m and buf[p + l] == buf[p - d + l]:
                l += 1
        if l >= 3:
            toks.append(0x80000000 | (l << 16) | cand[p])
            p += l
        else:
            toks.append(buf[p])
            p += 1
    return toks


@pytest.mark.parametrize("level,bits", [(1, 12), (3, 14)])
@pytest.mark.parametrize("kind", corpus.GENERATORS)
@pytest.mark.parametrize("halo", [0, 5, 4000, 32768])
def test_l1_dict_tokens_bruteforce(level, bits, kind, halo):
    dic = corpus.generate("text", halo, 3).tobytes()
    body = corpus.generate(kind, 30_000, 4).tobytes()
    buf = dic + body[:1000] + dic[-min(halo, 2000):] + body[1000:]
    got = [int(t) for t in orc.l1_lz77_h(buf, halo, level)]
    assert got == _greedy_from(buf, _brute_candidates(buf, bits), halo)
    if halo == 0:
        assert got == [int(t) for t in orc.lz77(buf, level)]


def test_l1_dict_mode_stream():
    """Dictionary mode: every chunk after the first starts with the 32 KiB
    before it; the stream inflates, is smaller than orot's at small blocks,
    and equals orot's bytes for a single chunk."""
    data = np.concatenate([corpus.generate(k, 400_000, 5 + i) for i, k in enumerate(corpus.GENERATORS)])
    for level in (1, 2):
        d = orc.l1_dict_compress(data, level, 2, 65536, threads=4)
        p = orc.parallel_compress(data, level, 2, 65536, threads=4)
        assert zlib.decompress(d, 31) == data.tobytes()
        assert len(d) < len(p)
        one = data[:60_000]
        assert orc.l1_dict_compress(one, level, 1, 65536) == orc.parallel_compress(one, level, 1, 65536)
