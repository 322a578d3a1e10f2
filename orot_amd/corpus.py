"""Seeded synthetic corpora for the compress benchmarks and parity tests.

Four content generators, all deterministic in (seed, size):

* ``zeros``   -- all-zero bytes.
* ``text``    -- space-separated words drawn Zipf(s=1.1) from a 4096-word
  vocabulary of 3-10 random lowercase letters.  A word's separator is a
  newline instead of a space when the text position crosses a multiple of 80
  columns, giving a newline about every 80 characters.
* ``random``  -- uniform random bytes.
* ``records`` -- 64-byte records of 8 little-endian u64 fields; fields 0-3 are
  drawn from a 16-value palette, fields 4-7 are uniform random.

The PRNG is splitmix64 (Vigna, public domain) used as a counter-based
generator, so every stream is vectorised with numpy.
"""
from __future__ import annotations

import numpy as np

GENERATORS = ("zeros", "text", "random", "records")

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    """splitmix64 finaliser applied to state x (numpy uint64 array or scalar)."""
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def stream(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """count pseudo-random u64 values of the stream keyed by seed."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
        ctr = np.arange(offset, offset + count, dtype=np.uint64) * _GOLDEN + base
    return splitmix64(ctr)


def buffer_seed(master: int, index: int) -> int:
    """Per-buffer seed = splitmix64(master xor index)."""
    return int(splitmix64(np.uint64((master ^ index) & 0xFFFFFFFFFFFFFFFF)))


_VOCAB_CACHE: dict = {}


def _vocab(seed: int):
    if seed in _VOCAB_CACHE:
        return _VOCAB_CACHE[seed]
    r = stream(seed ^ 0x5A5A, 4096 * 11)
    lens = (3 + (r[:4096] % np.uint64(8))).astype(np.int64)
    letters = (np.uint64(ord("a")) + r[4096:].reshape(4096, 10) % np.uint64(26)).astype(np.uint8)
    ranks = np.arange(1, 4097, dtype=np.float64)
    w = ranks ** -1.1
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    _VOCAB_CACHE[seed] = (lens, letters, cdf)
    return _VOCAB_CACHE[seed]


def zipf_text(n: int, seed: int) -> np.ndarray:
    lens, letters, cdf = _vocab(0x7E47)  # one vocabulary for every buffer
    out = np.empty(n, dtype=np.uint8)
    pos = 0
    offset = 0
    while pos < n:
        nw = max(1024, (n - pos) // 5 + 64)
        r = stream(seed, nw, offset)
        offset += nw
        u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        wi = np.minimum(np.searchsorted(cdf, u, side="right"), 4095)
        wl = lens[wi]
        starts = np.empty(nw + 1, dtype=np.int64)
        starts[0] = pos
        np.cumsum(wl + 1, out=starts[1:])
        starts[1:] += pos
        total = int(starts[-1] - pos)
        idx = np.repeat(np.arange(nw), wl + 1)
        col = np.arange(total, dtype=np.int64) - (starts[:-1] - pos)[idx]
        wl_rep = wl[idx]
        is_sep = col == wl_rep
        piece = np.empty(total, dtype=np.uint8)
        ch = letters[wi[idx[~is_sep]], col[~is_sep]]
        piece[~is_sep] = ch
        # separator after word k sits at absolute position starts[k+1]-1
        sep_pos = starts[1:] - 1
        nl = (sep_pos // 80) != (starts[:-1] // 80)
        piece[is_sep] = np.where(nl, ord("\n"), ord(" ")).astype(np.uint8)
        take = min(total, n - pos)
        out[pos:pos + take] = piece[:take]
        pos += take
    return out


def random_bytes(n: int, seed: int) -> np.ndarray:
    words = stream(seed, (n + 7) // 8)
    return words.view(np.uint8)[:n].copy()


def records(n: int, seed: int) -> np.ndarray:
    nrec = (n + 63) // 64
    palette = stream(seed ^ 0xC0FFEE, 16)
    r = stream(seed, nrec * 8).reshape(nrec, 8)
    rec = np.empty((nrec, 8), dtype=np.uint64)
    rec[:, :4] = palette[(r[:, :4] % np.uint64(16)).astype(np.int64)]
    rec[:, 4:] = r[:, 4:]
    return rec.view(np.uint8).reshape(-1)[:n].copy()


def generate(kind: str, n: int, seed: int) -> np.ndarray:
    if kind == "zeros":
        return np.zeros(n, dtype=np.uint8)
    if kind == "text":
        return zipf_text(n, seed)
    if kind == "random":
        return random_bytes(n, seed)
    if kind == "records":
        return records(n, seed)
    raise ValueError(f"unknown generator {kind!r}")


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """floor(exp(U(ln lo, ln hi))) sizes."""
    r = stream(seed, count)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
