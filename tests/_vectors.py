"""Helpers shared by the CPU and GPU tests: token streams built from the
symbol-frequency fixtures in golden.json, a raw-DEFLATE decode against a
preset window, and an independent heap statement of the reference's Huffman
code-length builder (codec-gpu/src/gpu/huffman.rs:26-117)."""
from __future__ import annotations

import heapq
import zlib

import numpy as np

# RFC 1951 3.2.5 base values of the length codes 257..285 and distance codes 0..29
LEN_BASE = [3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163,
            195, 227, 258]
DIST_BASE = [1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025, 1537, 2049,
             3073, 4097, 6145, 8193, 12289, 16385, 24577]


def expand_tokens(ll, dc) -> np.ndarray:
    """Walked tokens with exactly the given symbol counts: ll[s] literals of
    byte s (s < 256), ll[s] matches at the base length of code s (s > 256),
    paired in order with dc[c] distances at the base of distance code c."""
    toks = []
    for s in range(256):
        toks += [s] * int(ll[s])
    lens = []
    for s in range(257, 286):
        lens += [LEN_BASE[s - 257]] * int(ll[s])
    dists = []
    for c, k in enumerate(dc):
        dists += [DIST_BASE[c]] * int(k)
    assert len(lens) == len(dists), "length and distance counts differ"
    toks += [(l << 16) | d for l, d in zip(lens, dists)]
    return np.array(toks, np.uint32)


def inflate_with_zero_window(block: bytes) -> bytes:
    """Raw DEFLATE decode with 32 KiB of zeros as the preset window, so that
    every distance up to 32768 is valid from the first token on."""
    d = zlib.decompressobj(-15, zdict=bytes(32768))
    return d.decompress(block) + d.flush()


def reconstruct_with_zero_window(tokens) -> bytes:
    buf = bytearray(32768)
    for t in tokens:
        l, v = int(t) >> 16, int(t) & 0xFFFF
        if l == 0:
            buf.append(v)
        else:
            for _ in range(l):
                buf.append(buf[-v])
    return bytes(buf[32768:])


def heap_depths(freq) -> list[int]:
    """standard_huffman + walk (huffman.rs:63-117): pop the two least (freq,
    id) pairs, internal nodes numbered from len(freq) up; depth 0 becomes 1."""
    n = len(freq)
    heap = [(int(f), i) for i, f in enumerate(freq) if f > 0]
    heapq.heapify(heap)
    parent = {}
    nxt = n
    while len(heap) > 1:
        f1, i1 = heapq.heappop(heap)
        f2, i2 = heapq.heappop(heap)
        parent[i1] = parent[i2] = nxt
        heapq.heappush(heap, (f1 + f2, nxt))
        nxt += 1
    out = [0] * n
    for i, f in enumerate(freq):
        if f > 0:
            d, x = 0, i
            while x in parent:
                x = parent[x]
                d += 1
            out[i] = max(d, 1)
    return out


def heap_code_lengths(freq, limit: int) -> list[int]:
    """try_build_code_lengths (huffman.rs:26-60) over heap_depths."""
    work = [int(f) for f in freq]
    if sum(1 for f in work if f) <= 1:
        return [1 if f else 0 for f in work]
    while True:
        lens = heap_depths(work)
        if max(lens) <= limit:
            return lens
        work = [f // 2 if f > 1 else f for f in work]
