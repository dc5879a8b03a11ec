"""Writes tests/golden/golden.json: known-answer vectors for the compress path.

Nothing here runs the reference (it cannot run in this image: Rust + WGSL on
wgpu).  The vectors are
  * hand-derived match-finder answers on constructed inputs, each chosen so that
    a plausible mistake in the candidate rule changes the answer (the reasoning
    is written next to each case);
  * the reference's own walk test cases (codec-gpu/src/gpu/walk.rs:575-695),
    whose expected output the reference defines as host_strict_greedy, and
    hand-written lazy-walk cases (codec-gpu/src/gpu/lz77.rs:260-284);
  * Huffman cases from codec-gpu/src/gpu/huffman.rs:158-232 plus a tie-break
    case and length-limit halving cases (two derived by hand, one from the
    heap statement in tests/_vectors.py);
  * symbol-frequency vectors of token streams whose litlen, distance and
    code-length codes exceed their limits (15, 15, 7), so that the halving
    loop runs in each;
  * the empty-token-stream DEFLATE block derived by hand and checked with zlib.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import base64
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _vectors import heap_code_lengths  # noqa: E402


def hash3(b: bytes, p: int) -> int:
    # lz77_hash_build.wgsl:36-43: ((b0<<16|b1<<8|b2) * 0x9E3779B1 mod 2^32) >> 16
    x = (b[p] << 16) | (b[p + 1] << 8) | b[p + 2]
    return ((x * 0x9E3779B1) & 0xFFFFFFFF) >> 16


def lit(b: int):
    return [0, b]


def construct(n: int, plants: dict[int, bytes], query_hashes_from: list[int], seed: int,
              allow_internal: bool = False) -> bytes:
    """n filler bytes (0x80..0xff) with `plants` written at their offsets.  The
    filler is re-drawn until no 3-gram that touches filler hashes like the
    3-grams at `query_hashes_from`; 3-grams inside plants may share that hash
    only at plant starts (or anywhere when allow_internal) -- so the candidates
    of the query positions are exactly the ones the case reasons about."""
    rng = random.Random(seed)
    buf = bytearray(rng.randrange(0x80, 0x100) for _ in range(n))
    for off, pat in plants.items():
        buf[off:off + len(pat)] = pat
    protected = set()
    for off, pat in plants.items():
        protected.update(range(off, off + len(pat)))
    starts = set(plants)
    for _ in range(1000):
        targets = {hash3(buf, q) for q in query_hashes_from}
        bad = []
        for q in range(n - 2):
            if q in starts or hash3(buf, q) not in targets:
                continue
            inside = all(q + k in protected for k in range(3))
            if inside and not allow_internal:
                raise RuntimeError(f"planted 3-gram at {q} collides with a query hash")
            if not inside:
                bad.append(q)
        if not bad:
            return bytes(buf)
        for q in bad:
            for k in range(3):
                if q + k not in protected:
                    buf[q + k] = rng.randrange(0x80, 0x100)
    raise RuntimeError("could not isolate the planted hashes")


def match_kats():
    kats = []
    # hand-derived KATs: q x 64 -> [lit q, (64-p, p) for 1<=p<=61, lit, lit].  Every
    # position hashes alike (one segment): oldest = 0, newest = 61; newest >= p
    # is rejected for p <= 61 so the oldest candidate (p=0) gives distance p.
    q = b"q" * 64
    exp = {0: lit(0x71)}
    for p in range(1, 62):
        exp[p] = [64 - p, p]
    exp[62] = lit(0x71)
    exp[63] = lit(0x71)
    kats.append({"name": "q_x64", "input": q, "expect": exp})
    # "abc" x 8, p = 3: newest "abc" is 21 >= 3 (rejected), oldest 0: dist 3,
    # length runs to the end: 24 - 3 = 21.  p = 21: newest = 21 = p (rejected),
    # oldest 0: dist 21, 3 bytes left.
    kats.append({"name": "abc_x8", "input": b"abc" * 8, "expect": {3: [21, 3], 21: [3, 21], 0: lit(0x61)}})
    # short inputs are all literals (p + 2 >= n)
    kats.append({"name": "two_bytes", "input": b"zz", "expect": {0: lit(0x7A), 1: lit(0x7A)}})
    kats.append({"name": "distinct32", "input": bytes(range(32)), "expect": {p: lit(p) for p in range(32)}})

    # A: equal lengths -> the newest candidate of a segment wins; the oldest is
    # tried (best < 16) but must be strictly longer.  Swapping newest/oldest
    # would give distance 4190.
    plants = {10: b"WXYZa", 100: b"WXYZb", 4200: b"WXYZc"}
    buf = construct(8192, plants, [4200], 1)
    kats.append({"name": "newest_wins_tie", "input": buf, "expect": {4200: [4, 4100]}})
    # B: newest is short (4 < 16) -> the oldest is tried and is longer.
    pat = b"PQRSTUVWXYZpqrstuvw"  # 19 bytes
    plants = {10: pat + b"!", 100: pat[:4] + b"x", 4200: pat + b"?"}
    buf = construct(8192, plants, [4200], 2)
    kats.append({"name": "oldest_when_short", "input": buf, "expect": {4200: [19, 4190]}})
    # C: newest reaches 17 >= 16 -> the oldest (30 bytes) is skipped.
    pat = bytes(range(0x41, 0x41 + 30))  # 'A'.. 30 distinct ASCII bytes
    plants = {10: pat + b"!", 100: pat[:17] + b"x", 4200: pat + b"?"}
    buf = construct(8192, plants, [4200], 3)
    kats.append({"name": "oldest_skipped_at_16", "input": buf, "expect": {4200: [17, 4100]}})
    # D: window edge: dist 32768 accepted, 32769 rejected.
    plants = {100: b"ABCD1", 100 + 32768: b"ABCD2"}
    buf = construct(40000, plants, [100 + 32768], 4)
    kats.append({"name": "window_32768", "input": buf, "expect": {100 + 32768: [4, 32768]}})
    plants = {100: b"ABCD1", 100 + 32769: b"ABCD2"}
    buf = construct(40000, plants, [100 + 32769], 5)
    kats.append({"name": "window_32769", "input": buf, "expect": {100 + 32769: lit(ord("A"))}})
    # E: segments are walked newest first; an equal-length candidate in an
    # older segment does not replace the nearer one.
    plants = {50: b"LMNO1", 4150: b"LMNO2", 8300: b"LMNO3"}
    buf = construct(12288, plants, [8300], 6)
    kats.append({"name": "nearer_segment_wins", "input": buf, "expect": {8300: [4, 4150]}})
    # F: a longer match in an older segment replaces a shorter nearer one.
    plants = {50: b"LMNOPQ1", 4150: b"LMNO2", 8300: b"LMNOPQ3"}
    buf = construct(12288, plants, [8300], 7)
    kats.append({"name": "older_longer_wins", "input": buf, "expect": {8300: [6, 8250]}})
    # G: extension stops at the end of the input (p + len < n).
    plants = {10: b"EFGH" * 4, 200: b"EFGH" * 2}
    buf = construct(208, plants, [200], 8, allow_internal=True)
    kats.append({"name": "ends_at_input_end", "input": buf, "expect": {200: [8, 190]}})
    # H: matches cap at 258 (long repeat of a 5-byte period, 600 bytes).
    body = (b"vwxyz" * 200)[:600]
    buf = construct(1000, {100: body}, [110], 9, allow_internal=True)
    kats.append({"name": "cap_258", "input": buf, "expect": {110: [258, 10]}})
    # I: hash collisions shadow the real candidate.  "EHN" and "_gk" hash like
    # "ABC" under lz77_hash_build.wgsl:36-43 (checked below).  Segment 0 holds
    # EHN at 10 (its oldest entry for that hash), the real ABCD at 500 and _gk
    # at 1000 (its newest).  The query at 4200 (segment 1) is the only such
    # hash in its own segment, so it rejects itself (cand >= p); in segment 0
    # newest (_gk) and then oldest (EHN, since best is 0 < 16) fail the
    # three-byte check, and the real occurrence between them is never tried:
    # a literal.  A wrong multiplier or byte order makes EHN and _gk hash
    # elsewhere, so 500 becomes both oldest and newest and gives (4, 3700).
    assert COLLIDERS_OF_ABC == [b"EHN", b"_gk"]
    plants = {10: b"EHN1", 500: b"ABCD", 1000: b"_gk2", 4200: b"ABCDQ"}
    buf = construct(8192, plants, [4200], 10)
    kats.append({"name": "hash3_collisions_shadow", "input": buf, "expect": {4200: lit(ord("A"))}})
    # J: the colliding 3-gram is the newest entry and fails the three-byte
    # check; best stays 0 < 16, so the oldest entry (the real ABCDE at 10) is
    # tried: (5, 4190).
    plants = {10: b"ABCDE", 1000: b"_gk3", 4200: b"ABCDEx"}
    buf = construct(8192, plants, [4200], 11)
    kats.append({"name": "hash3_collision_newest_then_oldest", "input": buf, "expect": {4200: [5, 4190]}})
    return kats


def _colliders(target: bytes, alphabet: range) -> list[bytes]:
    h = hash3(target, 0)
    out = []
    for b0 in alphabet:
        for b1 in alphabet:
            for b2 in alphabet:
                g = bytes([b0, b1, b2])
                if g != target and hash3(g, 0) == h:
                    out.append(g)
    return out


# printable 3-grams (0x41..0x7a) that collide with "ABC": the first two, sorted
COLLIDERS_OF_ABC = sorted(_colliders(b"ABC", range(0x41, 0x7B)))[:2]


def strict_walk(tokens):
    # host_strict_greedy, codec-gpu/src/gpu/walk.rs:512-525
    out, p = [], 0
    while p < len(tokens):
        t = tokens[p]
        out.append(t)
        p += 1 if t[0] == 0 else t[0]
    return out


def walk_cases():
    cases = []
    L = lambda i: [0, i & 0xFF]
    cases.append(("empty", []))
    cases.append(("all_literals", [L(i) for i in range(32)]))
    t = [L(i) for i in range(8)] + [[5, 3]] + [L(i) for i in range(9, 16)]
    cases.append(("single_back_ref", t))
    t = [[7, 1] if (i % 13 == 0 and i > 0) else L(i) for i in range(4096)]
    cases.append(("many_blocks", t))
    t = [L(i) for i in range(200)] + [[50, 100]] + [L(i) for i in range(251, 400)]
    cases.append(("back_ref_crosses_boundary", t))
    cases.append(("realistic_size_all_literals", [L(i) for i in range(131072)]))
    t = [[100, 50]] + [L(i) for i in range(1, 200)]
    cases.append(("long_match_skips_positions", t))
    t = [[5, 1] if (i % 7 == 0 and i + 10 < 256) else L(i) for i in range(256)]
    cases.append(("realistic_mix", t))
    out = []
    for name, toks in cases:
        exp = strict_walk(toks)
        if name == "realistic_size_all_literals":
            # 131072 literals walk to themselves; stored as a rule to keep the file small
            assert exp == toks
            out.append({"name": name, "all_literals": 131072})
            continue
        out.append({"name": name, "tokens": toks, "expect": exp})
    return out


def lazy_cases():
    # greedy_walk (codec-gpu/src/gpu/lz77.rs:260-284), derived by hand:
    #  p0 match(3,1) but p1 match(4,2) is longer -> literal in[0], then p1 taken,
    #  jumps to 5; p5 match(3,5), p6 literal -> taken, jump to 8; p8.. literals.
    data = bytes(b"abcdefghij")
    toks = [[3, 1], [4, 2], [0, 0x63], [0, 0x64], [0, 0x65], [3, 5], [0, 0x67], [0, 0x68], [0, 0x69], [0, 0x6A]]
    exp = [[0, 0x61], [4, 2], [3, 5], [0, 0x69], [0, 0x6A]]
    cases = [{"name": "demote_then_take", "data": data, "tokens": toks, "expect": exp}]
    #  equal next length does not demote; a literal next never demotes.
    toks = [[4, 1], [4, 1], [0, 0x63], [0, 0x64], [3, 2], [0, 0x66], [0, 0x67], [0, 0x68]]
    exp = [[4, 1], [3, 2], [0, 0x68]]
    cases.append({"name": "no_demote_on_equal", "data": b"abcdefgh", "tokens": toks, "expect": exp})
    #  chain of increasing lengths demotes every step; the last position has no
    #  successor and keeps its match.
    toks = [[3, 1], [4, 1], [5, 1], [6, 1]]
    exp = [[0, ord("w")], [0, ord("x")], [0, ord("y")], [6, 1]]
    cases.append({"name": "demote_chain", "data": b"wxyz", "tokens": toks, "expect": exp})
    return cases


def huffman_cases():
    return [
        {"name": "empty", "freq": [0] * 10, "limit": 15, "expect": [0] * 10},
        {"name": "single", "freq": [0, 0, 0, 100, 0, 0, 0, 0, 0, 0], "limit": 15,
         "expect": [0, 0, 0, 1, 0, 0, 0, 0, 0, 0]},
        {"name": "balanced8", "freq": [1] * 8, "limit": 15, "expect": [3] * 8},
        # tie-break: heap pops (freq, id) with internal ids above every leaf id.
        # (1,0)+(1,1) -> n4:2; then leaves (2,2),(2,3) before n4 -> n5:4;
        # then n4, n5 -> root.  All depths 2.  Popping internal nodes first on
        # ties would give [3,3,2,1].
        {"name": "tie_leaf_first", "freq": [1, 1, 2, 2], "limit": 15, "expect": [2, 2, 2, 2]},
        # (1,0)+(2,1) -> n3:3, (3,2)+(3,n3) -> root: sym2 depth 1, syms 0,1 depth 2
        {"name": "three", "freq": [1, 2, 3], "limit": 15, "expect": [2, 2, 1]},
        # Length-limit halving (huffman.rs:44-58: every f > 1 becomes f / 2,
        # rounded down), derived by hand.  Fibonacci frequencies give the chain
        # tree [8,8,7,6,5,4,3,2,1]: 8 > 7, so halve to [1,1,1,1,2,4,6,10,17].
        # Heap pops: (1,0)+(1,1)=n9:2, (1,2)+(1,3)=n10:2, (2,4)+(2,n9)=n11:4,
        # (2,n10)+(4,5)=n12:6, (4,n11)+(6,6)=n13:10, (6,n12)+(10,7)=n14:16,
        # (10,n13)+(16,n14)=n15:26, (17,8)+(26,n15)=root.  Depths: sym8 1;
        # n13, n14 2; syms 6, 7 and n11, n12 3; syms 4, 5 and n9, n10 4;
        # syms 0-3 5.  Rounding up instead ((f+1)/2 -> [1,1,1,2,3,4,7,11,17])
        # gives [5,5,4,4,4,3,3,2,2].
        {"name": "halve_once_limit7", "freq": [1, 1, 2, 3, 5, 8, 13, 21, 34], "limit": 7,
         "expect": [5, 5, 5, 5, 4, 4, 3, 3, 1]},
        # The same at limit 4: [5,5,5,5,4,4,3,3,1] still exceeds 4, so halve
        # again to [1,1,1,1,1,2,3,5,8]: (1,0)+(1,1)=n9:2, (1,2)+(1,3)=n10:2,
        # (1,4)+(2,5)=n11:3, (2,n9)+(2,n10)=n12:4, (3,6)+(3,n11)=n13:6,
        # (4,n12)+(5,7)=n14:9, (6,n13)+(8,8)=n15:14, n14+n15=root.  Depths:
        # n14, n15 1; n12, sym7, n13, sym8 2; n9, n10, sym6, n11 3; syms 0-5 4.
        {"name": "halve_twice_limit4", "freq": [1, 1, 2, 3, 5, 8, 13, 21, 34], "limit": 4,
         "expect": [4, 4, 4, 4, 4, 4, 3, 2, 2]},
        # Limit 15 on 17 Fibonacci frequencies (natural depth 16), expected
        # lengths from heap_code_lengths (tests/_vectors.py: a direct statement
        # of the reference's BinaryHeap loop, independent of the oracle's C code).
        {"name": "halve_limit15_fib17", "freq": FIB[:17], "limit": 15,
         "expect": heap_code_lengths(FIB[:17], 15)},
    ]


FIB = [1, 1]
while len(FIB) < 24:
    FIB.append(FIB[-1] + FIB[-2])


# Token streams whose Huffman trees exceed the length limits, so that the
# halving loop runs for the litlen code (limit 15), the distance code (limit
# 15) and the code-length code (limit 7).  Stored as symbol frequencies:
# `ll` counts per litlen symbol (0-255 literal bytes, 257-285 length codes;
# 256 is the end of block, added by the encoder) and `dc` counts per
# distance code; tests expand them with expand_tokens() in tests/_vectors.py.
def encode_halving_cases():
    cases = []
    # litlen: bytes 0..18 with counts 1, 2, 3, 5, ...; with the end-of-block
    # symbol (count 1) the litlen counts are Fibonacci numbers (natural depth 18)
    ll = [0] * 286
    for i in range(19):
        ll[i] = FIB[i + 1]
    cases.append({"name": "litlen_depth_gt15", "ll": ll, "dc": [0] * 30})
    # distance: 20 distance codes with Fibonacci counts, all matches of length 3
    # (length code 257), plus one literal
    dc = [0] * 30
    for i in range(20):
        dc[i] = FIB[i]
    ll = [0] * 286
    ll[257] = sum(dc)
    ll[65] = 1
    cases.append({"name": "dist_depth_gt15", "ll": ll, "dc": dc})
    # code-length code: a stream whose combined litlen/distance code lengths
    # RLE into meta-symbol counts of natural depth 8 (> 7).  Found by a seeded
    # search: litlen counts 2000 * r^-j (floored, at least 1) over a random
    # order of the litlen symbols, distance counts from a Dirichlet draw.
    ll, dc = meta_depth8_case()
    cases.append({"name": "meta_depth_gt7", "ll": ll, "dc": dc})
    return cases


def meta_depth8_case():
    """Trial 17 of the seeded search (numpy PCG64, seed 0); earlier trials are
    replayed only to advance the generator.  The property (meta code of
    natural depth > 7) is checked by tests/test_oracle.py on the encoded
    header."""
    import numpy as np

    rng = np.random.default_rng(0)
    for trial in range(18):
        r = rng.uniform(1.02, 1.3)
        nused = rng.integers(40, 286)
        ranks = rng.permutation(286)
        ll = np.zeros(286, np.int64)
        for j, s in enumerate(ranks[:nused]):
            if s != 256:
                ll[s] = max(1, int(2000 * r ** (-j)))
        nm = int(ll[257:286].sum())
        dc = np.zeros(30, np.int64)
        if nm:
            w = rng.dirichlet(np.ones(30) * 0.3)
            dc = np.floor(w * nm).astype(np.int64)
            dc[0] += nm - dc.sum()
    return [int(x) for x in ll], [int(x) for x in dc]


def main():
    out = {
        "match_kats": [
            {"name": k["name"], "input_b64": base64.b64encode(k["input"]).decode(),
             "expect": {str(p): v for p, v in k["expect"].items()}}
            for k in match_kats()
        ],
        "walk_strict": walk_cases(),
        "walk_lazy": [
            {"name": c["name"], "data_b64": base64.b64encode(c["data"]).decode(), "tokens": c["tokens"],
             "expect": c["expect"]}
            for c in lazy_cases()
        ],
        "huffman": huffman_cases(),
        "encode_halving": encode_halving_cases(),
        # empty token stream: litlen {256:1} -> lens[0]=lens[256]=1 (fix-up),
        # dist none -> lens[0]=lens[1]=1; HLIT 0, HDIST 1, RLE 1,18(127),18(127),
        # 17(..)... ; verified with zlib inflate and gzip -t.
        "encode_empty_hex": "05c181000000000010ffd508",
    }
    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
