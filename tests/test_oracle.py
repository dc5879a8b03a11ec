"""The CPU oracle pinned against known answers, an independent statement of the
candidate rule, decoders and invariants (no GPU needed)."""
import base64
import gzip
import itertools
import json
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def tok_pairs(t):
    return [(int(x) >> 16, int(x) & 0xFFFF) for x in t]


def pack(pairs):
    return np.array([(l << 16) | v for l, v in pairs], dtype=np.uint32)


@pytest.mark.parametrize("kat", GOLD["match_kats"], ids=lambda k: k["name"])
def test_match_find_known_answers(kat):
    data = base64.b64decode(kat["input_b64"])
    t = O.match_find(data)
    for p, (l, v) in kat["expect"].items():
        got = (int(t[int(p)]) >> 16, int(t[int(p)]) & 0xFFFF)
        assert got == (l, v), f"{kat['name']} position {p}"


def _brute_inputs():
    rng = np.random.default_rng(11)
    yield b""
    yield b"a"
    yield b"ab"
    yield b"abc"
    yield b"aaaa"
    yield bytes(rng.integers(0, 4, 9000, dtype=np.uint8))  # small alphabet, many collisions
    yield bytes(rng.integers(0, 256, 9000, dtype=np.uint8))
    yield (b"the quick brown fox jumps over the lazy dog " * 300)[:12000]
    yield bytes(rng.integers(97, 100, 13000, dtype=np.uint8))


@pytest.mark.parametrize("i,data", list(enumerate(_brute_inputs())))
def test_match_find_table_equals_brute(i, data):
    """Dense-table restatement == table-free scan of each segment."""
    assert (O.match_find(data) == O.match_find_brute(data)).all()


def _invariants(data: bytes, t: np.ndarray):
    n = len(data)
    for p, x in enumerate(t):
        l, v = int(x) >> 16, int(x) & 0xFFFF
        if l == 0:
            assert v == data[p]
            continue
        assert 3 <= l <= 258 and 1 <= v <= 32768 and v <= p and p + l <= n
        assert data[p - v:p - v + l] == data[p:p + l]
        # the match length is maximal for its candidate (stops at a mismatch, 258 or the end)
        if l < 258 and p + l < n:
            assert data[p - v + l] != data[p + l]


@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
def test_match_find_invariants_and_round_trip(kind):
    import gpzip_amd

    data = gpzip_amd.synth(kind, 1 << 17, 5).tobytes()
    t = O.match_find(data)
    _invariants(data, t)
    w = O.greedy_walk(t, data)
    assert O.reconstruct(w, len(data)) == data
    s = O.strict_walk(t)
    assert O.reconstruct(s, len(data)) == data


@pytest.mark.parametrize("case", GOLD["walk_strict"], ids=lambda c: c["name"])
def test_strict_walk_reference_cases(case):
    if "all_literals" in case:
        toks = pack([(0, i & 0xFF) for i in range(case["all_literals"])])
        exp = toks
    else:
        toks, exp = pack(case["tokens"]), pack(case["expect"])
    assert (O.strict_walk(toks) == exp).all()


@pytest.mark.parametrize("case", GOLD["walk_lazy"], ids=lambda c: c["name"])
def test_lazy_walk_cases(case):
    data = base64.b64decode(case["data_b64"])
    assert (O.greedy_walk(pack(case["tokens"]), data) == pack(case["expect"])).all()


@pytest.mark.parametrize("case", GOLD["huffman"], ids=lambda c: c["name"])
def test_huffman_cases(case):
    lens = O.build_code_lengths(case["freq"], case["limit"])
    assert list(lens) == case["expect"]


def test_huffman_reference_properties():
    # huffman.rs:158-232
    lens = O.build_code_lengths([100, 50, 25, 10, 5, 2, 1, 1], 15)
    assert lens[0] <= lens[7]
    lens = O.build_code_lengths(list(range(1, 21)), 6)
    assert max(lens) <= 6 and sum(2.0 ** -int(l) for l in lens if l) <= 1 + 1e-9
    lens = [3, 3, 3, 3, 3, 2, 4, 4]
    codes = O.canonical_codes(lens)
    for i, j in itertools.permutations(range(8), 2):
        if lens[j] < lens[i]:
            assert codes[i] >> (lens[i] - lens[j]) != codes[j]


def _optimal_cost(freq):
    """Minimum sum f*l over all prefix codes (brute force over length vectors)."""
    used = [f for f in freq if f]
    best = None
    for ls in itertools.product(range(1, len(used)), repeat=len(used)):
        if sum(2.0 ** -l for l in ls) <= 1 + 1e-12:
            c = sum(f * l for f, l in zip(used, ls))
            best = c if best is None else min(best, c)
    return best


def test_huffman_optimal_on_small_alphabets():
    rng = np.random.default_rng(3)
    for _ in range(60):
        k = int(rng.integers(2, 7))
        freq = [int(x) for x in rng.integers(1, 40, k)]
        lens = O.build_code_lengths(freq, 15)
        assert sum(f * int(l) for f, l in zip(freq, lens)) == _optimal_cost(freq)


def test_huffman_length_limit_kicks_in():
    # Fibonacci frequencies give a depth-(n-1) tree; limit 15 forces halving
    fib = [1, 1]
    while len(fib) < 24:
        fib.append(fib[-1] + fib[-2])
    lens = O.build_code_lengths(fib, 15)
    assert max(lens) <= 15 and sum(2.0 ** -int(l) for l in lens) <= 1 + 1e-12
    lens7 = O.build_code_lengths([1 << i for i in range(19)], 7)
    assert max(lens7) <= 7


def test_encode_empty_known_answer():
    b = O.encode_block(np.zeros(0, np.uint32))
    assert b.hex() == GOLD["encode_empty_hex"]
    assert zlib.decompress(b, -15) == b""


def test_crc32_matches_zlib():
    rng = np.random.default_rng(1)
    for n in [0, 1, 7, 4096, 100003]:
        d = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.crc32(d) == zlib.crc32(d)


def _parse_dynamic_header(block: bytes):
    """Independent reader of a DEFLATE dynamic block header: returns (hlit,
    hdist, meta_lens, rle symbols, code lengths)."""
    bits = int.from_bytes(block, "little")
    pos = 0

    def get(n):
        nonlocal pos
        v = (bits >> pos) & ((1 << n) - 1)
        pos += n
        return v

    assert get(1) == 1 and get(2) == 2
    hlit, hdist, hclen = get(5) + 257, get(5) + 1, get(4) + 4
    order = [16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15]
    mlens = [0] * 19
    for i in range(hclen):
        mlens[order[i]] = get(3)
    # canonical decode table for the meta code
    bl_count = [0] * 8
    for l in mlens:
        if l:
            bl_count[l] += 1
    code, nextc, table = 0, {}, {}
    for b in range(1, 8):
        code = (code + bl_count[b - 1]) << 1
        nextc[b] = code
    for s in range(19):
        if mlens[s]:
            table[(mlens[s], nextc[mlens[s]])] = s
            nextc[mlens[s]] += 1
    syms, lens = [], []
    while len(lens) < hlit + hdist:
        c, l = 0, 0
        while True:
            c = (c << 1) | get(1)
            l += 1
            if (l, c) in table:
                s = table[(l, c)]
                break
        if s < 16:
            syms.append((s, 0))
            lens.append(s)
        elif s == 16:
            r = get(2) + 3
            syms.append((16, r))
            lens += [lens[-1]] * r
        elif s == 17:
            r = get(3) + 3
            syms.append((17, r))
            lens += [0] * r
        else:
            r = get(7) + 11
            syms.append((18, r))
            lens += [0] * r
    return hlit, hdist, mlens, syms, lens


@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
def test_encode_block_decodes_and_header_rules(kind):
    import gpzip_amd

    data = gpzip_amd.synth(kind, 1 << 17, 9).tobytes()
    t = O.match_find(data)
    w = O.greedy_walk(t, data)
    blk = O.encode_block(w)
    assert zlib.decompress(blk, -15) == data
    hlit, hdist, mlens, syms, lens = _parse_dynamic_header(blk)
    # HLIT/HDIST trimmed to the last used code (>= 257 / >= 1)
    assert hlit == 257 or lens[hlit - 1] != 0
    assert hdist == 1 or lens[hlit + hdist - 1] != 0
    # RLE rule of deflate.rs:254-326: zero runs use 18 (11..138) before 17
    # (3..10); repeats of a non-zero length use 16 only after a literal copy.
    seq = lens
    i = 0
    for s, r in syms:
        if s == 18:
            assert 11 <= r <= 138
        if s == 17:
            assert 3 <= r <= 10
            # an 18 would have been used if >= 11 zeros remained in the run
            run = 0
            while i + run < len(seq) and seq[i + run] == 0:
                run += 1
            assert run < 11
        i += 1 if s < 16 else r
    # code lengths respect the limits; meta lengths <= 7
    assert max(lens) <= 15 and max(mlens) <= 7


def test_gzip_chunk_valid_and_split():
    import gpzip_amd

    data = gpzip_amd.synth("binmix", 3 << 20, 2).tobytes()[: (3 << 20) - 12345]
    gz = O.gzip_chunk(data)
    assert gzip.decompress(gz) == data
    # one member per 128 KiB piece, each with the fixed 10-byte header
    assert gz.count(bytes.fromhex("1f8b08000000000000ff")) >= (len(data) + 131071) // 131072
    if shutil.which("gzip"):
        r = subprocess.run(["gzip", "-t"], input=gz, capture_output=True)
        assert r.returncode == 0, r.stderr


def test_gzip_empty_member():
    gz = O.gzip_chunk(b"")
    assert gzip.decompress(gz) == b""
    assert gz[:10] == bytes.fromhex("1f8b08000000000000ff")
    assert gz[10:22].hex() == GOLD["encode_empty_hex"]
