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


# ---------------------------------------------------------------------------
# Huffman halving and hash3, pinned so that plausible mistakes fail

from _vectors import expand_tokens, heap_code_lengths, heap_depths, inflate_with_zero_window, \
    reconstruct_with_zero_window  # noqa: E402


def test_huffman_matches_heap_model_random():
    """The oracle's code lengths equal a heap statement of huffman.rs on
    random and skewed vectors, many of which need the halving loop."""
    rng = np.random.default_rng(12)
    halved = 0
    for i in range(1500):
        n = int(rng.integers(2, 287))
        kind = i % 3
        if kind == 0:
            f = rng.integers(0, 1000, n)
        elif kind == 1:  # geometric: deep trees
            f = np.floor(1e6 * rng.uniform(1.2, 2.5) ** -rng.permutation(n).astype(float)).astype(np.int64) + 1
        else:  # sparse with zeros
            f = rng.integers(0, 50, n) * (rng.random(n) < 0.3)
        limit = 7 if (i % 5 == 0 and n <= 19) else 15
        if limit == 7:
            f = f[:19]
        f = [int(x) for x in f]
        exp = heap_code_lengths(f, limit)
        if sum(1 for x in f if x) > 1 and max(heap_depths(f)) > limit:
            halved += 1
        assert list(O.build_code_lengths(f, limit)) == exp, (i, f, limit)
    assert halved > 100


def _meta_freq(syms):
    mf = [0] * 19
    for s, _ in syms:
        mf[s] += 1
    return mf


@pytest.mark.parametrize("case", GOLD["encode_halving"], ids=lambda c: c["name"])
def test_encode_halving_cases(case):
    """Each stream needs the halving loop in the code it targets (natural
    depth above the limit); the oracle's header carries the heap model's
    limited lengths; the block decodes to the tokens."""
    toks = expand_tokens(case["ll"], case["dc"])
    blk = O.encode_block(toks)
    hlit, hdist, mlens, syms, lens = _parse_dynamic_header(blk)
    ll = list(case["ll"])
    ll[256] += 1  # end of block
    ll_lens = heap_code_lengths(ll, 15)
    if sum(1 for x in ll if x) < 2:
        ll_lens[0] = ll_lens[256] = 1
    d_lens = heap_code_lengths(case["dc"], 15)
    used_d = [i for i, x in enumerate(case["dc"]) if x]
    if not used_d:
        d_lens[0] = d_lens[1] = 1
    elif len(used_d) == 1:
        d_lens[used_d[0]] = 1
        d_lens[1 - min(used_d[0], 1)] = 1
    assert lens[:hlit] == ll_lens[:hlit] and not any(ll_lens[hlit:])
    assert lens[hlit:] == d_lens[:hdist] and not any(d_lens[hdist:])
    mf = _meta_freq(syms)
    assert [mlens[s] for s in range(19)] == heap_code_lengths(mf, 7)
    natural = {"litlen_depth_gt15": (max(heap_depths(ll)), 15),
               "dist_depth_gt15": (max(heap_depths(case["dc"])), 15),
               "meta_depth_gt7": (max(heap_depths(mf)), 7)}[case["name"]]
    assert natural[0] > natural[1], f"{case['name']}: natural depth {natural[0]} does not exceed {natural[1]}"
    assert inflate_with_zero_window(blk) == reconstruct_with_zero_window(toks)


def _build_mutant(tmp_path, name, defines):
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "gpzip_oracle.c")
    so = str(tmp_path / f"liboracle_{name}.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", *defines, "-o", so, src], check=True)
    import ctypes

    L = ctypes.CDLL(so)
    L.orc_match_find.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p]
    L.orc_build_code_lengths.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p]
    return L


def _kat_ok(L, kat) -> bool:
    data = np.frombuffer(base64.b64decode(kat["input_b64"]), np.uint8)
    t = np.zeros(data.size, np.uint32)
    L.orc_match_find(data.ctypes.data, data.size, 32768, t.ctypes.data)
    return all((int(t[int(p)]) >> 16, int(t[int(p)]) & 0xFFFF) == tuple(v) for p, v in kat["expect"].items())


@pytest.mark.parametrize("name,defines", [("mul", ["-DORC_HASH_MUL=0x9E3779B9u"]),
                                          ("reversed", ["-DORC_HASH_REVERSED"])])
def test_hash3_mutants_fail_collision_kats(tmp_path, name, defines):
    """A build of the oracle with another multiplier or byte order passes
    the collision-free known answers but fails the collision ones."""
    kats = {k["name"]: k for k in GOLD["match_kats"]}
    L = _build_mutant(tmp_path, name, defines)
    assert _kat_ok(L, kats["newest_wins_tie"])  # collision-free cases cannot tell
    assert not _kat_ok(L, kats["hash3_collisions_shadow"])
    good = _build_mutant(tmp_path, "good", [])
    assert _kat_ok(good, kats["hash3_collisions_shadow"])


def test_halving_mutant_fails_halving_kats(tmp_path):
    """A build that rounds the halving step up passes the cases that need no
    halving but fails the hand-derived halving ones."""
    L = _build_mutant(tmp_path, "halve_up", ["-DORC_HALVE(f)=(((f)+1u)/2u)"])
    cases = {c["name"]: c for c in GOLD["huffman"]}

    def lens(c):
        f = np.array(c["freq"], np.uint32)
        out = np.zeros(f.size, np.uint8)
        L.orc_build_code_lengths(f.ctypes.data, f.size, c["limit"], out.ctypes.data)
        return list(out)

    assert lens(cases["tie_leaf_first"]) == cases["tie_leaf_first"]["expect"]
    assert lens(cases["halve_once_limit7"]) != cases["halve_once_limit7"]["expect"]


# the ratio mode's walk (this library's opt-in rule, not the reference's)
LOOKAHEAD_CASES = [
    # p0 (3,1): the next position is a literal, but p2 holds (5,2), longer by 2:
    # p0 is demoted; p1 a literal; p2 (5,2) with literals after it is taken
    ("demote_for_two_ahead", b"abcdefghij",
     [[3, 1], [0, 0x62], [5, 2], [0, 0x64], [0, 0x65], [0, 0x66], [0, 0x67], [0, 0x68], [0, 0x69], [0, 0x6A]],
     [[0, 0x61], [0, 0x62], [5, 2], [0, 0x68], [0, 0x69], [0, 0x6A]]),
    # a match two ahead longer by only 1 does not demote
    ("longer_by_one_keeps", b"abcdefgh",
     [[3, 1], [0, 0x62], [4, 2], [0, 0x64], [0, 0x65], [0, 0x66], [0, 0x67], [0, 0x68]],
     [[3, 1], [0, 0x64], [0, 0x65], [0, 0x66], [0, 0x67], [0, 0x68]]),
    # the reference's rule still applies one ahead; near the end the positions
    # past the input count as literals
    ("end_of_input", b"wxyz",
     [[3, 1], [0, 0x78], [0, 0x79], [0, 0x7A]],
     [[3, 1], [0, 0x7A]]),
]


@pytest.mark.parametrize("name,data,toks,exp", LOOKAHEAD_CASES, ids=[c[0] for c in LOOKAHEAD_CASES])
def test_lookahead_walk_cases(name, data, toks, exp):
    assert (O.lookahead_walk(pack(toks), data) == pack(exp)).all()


@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
def test_lookahead_walk_round_trip_and_smaller(kind):
    import gpzip_amd

    data = gpzip_amd.synth(kind, 1 << 18, 8)
    t = O.match_find(data)
    w2 = O.lookahead_walk(t, data)
    assert O.reconstruct(w2, data.size) == data.tobytes()
    g0, g2 = O.gzip_chunk(data), O.gzip_chunk(data, rule=2)
    assert gzip.decompress(g2) == data.tobytes() and len(g2) <= len(g0)
