"""HIP path vs the CPU oracle, bit-exact, through the C ABI (needs a gfx950 GPU)."""
import base64
import gzip
import json
import os
import sys
import zlib

import numpy as np
import pytest

import gpzip_amd
from _vectors import expand_tokens
from gpzip_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pack(pairs):
    return np.array([(l << 16) | v for l, v in pairs], dtype=np.uint32)


def first_diff(a, b):
    a, b = np.asarray(a), np.asarray(b)
    n = min(a.size, b.size)
    d = np.nonzero(a[:n] != b[:n])[0]
    return int(d[0]) if d.size else (n if a.size != b.size else -1)


def assert_tokens_equal(got, exp, what):
    i = first_diff(got, exp)
    assert i < 0, f"{what}: first difference at {i}: got {int(got[i]) if i < got.size else None:#x} " \
                  f"expected {int(exp[i]) if i < exp.size else None:#x}"


@pytest.mark.parametrize("kat", GOLD["match_kats"], ids=lambda k: k["name"])
def test_match_find_known_answers(ctx, kat):
    data = base64.b64decode(kat["input_b64"])
    say("match_find KAT", kat["name"], len(data))
    t = ctx.match_find(data)
    assert_tokens_equal(t, O.match_find(data), kat["name"])
    for p, (l, v) in kat["expect"].items():
        assert (int(t[int(p)]) >> 16, int(t[int(p)]) & 0xFFFF) == (l, v)


SIZES = [1, 2, 3, 4, 5, 63, 4095, 4096, 4097, 8191, 12289, 40000, 131071, 131072]


@pytest.mark.parametrize("kind", ["text", "binmix", "log", "random", "zeros"])
@pytest.mark.parametrize("n", SIZES)
def test_match_find_equals_oracle(ctx, kind, n):
    if kind == "random":
        data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8)
    elif kind == "zeros":
        data = np.zeros(n, np.uint8)
    else:
        data = gpzip_amd.synth(kind, n, 100 + n)
    say("match_find", kind, n)
    assert_tokens_equal(ctx.match_find(data), O.match_find(data), f"{kind} {n}")


@pytest.mark.parametrize("case", GOLD["walk_strict"], ids=lambda c: c["name"])
def test_strict_walk_reference_cases(ctx, case):
    if "all_literals" in case:
        toks = pack([(0, i & 0xFF) for i in range(case["all_literals"])])
        exp = toks
    else:
        toks, exp = pack(case["tokens"]), pack(case["expect"])
    say("strict walk", case["name"], toks.size)
    assert_tokens_equal(ctx.walk(toks, strict=True), exp, case["name"])


@pytest.mark.parametrize("case", GOLD["walk_lazy"], ids=lambda c: c["name"])
def test_lazy_walk_cases(ctx, case):
    data = base64.b64decode(case["data_b64"])
    say("lazy walk", case["name"])
    assert_tokens_equal(ctx.walk(pack(case["tokens"]), data), pack(case["expect"]), case["name"])


@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
@pytest.mark.parametrize("n", [4096, 4097, 70000, 131072])
def test_walk_equals_oracle(ctx, kind, n):
    data = gpzip_amd.synth(kind, n, 7 + n)
    t = O.match_find(data)
    say("walk", kind, n)
    assert_tokens_equal(ctx.walk(t, data), O.greedy_walk(t, data), f"lazy {kind} {n}")
    assert_tokens_equal(ctx.walk(t, strict=True), O.strict_walk(t), f"strict {kind} {n}")


# ---- the per-member walk (k_walk_member), forced on every call of a context ----

@pytest.fixture(scope="module")
def ctx_mw():
    c = gpzip_amd.Context()
    t = gpzip_amd.Tune()
    t.walk_variant = 2
    c.set_tune(t)
    yield c
    c.close()


def test_member_walk_reference_cases(ctx_mw):
    """The reference's strict-walk cases and the hand-derived lazy cases
    through the per-member walk."""
    for case in GOLD["walk_strict"]:
        if "all_literals" in case:
            toks = pack([(0, i & 0xFF) for i in range(case["all_literals"])])
            exp = toks
        else:
            toks, exp = pack(case["tokens"]), pack(case["expect"])
        say("member walk strict", case["name"], toks.size)
        assert_tokens_equal(ctx_mw.walk(toks, strict=True), exp, case["name"])
    for case in GOLD["walk_lazy"]:
        data = base64.b64decode(case["data_b64"])
        say("member walk lazy", case["name"])
        assert_tokens_equal(ctx_mw.walk(pack(case["tokens"]), data), pack(case["expect"]), case["name"])


@pytest.mark.parametrize("kind", ["text", "binmix", "log", "zeros"])
@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 127, 511, 512, 513, 4097, 70000, 131072])
def test_member_walk_equals_oracle(ctx_mw, kind, n):
    """Lazy, strict and lookahead rules through the per-member walk: members
    ending inside a block, on a block or group boundary, one past it."""
    data = np.zeros(n, np.uint8) if kind == "zeros" else gpzip_amd.synth(kind, n, 11 + n)
    t = O.match_find(data)
    say("member walk", kind, n)
    assert_tokens_equal(ctx_mw.walk(t, data), O.greedy_walk(t, data), f"lazy {kind} {n}")
    assert_tokens_equal(ctx_mw.walk(t, strict=True), O.strict_walk(t), f"strict {kind} {n}")
    assert_tokens_equal(ctx_mw.walk(t, data, rule=2), O.lookahead_walk(t, data), f"lookahead {kind} {n}")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_walks_on_arbitrary_token_streams(ctx, ctx_mw, seed):
    """gpz_walk takes caller tokens: random streams of literals (some above
    255), matches of length 1-258 (some running past the end) and distances
    0-65535, in runs that make demotions common.  Both walks, all three rules,
    equal the oracle's walks."""
    rng = np.random.default_rng(seed)
    for n in (1, 2, 63, 64, 65, 700, 4097, 20000, 70001):
        kind = rng.integers(0, 4, n)
        lens = np.where(kind == 0, 0, np.where(kind == 1, rng.integers(1, 3, n), rng.integers(3, 259, n)))
        lens = np.where((kind == 3) & (rng.random(n) < 0.5), rng.integers(3, 9, n), lens)
        vals = np.where(lens == 0, rng.integers(0, 300, n), rng.integers(0, 65536, n))
        toks = ((lens.astype(np.uint32) << 16) | vals.astype(np.uint32)).astype(np.uint32)
        data = rng.integers(0, 256, n, dtype=np.uint8)
        say("arbitrary tokens", seed, n)
        for c, name in ((ctx, "segment"), (ctx_mw, "member")):
            assert_tokens_equal(c.walk(toks, data), O.greedy_walk(toks, data), f"{name} lazy {n}")
            assert_tokens_equal(c.walk(toks, strict=True), O.strict_walk(toks), f"{name} strict {n}")
            assert_tokens_equal(c.walk(toks, data, rule=2), O.lookahead_walk(toks, data), f"{name} lookahead {n}")


def test_member_walk_demotions_at_block_edges():
    """Lazy demotions on the last lanes of a block and of a group (the rule
    reads the next block's first tokens), decided by hand: position 64 k - 1
    holds a match of 3 and position 64 k one of 5, so the first is demoted
    to its literal byte."""
    c = gpzip_amd.Context()
    t = gpzip_amd.Tune()
    t.walk_variant = 2
    c.set_tune(t)
    try:
        n = 2048
        data = gpzip_amd.synth("binmix", n, 5)
        toks = data.astype(np.uint32)
        for k in (1, 7, 8, 9, 16, 31):
            p = 64 * k - 1
            toks[p] = (3 << 16) | 1
            toks[p + 1] = (5 << 16) | 2
        say("member walk demotions")
        assert_tokens_equal(c.walk(toks, data), O.greedy_walk(toks, data), "demotions")
        assert_tokens_equal(c.walk(toks, data, rule=2), O.lookahead_walk(toks, data), "lookahead demotions")
    finally:
        c.close()


def test_member_walk_forms_equal_oracle(ctx_mw):
    """Plain, batched and edge-size chunks with the per-member walk equal the
    oracle; a batch forced onto the per-segment walk gives the same bytes."""
    kinds = ["text", "binmix", "log"]
    sizes = [0, 1, 2, 3, 64, 4096, 131072, 131073, 700001, 2 << 20]
    chunks = [gpzip_amd.synth(kinds[i % 3], max(s, 1), 300 + i)[:s] for i, s in enumerate(sizes)]
    ref = [O.gzip_chunk(c) for c in chunks]
    say("member walk plain")
    for c, r in zip(chunks, ref):
        assert ctx_mw.compress_chunk(c) == r, c.size
    say("member walk batch")
    assert ctx_mw.batch(chunks) == ref
    seg = gpzip_amd.Context()
    t = gpzip_amd.Tune()
    t.walk_variant = 1
    seg.set_tune(t)
    try:
        say("segment walk batch")
        assert seg.batch(chunks) == ref
    finally:
        seg.close()
    bad = gpzip_amd.Tune()
    bad.walk_variant = 3
    with pytest.raises(gpzip_amd.GpzError):
        ctx_mw.set_tune(bad)


@pytest.mark.parametrize("parts", [1, 2, 4])
def test_member_walk_parts_equal_oracle(parts):
    """The per-member walk with each member cut into 1, 2 or 4 parts walked at
    once.  A stream of a literal followed by matches of length 2 never lets a
    walk from a block start (even) meet the true walk (odd), so every part's
    warm-up guess is wrong and every part is walked again from the true
    entry; the stream with the literal removed has right guesses.  Reference
    cases, random token streams, matched corpora and whole chunks (plain and
    batched) equal the oracle."""
    c = gpzip_amd.Context()
    t = gpzip_amd.Tune()
    t.walk_variant, t.walk_parts = 2, parts
    c.set_tune(t)
    try:
        for n in (1, 64, 65, 129, 700, 4097, 70001):
            for lead in (True, False):
                toks = np.full(n, (2 << 16) | 1, np.uint32)
                if lead:
                    toks[0] = 65
                data = np.full(n, 65, np.uint8)
                say("parts", parts, "len-2 stream", n, lead)
                exp = O.strict_walk(toks)
                assert_tokens_equal(c.walk(toks, strict=True), exp, f"strict {n}")
                assert_tokens_equal(c.walk(toks, data), O.greedy_walk(toks, data), f"lazy {n}")
                assert_tokens_equal(c.walk(toks, data, rule=2), O.lookahead_walk(toks, data), f"lookahead {n}")
        for case in GOLD["walk_strict"]:
            if "all_literals" not in case:
                assert_tokens_equal(c.walk(pack(case["tokens"]), strict=True), pack(case["expect"]), case["name"])
        rng = np.random.default_rng(parts)
        for n in (3, 200, 5000, 70001):
            lens = np.where(rng.random(n) < 0.4, 0, rng.integers(1, 259, n))
            vals = np.where(lens == 0, rng.integers(0, 256, n), rng.integers(1, 32769, n))
            toks = ((lens.astype(np.uint32) << 16) | vals.astype(np.uint32)).astype(np.uint32)
            data = rng.integers(0, 256, n, dtype=np.uint8)
            say("parts", parts, "random stream", n)
            assert_tokens_equal(c.walk(toks, data), O.greedy_walk(toks, data), f"lazy random {n}")
            assert_tokens_equal(c.walk(toks, strict=True), O.strict_walk(toks), f"strict random {n}")
            assert_tokens_equal(c.walk(toks, data, rule=2), O.lookahead_walk(toks, data), f"lookahead random {n}")
        for kind in ("text", "binmix", "log"):
            data = gpzip_amd.synth(kind, 131072, 17)
            tk = O.match_find(data)
            say("parts", parts, "corpus", kind)
            assert_tokens_equal(c.walk(tk, data), O.greedy_walk(tk, data), f"lazy {kind}")
            assert_tokens_equal(c.walk(tk, data, rule=2), O.lookahead_walk(tk, data), f"lookahead {kind}")
        sizes = [0, 1, 64, 4096, 131072, 131073, 700001]
        chunks = [gpzip_amd.synth(("text", "binmix", "log")[i % 3], max(s, 1), 500 + i)[:s]
                  for i, s in enumerate(sizes)]
        ref = [O.gzip_chunk(x) for x in chunks]
        say("parts", parts, "chunks")
        for x, r in zip(chunks, ref):
            assert c.compress_chunk(x) == r, x.size
        assert c.batch(chunks) == ref
    finally:
        c.close()
    bad = gpzip_amd.Tune()
    bad.walk_parts = 3
    with pytest.raises(gpzip_amd.GpzError):
        ctx_bad = gpzip_amd.Context()
        try:
            ctx_bad.set_tune(bad)
        finally:
            ctx_bad.close()


@pytest.mark.parametrize("kind", ["text", "binmix", "log", "random"])
def test_encode_block_equals_oracle(ctx, kind):
    n = 131072
    data = (np.random.default_rng(5).integers(0, 256, n, dtype=np.uint8) if kind == "random"
            else gpzip_amd.synth(kind, n, 21))
    w = O.greedy_walk(O.match_find(data), data)
    say("encode_block", kind, w.size)
    got = ctx.encode_block(w)
    exp = O.encode_block(w)
    assert got == exp, f"first byte difference at {first_diff(np.frombuffer(got, np.uint8), np.frombuffer(exp, np.uint8))}"
    assert zlib.decompress(got, -15) == data.tobytes()


def test_encode_block_small_cases(ctx):
    say("encode_block small")
    assert ctx.encode_block(np.zeros(0, np.uint32)).hex() == GOLD["encode_empty_hex"]
    for toks in ([(0, 97)], [(0, 0), (0, 255)], [(0, 120), (3, 1)], [(0, 1), (258, 1), (4, 2)]):
        t = pack(toks)
        assert ctx.encode_block(t) == O.encode_block(t), toks


def test_encode_block_rejects_bad_tokens(ctx):
    with pytest.raises(gpzip_amd.GpzError):
        ctx.encode_block(pack([(2, 1)]))
    with pytest.raises(gpzip_amd.GpzError):
        ctx.encode_block(pack([(0, 300)]))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 100, 4096, 65537, 131072])
def test_gzip_member_equals_oracle(ctx, n):
    data = gpzip_amd.synth("binmix", max(n, 1), 3)[:n]
    say("gzip_member", n)
    got = ctx.gzip_member(data)
    assert got == O.gzip_member(data)
    assert gzip.decompress(got) == data.tobytes()


@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
def test_compress_chunk_2mib_equals_oracle(ctx, kind):
    data = gpzip_amd.synth(kind, 2 << 20, {"text": 0, "binmix": 1, "log": 2}[kind])
    say("compress_chunk", kind)
    got = ctx.compress_chunk(data)
    exp = O.gzip_chunk(data)
    assert got == exp, f"first byte difference at {first_diff(np.frombuffer(got, np.uint8), np.frombuffer(exp, np.uint8))}"
    assert gzip.decompress(got) == data.tobytes()


@pytest.mark.parametrize("n", [0, 1, 131073, 262144, 300001])
def test_compress_chunk_edge_sizes(ctx, n):
    data = gpzip_amd.synth("log", max(n, 1), n)[:n]
    say("compress_chunk edge", n)
    got = ctx.compress_chunk(data)
    assert got == O.gzip_chunk(data)
    assert gzip.decompress(got) == data.tobytes()


def test_batch_equals_plain(ctx):
    rng = np.random.default_rng(9)
    sizes = [0, 1, 5000, 131072, 131073, 700000, 2 << 20, 3, 4096 * 33 + 17]
    kinds = ["text", "binmix", "log"]
    chunks = [gpzip_amd.synth(kinds[i % 3], max(s, 1), 40 + i)[:s] for i, s in enumerate(sizes)]
    say("batch", len(chunks))
    res = ctx.batch(chunks)
    for c, r in zip(chunks, res):
        assert r == ctx.compress_chunk(c)
        assert r == O.gzip_chunk(c)
    # the same batch twice (graph replay) gives the same bytes
    assert ctx.batch(chunks) == res


def test_device_plan_across_tiles_and_smaller_batches(ctx):
    """The batched form builds its member and segment tables on the device
    from the chunk descriptors (k_plan_batch, tiles of 1024 chunks): 1,500
    chunks of 0-9000 bytes (two tiles, empty chunks among them), then a
    3-chunk batch replayed on the same lane, each equal to the plain form."""
    rng = np.random.default_rng(77)
    sizes = rng.integers(0, 9000, 1500)
    sizes[[0, 1023, 1024, 1499]] = [0, 131073, 0, 4096]
    kinds = ["text", "binmix", "log"]
    chunks = [gpzip_amd.synth(kinds[i % 3], max(int(s), 1), 700 + i)[:int(s)] for i, s in enumerate(sizes)]
    say("device plan", len(chunks))
    res = ctx.batch(chunks)
    for i, (c, r) in enumerate(zip(chunks, res)):
        assert r == ctx.compress_chunk(c), f"chunk {i} ({c.size} bytes)"
    small = [chunks[1023], chunks[5], chunks[1024]]
    assert ctx.batch(small) == [res[1023], res[5], res[1024]]


def _plan_outputs(ctx, chunks, tune, flags):
    """Outputs of one gpz_plan over `chunks` (resident on the device) with
    the given tuning record and plan flags, run twice (the second run checks
    that per-run state starts clean)."""
    import torch

    offs, o = [], 0
    for c in chunks:
        offs.append(o)
        o += (c.size + 255) // 256 * 256
    host = np.zeros(o + 256, np.uint8)
    for off, c in zip(offs, chunks):
        host[off:off + c.size] = c
    d_in = torch.from_numpy(host).to("cuda:0")
    arr = (gpzip_amd.Desc * len(chunks))()
    for i, (a, c) in enumerate(zip(offs, chunks)):
        arr[i].offset, arr[i].len = a, c.size
    L = ctx.L
    h = gpzip_amd.C.c_void_p()
    t = gpzip_amd.C.byref(tune) if tune is not None else None
    gpzip_amd._check(L.gpz_plan_create(ctx.h, arr, len(chunks), flags, t, gpzip_amd.C.byref(h)))
    try:
        cap = L.gpz_plan_out_capacity(h)
        runs = []
        for _ in range(2):
            d_out = torch.zeros(cap + 256, dtype=torch.uint8, device="cuda:0")
            gpzip_amd._check(L.gpz_plan_run(h, gpzip_amd.C.c_void_p(d_in.data_ptr()),
                                            gpzip_amd.C.c_void_p(d_out.data_ptr())))
            gpzip_amd._check(L.gpz_plan_sync(h))
            off = np.zeros(len(chunks), np.uint64)
            ln = np.zeros(len(chunks), np.uint64)
            tot = gpzip_amd.C.c_uint64()
            gpzip_amd._check(L.gpz_plan_result(h, off.ctypes.data_as(gpzip_amd.u64p),
                                               ln.ctypes.data_as(gpzip_amd.u64p), gpzip_amd.C.byref(tot)))
            out = d_out.cpu().numpy()
            runs.append([out[int(a):int(a) + int(b)].tobytes() for a, b in zip(off, ln)])
        assert runs[0] == runs[1]
        return runs[0]
    finally:
        L.gpz_plan_destroy(h)


@pytest.mark.parametrize("variant", [3, 4])
def test_tiled_emit_equals_oracle(ctx, variant):
    """Bit emission with one workgroup per 4,096-token slot (look-back
    offsets, words shared by two slots written by the second to arrive):
    empty and tiny members, members whose token count is an exact multiple
    of 4,096 (random bytes: one literal per position, so the last slot is
    full), members with a few tokens and many empty slots (zeros), and text /
    binmix / log members, in the persistent, direct and graph plan forms,
    equal the oracle."""
    rng = np.random.default_rng(variant)
    chunks = [np.zeros(0, np.uint8), np.array([7], np.uint8), rng.integers(0, 256, 100, dtype=np.uint8),
              rng.integers(0, 256, 4096, dtype=np.uint8), rng.integers(0, 256, 4097, dtype=np.uint8),
              rng.integers(0, 256, 8192, dtype=np.uint8), rng.integers(0, 256, 131072, dtype=np.uint8),
              np.zeros(131072, np.uint8), np.zeros(300001, np.uint8)]
    chunks += [gpzip_amd.synth(k, s, 40 + i) for i, (k, s) in
               enumerate([("text", 131072), ("binmix", 131073), ("log", 700001), ("text", 3 << 20),
                          ("binmix", 2 << 20)])]
    ref = [O.gzip_chunk(c) for c in chunks]
    t = gpzip_amd.Tune()
    t.emit_variant = variant
    for flags in (gpzip_amd.GPZ_PLAN_PERSISTENT, 0, gpzip_amd.GPZ_PLAN_GRAPH):
        say("tiled emit", variant, "flags", flags)
        got = _plan_outputs(ctx, chunks, t, flags)
        for i, (g, r) in enumerate(zip(got, ref)):
            assert g == r, f"chunk {i} ({chunks[i].size} bytes), plan flags {flags}"


def test_autotune_and_tuned_shapes_equal_builtin(ctx):
    """The autotuner returns a valid record; every launch shape it can pick
    (and both match-find occupancy variants) gives the same bytes."""
    import torch

    kinds = ["text", "binmix", "log"]
    sizes = [3 << 20, 1 << 20, 700001, 4096, 2 << 20]
    chunks = [gpzip_amd.synth(kinds[i % 3], s, 90 + i) for i, s in enumerate(sizes)]
    ref = [O.gzip_chunk(c) for c in chunks]
    offs, o = [], 0
    for c in chunks:
        offs.append(o)
        o += (c.size + 255) // 256 * 256
    host = np.zeros(o + 256, np.uint8)
    for off, c in zip(offs, chunks):
        host[off:off + c.size] = c
    d_in = torch.from_numpy(host).to("cuda:0")
    descs = list(zip(offs, [c.size for c in chunks]))
    arr = (gpzip_amd.Desc * len(descs))()
    for i, (a, n) in enumerate(descs):
        arr[i].offset, arr[i].len = a, n
    L = ctx.L

    def run(tune):
        h = gpzip_amd.C.c_void_p()
        t = gpzip_amd.C.byref(tune) if tune is not None else None
        gpzip_amd._check(L.gpz_plan_create(ctx.h, arr, len(descs), gpzip_amd.GPZ_PLAN_PERSISTENT, t,
                                           gpzip_amd.C.byref(h)))
        try:
            cap = L.gpz_plan_out_capacity(h)
            d_out = torch.zeros(cap + 256, dtype=torch.uint8, device="cuda:0")
            gpzip_amd._check(L.gpz_plan_run(h, gpzip_amd.C.c_void_p(d_in.data_ptr()),
                                            gpzip_amd.C.c_void_p(d_out.data_ptr())))
            gpzip_amd._check(L.gpz_plan_sync(h))
            off = np.zeros(len(descs), np.uint64)
            ln = np.zeros(len(descs), np.uint64)
            tot = gpzip_amd.C.c_uint64()
            gpzip_amd._check(L.gpz_plan_result(h, off.ctypes.data_as(gpzip_amd.u64p),
                                               ln.ctypes.data_as(gpzip_amd.u64p), gpzip_amd.C.byref(tot)))
            out = d_out.cpu().numpy()
            return [out[int(a):int(a) + int(b)].tobytes() for a, b in zip(off, ln)]
        finally:
            L.gpz_plan_destroy(h)

    say("autotune")
    probe = torch.zeros(4 * o + (1 << 20), dtype=torch.uint8, device="cuda:0")
    tune, t_ms, t_def = ctx.autotune(descs, d_in.data_ptr(), probe.data_ptr())
    assert tune.k1_variant in (1, 2) and tune.emit_variant in (0, 1, 2, 3, 4) and tune.k1_grid >= 1 and tune.k0_grid >= 1
    assert tune.k0_variant in (1, 2, 3) and tune.k1_seg_per_wg >= 1 and tune.k3_seg_per_wg >= 1
    assert tune.k3_grid >= 1 and tune.emit_grid >= 1 and t_ms > 0 and t_def > 0
    assert tune.walk_variant in (0, 1, 2) and tune.walk_parts in (0, 1, 2, 4)
    assert run(tune) == ref
    for v in (1, 2):
        for g in (1, 7, 256, 100000):
            t = gpzip_amd.Tune()
            t.k0_grid, t.k1_grid, t.k3_grid, t.emit_grid, t.k1_variant = g, g, g, max(1, g // 3), v
            t.emit_variant = (2, 1, 3, 4)[(g + v) % 4]
            t.walk_variant = 1 + (g & 1)
            t.walk_parts = (1, 2, 4)[g % 3]
            say("shape", v, g)
            assert run(t) == ref
    # segment-table workgroup sizes and segments per workgroup
    for k0v in (1, 2, 3):
        for spw in (1, 3, 64):
            t = gpzip_amd.Tune()
            t.k0_variant, t.k0_grid = k0v, 97 * k0v
            t.k1_seg_per_wg, t.k3_seg_per_wg = spw, 4 - k0v
            say("shape k0v", k0v, "spw", spw)
            assert run(t) == ref
    # the context-wide record drives the batch API (graph re-captured on change)
    t = gpzip_amd.Tune()
    t.k1_grid, t.k1_variant = 64, 2
    ctx.set_tune(t)
    try:
        assert ctx.get_tune().k1_variant == 2
        assert ctx.batch(chunks) == ref
    finally:
        ctx.set_tune(None)
    assert ctx.get_tune().k1_grid == 0
    for field, v in (("k1_variant", 3), ("emit_variant", 5), ("k0_variant", 4), ("walk_variant", 3),
                     ("walk_parts", 3), ("walk_parts", 8)):
        bad = gpzip_amd.Tune()
        setattr(bad, field, v)
        with pytest.raises(gpzip_amd.GpzError):
            ctx.set_tune(bad)


def test_plan_flags_and_timing(ctx):
    """Direct, timed and graph runs of one plan layout give the oracle's bytes;
    per-kernel timing needs GPZ_PLAN_TIMED; unknown flags are rejected."""
    import torch

    chunks = [gpzip_amd.synth(k, s, 70 + i) for i, (k, s) in enumerate([("binmix", 2 << 20), ("text", 300001)])]
    ref = [O.gzip_chunk(c) for c in chunks]
    offs, o = [], 0
    for c in chunks:
        offs.append(o)
        o += (c.size + 255) // 256 * 256
    host = np.zeros(o + 256, np.uint8)
    for off, c in zip(offs, chunks):
        host[off:off + c.size] = c
    d_in = torch.from_numpy(host).to("cuda:0")
    arr = (gpzip_amd.Desc * len(chunks))()
    for i, (a, c) in enumerate(zip(offs, chunks)):
        arr[i].offset, arr[i].len = a, c.size
    L = ctx.L
    G, P, T = gpzip_amd.GPZ_PLAN_GRAPH, gpzip_amd.GPZ_PLAN_PERSISTENT, gpzip_amd.GPZ_PLAN_TIMED
    h = gpzip_amd.C.c_void_p()
    assert L.gpz_plan_create(ctx.h, arr, len(chunks), 8, None, gpzip_amd.C.byref(h)) != 0
    for flags in (0, T, P | T, G, G | P | T):
        say("plan flags", flags)
        h = gpzip_amd.C.c_void_p()
        gpzip_amd._check(L.gpz_plan_create(ctx.h, arr, len(chunks), flags, None, gpzip_amd.C.byref(h)))
        try:
            d_out = torch.zeros(L.gpz_plan_out_capacity(h) + 256, dtype=torch.uint8, device="cuda:0")
            for _ in range(2):
                gpzip_amd._check(L.gpz_plan_run(h, gpzip_amd.C.c_void_p(d_in.data_ptr()),
                                                gpzip_amd.C.c_void_p(d_out.data_ptr())))
                gpzip_amd._check(L.gpz_plan_sync(h))
            off = np.zeros(len(chunks), np.uint64)
            ln = np.zeros(len(chunks), np.uint64)
            tot = gpzip_amd.C.c_uint64()
            gpzip_amd._check(L.gpz_plan_result(h, off.ctypes.data_as(gpzip_amd.u64p),
                                               ln.ctypes.data_as(gpzip_amd.u64p), gpzip_amd.C.byref(tot)))
            out = d_out.cpu().numpy()
            assert [out[int(a):int(a) + int(b)].tobytes() for a, b in zip(off, ln)] == ref
            ms = (gpzip_amd.C.c_float * 4)()
            rc = L.gpz_plan_timing(h, ms, 4)
            if flags & T:
                assert rc == 0 and ms[0] > 0
                if not flags & G:
                    assert 0 < ms[1] < ms[0] and 0 < ms[2] < ms[0] and 0 < ms[3] < ms[0]
            else:
                assert rc != 0
        finally:
            L.gpz_plan_destroy(h)


def test_server_alone_equals_mix(ctx):
    kinds = ["text", "binmix", "log"]
    reqs = [gpzip_amd.synth(kinds[i % 3], s, 70 + i)
            for i, s in enumerate([1000, 300000, 5 << 20, 4096, 2 << 20, 123457])]
    say("server alone")
    alone = {}
    srv = gpzip_amd.Server(ctx, streams=2, budget_bytes=2 << 30)
    try:
        for i, r in enumerate(reqs):
            srv.submit(r, i)
            if i == 0:
                # a tag not yet returned by gpz_request_next cannot be released
                assert ctx.L.gpz_request_release(srv.h, 0) != 0
            tag, data, lat = srv.next()
            assert tag == i
            alone[i] = data
        say("server mix")
        for i, r in enumerate(reqs):
            srv.submit(r, 100 + i)
        mixed = {}
        while len(mixed) < len(reqs):
            tag, data, lat = srv.next()
            mixed[tag - 100] = data
    finally:
        srv.stop()
    for i, r in enumerate(reqs):
        assert alone[i] == mixed[i]
        assert alone[i] == O.gzip_chunk(r, 131072)
        assert gzip.decompress(alone[i]) == r.tobytes()


def test_random_chunks_equal_oracle(ctx):
    """Seeded random chunks (sizes log-uniform over 1 B - 1 MiB, the synthetic
    corpora, uniform random bytes, zeros, and splices of two kinds) through
    the plain and the batched form, each equal to the oracle."""
    rng = np.random.default_rng(2024)
    kinds = ["text", "binmix", "log", "random", "zeros"]
    chunks = []
    for _ in range(60):
        n = int(np.exp(rng.uniform(0, np.log(1 << 20))))
        k = kinds[rng.integers(len(kinds))]
        if k == "random":
            c = rng.integers(0, 256, n, dtype=np.uint8)
        elif k == "zeros":
            c = np.zeros(n, np.uint8)
        else:
            c = gpzip_amd.synth(k, n, int(rng.integers(1 << 30)))
        if n > 16 and rng.random() < 0.25:
            cut = int(rng.integers(1, n))
            c = np.concatenate([c[:cut], gpzip_amd.synth(kinds[rng.integers(3)], n - cut, int(rng.integers(1 << 30)))])
        chunks.append(c)
    ref = [O.gzip_chunk(c) for c in chunks]
    say("random chunks batched")
    assert ctx.batch(chunks) == ref
    say("random chunks plain")
    for c, r in zip(chunks, ref):
        assert ctx.compress_chunk(c) == r, c.size


def test_determinism_repeat(ctx):
    data = gpzip_amd.synth("binmix", 1 << 20, 77)
    say("determinism")
    a = ctx.compress_chunk(data)
    for _ in range(3):
        assert ctx.compress_chunk(data) == a


@pytest.mark.parametrize("case", GOLD["encode_halving"], ids=lambda c: c["name"])
def test_encode_block_halving_cases(ctx, case):
    """Streams whose litlen, distance and code-length codes exceed their
    limits: k_huffman's halving loop runs for each, bytes equal the oracle."""
    toks = expand_tokens(case["ll"], case["dc"])
    say("encode_block halving", case["name"], toks.size)
    got = ctx.encode_block(toks)
    exp = O.encode_block(toks)
    assert got == exp, f"first byte difference at {first_diff(np.frombuffer(got, np.uint8), np.frombuffer(exp, np.uint8))}"


def test_encode_block_match_heavy_stream(ctx):
    """300,000 long far matches (lengths 131-257, distances 16385-32768: up to
    48 bits each) encode to more than 2 bytes per token; the output region is
    sized for 48 bits per token, and the bytes equal the oracle."""
    rng = np.random.default_rng(31)
    n = 300_000
    toks = ((rng.integers(131, 258, n).astype(np.uint32) << 16) | rng.integers(16385, 32769, n).astype(np.uint32))
    say("encode_block match-heavy", n)
    got = ctx.encode_block(toks)
    exp = O.encode_block(toks)
    assert len(exp) > 2 * n + 1024
    assert got == exp, f"first byte difference at {first_diff(np.frombuffer(got, np.uint8), np.frombuffer(exp, np.uint8))}"


@pytest.mark.parametrize("kind,n", [("text", 1 << 20), ("random", (1 << 22) - 2), ("log", (1 << 22) + 4097)])
def test_large_member_equals_oracle(ctx, kind, n):
    """Members above the 128 KiB split: 256+ segments (k_chain's uncached
    path), and on both sides of 2^22 - 1 tokens, where k_huffman's heap keys
    go from 32 to 64 bits."""
    data = (np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8) if kind == "random"
            else gpzip_amd.synth(kind, n, 5 + n))
    say("large member", kind, n)
    got = ctx.gzip_member(data)
    exp = O.gzip_member(data)
    assert got == exp, f"first byte difference at {first_diff(np.frombuffer(got, np.uint8), np.frombuffer(exp, np.uint8))}"


def test_member_at_the_16mib_cap(ctx):
    """A 16 MiB - 1 member (the largest any entry point accepts; the oracle
    would take minutes): it decodes to the input, its block equals
    gpz_encode_block of gpz_walk of gpz_match_find (the stage entry points),
    and its litlen code lengths are the limited heap-Huffman lengths of the
    walked tokens' histogram."""
    from _vectors import heap_code_lengths
    from test_oracle import _parse_dynamic_header

    n = (16 << 20) - 1
    data = gpzip_amd.synth("text", n, 16)
    say("member at the cap", n)
    gz = ctx.gzip_member(data)
    assert gzip.decompress(gz) == data.tobytes()
    walked = ctx.walk(ctx.match_find(data), data)
    blk = ctx.encode_block(walked)
    assert gz[10:-8] == blk
    lens = walked >> 16
    lit = walked[lens == 0] & 0xFF
    ll = np.bincount(lit, minlength=286)
    ml = lens[lens != 0].astype(np.int64)
    base = [3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195,
            227, 258]
    code = np.searchsorted(np.array(base), ml, side="right") - 1
    ll[257:286] += np.bincount(code, minlength=29)
    ll[256] += 1
    hlit, hdist, mlens, syms, hl = _parse_dynamic_header(blk)
    exp = heap_code_lengths([int(x) for x in ll], 15)
    assert hl[:hlit] == exp[:hlit] and not any(exp[hlit:])


def test_compress_chunk_with_1mib_members():
    """A context whose members are 1 MiB (256 segments each): plain and
    batched forms equal the oracle's 1 MiB split."""
    c = gpzip_amd.Context(sub_chunk=1 << 20)
    try:
        chunks = [gpzip_amd.synth(k, s, 60 + i) for i, (k, s) in
                  enumerate([("text", (3 << 20) + 777), ("binmix", 1 << 20), ("log", (2 << 20) - 5)])]
        say("1 MiB members")
        ref = [O.gzip_chunk(x, 1 << 20) for x in chunks]
        for x, r in zip(chunks, ref):
            assert c.compress_chunk(x) == r
        assert c.batch(chunks) == ref
    finally:
        c.close()


@pytest.mark.parametrize("sub", [1 << 20, 4 << 20, 100003])
def test_member_walk_large_and_odd_members(sub):
    """The per-member walk on members of 1 MiB and 4 MiB (2,048 and 8,192 groups
    of eight blocks per wave) and of 100,003 bytes (off the block grid): plain
    and batched forms equal the oracle's split."""
    c = gpzip_amd.Context(sub_chunk=sub)
    t = gpzip_amd.Tune()
    t.walk_variant = 2
    c.set_tune(t)
    try:
        chunks = [gpzip_amd.synth(k, s, 80 + i) for i, (k, s) in
                  enumerate([("text", (5 << 20) + 333), ("log", 4 << 20), ("binmix", (1 << 20) + 64)])]
        say("member walk, members of", sub)
        ref = [O.gzip_chunk(x, sub) for x in chunks]
        for x, r in zip(chunks, ref):
            assert c.compress_chunk(x) == r
        assert c.batch(chunks) == ref
    finally:
        c.close()


@pytest.mark.parametrize("sub", [100003, 4000, 64])
def test_member_sizes_off_the_segment_grid(sub):
    """Members of sizes that are not multiples of the 4 KiB segment (a
    partial last segment in every member; members shorter than a segment;
    64-byte members): plain and batched forms (the batched one building its
    tables on the device) equal the oracle's split at the same size."""
    c = gpzip_amd.Context(sub_chunk=sub)
    try:
        chunks = [gpzip_amd.synth(k, s, 80 + i)[:s] for i, (k, s) in
                  enumerate([("text", 300001), ("binmix", 4096), ("log", 70000), ("text", 1), ("binmix", 0)])]
        say("members of", sub, "bytes")
        ref = [O.gzip_chunk(x, sub) for x in chunks]
        for x, r in zip(chunks, ref):
            assert c.compress_chunk(x) == r
        assert c.batch(chunks) == ref
    finally:
        c.close()


def _oracle_many(chunks, sub=131072):
    import concurrent.futures as cf

    with cf.ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda x: O.gzip_chunk(x, sub), chunks))


def test_mixed_sizes_sample_equals_oracle(ctx):
    """96 chunks of the mixed-size workload (4 KiB-4 MiB, text/binmix/log),
    the 8 largest among them: the autotuned batch == the plain form == the
    oracle."""
    import torch

    data, sizes = workloads.mixed_sizes_workload()
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    order = np.argsort(sizes)
    pick = sorted(set(order[-8:].tolist()) | set(range(0, len(sizes), 46)))[:96]
    chunks = [np.ascontiguousarray(data[offs[i]:offs[i] + sizes[i]]) for i in pick]
    say("mixed sample", len(chunks), sum(c.size for c in chunks))
    ref = _oracle_many(chunks)
    # autotune on the sample, device-resident
    lay, o = [], 0
    for c in chunks:
        lay.append(o)
        o += (c.size + 255) // 256 * 256
    host = np.zeros(o + 256, np.uint8)
    for a, c in zip(lay, chunks):
        host[a:a + c.size] = c
    d_in = torch.from_numpy(host).to("cuda:0")
    d_out = torch.zeros(3 * o + (1 << 24), dtype=torch.uint8, device="cuda:0")
    tune, _, _ = ctx.autotune(list(zip(lay, [c.size for c in chunks])), d_in.data_ptr(), d_out.data_ptr())
    del d_in, d_out
    ctx.set_tune(tune)
    try:
        assert ctx.batch(chunks) == ref
    finally:
        ctx.set_tune(None)
    for c, r in zip(chunks, ref):
        assert ctx.compress_chunk(c) == r


def test_mixed_sizes_full_batch_equals_plain(ctx):
    """All 4,096 chunks (2.5 GB) as one batch equal the plain form chunk by chunk."""
    import concurrent.futures as cf

    data, sizes = workloads.mixed_sizes_workload()
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    chunks = [data[offs[i]:offs[i] + sizes[i]] for i in range(len(sizes))]
    say("mixed full batch", len(chunks))
    got = ctx.batch(chunks)
    say("mixed full plain")
    with cf.ThreadPoolExecutor(8) as ex:
        plain = list(ex.map(ctx.compress_chunk, chunks))
    bad = [i for i, (a, b) in enumerate(zip(got, plain)) if a != b]
    assert not bad, f"{len(bad)} chunks differ, first {bad[:5]}"


def test_serving_requests_equal_plain(ctx):
    """500 requests of the serving workload (log-normal sizes, Poisson
    arrivals at 4 GB/s) from 2 submitting threads into a 4-lane server: every
    result equals the plain form of that request, 24 of them equal the
    oracle."""
    import threading
    import time

    n = 500
    R = workloads.ServingRequests(n, 4.0)
    srv = gpzip_amd.Server(ctx, streams=4, budget_bytes=4 << 30)
    results = {}
    errors = []

    def submit(t):
        try:
            for i in range(t, n, 2):
                now = time.perf_counter()
                if t0 + R.arrive[i] > now:
                    time.sleep(t0 + R.arrive[i] - now)
                srv.submit(R.data(i), i)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    say("serving", n, int(R.sizes.sum()))
    t0 = time.perf_counter()
    subs = [threading.Thread(target=submit, args=(t,)) for t in range(2)]
    for s_ in subs:
        s_.start()
    try:
        while len(results) < n:
            r = srv.next(timeout_ms=10000)
            assert r is not None, "no result within 10 s"
            results[r[0]] = r[1]
    finally:
        for s_ in subs:
            s_.join()
        srv.stop()
    assert not errors, errors
    for i in range(n):
        assert results[i] == ctx.compress_chunk(R.data(i)), f"request {i} ({R.sizes[i]} bytes)"
    sample = list(np.argsort(R.sizes)[:16]) + list(range(0, n, 63))
    ref = _oracle_many([R.data(int(i)) for i in sample])
    for i, r in zip(sample, ref):
        assert results[int(i)] == r, f"request {i} differs from the oracle"


def test_server_admission_blocks_at_budget(ctx):
    """With an input pool of 12 MiB, 40 requests of 1-4 MiB submitted from two
    threads wait for pool space (admission blocks, first come first served);
    peak arena use stays within the budget; results equal the plain form; a
    request larger than the pool is refused; a tag in use is refused."""
    import threading

    probe = gpzip_amd.Server(ctx, streams=2, max_batch_bytes=4 << 20, budget_bytes=1 << 30)
    ws = probe.stats()["workspace_bytes"]
    probe.stop()
    budget = ws + (12 << 20)
    srv = gpzip_amd.Server(ctx, streams=2, max_batch_bytes=4 << 20, budget_bytes=budget)
    rng = np.random.default_rng(8)
    kinds = ["text", "binmix", "log"]
    reqs = [gpzip_amd.synth(kinds[i % 3], int(rng.integers(1 << 20, 4 << 20)), 300 + i) for i in range(40)]
    say("admission", budget, sum(r.size for r in reqs))
    with pytest.raises(gpzip_amd.GpzError) as e:
        srv.submit(np.zeros(16 << 20, np.uint8), 999)
    assert e.value.status == gpzip_amd.GPZ_E_BUSY
    results = {}
    errors = []

    def submit(t):
        try:
            for i in range(t, len(reqs), 2):
                srv.submit(reqs[i], i)
        except Exception as ex:  # noqa: BLE001
            errors.append(ex)

    subs = [threading.Thread(target=submit, args=(t,)) for t in range(2)]
    for s_ in subs:
        s_.start()
    try:
        while len(results) < len(reqs):
            r = srv.next(timeout_ms=20000)
            assert r is not None, "no result within 20 s"
            results[r[0]] = r[1]
        for s_ in subs:
            s_.join()
        # a tag that is live (submitted, not yet released) is refused
        srv.submit(reqs[0][:1000], 7000)
        with pytest.raises(gpzip_amd.GpzError):
            srv.submit(reqs[1][:1000], 7000)
        assert srv.next(timeout_ms=20000)[0] == 7000
        st = srv.stats()
    finally:
        for s_ in subs:
            s_.join()
        srv.stop()
    assert not errors, errors
    assert st["arena_bytes"] >= budget and st["peak_bytes"] <= st["arena_bytes"]
    assert st["waited"] > 0, "no submission waited for pool space"
    for i, r in enumerate(reqs):
        assert results[i] == ctx.compress_chunk(r), f"request {i}"


def test_gunzip_decodes_every_form(ctx):
    """gpzip x's decoder (gpz_gunzip: member scan, parallel inflate) decodes
    the plain, batched and served outputs back to their inputs."""
    kinds = ["text", "binmix", "log"]
    chunks = [gpzip_amd.synth(kinds[i % 3], s, 500 + i) for i, s in enumerate([0, 1, 4097, 131072, 3 << 20, 2 << 20])]
    say("gunzip plain / batched / served")
    for c, b in zip(chunks, ctx.batch(chunks)):
        assert gpzip_amd.gunzip(b) == c.tobytes()
        assert gpzip_amd.gunzip(ctx.compress_chunk(c)) == c.tobytes()
    srv = gpzip_amd.Server(ctx, streams=2, budget_bytes=2 << 30)
    try:
        for i, c in enumerate(chunks):
            srv.submit(c, i)
        got = {}
        while len(got) < len(chunks):
            tag, data, _ = srv.next()
            got[tag] = data
    finally:
        srv.stop()
    for i, c in enumerate(chunks):
        assert gpzip_amd.gunzip(got[i]) == c.tobytes()
    whole = b"".join(got[i] for i in range(len(chunks)))
    assert gpzip_amd.gunzip(whole) == b"".join(c.tobytes() for c in chunks)


# ---- the opt-in ratio mode (GPZ_CFG_RATIO; outside reference parity) ----------

@pytest.mark.parametrize("kind", ["text", "binmix", "log"])
@pytest.mark.parametrize("n", [4095, 4096, 4097, 8194, 70000, 131072])
def test_lookahead_walk_equals_oracle(ctx, kind, n):
    """The ratio mode's walk rule (strict = 2) on the GPU equals the oracle's
    statement of it, across segment boundaries (the rule reads two tokens
    ahead, so the summary runs after every token exists)."""
    data = gpzip_amd.synth(kind, n, 31 + n)
    t = O.match_find(data)
    say("lookahead walk", kind, n)
    assert_tokens_equal(ctx.walk(t, data, rule=2), O.lookahead_walk(t, data), f"{kind} {n}")


def test_ratio_mode_forms_equal_oracle_and_are_smaller():
    """A GPZ_CFG_RATIO context: plain, batched and no-alloc outputs equal the
    oracle with the lookahead walk, decode, and are no larger than the
    reference-parity output."""
    import ctypes as Cc

    c = gpzip_amd.Context(flags=gpzip_amd.GPZ_CFG_RATIO)
    base = gpzip_amd.Context()
    try:
        chunks = [gpzip_amd.synth(k, s, 60 + i) for i, (k, s) in
                  enumerate([("text", 2 << 20), ("binmix", 2 << 20), ("log", (2 << 20) - 9), ("text", 300001)])]
        say("ratio mode")
        ref = [O.gzip_chunk(x, 131072, rule=2) for x in chunks]
        for x, r in zip(chunks, ref):
            got = c.compress_chunk(x)
            assert got == r
            assert gzip.decompress(got) == x.tobytes()
            assert len(got) <= len(base.compress_chunk(x))
            out = np.empty(c.bound(x.size), np.uint8)
            nn = Cc.c_size_t()
            gpzip_amd._check(c.L.gpz_compress_chunk_noalloc(c.h, Cc.c_void_p(x.ctypes.data), x.size,
                                                            Cc.c_void_p(out.ctypes.data), out.size, Cc.byref(nn)))
            assert out[:nn.value].tobytes() == r
        assert c.batch(chunks) == ref
    finally:
        c.close()
        base.close()


def test_gpzip_gpu_device_selection(monkeypatch):
    """GPZIP_GPU as in the reference (codec-gpu/src/gpu/context.rs:119-155): an
    index past the device count is GPZ_E_NODEV; "high", "low" and "0" select
    the device, and an explicit gpz_config::device takes precedence."""
    n = gpzip_amd.device_count()
    monkeypatch.setenv("GPZIP_GPU", str(n + 3))
    with pytest.raises(gpzip_amd.GpzError) as e:
        gpzip_amd.Context()
    assert e.value.status == gpzip_amd.GPZ_E_NODEV
    c = gpzip_amd.Context(device=0)  # the config's device wins over the variable
    c.close()
    for v in ("high", "low", "0"):
        monkeypatch.setenv("GPZIP_GPU", v)
        c = gpzip_amd.Context()
        try:
            data = gpzip_amd.synth("text", 5000, 9)
            assert c.compress_chunk(data) == O.gzip_chunk(data)
        finally:
            c.close()
