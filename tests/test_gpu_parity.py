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
    assert tune.k1_variant in (1, 2) and tune.emit_variant in (1, 2) and tune.k1_grid >= 1 and tune.k0_grid >= 1
    assert tune.k3_grid >= 1 and tune.emit_grid >= 1 and t_ms > 0 and t_def > 0
    assert run(tune) == ref
    for v in (1, 2):
        for g in (1, 7, 256, 100000):
            t = gpzip_amd.Tune()
            t.k0_grid, t.k1_grid, t.k3_grid, t.emit_grid, t.k1_variant = g, g, g, max(1, g // 3), v
            t.emit_variant = 3 - v
            say("shape", v, g)
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
    for field in ("k1_variant", "emit_variant"):
        bad = gpzip_amd.Tune()
        setattr(bad, field, 3)
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
    srv = gpzip_amd.Server(ctx, streams=2)
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
