"""gpz_compress_chunk_noalloc (scratch in a file-scope device array) against
gpz_compress_chunk, through the C ABI."""
import ctypes as C
import sys

import numpy as np
import pytest

import gpzip_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(ctx, data, cap=None):
    a = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.empty(cap if cap is not None else ctx.bound(a.size), np.uint8)
    n = C.c_size_t()
    rc = ctx.L.gpz_compress_chunk_noalloc(ctx.h, C.c_void_p(a.ctypes.data) if a.size else None, a.size,
                                          C.c_void_p(out.ctypes.data) if out.size else None, out.size, C.byref(n))
    return rc, out[: n.value].tobytes() if rc == 0 else n.value


def test_noalloc_equals_compress_chunk_twice_smaller_second(ctx):
    big = gpzip_amd.synth("binmix", 2 << 20, 41)  # the limit
    small = gpzip_amd.synth("text", 300001, 42)
    say("noalloc big", big.size)
    rc, got_big = noalloc(ctx, big)
    assert rc == 0, ctx.L.gpz_last_error()
    say("noalloc small", small.size)
    rc, got_small = noalloc(ctx, small)
    assert rc == 0, ctx.L.gpz_last_error()
    assert got_big == ctx.compress_chunk(big)
    assert got_small == ctx.compress_chunk(small)


def test_noalloc_edges(ctx):
    for n in (0, 1, 131072, 131073):
        d = gpzip_amd.synth("log", max(n, 1), 50 + n)[:n]
        say("noalloc edge", n)
        rc, got = noalloc(ctx, d)
        assert rc == 0, ctx.L.gpz_last_error()
        assert got == ctx.compress_chunk(d)
    # above the limit: a clear error, nothing launched
    rc, _ = noalloc(ctx, np.zeros((2 << 20) + 1, np.uint8))
    assert rc == gpzip_amd.GPZ_E_INVAL
    assert b"no-alloc limit" in ctx.L.gpz_last_error()
    # too small an output buffer: the needed size comes back
    d = gpzip_amd.synth("text", 5000, 3)
    rc, need = noalloc(ctx, d, cap=10)
    assert rc == gpzip_amd.GPZ_E_NOSPACE and need == len(ctx.compress_chunk(d))


def test_noalloc_device_buffers(ctx):
    import torch

    d = gpzip_amd.synth("binmix", 777777, 9)
    x = torch.from_numpy(d.copy()).to("cuda:0")
    y = torch.zeros(ctx.bound(d.size), dtype=torch.uint8, device="cuda:0")
    n = C.c_size_t()
    say("noalloc device buffers")
    rc = ctx.L.gpz_compress_chunk_noalloc(ctx.h, C.c_void_p(x.data_ptr()), d.size, C.c_void_p(y.data_ptr()),
                                          y.numel(), C.byref(n))
    assert rc == 0, ctx.L.gpz_last_error()
    assert y[: n.value].cpu().numpy().tobytes() == ctx.compress_chunk(d)
