"""gpzip_amd -- MI355X (gfx950) gzip compress path with gpzip's GPU-backend surface.

The compute path is libgpzip_hip.so (HIP kernels + C++ runtime, C ABI in
include/gpzip_hip.h).  This module is a thin ctypes binding used by the tests
and bench.py; it never falls back to a CPU implementation: if the native
library is missing or no gfx950 device is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(HERE, "libgpzip_hip.so")
SYNTH_LIB_PATH = os.path.join(HERE, "libgpzip_synth.so")

GPZ_OK, GPZ_E_NODEV, GPZ_E_OOM, GPZ_E_INVAL, GPZ_E_DEVICE, GPZ_E_NOSPACE, GPZ_E_BUSY, GPZ_E_AGAIN = range(8)
GPZ_PLAN_GRAPH, GPZ_PLAN_PERSISTENT, GPZ_PLAN_TIMED = 1, 2, 4
GPZ_CFG_RATIO = 1

# Every entry point include/gpzip_hip.h declares.
ABI_SYMBOLS = [
    "gpz_config_default", "gpz_init", "gpz_destroy", "gpz_last_error", "gpz_status_string",
    "gpz_device_count", "gpz_device_label", "gpz_compress_bound", "gpz_compress_chunk",
    "gpz_compress_chunk_noalloc",
    "gpz_match_find", "gpz_walk", "gpz_encode_block", "gpz_gzip_member",
    "gpz_batch_submit", "gpz_batch_wait", "gpz_batch_result", "gpz_batch_free",
    "gpz_plan_create", "gpz_plan_out_capacity", "gpz_plan_run", "gpz_plan_sync", "gpz_plan_result",
    "gpz_plan_timing", "gpz_plan_destroy", "gpz_set_tune", "gpz_get_tune", "gpz_autotune",
    "gpz_server_config_default", "gpz_server_start", "gpz_request_submit", "gpz_request_next",
    "gpz_request_release", "gpz_server_stats", "gpz_server_stop",
    "gpz_gzip_member_starts", "gpz_gunzip",
]


class GpzError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gpz status {status}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("sub_chunk", C.c_uint32), ("window", C.c_uint32), ("lanes", C.c_uint32),
                ("max_in_flight", C.c_uint32), ("flags", C.c_uint32), ("lane_bytes", C.c_uint64)]


class Desc(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("len", C.c_uint64)]


class Tune(C.Structure):
    _fields_ = [("k0_grid", C.c_uint32), ("k1_grid", C.c_uint32), ("k3_grid", C.c_uint32), ("emit_grid", C.c_uint32),
                ("k1_variant", C.c_uint32), ("emit_variant", C.c_uint32), ("k0_variant", C.c_uint32),
                ("k1_seg_per_wg", C.c_uint32), ("k3_seg_per_wg", C.c_uint32), ("walk_variant", C.c_uint32),
                ("walk_parts", C.c_uint32), ("reserved", C.c_uint32 * 1)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_batch_chunks", C.c_uint32), ("max_batch_bytes", C.c_uint32),
                ("chunk_bytes", C.c_uint32), ("streams", C.c_uint32)]


_lib = None
_synth = None
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
szp = C.POINTER(C.c_size_t)


def load_library() -> C.CDLL:
    """Load libgpzip_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HIP_LIB_PATH):
        raise RuntimeError(f"{HIP_LIB_PATH} is missing: run `python -m gpzip_amd.build` first")
    L = C.CDLL(HIP_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "gpz_config_default": (None, [C.POINTER(Config)]),
        "gpz_init": (C.c_int, [C.POINTER(Config), C.POINTER(vp)]),
        "gpz_destroy": (None, [vp]),
        "gpz_last_error": (C.c_char_p, []),
        "gpz_status_string": (C.c_char_p, [C.c_int]),
        "gpz_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "gpz_device_label": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
        "gpz_compress_bound": (C.c_size_t, [vp, C.c_size_t]),
        "gpz_compress_chunk": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t, szp]),
        "gpz_compress_chunk_noalloc": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t, szp]),
        "gpz_match_find": (C.c_int, [vp, vp, C.c_size_t, vp]),
        "gpz_walk": (C.c_int, [vp, vp, vp, C.c_size_t, C.c_int, vp, szp]),
        "gpz_encode_block": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t, szp]),
        "gpz_gzip_member": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t, szp]),
        "gpz_batch_submit": (C.c_int, [vp, vp, C.POINTER(Desc), C.c_size_t, C.POINTER(vp)]),
        "gpz_batch_wait": (C.c_int, [vp]),
        "gpz_batch_result": (C.c_int, [vp, C.c_size_t, C.POINTER(u8p), szp]),
        "gpz_batch_free": (None, [vp]),
        "gpz_plan_create": (C.c_int, [vp, C.POINTER(Desc), C.c_size_t, C.c_uint32, C.POINTER(Tune), C.POINTER(vp)]),
        "gpz_plan_out_capacity": (C.c_size_t, [vp]),
        "gpz_plan_run": (C.c_int, [vp, vp, vp]),
        "gpz_plan_sync": (C.c_int, [vp]),
        "gpz_plan_result": (C.c_int, [vp, u64p, u64p, u64p]),
        "gpz_plan_timing": (C.c_int, [vp, C.POINTER(C.c_float), C.c_size_t]),
        "gpz_set_tune": (C.c_int, [vp, C.POINTER(Tune)]),
        "gpz_get_tune": (C.c_int, [vp, C.POINTER(Tune)]),
        "gpz_autotune": (C.c_int, [vp, C.POINTER(Desc), C.c_size_t, vp, vp, C.POINTER(Tune), C.POINTER(C.c_float)]),
        "gpz_plan_destroy": (None, [vp]),
        "gpz_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "gpz_server_start": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
        "gpz_request_submit": (C.c_int, [vp, vp, C.c_size_t, C.c_uint64]),
        "gpz_request_next": (C.c_int, [vp, C.c_int, u64p, C.POINTER(u8p), szp, C.POINTER(C.c_double)]),
        "gpz_request_release": (C.c_int, [vp, C.c_uint64]),
        "gpz_server_stats": (C.c_int, [vp, u64p, C.c_size_t]),
        "gpz_server_stop": (None, [vp]),
        "gpz_gzip_member_starts": (C.c_int, [vp, C.c_size_t, vp, C.c_size_t, szp]),
        "gpz_gunzip": (C.c_int, [vp, C.c_size_t, vp, C.c_size_t, szp, C.c_uint]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def load_synth() -> C.CDLL:
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(SYNTH_LIB_PATH):
        raise RuntimeError(f"{SYNTH_LIB_PATH} is missing: run `python -m gpzip_amd.build` first")
    S = C.CDLL(SYNTH_LIB_PATH)
    S.gpz_synth_fill.argtypes = [C.c_int, C.c_uint64, C.c_void_p, C.c_size_t]
    S.gpz_synth_loguniform_sizes.argtypes = [C.c_uint64, C.c_size_t, C.c_uint64, C.c_uint64, C.c_void_p]
    S.gpz_synth_lognormal_sizes.argtypes = [C.c_uint64, C.c_size_t, C.c_double, C.c_double, C.c_uint64,
                                            C.c_uint64, C.c_void_p]
    S.gpz_synth_uniform.argtypes = [C.c_uint64, C.c_size_t, C.c_void_p]
    _synth = S
    return S


def _check(rc: int) -> None:
    if rc != GPZ_OK:
        raise GpzError(rc, load_library().gpz_last_error().decode(errors="replace"))


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def _u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def device_count() -> int:
    n = C.c_int(0)
    _check(load_library().gpz_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU context (gpz_ctx)."""

    def __init__(self, device: int = -1, sub_chunk: int = 128 * 1024, lanes: int = 4, flags: int = 0):
        """flags: GPZ_CFG_RATIO for the opt-in ratio mode (outside reference parity)."""
        L = load_library()
        cfg = Config()
        L.gpz_config_default(C.byref(cfg))
        cfg.device = device
        cfg.sub_chunk = sub_chunk
        cfg.lanes = lanes
        cfg.flags = flags
        self.cfg = cfg
        h = C.c_void_p()
        _check(L.gpz_init(C.byref(cfg), C.byref(h)))
        self.h = h
        self.L = L

    def close(self) -> None:
        if self.h:
            self.L.gpz_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def label(self) -> str:
        buf = C.create_string_buffer(256)
        _check(self.L.gpz_device_label(self.h, buf, 256))
        return buf.value.decode()

    def bound(self, n: int) -> int:
        return self.L.gpz_compress_bound(self.h, n)

    def compress_chunk(self, data) -> bytes:
        a = _u8(data)
        out = np.empty(self.bound(a.size), np.uint8)
        n = C.c_size_t()
        _check(self.L.gpz_compress_chunk(self.h, _ptr(a), a.size, _ptr(out), out.size, C.byref(n)))
        return out[: n.value].tobytes()

    def gzip_member(self, data) -> bytes:
        a = _u8(data)
        out = np.empty(2 * a.size + 2048, np.uint8)
        n = C.c_size_t()
        _check(self.L.gpz_gzip_member(self.h, _ptr(a), a.size, _ptr(out), out.size, C.byref(n)))
        return out[: n.value].tobytes()

    def match_find(self, data) -> np.ndarray:
        a = _u8(data)
        tok = np.zeros(a.size, np.uint32)
        _check(self.L.gpz_match_find(self.h, _ptr(a), a.size, _ptr(tok)))
        return tok

    def walk(self, tokens, data=None, strict: bool = False, rule: int | None = None) -> np.ndarray:
        """rule: 0 the reference's lazy walk, 1 strict, 2 the ratio mode's
        lookahead walk (default: 1 if strict else 0)."""
        t = np.ascontiguousarray(tokens, dtype=np.uint32)
        a = _u8(data) if data is not None else None
        out = np.zeros(max(t.size, 1), np.uint32)
        n = C.c_size_t()
        r = int(strict) if rule is None else rule
        _check(self.L.gpz_walk(self.h, _ptr(a), _ptr(t), t.size, r, _ptr(out), C.byref(n)))
        return out[: n.value].copy()

    def encode_block(self, walked) -> bytes:
        t = np.ascontiguousarray(walked, dtype=np.uint32)
        out = np.empty(2 * 3 * t.size + 4096, np.uint8)
        n = C.c_size_t()
        _check(self.L.gpz_encode_block(self.h, _ptr(t), t.size, _ptr(out), out.size, C.byref(n)))
        return out[: n.value].tobytes()

    def batch(self, chunks) -> list[bytes]:
        """Compress a list of chunks as one batch; returns each chunk's members."""
        arrs = [_u8(c) for c in chunks]
        base = np.concatenate(arrs) if arrs else np.zeros(1, np.uint8)
        if base.size == 0:
            base = np.zeros(1, np.uint8)
        descs = (Desc * max(len(arrs), 1))()
        off = 0
        for i, a in enumerate(arrs):
            descs[i].offset = off
            descs[i].len = a.size
            off += a.size
        b = C.c_void_p()
        _check(self.L.gpz_batch_submit(self.h, _ptr(base), descs, len(arrs), C.byref(b)))
        try:
            _check(self.L.gpz_batch_wait(b))
            res = []
            for i in range(len(arrs)):
                p = u8p()
                n = C.c_size_t()
                _check(self.L.gpz_batch_result(b, i, C.byref(p), C.byref(n)))
                res.append(C.string_at(p, n.value) if n.value else b"")
            return res
        finally:
            self.L.gpz_batch_free(b)

    def set_tune(self, tune: "Tune | None") -> None:
        """Launch shapes for every batched launch of this context (None: built-in)."""
        _check(self.L.gpz_set_tune(self.h, C.byref(tune) if tune is not None else None))

    def get_tune(self) -> Tune:
        t = Tune()
        _check(self.L.gpz_get_tune(self.h, C.byref(t)))
        return t

    def autotune(self, descs, d_in: int, d_out: int) -> tuple[Tune, float, float]:
        """Tune on a device-resident batch: descs = [(offset, len)], d_in / d_out
        device pointers as for gpz_plan_run.  Returns (tune, tuned ms, built-in ms)."""
        arr = (Desc * len(descs))()
        for i, (o, n) in enumerate(descs):
            arr[i].offset, arr[i].len = o, n
        t = Tune()
        ms = (C.c_float * 2)()
        _check(self.L.gpz_autotune(self.h, arr, len(descs), C.c_void_p(d_in), C.c_void_p(d_out), C.byref(t), ms))
        return t, float(ms[0]), float(ms[1])


class Server:
    """Serving form (gpz_server_*)."""

    def __init__(self, ctx: Context, budget_bytes: int | None = None, max_batch_chunks: int | None = None,
                 max_batch_bytes: int | None = None, chunk_bytes: int | None = None, streams: int | None = None):
        """Fields left at None keep gpz_server_config_default's values."""
        self.L = ctx.L
        cfg = ServerConfig()
        self.L.gpz_server_config_default(C.byref(cfg))
        for name, v in (("budget_bytes", budget_bytes), ("max_batch_chunks", max_batch_chunks),
                        ("max_batch_bytes", max_batch_bytes), ("chunk_bytes", chunk_bytes), ("streams", streams)):
            if v is not None:
                setattr(cfg, name, v)
        h = C.c_void_p()
        _check(self.L.gpz_server_start(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, data, tag: int) -> None:
        a = _u8(data)
        _check(self.L.gpz_request_submit(self.h, _ptr(a), a.size, tag))

    def next(self, timeout_ms: int = -1, copy: bool = True):
        """(tag, bytes, latency_s) of the next finished request, or None on timeout.
        With copy=False the result is released unread and its length stands in
        for the bytes (a load generator that only counts output)."""
        tag = C.c_uint64()
        p = u8p()
        n = C.c_size_t()
        lat = C.c_double()
        rc = self.L.gpz_request_next(self.h, timeout_ms, C.byref(tag), C.byref(p), C.byref(n), C.byref(lat))
        if rc == GPZ_E_AGAIN:
            return None
        _check(rc)
        data = (C.string_at(p, n.value) if n.value else b"") if copy else n.value
        _check(self.L.gpz_request_release(self.h, tag.value))
        return tag.value, data, lat.value

    def stats(self) -> dict:
        v = (C.c_uint64 * 7)()
        _check(self.L.gpz_server_stats(self.h, v, 7))
        keys = ["done", "bytes_done", "steps", "peak_bytes", "arena_bytes", "waited", "workspace_bytes"]
        return dict(zip(keys, list(v)))

    def stop(self) -> None:
        if self.h:
            self.L.gpz_server_stop(self.h)
            self.h = None


def member_starts(data) -> list[int]:
    """Offsets of the fixed 10-byte gzip member headers (gpz_gzip_member_starts)."""
    L = load_library()
    a = _u8(data)
    n = C.c_size_t()
    rc = L.gpz_gzip_member_starts(_ptr(a), a.size, None, 0, C.byref(n))
    if rc not in (GPZ_OK, GPZ_E_NOSPACE):
        _check(rc)
    out = np.zeros(max(n.value, 1), np.uint64)
    _check(L.gpz_gzip_member_starts(_ptr(a), a.size, _ptr(out), out.size, C.byref(n)))
    return [int(x) for x in out[: n.value]]


def gunzip(data, threads: int = 0) -> bytes:
    """Decode a gzip stream on the host cores (gpz_gunzip): members in parallel
    when the header scan finds two or more, else a serial decode."""
    L = load_library()
    a = _u8(data)
    n = C.c_size_t()
    rc = L.gpz_gunzip(_ptr(a), a.size, None, 0, C.byref(n), threads)
    if rc == GPZ_OK:
        return b""
    if rc != GPZ_E_NOSPACE:
        _check(rc)
    out = np.empty(n.value, np.uint8)
    _check(L.gpz_gunzip(_ptr(a), a.size, _ptr(out), out.size, C.byref(n), threads))
    return out[: n.value].tobytes()


def synth(kind: str, n: int, seed: int) -> np.ndarray:
    """Seeded synthetic workload bytes: kind in {'text', 'binmix', 'log'}."""
    S = load_synth()
    k = {"text": 0, "binmix": 1, "log": 2}[kind]
    out = np.empty(n, np.uint8)
    if n:
        S.gpz_synth_fill(k, seed, out.ctypes.data, n)
    return out


def loguniform_sizes(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(count, np.uint64)
    load_synth().gpz_synth_loguniform_sizes(seed, count, lo, hi, out.ctypes.data)
    return out


def lognormal_sizes(seed: int, count: int, median: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    out = np.empty(count, np.uint64)
    load_synth().gpz_synth_lognormal_sizes(seed, count, median, sigma, lo, hi, out.ctypes.data)
    return out


def uniform(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    load_synth().gpz_synth_uniform(seed, count, out.ctypes.data)
    return out
