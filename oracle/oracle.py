"""ctypes binding of the CPU oracle (liboracle.so) and the zlib CPU baseline
(libcpuzlib.so).  TEST AND BENCH INFRASTRUCTURE ONLY: the tests compare the HIP
library against it, bench.py's cpu_baseline leg times it; the product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
ZLIB_LIB = os.path.join(HERE, "libcpuzlib.so")

_o = None
_z = None
SIZE_MAX = (1 << 64) - 1


def build() -> None:
    """Compile the oracle and the zlib baseline (make, in-tree)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib() -> C.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        vp, sz = C.c_void_p, C.c_size_t
        L.orc_hash3.restype = C.c_uint32
        L.orc_hash3.argtypes = [vp, sz]
        L.orc_match_find.argtypes = [vp, sz, C.c_uint32, vp]
        L.orc_match_find_brute.argtypes = [vp, sz, C.c_uint32, vp]
        L.orc_greedy_walk.restype = sz
        L.orc_greedy_walk.argtypes = [vp, vp, sz, vp]
        L.orc_strict_walk.restype = sz
        L.orc_strict_walk.argtypes = [vp, sz, vp]
        L.orc_lookahead_walk.restype = sz
        L.orc_lookahead_walk.argtypes = [vp, vp, sz, vp]
        L.orc_gzip_chunk_rule.restype = sz
        L.orc_gzip_chunk_rule.argtypes = [vp, sz, sz, C.c_int, vp, sz]
        L.orc_reconstruct.restype = sz
        L.orc_reconstruct.argtypes = [vp, sz, vp, sz]
        L.orc_build_code_lengths.argtypes = [vp, sz, C.c_uint32, vp]
        L.orc_canonical_codes.argtypes = [vp, sz, vp]
        L.orc_encode_block_fast.restype = sz
        L.orc_encode_block_fast.argtypes = [vp, sz, vp, sz]
        L.orc_crc32.restype = C.c_uint32
        L.orc_crc32.argtypes = [vp, sz]
        L.orc_gzip_member.restype = sz
        L.orc_gzip_member.argtypes = [vp, sz, vp, sz]
        L.orc_gzip_chunk.restype = sz
        L.orc_gzip_chunk.argtypes = [vp, sz, sz, vp, sz]
        L.orc_chunk_bound.restype = sz
        L.orc_chunk_bound.argtypes = [sz, sz]
        _o = L
    return _o


def _u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def _p(a):
    return C.c_void_p(a.ctypes.data) if a.size else None


def hash3(data, p: int) -> int:
    a = _u8(data)
    return _lib().orc_hash3(_p(a), p)


def match_find(data, window: int = 32768) -> np.ndarray:
    a = _u8(data)
    t = np.zeros(max(a.size, 1), np.uint32)
    if a.size:
        rc = _lib().orc_match_find(_p(a), a.size, window, _p(t))
        assert rc == 0
    return t[: a.size]


def match_find_brute(data, window: int = 32768) -> np.ndarray:
    a = _u8(data)
    t = np.zeros(max(a.size, 1), np.uint32)
    _lib().orc_match_find_brute(_p(a), a.size, window, _p(t))
    return t[: a.size]


def greedy_walk(tokens, data) -> np.ndarray:
    t = np.ascontiguousarray(tokens, np.uint32)
    a = _u8(data)
    out = np.zeros(max(t.size, 1), np.uint32)
    n = _lib().orc_greedy_walk(_p(t), _p(a), t.size, _p(out))
    return out[:n].copy()


def lookahead_walk(tokens, data) -> np.ndarray:
    """The ratio mode's walk (this library's opt-in rule, not the reference's)."""
    t = np.ascontiguousarray(tokens, np.uint32)
    a = _u8(data)
    out = np.zeros(max(t.size, 1), np.uint32)
    n = _lib().orc_lookahead_walk(_p(t), _p(a), t.size, _p(out))
    return out[:n].copy()


def strict_walk(tokens) -> np.ndarray:
    t = np.ascontiguousarray(tokens, np.uint32)
    out = np.zeros(max(t.size, 1), np.uint32)
    n = _lib().orc_strict_walk(_p(t), t.size, _p(out))
    return out[:n].copy()


def reconstruct(tokens, cap: int) -> bytes | None:
    t = np.ascontiguousarray(tokens, np.uint32)
    out = np.zeros(max(cap, 1), np.uint8)
    n = _lib().orc_reconstruct(_p(t), t.size, _p(out), cap)
    return None if n == SIZE_MAX else out[:n].tobytes()


def build_code_lengths(freq, max_bits: int):
    f = np.ascontiguousarray(freq, np.uint32)
    lens = np.zeros(f.size, np.uint8)
    ok = _lib().orc_build_code_lengths(_p(f), f.size, max_bits, _p(lens))
    return lens if ok else None


def canonical_codes(lens) -> np.ndarray:
    l = np.ascontiguousarray(lens, np.uint8)
    c = np.zeros(l.size, np.uint32)
    _lib().orc_canonical_codes(_p(l), l.size, _p(c))
    return c


def encode_block(tokens) -> bytes:
    t = np.ascontiguousarray(tokens, np.uint32)
    cap = 6 * t.size + 4096
    out = np.zeros(cap, np.uint8)
    n = _lib().orc_encode_block_fast(_p(t), t.size, _p(out), cap)
    if n == SIZE_MAX:
        raise ValueError("token outside the DEFLATE alphabet")
    return out[:n].tobytes()


def crc32(data) -> int:
    a = _u8(data)
    return _lib().orc_crc32(_p(a), a.size)


def gzip_member(data) -> bytes:
    a = _u8(data)
    cap = _lib().orc_chunk_bound(a.size, max(a.size, 1))
    out = np.zeros(cap, np.uint8)
    n = _lib().orc_gzip_member(_p(a), a.size, _p(out), cap)
    assert n != SIZE_MAX
    return out[:n].tobytes()


def gzip_chunk(data, sub_chunk: int = 131072, rule: int = 0) -> bytes:
    """rule 0: the reference's walk; rule 2: the ratio mode's lookahead walk."""
    a = _u8(data)
    cap = _lib().orc_chunk_bound(a.size, sub_chunk)
    out = np.zeros(cap, np.uint8)
    n = _lib().orc_gzip_chunk_rule(_p(a), a.size, sub_chunk, rule, _p(out), cap)
    assert n != SIZE_MAX
    return out[:n].tobytes()


def gzip_chunk_into(a: np.ndarray, sub_chunk: int = 131072) -> int:
    """Compress a contiguous uint8 array, discarding the bytes; returns the size.
    (CPU-baseline timing: the ctypes call releases the GIL.)"""
    cap = _lib().orc_chunk_bound(a.size, sub_chunk)
    out = np.empty(cap, np.uint8)
    n = _lib().orc_gzip_chunk(_p(a), a.size, sub_chunk, _p(out), cap)
    assert n != SIZE_MAX
    return int(n)


def _zlib() -> C.CDLL:
    global _z
    if _z is None:
        if not os.path.exists(ZLIB_LIB):
            build()
        Z = C.CDLL(ZLIB_LIB)
        Z.cpuz_compress_chunks.restype = C.c_int64
        Z.cpuz_compress_chunks.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int,
                                           C.c_void_p, C.c_uint64, C.c_void_p]
        Z.cpuz_zlib_version.restype = C.c_char_p
        _z = Z
    return _z


def zlib_chunks(data: np.ndarray, chunk: int, level: int, threads: int, keep: bool = False):
    """gzip-codec-cpu structure on system zlib: returns (total_bytes, bytes|None)."""
    a = _u8(data)
    n = (a.size + chunk - 1) // chunk
    off = np.arange(n, dtype=np.uint64) * chunk
    lens = np.minimum(chunk, a.size - off).astype(np.uint64)
    outl = np.zeros(n, np.uint64)
    out = np.empty(a.size + a.size // 8 + 1024 * n, np.uint8) if keep else None
    tot = _zlib().cpuz_compress_chunks(_p(a), _p(off), _p(lens), n, level, threads,
                                       _p(out) if keep else None, out.size if keep else 0, _p(outl))
    return int(tot), (out[:tot].tobytes() if keep else None)


def zlib_version() -> str:
    return _zlib().cpuz_zlib_version().decode()
