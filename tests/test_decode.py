"""The decode side (gpz_gunzip, gpz_gzip_member_starts): gpzip x's member scan
and parallel inflate (codec-cpu/src/parallel_decompress.rs:35-113) with its
serial fallback, on the host cores.  No GPU needed: the streams come from the
oracle (the GPU path's bytes), zlib and Python's gzip module."""
import gzip
import zlib

import numpy as np
import pytest

import gpzip_amd
from oracle import oracle as O

HDR = bytes.fromhex("1f8b08000000000000ff")


def test_multi_member_stream_decodes_in_parallel():
    data = gpzip_amd.synth("binmix", (3 << 20) + 4321, 3).tobytes()
    gz = O.gzip_chunk(data)  # 25 members of 128 KiB, the GPU path's bytes
    starts = gpzip_amd.member_starts(gz)
    assert len(starts) == (len(data) + 131071) // 131072 and starts[0] == 0
    for t in (1, 4, 0):
        assert gpzip_amd.gunzip(gz, threads=t) == data


def test_reference_style_cases():
    # parallel_decompress.rs:124-158: one member falls back to the serial decoder;
    # 16 members are found and decoded
    one = b"hello world hello world hello world" * 100
    gz1 = O.gzip_chunk(one)
    assert len(gpzip_amd.member_starts(gz1)) == 1
    assert gpzip_amd.gunzip(gz1) == one
    parts = [(f"chunk {i} " * 2000).encode() for i in range(16)]
    gz16 = b"".join(O.gzip_chunk(p) for p in parts)
    assert len(gpzip_amd.member_starts(gz16)) == 16
    assert gpzip_amd.gunzip(gz16) == b"".join(parts)


def test_foreign_and_empty_streams_take_the_serial_path():
    data = gpzip_amd.synth("text", 1 << 20, 4).tobytes()
    # Python's gzip writes an mtime, so no fixed header is found
    g = gzip.compress(data, compresslevel=6)
    assert gpzip_amd.member_starts(g) == []
    assert gpzip_amd.gunzip(g) == data
    # concatenation of foreign members (RFC 1952 2.2)
    assert gpzip_amd.gunzip(g + gzip.compress(b"tail", mtime=123)) == data + b"tail"
    assert gpzip_amd.gunzip(b"") == b""
    assert gpzip_amd.gunzip(O.gzip_chunk(b"")) == b""


def test_false_header_inside_a_member_falls_back():
    # a stored (level 0) member whose payload holds the header bytes: the scan
    # finds a start inside DEFLATE data, the parallel decode fails, and the
    # serial decoder gives the right bytes
    payload = b"A" * 5000 + HDR + b"B" * 5000
    d = zlib.compressobj(0, zlib.DEFLATED, -15)
    raw = d.compress(payload) + d.flush()
    trailer = zlib.crc32(payload).to_bytes(4, "little") + len(payload).to_bytes(4, "little")
    stream = HDR + raw + trailer + O.gzip_chunk(b"second member " * 100)
    assert len(gpzip_amd.member_starts(stream)) == 3
    assert gpzip_amd.gunzip(stream) == payload + b"second member " * 100


def test_corrupt_and_truncated_streams_fail():
    data = gpzip_amd.synth("log", 600_000, 5).tobytes()
    gz = bytearray(O.gzip_chunk(data))
    with pytest.raises(gpzip_amd.GpzError):
        gpzip_amd.gunzip(bytes(gz[:-5]))
    gz[len(gz) // 2] ^= 0xFF
    with pytest.raises(gpzip_amd.GpzError):
        gpzip_amd.gunzip(bytes(gz))
    with pytest.raises(gpzip_amd.GpzError):
        gpzip_amd.gunzip(b"not a gzip stream")


def test_cpu_zlib_members_decode():
    data = gpzip_amd.synth("binmix", 5 << 20, 6)
    _, z = O.zlib_chunks(data, 2 << 20, 5, 4, keep=True)
    assert len(gpzip_amd.member_starts(z)) == 3
    assert gpzip_amd.gunzip(z) == data.tobytes()
