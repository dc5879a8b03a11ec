"""C++ mirror of the reference backend surface (include/gpzip_backend.hpp)."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import gpzip_amd

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(HERE, "cpp", "test_backend")


def build_exe():
    src = os.path.join(HERE, "cpp", "test_backend.cpp")
    if os.path.exists(EXE) and os.path.getmtime(EXE) >= max(os.path.getmtime(src), os.path.getmtime(gpzip_amd.HIP_LIB_PATH)):
        return EXE
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", EXE,
                    gpzip_amd.HIP_LIB_PATH, "-Wl,-rpath," + os.path.dirname(gpzip_amd.HIP_LIB_PATH), "-lpthread"],
                   check=True)
    return EXE


def test_parallel_chunked_writer_and_lazy_backend():
    exe = build_exe()
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "cpu ok" in r.stdout


@pytest.mark.gpu
def test_gpu_backend_writer_equals_oracle(tmp_path):
    from oracle import oracle as O

    exe = build_exe()
    data = gpzip_amd.synth("text", (3 << 20) + 12345, 5)
    f = tmp_path / "in.bin"
    f.write_bytes(data.tobytes())
    r = subprocess.run([exe, "gpu", str(f)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    assert gzip.decompress(r.stdout) == data.tobytes()
    assert r.stdout == O.gzip_chunk(data, 131072)


def _split_members(stream: bytes, chunks, cpu_members, gpu_members):
    """Which encoding each chunk's output is: 'cpu', 'gpu' or None."""
    pos, kinds = 0, []
    for c, m_cpu, m_gpu in zip(chunks, cpu_members, gpu_members):
        if stream.startswith(m_cpu, pos):
            kinds.append("cpu")
            pos += len(m_cpu)
        elif stream.startswith(m_gpu, pos):
            kinds.append("gpu")
            pos += len(m_gpu)
        else:
            kinds.append(None)
            break
    return kinds, pos


@pytest.mark.gpu
@pytest.mark.parametrize("gpu_workers,warmup", [(1, 4), (8, 4), (16, 0)])
def test_hybrid_members_are_cpu_or_gpu_encodings(tmp_path, gpu_workers, warmup):
    """`--backend auto` (cli/src/hybrid.rs): the stream decodes, and each 2 MiB
    chunk's output is exactly the CPU path's zlib member or the GPU path's
    members; the warm-up chunks are on the CPU."""
    from oracle import oracle as O

    exe = build_exe()
    kinds = ["text", "binmix", "log"]
    data = np.concatenate([gpzip_amd.synth(kinds[i % 3], 2 << 20, 80 + i) for i in range(11)] +
                          [gpzip_amd.synth("log", 777777, 99)])
    f = tmp_path / "in.bin"
    f.write_bytes(data.tobytes())
    r = subprocess.run([exe, "hybrid", str(f), str(gpu_workers), str(warmup)], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    stream = r.stdout
    assert gpzip_amd.gunzip(stream) == data.tobytes()
    chunks = [data[i:i + (2 << 20)] for i in range(0, data.size, 2 << 20)]
    cpu_m = [O.zlib_chunks(c, 2 << 20, 5, 1, keep=True)[1] for c in chunks]
    gpu_m = [O.gzip_chunk(c, 131072) for c in chunks]
    got, pos = _split_members(stream, chunks, cpu_m, gpu_m)
    assert None not in got and pos == len(stream), got
    assert got[:warmup] == ["cpu"] * warmup
    st = dict(zip(r.stderr.decode().split()[0::2], r.stderr.decode().split()[1::2]))
    assert int(st["gpu_chunks"]) == got.count("gpu") >= 1 and int(st["cpu_chunks"]) == got.count("cpu")


@pytest.mark.gpu
def test_decoder_program_decodes_gpu_stream(tmp_path):
    exe = build_exe()
    data = gpzip_amd.synth("binmix", (5 << 20) + 3, 12)
    with gpzip_amd.Context() as ctx:
        gz = ctx.compress_chunk(data)
    f = tmp_path / "in.gz"
    f.write_bytes(gz)
    r = subprocess.run([exe, "decode", str(f)], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    assert r.stdout == data.tobytes()
