"""C++ mirror of the reference backend surface (include/gpzip_backend.hpp)."""
import gzip
import os
import subprocess

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
