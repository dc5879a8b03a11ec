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


def _tree(root):
    """A small tree: nested directories, empty and multi-MiB files, a name
    longer than the 100-byte header field, a symlink."""
    (root / "src" / "deep" / "er").mkdir(parents=True)
    (root / "src" / "empty.txt").write_bytes(b"")
    (root / "src" / "a.log").write_bytes(gpzip_amd.synth("log", (5 << 20) + 123, 3).tobytes())
    (root / "src" / "deep" / "b.bin").write_bytes(gpzip_amd.synth("binmix", 700001, 4).tobytes())
    (root / "src" / "deep" / "er" / ("n" * 120 + ".txt")).write_bytes(gpzip_amd.synth("text", 3000, 5).tobytes())
    (root / "src" / "link").symlink_to("a.log")
    (root / "top.txt").write_bytes(b"top level file\n" * 1000)
    return [str(root / "src"), str(root / "top.txt")]


def _check_tar(tarbytes: bytes, root):
    import io
    import tarfile

    with tarfile.open(fileobj=io.BytesIO(tarbytes)) as t:
        names = t.getnames()
        assert names[0] == "src" and "top.txt" in names
        assert names.index("src/a.log") < names.index("src/deep") < names.index("src/deep/b.bin")
        long_name = "src/deep/er/" + "n" * 120 + ".txt"
        for n in ("src/empty.txt", "src/a.log", "src/deep/b.bin", long_name, "top.txt"):
            assert t.extractfile(n).read() == (root / n).read_bytes(), n
        assert t.getmember("src/link").issym() and t.getmember("src/link").linkname == "a.log"


def test_tar_builder_reads_back(tmp_path):
    exe = build_exe()
    inputs = _tree(tmp_path)
    out = tmp_path / "x.tar"
    r = subprocess.run([exe, "tar", str(out), *inputs], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    data = out.read_bytes()
    assert len(data) % 512 == 0 and data[-1024:] == bytes(1024)
    _check_tar(data, tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk,batch", [(2 << 20, 64), (2 << 20, 1), (131072, 5)])
def test_tar_gz_on_batched_form(tmp_path, chunk, batch):
    """`gpzip a x.tar.gz` on the batched form: the archive decodes to the tar
    stream, the entries read back, and the bytes equal the per-chunk GPU form
    over the same chunks."""
    exe = build_exe()
    inputs = _tree(tmp_path)
    tar = tmp_path / "x.tar"
    r = subprocess.run([exe, "tar", str(tar), *inputs], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    tgz = tmp_path / "x.tar.gz"
    r = subprocess.run([exe, "targz", str(tgz), str(chunk), str(batch), *inputs], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    raw = tar.read_bytes()
    gz = tgz.read_bytes()
    assert (tmp_path / "x.tar.gz.2").read_bytes() == gz  # pack_tar_gz itself
    assert gpzip_amd.gunzip(gz) == raw
    _check_tar(raw, tmp_path)
    nchunks = (len(raw) + chunk - 1) // chunk
    kv = r.stderr.decode().split()
    st = dict(zip(kv[0::2], kv[1::2]))
    assert int(st["chunks"]) == nchunks and int(st["batches"]) == (nchunks + batch - 1) // batch
    with gpzip_amd.Context() as ctx:
        plain = b"".join(ctx.compress_chunk(raw[o:o + chunk]) for o in range(0, len(raw), chunk))
    assert gz == plain
