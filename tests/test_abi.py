"""The C ABI library loads and exports exactly what include/gpzip_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpzip_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "gpzip_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gpz_[a-z_0-9]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(gpzip_amd.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = gpzip_amd.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_status_strings_and_defaults():
    L = gpzip_amd.load_library()
    assert L.gpz_status_string(0) == b"ok"
    cfg = gpzip_amd.Config()
    L.gpz_config_default(ctypes.byref(cfg))
    # the reference's fixed tunables (codec-gpu/src/gpu.rs:160-165, lz77_hash.rs:48)
    assert cfg.sub_chunk == 131072 and cfg.window == 32768 and cfg.max_in_flight == 16


def test_init_without_gpu_reports_nodev():
    if gpzip_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(gpzip_amd.GpzError) as e:
        gpzip_amd.Context()
    assert e.value.status == gpzip_amd.GPZ_E_NODEV


def test_bad_arguments_fail_cleanly():
    L = gpzip_amd.load_library()
    n = ctypes.c_size_t()
    assert L.gpz_compress_chunk(None, None, 0, None, 0, ctypes.byref(n)) == gpzip_amd.GPZ_E_INVAL
    assert L.gpz_init(None, None) == gpzip_amd.GPZ_E_INVAL
    assert L.gpz_last_error() != b""


def test_synth_deterministic():
    a = gpzip_amd.synth("binmix", 3 << 20, 7)
    b = gpzip_amd.synth("binmix", 3 << 20, 7)
    c = gpzip_amd.synth("binmix", 3 << 20, 8)
    assert (a == b).all() and not (a == c).all()
    t = gpzip_amd.synth("text", 1 << 16, 1).tobytes()
    assert set(t) <= set(b"abcdefghijklmnopqrstuvwxyz \n")
    lg = gpzip_amd.synth("log", 1 << 16, 1).tobytes()
    assert lg.count(b"\n") > 300 and max(lg) < 128
    s = gpzip_amd.loguniform_sizes(3, 4096, 4096, 4 << 20)
    assert s.min() >= 4096 and s.max() <= 4 << 20


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the header's structs have the C sizes and field
    offsets (a small C program built with gcc against include/gpzip_hip.h
    prints them)."""
    import shutil
    import subprocess

    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    mirrors = {"gpz_config": gpzip_amd.Config, "gpz_desc": gpzip_amd.Desc, "gpz_tune": gpzip_amd.Tune,
               "gpz_server_config": gpzip_amd.ServerConfig}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "gpzip_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
