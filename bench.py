"""gpzip_amd benchmark: input GB/s of the gzip compress path on one MI355X.

Default workload (the configuration the headline metric is quoted on): the
'binmix' corpus, 256 MiB of seeded synthetic data, 2 MiB chunks, each chunk
compressed by the plain per-chunk form (one pipeline launch sequence per chunk,
all chunks submitted on their own streams), inputs resident in HBM.  A step is
one pass over the whole corpus.  Beside it: the batched form (one graph replay
over all 128 chunks) on the same data, the roofline of the dominant kernel
(match-find) from HIP events on its stream, the end-to-end rates through the
drop-in boundary with host buffers (PCIe both ways; never the headline), and
the CPU baseline.  The same JSON line carries the mixed-size comparison
(4,096 chunks of 4 KiB-4 MiB, batched form against the plain form on the same
GPU) and the serving form (20,000 Poisson requests at a stated offered load:
p50/p99 latency and sustained GB/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-mixed] [--no-serve]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes as C
import gzip
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpzip_amd  # noqa: E402
from gpzip_amd import workloads  # noqa: E402

HBM_PEAK_GBPS = 8000.0
MiB = 1 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DevicePlan:
    """gpz_plan over chunks laid out in a device buffer."""

    def __init__(self, ctx, descs, flags, tune=None):
        self.L = ctx.L
        arr = (gpzip_amd.Desc * len(descs))()
        for i, (o, n) in enumerate(descs):
            arr[i].offset, arr[i].len = o, n
        h = C.c_void_p()
        t = C.byref(tune) if tune is not None else None
        gpzip_amd._check(self.L.gpz_plan_create(ctx.h, arr, len(descs), flags, t, C.byref(h)))
        self.h = h
        self.n = len(descs)
        self.out_cap = self.L.gpz_plan_out_capacity(h)

    def run(self, d_in: int, d_out: int):
        gpzip_amd._check(self.L.gpz_plan_run(self.h, C.c_void_p(d_in), C.c_void_p(d_out)))

    def sync(self):
        gpzip_amd._check(self.L.gpz_plan_sync(self.h))

    def result(self):
        off = np.zeros(self.n, np.uint64)
        ln = np.zeros(self.n, np.uint64)
        tot = C.c_uint64()
        gpzip_amd._check(self.L.gpz_plan_result(self.h, off.ctypes.data_as(gpzip_amd.u64p),
                                                ln.ctypes.data_as(gpzip_amd.u64p), C.byref(tot)))
        return off, ln, tot.value

    def timing(self):
        v = (C.c_float * 4)()
        gpzip_amd._check(self.L.gpz_plan_timing(self.h, v, 4))
        return list(v)

    def close(self):
        if self.h:
            self.L.gpz_plan_destroy(self.h)
            self.h = None


def layout(sizes, align=256):
    offs, o = [], 0
    for s in sizes:
        offs.append(o)
        o += (int(s) + align - 1) // align * align
    return offs, o


def fetch(torch, d_out, off, ln):
    return d_out[int(off):int(off) + int(ln)].cpu().numpy().tobytes()


def plain_and_batched(torch, ctx, data: np.ndarray, sizes, steps, warmup, check_oracle=2, autotune=True):
    """Returns dict with plain / batched timings for chunks `sizes` of `data`."""
    offs, total_in = layout(sizes)
    host = np.zeros(total_in + 256, np.uint8)
    src = 0
    for o, s in zip(offs, sizes):
        host[o:o + s] = data[src:src + s]
        src += s
    d_in = torch.from_numpy(host).to("cuda:0")
    n_in = int(sum(sizes))

    # plain per-chunk form: one plan (own stream, one workgroup per segment /
    # member, direct launches) per chunk
    plans = [DevicePlan(ctx, [(o, s)], 0) for o, s in zip(offs, sizes)]
    out_offs = []
    acc = 0
    for p in plans:
        out_offs.append(acc)
        acc += (p.out_cap + 255) // 256 * 256
    d_out_plain = torch.zeros(acc + 256, dtype=torch.uint8, device="cuda:0")
    base_in, base_out = d_in.data_ptr(), d_out_plain.data_ptr()

    def run_plain():
        for p, oo in zip(plans, out_offs):
            p.run(base_in, base_out + oo)
        for p in plans:
            p.sync()

    # batched form: autotuned launch shapes, one plan over every chunk, graph
    # replay, persistent grids
    tune = None
    if autotune:
        probe = DevicePlan(ctx, list(zip(offs, sizes)), gpzip_amd.GPZ_PLAN_PERSISTENT)
        d_probe = torch.zeros(probe.out_cap + 256, dtype=torch.uint8, device="cuda:0")
        probe.close()
        tune, t_ms, t_def = ctx.autotune(list(zip(offs, sizes)), base_in, d_probe.data_ptr())
        del d_probe
        ctx.set_tune(tune)
        log(f"  autotune: {tune.as_dict()}  pipeline {t_def:.3f} -> {t_ms:.3f} ms")
    bplan = DevicePlan(ctx, list(zip(offs, sizes)), gpzip_amd.GPZ_PLAN_GRAPH | gpzip_amd.GPZ_PLAN_PERSISTENT)
    d_out_b = torch.zeros(bplan.out_cap + 256, dtype=torch.uint8, device="cuda:0")

    def run_batched():
        bplan.run(base_in, d_out_b.data_ptr())
        bplan.sync()

    res = {}
    for name, fn in (("plain", run_plain), ("batched", run_batched)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[name] = {"s_per_step": dt, "gbps": n_in / dt / 1e9}
        log(f"  {name:8s}: {dt * 1e3:9.2f} ms/step  {n_in / dt / 1e9:8.2f} GB/s")

    # outputs: plain == batched chunk by chunk; decode; oracle on a few chunks
    boff, blen, btot = bplan.result()
    total_out = 0
    chunks_out = []
    for i, (p, oo) in enumerate(zip(plans, out_offs)):
        poff, plen, ptot = p.result()
        a = fetch(torch, d_out_plain, oo + poff[0], plen[0])
        b = fetch(torch, d_out_b, boff[i], blen[i])
        assert a == b, f"chunk {i}: plain and batched outputs differ"
        total_out += len(a)
        chunks_out.append(a)
    full = b"".join(chunks_out)
    assert gzip.decompress(full) == data[:n_in].tobytes(), "output does not decode to the input"
    if check_oracle:
        from oracle import oracle as O  # expected outputs only

        src = 0
        for i, s in enumerate(sizes[:check_oracle]):
            assert chunks_out[i] == O.gzip_chunk(data[src:src + s]), f"chunk {i} differs from the oracle"
            src += s
    res["ratio"] = total_out / n_in
    res["chunks_out"] = chunks_out
    res["n_in"] = n_in
    res["n_out"] = total_out
    res["members"] = int(sum((int(s) + 131071) // 131072 if s > 131072 else 1 for s in sizes))

    # dominant-kernel timing: batched plan, direct launches on its stream, HIP
    # events bracket the match-find and emit kernels
    tplan = DevicePlan(ctx, list(zip(offs, sizes)), gpzip_amd.GPZ_PLAN_PERSISTENT | gpzip_amd.GPZ_PLAN_TIMED)
    d_out_t = torch.zeros(tplan.out_cap + 256, dtype=torch.uint8, device="cuda:0")
    tplan.run(base_in, d_out_t.data_ptr())
    tplan.sync()
    k1, emit, whole, k0 = [], [], [], []
    for _ in range(max(steps, 3)):
        tplan.run(base_in, d_out_t.data_ptr())
        tplan.sync()
        w_, a_, e_, z_ = tplan.timing()
        whole.append(w_)
        k1.append(a_)
        emit.append(e_)
        k0.append(z_)
    res["tune"] = tune.as_dict() if tune is not None else None
    res["k1_ms"] = float(np.mean(k1))
    res["k0_ms"] = float(np.mean(k0))
    res["emit_ms"] = float(np.mean(emit))
    res["pipeline_ms"] = float(np.mean(whole))
    for p in plans:
        p.close()
    bplan.close()
    tplan.close()
    del d_in, d_out_plain, d_out_b, d_out_t
    torch.cuda.empty_cache()
    return res


def host_cores() -> tuple[int, dict]:
    """Cores this process may use: its CPU affinity, capped by a cgroup CPU
    quota when one is set (a share of a larger host)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    n = min(aff, quota) if quota else aff
    return n, {"affinity_cpus": aff, "cgroup_cpu_quota": quota}


def cpu_baseline(data: np.ndarray, chunk: int, reps: int = 3):
    """gpzip-codec-cpu's structure (2 MiB chunks, one zlib level-5 gzip member
    each, in chunk order, all host cores) on the whole workload, median of
    `reps` runs; and one core on the first 32 MiB."""
    from oracle import oracle as O  # CPU baseline only

    cores, info = host_cores()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ztot, _ = O.zlib_chunks(data, chunk, 5, cores)
        times.append(time.perf_counter() - t0)
    dz = float(np.median(times))
    one = data[:32 * MiB]
    t0 = time.perf_counter()
    O.zlib_chunks(one, chunk, 5, 1)
    d1 = time.perf_counter() - t0
    return {"value": data.size / dz / 1e9, "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"the whole workload ({data.size // MiB} MiB, {data.size // chunk} x {chunk // MiB} MiB chunks), "
                      f"median of {reps} runs",
            "program": "gpzip-codec-cpu structure (codec-cpu/src/compress.rs:52-59, backend.rs:14-35): one gzip "
                       "member per 2 MiB chunk, zlib level 5, chunk order, on system zlib "
                       + O.zlib_version() + " (not zlib-rs 0.6.3; member bytes unpinned)",
            "ratio": ztot / data.size, "single_core_gbps": one.size / d1 / 1e9, **info}


def cpu_port(data: np.ndarray, sample_bytes: int, chunk: int):
    """The oracle (a CPU restatement of the GPU algorithm) on the host cores."""
    from oracle import oracle as O  # CPU baseline only

    cores, _ = host_cores()
    sample = data[:sample_bytes]
    pieces = [np.ascontiguousarray(sample[i:i + chunk]) for i in range(0, sample.size, chunk)]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        outs = list(ex.map(O.gzip_chunk_into, pieces))
    dt = time.perf_counter() - t0
    return {"value": sample.size / dt / 1e9, "unit": "GB/s", "cores": cores,
            "sample": f"first {sample.size // MiB} MiB of the workload ({len(pieces)} x {chunk // MiB} MiB chunks)",
            "program": "the oracle: CPU restatement of the GPU algorithm (segmented hash, lazy walk, "
                       "encode_block_fast), one chunk per thread",
            "ratio": sum(outs) / sample.size}


def end_to_end(ctx, data: np.ndarray, sizes, expect, threads: int = 16, reps: int = 3):
    """Host buffers through the drop-in boundary: gpz_compress_chunk called
    from `threads` workers at once (the reference's rayon ChunkFn callers), and
    the whole list as one gpz_batch_submit.  PCIe both ways; never `value`."""
    pieces, src = [], 0
    for s in sizes:
        pieces.append(np.ascontiguousarray(data[src:src + s]))
        src += s
    n_in = sum(sizes)
    with cf.ThreadPoolExecutor(threads) as ex:
        out = list(ex.map(ctx.compress_chunk, pieces))  # warm-up, and the check
        assert out == expect, "gpz_compress_chunk differs from the device-resident plan"
        t0 = time.perf_counter()
        for _ in range(reps):
            list(ex.map(ctx.compress_chunk, pieces))
        dp = (time.perf_counter() - t0) / reps
    # the batch API on one contiguous host buffer (as a service holds it);
    # results are pointers into the batch, read after the timed region
    L = ctx.L
    base = np.concatenate(pieces)
    descs = (gpzip_amd.Desc * len(pieces))()
    o = 0
    for i, s in enumerate(sizes):
        descs[i].offset, descs[i].len = o, s
        o += s

    def batch_once():
        b = C.c_void_p()
        gpzip_amd._check(L.gpz_batch_submit(ctx.h, C.c_void_p(base.ctypes.data), descs, len(pieces), C.byref(b)))
        gpzip_amd._check(L.gpz_batch_wait(b))
        return b

    b = batch_once()
    try:
        got = []
        for i in range(len(pieces)):
            ptr, ln = C.POINTER(C.c_uint8)(), C.c_size_t()
            gpzip_amd._check(L.gpz_batch_result(b, i, C.byref(ptr), C.byref(ln)))
            got.append(C.string_at(ptr, ln.value))
    finally:
        L.gpz_batch_free(b)
    assert got == expect, "gpz_batch_submit differs from the device-resident plan"
    t0 = time.perf_counter()
    for _ in range(reps):
        L.gpz_batch_free(batch_once())
    db = (time.perf_counter() - t0) / reps
    return {"plain_gbps": round(n_in / dp / 1e9, 3), "batched_gbps": round(n_in / db / 1e9, 3), "unit": "GB/s",
            "plain_api": f"gpz_compress_chunk from {threads} threads, 2 MiB chunks, host buffers (PCIe both ways)",
            "batched_api": "gpz_batch_submit + gpz_batch_wait over the 128 chunks in one host buffer, "
                           "results in host memory (PCIe both ways)"}


def kernel_source_hash() -> str:
    import hashlib

    with open(os.path.join(ROOT, "gpzip_amd", "csrc", "gpz_kernels.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profiled_counters(n_in: int, k1_grid: int = 0) -> dict:
    """k_match_find counter figures as last measured with rocprofv3 --pmc (the
    counters cannot be read from inside this process): HBM bytes per launch
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by a copy-kernel
    calibration) for the launch shape of k1_grid workgroups when the profile
    has it, VALU issue and lane-use fractions.  Empty if the committed profile
    is for another input size or for other kernel sources (its
    kernel_source_sha256 line differs from the current gpz_kernels.hip)."""
    path = os.path.join(ROOT, "profiles", "k_match_find_binmix256_traffic.md")
    vals = {}
    try:
        for line in open(path):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("traffic_bytes_per_launch", "input_bytes") or k.startswith("traffic_bytes_per_launch_wg"):
                vals[k] = int(v.strip())
            elif k in ("valu_issue_frac", "valu_lane_frac"):
                vals[k] = float(v.strip())
            elif k == "kernel_source_sha256":
                vals[k] = v.strip()
    except (OSError, ValueError):
        return {}
    if vals.get("input_bytes") != n_in or vals.get("kernel_source_sha256") != kernel_source_hash():
        return {}
    shaped = vals.get(f"traffic_bytes_per_launch_wg{k1_grid}")
    if shaped is not None:
        vals["traffic_bytes_per_launch"] = shaped
        vals["traffic_shape"] = f"{k1_grid} workgroups (the launch shape of this run)"
    else:
        vals["traffic_shape"] = "the shape the profile's own autotune picked"
    return vals


def serve(ctx, n_requests, load_gbps, seed=4, submitters=4, budget_bytes=64 * 10**9, **server_cfg):
    """Config 4: Poisson arrivals at load_gbps, log-normal request sizes.  The
    schedule is shared by `submitters` threads (request i goes to thread
    i % submitters, each sleeping until its arrival time), so one caller's
    copy into the server does not cap the offered rate; the drain counts
    output bytes without copying them."""
    import threading

    R = workloads.ServingRequests(n_requests, load_gbps, seed)
    srv = gpzip_amd.Server(ctx, budget_bytes=budget_bytes, **server_cfg)
    lat = []
    done = 0
    out_bytes = 0

    def drain():
        nonlocal done, out_bytes
        while done < n_requests:
            r = srv.next(timeout_ms=100, copy=False)
            if r is None:
                continue
            tag, nbytes, l = r
            lat.append(l)
            out_bytes += nbytes
            done += 1

    t_last = [0.0] * submitters

    def submit(t):
        for i in range(t, n_requests, submitters):
            now = time.perf_counter()
            if t0 + R.arrive[i] > now:
                time.sleep(t0 + R.arrive[i] - now)
            srv.submit(R.data(i), i)
        t_last[t] = time.perf_counter()

    th = threading.Thread(target=drain)
    t0 = time.perf_counter()
    th.start()
    subs = [threading.Thread(target=submit, args=(t,)) for t in range(submitters)]
    for s_ in subs:
        s_.start()
    for s_ in subs:
        s_.join()
    t_arrived = max(t_last)
    th.join()
    dt = time.perf_counter() - t0
    stats = srv.stats()
    srv.stop()
    lat = np.array(lat)
    total = int(R.sizes.sum())
    # arrived_gbps: the rate the submitters actually reached; below offered_gbps when
    # submission itself fell behind the schedule
    return {"requests": n_requests, "offered_gbps": load_gbps,
            "arrived_gbps": round(total / (t_arrived - t0) / 1e9, 3),
            "sustained_gbps": round(total / dt / 1e9, 3),
            "p50_ms": round(float(np.percentile(lat, 50) * 1e3), 3),
            "p99_ms": round(float(np.percentile(lat, 99) * 1e3), 3),
            "ratio": round(out_bytes / float(total), 5), "device_steps": stats["steps"],
            "input_bytes": total, "submitters": submitters, "budget_bytes": budget_bytes,
            "peak_arena_bytes": stats["peak_bytes"], "admissions_waited": stats["waited"]}


def hybrid_sweep(mib: int = 256, settings=((0, 0, 4), (1, 0, 4), (8, 0, 4), (16, 0, 4), (8, 1000, 4), (4, 1000, 0),
                                           (8, 1000, 0), (16, 1000, 0), (8, 0, 0))) -> dict:
    """`--backend auto` end to end: the C++ HybridBackend (cli/src/hybrid.rs
    mirror) compresses the binmix corpus from host memory through a
    ParallelChunkedWriter of 2 MiB chunks, CPU workers = host cores, with 0..16
    GPU permits; every stream is decoded and checked."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "tests", "cpp", "test_backend")
    src = exe + ".cpp"
    lib = gpzip_amd.HIP_LIB_PATH
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(src), os.path.getmtime(lib)):
        subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", exe, lib,
                        "-Wl,-rpath," + os.path.dirname(lib), "-lpthread"], check=True)
    data = gpzip_amd.synth("binmix", mib * MiB, 1)
    res = []
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "in.bin")
        data.tofile(f)
        for p, wait, warmup in settings:
            log(f"hybrid: {p} GPU permits, wait {wait} ms, warm-up {warmup} ...")
            r = subprocess.run([exe, "hybrid", f, str(p), str(warmup), str(wait)], capture_output=True, timeout=600)
            if r.returncode:
                raise RuntimeError(r.stderr.decode())
            kv = r.stderr.decode().split()
            st = dict(zip(kv[0::2], kv[1::2]))
            assert gpzip_amd.gunzip(r.stdout) == data.tobytes(), "hybrid stream does not decode"
            g, c = int(st["gpu_chunks"]), int(st["cpu_chunks"])
            res.append({"gpu_permits": p, "gpu_wait_ms": wait, "warmup_chunks": warmup, "gbps": round(data.size / float(st["seconds"]) / 1e9, 3),
                        "cold_gbps": round(data.size / float(st["cold_seconds"]) / 1e9, 3),
                        "gpu_chunk_share": round(g / max(1, g + c), 3), "ratio": round(len(r.stdout) / data.size, 5)})
        # `--backend gpu`: GpuBackend's own writer (128 KiB members, 16 in flight, gpu.rs:160-166)
        log("backend gpu ...")
        r = subprocess.run([exe, "gpu", f], capture_output=True, timeout=600)
        if r.returncode:
            raise RuntimeError(r.stderr.decode())
        kv = r.stderr.decode().split()
        st = dict(zip(kv[0::2], kv[1::2]))
        assert gpzip_amd.gunzip(r.stdout) == data.tobytes(), "backend gpu stream does not decode"
        gpu_only = {"gbps": round(data.size / float(st["seconds"]) / 1e9, 3),
                    "cold_gbps": round(data.size / float(st["cold_seconds"]) / 1e9, 3), "ratio": round(len(r.stdout) / data.size, 5),
                    "writer": "GpuBackend compressor: ParallelChunkedWriter of 2 MiB chunks (16 members of 128 KiB, "
                              "the bytes of 128 KiB chunks), 16 in flight on 16 lanes"}
    cores, _ = host_cores()
    return {"workload": f"`--backend auto`: binmix {mib} MiB from host memory, 2 MiB chunks, {cores} CPU workers "
                        f"(zlib level 5) + GPU permits", "runs": res, "backend_gpu": gpu_only}


def mixed_sizes(torch, ctx, steps):
    """Config 3: the batched form against the plain form on 4,096 chunks of
    4 KiB-4 MiB (outputs equal chunk by chunk, the whole output decodes, the
    first chunks equal the oracle)."""
    mdata, msizes = workloads.mixed_sizes_workload()
    rm = plain_and_batched(torch, ctx, mdata, msizes, steps, 1, check_oracle=2)
    n_in, n_out = rm["n_in"], rm["n_out"]
    return {"workload": "4,096 chunks, sizes log-uniform over 4 KiB-4 MiB (seed 3), content 1/3 each "
                        "text / binmix / log, inputs resident in HBM",
            "input_bytes": n_in, "chunks": len(msizes), "steps": steps,
            "plain_gbps": round(rm["plain"]["gbps"], 3), "batched_gbps": round(rm["batched"]["gbps"], 3),
            "batched_vs_plain": round(rm["batched"]["gbps"] / rm["plain"]["gbps"], 3),
            "target": ">= 3x plain", "ratio": round(rm["ratio"], 5), "autotune": rm["tune"],
            "k_match_find_ms": round(rm["k1_ms"], 3), "pipeline_ms": round(rm["pipeline_ms"], 3),
            "k_match_find_frac_hbm": round((n_in + n_out) / (rm["k1_ms"] / 1e3) / 1e9 / HBM_PEAK_GBPS, 5),
            "check": "plain == batched chunk by chunk; whole output decodes; first 2 chunks == oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-mixed", action="store_true", help="leave out the mixed-size comparison")
    ap.add_argument("--corpus", nargs="+", default=[], metavar="KIND:MIB",
                    help="also run the plain and batched forms on KIND (text, binmix, log) of MIB MiB in "
                         "2 MiB chunks, e.g. text:16 log:1024 (extra JSON lines)")
    ap.add_argument("--no-serve", action="store_true", help="leave out the serving form")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-submitters", type=int, default=4, help="threads sharing the arrival schedule")
    ap.add_argument("--serve-load", type=float, nargs="+", default=[8.0],
                    help="offered loads (GB/s) of the serving runs; the first goes into the headline line, "
                         "the others are extra JSON lines")
    ap.add_argument("--hybrid", action="store_true",
                    help="only time `--backend auto` (the C++ HybridBackend, host buffers) on the binmix corpus "
                         "for several GPU permit counts, one JSON line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="leave out the host-buffer end-to-end rates")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain elementwise copy (y = x + 1, uint8) of MIB MiB three times; "
                         "profiling its FETCH_SIZE / WRITE_SIZE gives the counter-to-bytes factor")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (--gpus {args.gpus} requested)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        nb = args.calibrate_copy * MiB
        x = torch.zeros(nb, dtype=torch.uint8, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 1, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "elementwise uint8 copy y = x + 1", "read_bytes": nb,
                          "write_bytes": nb, "launches": 3}), flush=True)
        return

    if args.hybrid:
        print(json.dumps(hybrid_sweep()), flush=True)
        return

    log("generating binmix 256 MiB ...")
    n = 256 * MiB
    data = gpzip_amd.synth("binmix", n, 1)
    ctx = gpzip_amd.Context(device=0)
    log("device:", ctx.label())
    sizes = [2 * MiB] * (n // (2 * MiB))
    r = plain_and_batched(torch, ctx, data, sizes, args.steps, args.warmup)

    members = r["members"]
    alg_bytes = r["n_in"] + r["n_out"]  # per launch: sum over members of (n + c)
    k1_s = r["k1_ms"] / 1e3
    achieved = alg_bytes / k1_s / 1e9
    tune = r["tune"] or {}
    # the per-shape figures are of the two-workgroups-per-CU variant
    pc = profiled_counters(r["n_in"], tune.get("k1_grid", 0) if tune.get("k1_variant", 1) in (0, 1) else 0)
    line = {
        "metric": "input GB/s on 1 MI355X (% HBM roofline) + ratio; p50/p99 request latency",
        "value": round(r["plain"]["gbps"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["plain"]["s_per_step"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "binmix 256 MiB, 2 MiB chunks, level 5, plain per-chunk GPU form",
                   "input_bytes": r["n_in"], "chunk_bytes": 2 * MiB, "member_bytes": 131072,
                   "members": members, "seed": 1, "inputs": "resident in HBM"},
        "ratio": round(r["ratio"], 5),
        "hbm_fraction_of_value": round(r["plain"]["gbps"] / HBM_PEAK_GBPS, 5),
        "batched": {"value": round(r["batched"]["gbps"], 3), "unit": "GB/s",
                    "ms_per_step": round(r["batched"]["s_per_step"] * 1e3, 3),
                    "vs_plain": round(r["batched"]["gbps"] / r["plain"]["gbps"], 3), "autotune": r["tune"]},
        "roofline": {"bound": "hbm", "kernel": "k_match_find", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                     "traffic": pc.get("traffic_bytes_per_launch"), "alg_bytes_per_launch": alg_bytes,
                     "valu_issue_frac": pc.get("valu_issue_frac"), "valu_lane_frac": pc.get("valu_lane_frac"),
                     "traffic_source": ("profiles/k_match_find_binmix256_traffic.md (rocprofv3 --pmc, "
                                        "same kernel sources), at " + pc["traffic_shape"]) if pc else
                                       "null: the committed counter profile is of other kernel sources",
                     "avg_launch_ms": round(r["k1_ms"], 4), "seg_table_ms": round(r["k0_ms"], 4),
                     "emit_ms": round(r["emit_ms"], 4),
                     "pipeline_ms": round(r["pipeline_ms"], 4)},
    }
    if not args.no_e2e:
        log("end to end (host buffers) ...")
        line["end_to_end"] = end_to_end(ctx, data, sizes, r["chunks_out"])
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(data, 2 * MiB)
        line["cpu_baseline"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in cb.items()}
        port = cpu_port(data, 64 * MiB, 2 * MiB)
        line["cpu_port"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in port.items()}
    del r
    if not args.no_mixed:
        log("mixed sizes: 4096 chunks, 4 KiB - 4 MiB ...")
        line["mixed"] = mixed_sizes(torch, ctx, max(3, args.steps // 2))
        torch.cuda.empty_cache()
    extra = []
    if not args.no_serve:
        for i, load in enumerate(args.serve_load):
            log(f"serving: {args.serve_requests} requests at {load} GB/s offered ...")
            rs = serve(ctx, args.serve_requests, load, submitters=args.serve_submitters)
            rs = {"workload": f"{args.serve_requests} requests, Poisson arrivals at {load} GB/s offered, sizes "
                              "log-normal (median 256 KiB, sigma 2.0, 1 KiB-64 MiB), content 1/3 each text / "
                              "binmix / log, from host buffers (PCIe both ways), 64 GB device arena", **rs}
            if i == 0:
                line["serving"] = rs
            else:
                extra.append(rs)
    print(json.dumps(line), flush=True)
    for rs in extra:
        print(json.dumps(rs), flush=True)

    for spec in args.corpus:
        kind, mib = spec.split(":")
        n = int(mib) << 20
        log(f"corpus {kind}, {mib} MiB ...")
        cdata = gpzip_amd.synth(kind, n, 1)
        csizes = [min(2 * MiB, n - o) for o in range(0, n, 2 * MiB)]
        rc = plain_and_batched(torch, ctx, cdata, csizes, max(3, args.steps // 2), 1, check_oracle=min(8, len(csizes)))
        print(json.dumps({"workload": f"{kind} {mib} MiB, 2 MiB chunks", "input_bytes": rc["n_in"],
                          "plain_gbps": round(rc["plain"]["gbps"], 3), "batched_gbps": round(rc["batched"]["gbps"], 3),
                          "ratio": round(rc["ratio"], 5), "autotune": rc["tune"]}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
