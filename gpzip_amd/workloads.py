"""Seeded synthetic workloads of the benchmark configurations, shared by
bench.py and the tests (content from libgpzip_synth.so).

    mixed_sizes_workload()   4,096 chunks, sizes log-uniform over 4 KiB-4 MiB,
                             content 1/3 each text / binmix / log
    serving_requests(n)      request sizes log-normal (median 256 KiB,
                             sigma 2.0, clamped to 1 KiB-64 MiB), content as
                             above, Poisson arrival times at a given load
"""
from __future__ import annotations

import numpy as np

from . import lognormal_sizes, loguniform_sizes, synth, uniform

KINDS = ["text", "binmix", "log"]


def mixed_sizes_workload(seed: int = 3, count: int = 4096, lo: int = 4096, hi: int = 4 << 20):
    """(bytes, sizes): `count` chunks laid end to end.  Content per chunk is
    drawn 1/3 each from the three generators, as slices of three shared
    pools so that 2.5 GiB need not be generated chunk by chunk."""
    sizes = loguniform_sizes(seed, count, lo, hi).astype(np.int64)
    choose = uniform(seed + 1, count)
    pool_n = int(sizes.max()) + (256 << 20)
    pools = [synth(k, pool_n, 10 + i) for i, k in enumerate(KINDS)]
    offs = (uniform(seed + 2, count) * (pool_n - sizes)).astype(np.int64) // 16 * 16
    parts = [pools[int(c * 3)][o:o + s] for c, o, s in zip(choose, offs, sizes)]
    return np.concatenate(parts), [int(s) for s in sizes]


class ServingRequests:
    """Request i: bytes pools[kind[i]][off[i]:off[i] + size[i]], arriving at
    arrive[i] seconds (Poisson process whose mean rate in bytes is load_gbps)."""

    def __init__(self, n_requests: int, load_gbps: float, seed: int = 4, pool_mib: int = 256):
        self.sizes = lognormal_sizes(seed, n_requests, 256 * 1024, 2.0, 1024, 64 << 20).astype(np.int64)
        pool_n = max(int(self.sizes.max()), 64 << 20) + (pool_mib << 20) - (64 << 20)
        self.pools = [synth(k, pool_n, 20 + i) for i, k in enumerate(KINDS)]
        self.kind = (uniform(seed + 1, n_requests) * 3).astype(np.int64)
        self.offs = (uniform(seed + 2, n_requests) * (pool_n - self.sizes)).astype(np.int64)
        gaps = -np.log(1.0 - uniform(seed + 3, n_requests)) * (self.sizes.mean() / (load_gbps * 1e9))
        self.arrive = np.cumsum(gaps)
        self.n = n_requests

    def data(self, i: int) -> np.ndarray:
        return self.pools[int(self.kind[i])][int(self.offs[i]):int(self.offs[i]) + int(self.sizes[i])]
