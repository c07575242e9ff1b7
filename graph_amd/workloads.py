"""The benchmark workloads as seeded synthetic inputs.

Each builder returns a dict with positions, links, per-link rest-length draws
(IN and OUT), jitter, the simulation parameters and the space size.  Defaults
not named by a workload take the reference defaults (src/variables.ts:68-83).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib, datagen

# shared simulation parameters of every workload
BASE = dict(repulsion=0.5, gravity=0.25, friction=0.85, link_spring=1.0, link_distance=10.0, decay=5000.0)


def sim_params(S: float, **over) -> _lib.CgSimParams:
    p = _lib.default_params()
    p.space_size = S
    kw = dict(BASE)
    kw.update(over)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _draws(m: int, seed: int):
    return datagen.uniform(m, 0.0, 1.0, seed * 7 + 1), datagen.uniform(m, 0.0, 1.0, seed * 7 + 2)


def cfg0() -> dict:
    """10,000 points, Barabasi-Albert m=2 (seed 42), uniform in [0,4096)^2, S = 4096."""
    n, S = 10_000, 4096.0
    links = datagen.barabasi_albert(n, 2, 42)
    r_in, r_out = _draws(links.size // 2, 42)
    return dict(name="cfg0: BA n=10k m=2, uniform square, S=4096", n=n, S=S,
                xy=datagen.uniform_square(n, S, 42), links=links, r_in=r_in, r_out=r_out,
                jitter=datagen.jitter(n, 42), params=sim_params(S), ticks=300)


def cfg1(n: int = 1_000_000, m: int = 5_000_000) -> dict:
    """1M points, ~5M Chung-Lu power-law links (exponent 2.1, min degree 1, seed 7),
    uniform in a disk of radius 2048 centred in S = 8192."""
    S = 8192.0
    links = datagen.chung_lu(n, 2.1, 1.0, m, 7)
    r_in, r_out = _draws(m, 7)
    return dict(name="cfg1: Chung-Lu n=1M m=5M gamma=2.1, disk r=2048, S=8192", n=n, S=S,
                xy=datagen.uniform_disk(n, S / 2, S / 2, 2048.0, 7), links=links, r_in=r_in, r_out=r_out,
                jitter=datagen.jitter(n, 7), params=sim_params(S), ticks=1000)


def cfg2(n: int = 4_000_000) -> dict:
    """4M points, SBM with 1000 Zipf(1.1) blocks, p_in*n_b = 5, 10% cross links
    (seed 11), log-normal radii (mu = ln 2, sigma = 0.5), collision on, S = 16384."""
    S = 16384.0
    links = datagen.sbm(n, 1000, 1.1, 5.0, 0.1, 11)
    r_in, r_out = _draws(links.size // 2, 11)
    sizes = 2.0 * datagen.lognormal(n, math.log(2.0), 0.5, 11)
    return dict(name="cfg2: SBM n=4M 1000 blocks, collision, S=16384", n=n, S=S,
                xy=datagen.uniform_square(n, S, 11), links=links, r_in=r_in, r_out=r_out,
                jitter=datagen.jitter(n, 11), sizes=sizes.astype(np.float32),
                params=sim_params(S, collision=1.0), ticks=500)


def next_pow2(x: float) -> int:
    return 1 << max(0, math.ceil(math.log2(x)))


def cfg3_graph(i: int, n: int, m_ba: int) -> dict:
    S = float(next_pow2(4 * math.sqrt(n) * 8))
    links = datagen.barabasi_albert(n, m_ba, 23_000 + i)
    r_in, r_out = _draws(links.size // 2, 23_000 + i)
    return dict(n=n, S=S, xy=datagen.uniform_square(n, S, 23_000 + i), links=links, r_in=r_in, r_out=r_out,
                jitter=datagen.jitter(n, 23_000 + i), params=sim_params(S))


def cfg3_sizes(count: int = 2048, seed: int = 23):
    """Point counts log-uniform in [1000, 200000] and BA m uniform in {1..4}."""
    u = datagen.uniform(2 * count, 0.0, 1.0, seed)
    n = np.exp(np.log(1000) + u[:count] * (np.log(200_000) - np.log(1000))).astype(np.int64)
    m = 1 + np.minimum((u[count:] * 4).astype(np.int64), 3)
    return n, m


def cfg3_plain_sample(count: int, sample: int) -> list[int]:
    """Graph indices of the plain one-graph-at-a-time timing sample: every
    (count // sample)-th graph.  The point counts are i.i.d. log-uniform, so an
    evenly strided sample is a uniform random sample of the batch."""
    stride = max(1, count // max(1, sample))
    return list(range(0, count, stride))[:sample]


def cfg4_jobs(rate: float = 40.0, duration: float = 120.0, seed: int = 31, n_lo: int = 500,
              n_hi: int = 500_000, collision_frac: float = 0.1):
    """Layout requests: Poisson arrivals at `rate`/s over `duration` s; point
    counts log-uniform in [n_lo, n_hi]; links 2-8x points from Chung-Lu with
    exponent uniform in [2.0, 2.8]; a `collision_frac` share with collision on
    and log-normal radii (mu = ln 2, sigma = 0.5); positions uniform in a
    spaceSize of next_pow2(4 sqrt(n) 8); parameters as in the 10k workload."""
    from .serving import Job

    # draws sized for the full 120 s window whatever `duration` is, so a shorter
    # window is the first part of the same request stream
    k = int(rate * max(duration, 120.0) * 3) + 16
    u = datagen.uniform(6 * k, 0.0, 1.0, seed).astype(np.float64).reshape(6, k)
    gaps = -np.log(1.0 - u[0]) / rate
    arrivals = np.cumsum(gaps)
    jobs = []
    for i in range(k):
        if arrivals[i] > duration:
            break
        n = int(math.exp(math.log(n_lo) + u[1, i] * (math.log(n_hi) - math.log(n_lo))))
        ratio = 2.0 + 6.0 * u[2, i]
        gamma = 2.0 + 0.8 * u[3, i]
        collide = u[4, i] < collision_frac
        s = seed * 100_003 + i
        S = float(next_pow2(4 * math.sqrt(n) * 8))
        m = int(ratio * n)
        links = datagen.chung_lu(n, gamma, 1.0, m, s)
        r_in, r_out = _draws(m, s)
        sizes = (2.0 * datagen.lognormal(n, math.log(2.0), 0.5, s)).astype(np.float32) if collide else None
        p = sim_params(S, collision=1.0) if collide else sim_params(S)
        jobs.append(Job(jid=i, arrival=float(arrivals[i]), xy=datagen.uniform_square(n, S, s), links=links,
                        params=p, r_in=r_in, r_out=r_out, jitter=datagen.jitter(n, s), sizes=sizes))
    return jobs
