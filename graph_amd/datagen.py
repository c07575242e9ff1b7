"""Seeded synthetic inputs (``graph_amd/libcgdata.so``): graphs, positions,
radii, jitter and uniform draws for the force-simulation workloads."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PATH = Path(__file__).resolve().parent / "libcgdata.so"
_lib = None


def _l():
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise ImportError(f"{_PATH} is missing: run `python -m graph_amd.build`")
        l = C.CDLL(str(_PATH))
        F = C.POINTER(C.c_float)
        l.cgd_uniform_square.argtypes = [C.c_uint32, C.c_float, C.c_uint64, F]
        l.cgd_uniform_disk.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_uint64, F]
        l.cgd_barabasi_albert.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, F]
        l.cgd_barabasi_albert.restype = C.c_uint64
        l.cgd_barabasi_albert_count.argtypes = [C.c_uint32, C.c_uint32]
        l.cgd_barabasi_albert_count.restype = C.c_uint64
        l.cgd_chung_lu.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_uint64, C.c_uint64, F]
        l.cgd_sbm.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64, F, C.c_uint64]
        l.cgd_sbm.restype = C.c_uint64
        l.cgd_lognormal.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_uint64, F]
        l.cgd_uniform.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_uint64, F]
        l.cgd_jitter.argtypes = [C.c_uint32, C.c_uint64, F]
        _lib = l
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uniform_square(n: int, S: float, seed: int) -> np.ndarray:
    out = np.empty(2 * n, np.float32)
    _l().cgd_uniform_square(n, S, seed, _fp(out))
    return out


def uniform_disk(n: int, cx: float, cy: float, R: float, seed: int) -> np.ndarray:
    out = np.empty(2 * n, np.float32)
    _l().cgd_uniform_disk(n, cx, cy, R, seed, _fp(out))
    return out


def barabasi_albert(n: int, m: int, seed: int) -> np.ndarray:
    cnt = _l().cgd_barabasi_albert_count(n, m)
    out = np.empty(2 * max(cnt, 1), np.float32)
    got = _l().cgd_barabasi_albert(n, m, seed, _fp(out))
    return out[: 2 * got]


def chung_lu(n: int, gamma: float, kmin: float, m: int, seed: int) -> np.ndarray:
    out = np.empty(2 * max(m, 1), np.float32)
    _l().cgd_chung_lu(n, gamma, kmin, m, seed, _fp(out))
    return out[: 2 * m]


def sbm(n: int, blocks: int, zipf_s: float, deg_in: float, cross_frac: float, seed: int) -> np.ndarray:
    cap = int(n * deg_in / 2 / max(1e-9, 1 - cross_frac) * 1.1) + 1024
    out = np.empty(2 * cap, np.float32)
    got = _l().cgd_sbm(n, blocks, zipf_s, deg_in, cross_frac, seed, _fp(out), cap)
    return out[: 2 * got]


def lognormal(n: int, mu: float, sigma: float, seed: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    _l().cgd_lognormal(n, mu, sigma, seed, _fp(out))
    return out


def uniform(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    _l().cgd_uniform(n, lo, hi, seed, _fp(out))
    return out


def jitter(n: int, seed: int) -> np.ndarray:
    out = np.empty(2 * n, np.float32)
    _l().cgd_jitter(n, seed, _fp(out))
    return out
