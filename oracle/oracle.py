"""ctypes binding of the oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg, as the answer the GPU results are compared with.  Nothing in
graph_amd/ imports it.

Functions are thin wrappers over ``cosmos_oracle.c``; see that file for the
reference file:line each one restates.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_PATH = HERE / "liboracle.so"
_lib = None

F = C.POINTER(C.c_float)
D = C.POINTER(C.c_double)
U = C.POINTER(C.c_uint32)
I = C.POINTER(C.c_int32)
B = C.POINTER(C.c_uint8)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return _PATH


def lib():
    global _lib
    if _lib is None:
        if not _PATH.exists():
            build()
        l = C.CDLL(str(_PATH))
        l.or_lowbias32.argtypes = [C.c_uint32]
        l.or_lowbias32.restype = C.c_uint32
        l.or_ticket.argtypes = [C.c_uint32, C.c_float]
        l.or_ticket.restype = C.c_float
        l.or_geometry.argtypes = [C.c_uint32, U, U]
        l.or_cell.argtypes = [C.c_float, C.c_float, C.c_int32]
        l.or_cell.restype = C.c_int32
        l.or_cells.argtypes = [C.c_uint32, F, B, C.c_float, C.c_int32, C.c_float, C.c_float, I]
        l.or_cell_size.argtypes = [C.c_float, C.c_uint32]
        l.or_cell_size.restype = C.c_float
        l.or_slots.argtypes = [C.c_uint32, F, B, C.c_float, C.c_float, I]
        l.or_link_csr.argtypes = [C.c_uint32, C.c_uint32, F, C.c_int, U, U, U, U]
        l.or_link_csr.restype = C.c_uint32
        l.or_link_props.argtypes = [C.c_uint32, C.c_uint32, U, U, U, U, F, F, F]
        l.or_gravity.argtypes = [C.c_uint32, D, C.c_double, C.c_double, C.c_double, D]
        l.or_center.argtypes = [C.c_uint32, D, B, C.c_double, C.c_double, D]
        l.or_mouse.argtypes = [C.c_uint32, D, C.c_double, C.c_double, C.c_double, D]
        l.or_manybody.argtypes = [C.c_uint32, D, B, F, C.c_float, C.c_double, C.c_double, C.c_float,
                                  U, C.c_uint32, D]
        l.or_link.argtypes = [C.c_uint32, D, B, U, U, F, F, F, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_double, D]
        l.or_collision_grid.argtypes = [C.c_float, C.c_double, C.c_double, C.c_double, U, F]
        l.or_collision.argtypes = [C.c_uint32, D, B, F, C.c_float, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, D]
        l.or_integrate.argtypes = [C.c_uint32, D, D, B, B, C.c_double, C.c_double]
        l.or_cluster.argtypes = [C.c_uint32, D, B, I, C.c_uint32, D, F, C.c_double, C.c_double, D]
        l.or_alpha_step.argtypes = [C.c_double, C.c_double]
        l.or_alpha_step.restype = C.c_double
        l.or_set_threads.argtypes = [C.c_int]
        l.or_get_threads.restype = C.c_int
        _lib = l
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u8(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint8)


def absent_mask(xy: np.ndarray) -> np.ndarray:
    xy = np.asarray(xy).reshape(-1, 2)
    return np.isnan(xy).any(axis=1).astype(np.uint8)


def lowbias32(h: int) -> int:
    return lib().or_lowbias32(h)


def ticket(index: int, seed: float) -> float:
    return lib().or_ticket(index, seed)


def geometry(n: int) -> tuple[int, int]:
    gf, L = C.c_uint32(), C.c_uint32()
    lib().or_geometry(n, C.byref(gf), C.byref(L))
    return gf.value, L.value


def cell(p: float, h: float, g: int) -> int:
    return lib().or_cell(p, h, g)


def cell_size(S: float, g: int) -> float:
    return lib().or_cell_size(S, g)


def cells(xy, h: float, g: int, absent=None, offset=(0.0, 0.0)) -> np.ndarray:
    """Per-point (cx, cy) at a grid of g cells of size h (or_cells); -1 for absent."""
    xy = _f32(xy).reshape(-1)
    n = xy.size // 2
    out = np.empty(2 * max(n, 1), np.int32)
    lib().or_cells(n, _p(xy, F), _p(_u8(absent), B), h, g, offset[0], offset[1], _p(out, I))
    return out[: 2 * n].reshape(-1, 2)


def slots(xy, S: float, seed: float, absent=None) -> np.ndarray:
    xy = _f32(xy).reshape(-1)
    n = xy.size // 2
    gf, _ = geometry(n)
    out = np.empty(gf * gf * 8, np.int32)
    lib().or_slots(n, _p(xy, F), _p(_u8(absent), B), S, seed, _p(out, I))
    return out


def link_csr(n: int, links, direction: int):
    links = _f32(links).reshape(-1)
    m = links.size // 2
    off = np.empty(n + 1, np.uint32)
    nbr = np.empty(max(m, 1), np.uint32)
    lof = np.empty(max(m, 1), np.uint32)
    deg = np.empty(max(n, 1), np.uint32)
    E = lib().or_link_csr(n, m, _p(links, F), direction, _p(off, U), _p(nbr, U), _p(lof, U), _p(deg, U))
    return off, nbr[:E], lof[:E], deg[:n]


def link_props(n: int, off, nbr, lof, deg, strength=None):
    E = nbr.size
    bias = np.empty(max(E, 1), np.float32)
    st = np.empty(max(E, 1), np.float32)
    s_in = _f32(strength)
    lib().or_link_props(n, E, _p(off, U), _p(nbr, U), _p(lof, U), _p(deg, U), _p(s_in, F), _p(bias, F),
                        _p(st, F))
    return bias[:E], st[:E]


def gravity(xy, S, alpha, g):
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    v = np.zeros(2 * n)
    lib().or_gravity(n, _p(pos, D), S, alpha, g, _p(v, D))
    return v


def center(xy, alpha, c, absent=None):
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    v = np.zeros(2 * n)
    lib().or_center(n, _p(pos, D), _p(_u8(absent), B), alpha, c, _p(v, D))
    return v


def cluster(xy, clusters, alpha, coef, cluster_pos=None, strength=None, absent=None):
    """Cluster pull (or_cluster); clusters: per-point index, negative = none;
    cluster_pos: [x, y] per cluster (a negative coordinate = centroid) or None."""
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    cl = np.ascontiguousarray(clusters, dtype=np.int32)
    ncl = int(max(0, cl.max(initial=0))) + 1
    cp = None
    if cluster_pos is not None:
        cp = np.full(2 * ncl, -1.0)
        src = np.asarray(cluster_pos, dtype=np.float64).reshape(-1)[: 2 * ncl]
        cp[: src.size] = src
    s = None if strength is None else _f32(strength)
    v = np.zeros(2 * n)
    lib().or_cluster(n, _p(pos, D), _p(_u8(absent), B), _p(cl, I), ncl, _p(cp, D), _p(s, F), alpha, coef,
                     _p(v, D))
    return v


def mouse(xy, mx, my, k):
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    v = np.zeros(2 * n)
    lib().or_mouse(n, _p(pos, D), mx, my, k, _p(v, D))
    return v


def manybody(xy, S, alpha, k, seed, jitter=None, absent=None, eval_idx=None):
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    v = np.zeros(2 * n)
    ev = None if eval_idx is None else np.ascontiguousarray(eval_idx, dtype=np.uint32)
    lib().or_manybody(n, _p(pos, D), _p(_u8(absent), B), _p(_f32(jitter), F), S, alpha, k, seed,
                      _p(ev, U), 0 if ev is None else ev.size, _p(v, D))
    return v


def link(xy, links, direction, rand_by_link, alpha, spring=1.0, distance=10.0, lo=1.0, hi=1.2,
         strength=None, absent=None):
    """Velocity of one link pass; rand_by_link[l] is link l's draw for this direction."""
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    off, nbr, lof, deg = link_csr(n, links, direction)
    bias, st = link_props(n, off, nbr, lof, deg, strength)
    r = _f32(np.asarray(rand_by_link, dtype=np.float32)[lof]) if nbr.size else np.zeros(1, np.float32)
    v = np.zeros(2 * n)
    lib().or_link(n, _p(pos, D), _p(_u8(absent), B), _p(off, U), _p(nbr, U), _p(r, F), _p(bias, F),
                  _p(st, F), spring, distance, lo, hi, alpha, _p(v, D))
    return v


def collision_grid(S, radius, padding, max_size):
    G, cs = C.c_uint32(), C.c_float()
    lib().or_collision_grid(S, radius, padding, max_size, C.byref(G), C.byref(cs))
    return G.value, cs.value


def collision(xy, sizes, S, alpha, strength, radius=0.0, padding=0.0, max_size=None, absent=None):
    pos = _f64(xy).reshape(-1)
    n = pos.size // 2
    sizes = _f32(sizes)
    if max_size is None:
        max_size = float(max(4.0, float(np.max(sizes)) if n else 4.0))
    v = np.zeros(2 * n)
    lib().or_collision(n, _p(pos, D), _p(_u8(absent), B), _p(sizes, F), S, alpha, strength, radius, padding,
                       max_size, _p(v, D))
    return v


def integrate(xy, v, friction, S, pinned=None, absent=None):
    pos = _f64(xy).reshape(-1).copy()
    n = pos.size // 2
    vv = _f64(v)
    lib().or_integrate(n, _p(pos, D), _p(vv, D), _p(_u8(pinned), B), _p(_u8(absent), B), friction, S)
    return pos


def alpha_step(alpha: float, decay: float) -> float:
    return lib().or_alpha_step(alpha, decay)


class Links:
    """Both directions' CSR with per-entry draws, built once (GraphData/index.ts:525-568,
    ForceLink/index.ts:40-83)."""

    def __init__(self, n, links, rand_in, rand_out, strength=None):
        self.n = n
        self.m = np.asarray(links).size // 2
        self.dirs = []
        for d, rnd in ((0, rand_in), (1, rand_out)):
            off, nbr, lof, deg = link_csr(n, links, d)
            bias, st = link_props(n, off, nbr, lof, deg, strength)
            r = np.asarray(rnd, np.float32)[lof] if nbr.size else np.zeros(0, np.float32)
            self.dirs.append((off, nbr, np.ascontiguousarray(r), bias, st))

    def force(self, pos, d, alpha, spring, distance, lo, hi, absent=None):
        off, nbr, r, bias, st = self.dirs[d]
        pos = _f64(pos).reshape(-1)
        v = np.zeros(2 * self.n)
        rr = r if r.size else np.zeros(1, np.float32)
        nb = nbr if nbr.size else np.zeros(1, np.uint32)
        bb = bias if bias.size else np.zeros(1, np.float32)
        ss = st if st.size else np.zeros(1, np.float32)
        lib().or_link(self.n, _p(pos, D), _p(_u8(absent), B), _p(off, U), _p(nb, U), _p(rr, F), _p(bb, F),
                      _p(ss, F), spring, distance, lo, hi, alpha, _p(v, D))
        return v


def tick(pos, alpha, seed, p, links=None, jitter=None, absent=None, pinned=None, sizes=None,
         clusters=None):
    """One FP64 tick in the order of runSimulationStep (src/index.ts:2031-2087);
    p is a dict of the simulation* keys (space_size, gravity, center, repulsion,
    link_spring, link_distance, lo, hi, friction, decay, collision, collision_radius,
    collision_padding, point_default_size).  Returns (new positions, new alpha)."""
    S = p["space_size"]
    fr = p["friction"]
    pos = _f64(pos).reshape(-1).copy()
    if p.get("gravity", 0):
        pos = integrate(pos, gravity(pos, S, alpha, p["gravity"]), fr, S, pinned, absent)
    if p.get("center", 0):
        pos = integrate(pos, center(pos, alpha, p["center"], absent), fr, S, pinned, absent)
    pos = integrate(pos, manybody(pos, S, alpha, p["repulsion"], seed, jitter, absent), fr, S, pinned, absent)
    if links is not None and links.m > 0:
        for d in (0, 1):
            v = links.force(pos, d, alpha, p["link_spring"], p["link_distance"], p["lo"], p["hi"], absent)
            pos = integrate(pos, v, fr, S, pinned, absent)
    if clusters is not None:  # {"index": per point, "positions": or None, "strength": or None}
        v = cluster(pos, clusters["index"], alpha, p.get("cluster", 0.1), clusters.get("positions"),
                    clusters.get("strength"), absent)
        pos = integrate(pos, v, fr, S, pinned, absent)
    if p.get("collision", 0):
        v = collision(pos, sizes, S, alpha, p["collision"], p.get("collision_radius", 0.0),
                      p.get("collision_padding", 0.0), absent=absent)
        pos = integrate(pos, v, fr, S, pinned, absent)
    return pos, alpha_step(alpha, p["decay"])


def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (results do not depend on it);
    0 restores the OpenMP default (OMP_NUM_THREADS)."""
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())
