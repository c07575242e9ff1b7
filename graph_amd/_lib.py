"""ctypes binding of ``include/cosmosgl.h`` (``graph_amd/libcosmosgl.so``).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
HEADER = PKG.parent / "include" / "cosmosgl.h"
LIB_PATH = PKG / "libcosmosgl.so"

CG_OK = 0
CG_ERR_INVALID = -1
CG_ERR_DEVICE = -2
CG_ERR_BUDGET = -3
CG_ERR_STATE = -4

FORCE_GRAVITY = 0
FORCE_CENTER = 1
FORCE_MOUSE = 2
FORCE_MANYBODY = 3
FORCE_LINK_IN = 4
FORCE_LINK_OUT = 5
FORCE_COLLISION = 6
FORCE_CLUSTER = 7


class CgCreateParams(C.Structure):
    _fields_ = [("device", C.c_int32), ("hbm_budget", C.c_uint64)]


class CgSimParams(C.Structure):
    _fields_ = [
        ("space_size", C.c_float),
        ("decay", C.c_float),
        ("gravity", C.c_float),
        ("center", C.c_float),
        ("repulsion", C.c_float),
        ("link_spring", C.c_float),
        ("link_distance", C.c_float),
        ("link_dist_random_variation_range", C.c_float * 2),
        ("repulsion_from_mouse", C.c_float),
        ("friction", C.c_float),
        ("collision", C.c_float),
        ("collision_radius", C.c_float),
        ("collision_padding", C.c_float),
        ("point_default_size", C.c_float),
        ("enable_simulation", C.c_int32),
        ("run_forces", C.c_int32),
        ("mouse_repulsion", C.c_int32),
        ("mouse_position", C.c_float * 2),
        ("cluster", C.c_float),
    ]


class CgTuning(C.Structure):
    _fields_ = [
        ("graph_ticks", C.c_uint32),
        ("cell_items", C.c_uint32),
        ("force_tile", C.c_uint32),
        ("link_lane_max", C.c_uint32),
        ("link_group_max", C.c_uint32),
        ("link_hub_entries", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class CosmosGLError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cosmosgl error {code}: {msg}")
        self.code = code


_lib = None

_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_I = C.POINTER(C.c_int32)
_D = C.POINTER(C.c_double)
_CTX = C.c_void_p

_SIGNATURES = {
    "cg_default_params": (None, [C.POINTER(CgSimParams)]),
    "cg_default_tuning": (None, [C.POINTER(CgTuning)]),
    "cg_set_tuning": (C.c_int, [_CTX, C.POINTER(CgTuning)]),
    "cg_get_tuning": (C.c_int, [_CTX, C.POINTER(CgTuning)]),
    "cg_create": (C.c_int, [C.POINTER(_CTX), C.POINTER(CgCreateParams)]),
    "cg_destroy": (C.c_int, [_CTX]),
    "cg_set_positions": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_set_links": (C.c_int, [_CTX, _F, C.c_uint32, _F, _F, _F]),
    "cg_set_jitter": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_set_point_sizes": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_set_pinned": (C.c_int, [_CTX, _U, C.c_uint32]),
    "cg_set_point_clusters": (C.c_int, [_CTX, _I, C.c_uint32]),
    "cg_set_cluster_positions": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_set_point_cluster_strength": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_graph_set_point_clusters": (C.c_int, [_CTX, C.c_uint32, _I, C.c_uint32]),
    "cg_graph_set_cluster_positions": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_graph_set_point_cluster_strength": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_set_alpha": (C.c_int, [_CTX, C.c_double]),
    "cg_get_alpha": (C.c_int, [_CTX, _D]),
    "cg_step": (C.c_int, [_CTX, C.POINTER(CgSimParams), C.c_uint32, _F, _D]),
    "cg_get_positions": (C.c_int, [_CTX, _F, C.c_uint32]),
    "cg_last_error": (C.c_char_p, [_CTX]),
    "cg_eval_force": (C.c_int, [_CTX, C.POINTER(CgSimParams), C.c_int32, C.c_float, C.c_float, _F]),
    "cg_debug_slots": (C.c_int, [_CTX, _I, C.c_uint32, _U, _U]),
    "cg_debug_link_csr": (C.c_int, [_CTX, C.c_int32, _U, _U, _F, _F, _U, C.c_uint32, _U]),
    "cg_debug_cells": (C.c_int, [_CTX, C.c_uint32, _U, _F, _U, C.c_uint32]),
    "cg_debug_collision_cells": (C.c_int, [_CTX, _U, _U, _U, _U, C.c_uint32, _U, _U]),
    "cg_device_bytes": (C.c_uint64, [_CTX]),
    "cg_set_profiling": (C.c_int, [_CTX, C.c_uint32]),
    "cg_debug_link_stats": (C.c_int, [_CTX, _U]),
    "cg_get_profile": (C.c_int, [_CTX, _D, _U, C.c_uint32]),
    "cg_graph_add": (C.c_int, [_CTX, _U]),
    "cg_graph_remove": (C.c_int, [_CTX, C.c_uint32]),
    "cg_graph_set_params": (C.c_int, [_CTX, C.c_uint32, C.POINTER(CgSimParams)]),
    "cg_graph_set_seed": (C.c_int, [_CTX, C.c_uint32, C.c_uint64]),
    "cg_graph_set_positions": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_graph_set_links": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32, _F, _F, _F]),
    "cg_graph_set_jitter": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_graph_set_point_sizes": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_graph_set_pinned": (C.c_int, [_CTX, C.c_uint32, _U, C.c_uint32]),
    "cg_graph_set_alpha": (C.c_int, [_CTX, C.c_uint32, C.c_double]),
    "cg_graph_get_alpha": (C.c_int, [_CTX, C.c_uint32, _D]),
    "cg_graph_get_positions": (C.c_int, [_CTX, C.c_uint32, _F, C.c_uint32]),
    "cg_step_graphs": (C.c_int, [_CTX, _U, C.c_uint32, C.c_uint32]),
    "cg_sync": (C.c_int, [_CTX]),
    "cg_prepare": (C.c_int, [_CTX, C.POINTER(CgSimParams)]),
    "cg_graph_eval_force": (C.c_int, [_CTX, C.c_uint32, C.c_int32, C.c_float, C.c_float, _F]),
    "cg_eval_manybody_noalloc": (C.c_int, [_CTX, C.POINTER(CgSimParams), C.c_float, C.c_float, _F, _F,
                                           C.c_uint32, _F]),
    "cg_set_capture": (C.c_int, [_CTX, C.c_int32]),
    "cg_drag_points": (C.c_int, [_CTX, _U, _F, C.c_uint32]),
    "cg_get_positions_at": (C.c_int, [_CTX, _U, C.c_uint32, _F]),
    "cg_graph_drag_points": (C.c_int, [_CTX, C.c_uint32, _U, _F, C.c_uint32]),
    "cg_graph_get_positions_at": (C.c_int, [_CTX, C.c_uint32, _U, C.c_uint32, _F]),
}

STAGES = ["mouse", "gravity", "center", "mb_sort", "mb_cells", "mb_pyramid", "mb_force", "link_in",
          "link_out", "collision", "cluster"]


def header_symbols() -> list[str]:
    """Function names declared in include/cosmosgl.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(cg_\w+)\s*\(", text, re.M)))


def lib() -> C.CDLL:
    """Loads libcosmosgl.so (once)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m graph_amd.build` "
                "(the simulation has no CPU fallback)"
            )
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def uptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def default_tuning() -> dict:
    t = CgTuning()
    lib().cg_default_tuning(C.byref(t))
    return t.as_dict()


def default_params() -> CgSimParams:
    p = CgSimParams()
    lib().cg_default_params(C.byref(p))
    return p


class Context:
    """A set of simulated graphs on one GPU (owner of a ``cg_ctx``).

    The single-graph methods (set_positions ... get_positions) act on graph 0;
    the ``graph_*`` methods manage further graphs packed beside it."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._l = lib()
        self._ctx = _CTX()
        cp = CgCreateParams(device, hbm_budget)
        rc = self._l.cg_create(C.byref(self._ctx), C.byref(cp))
        if rc != CG_OK:
            raise CosmosGLError(rc, "cg_create failed")
        self.n = 0
        self._sizes: dict[int, int] = {}  # point count of each graph_* graph

    def _check(self, rc: int) -> None:
        if rc != CG_OK:
            raise CosmosGLError(rc, (self._l.cg_last_error(self._ctx) or b"").decode())

    def close(self) -> None:
        if self._ctx:
            self._l.cg_destroy(self._ctx)
            self._ctx = _CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- ingest ---------------------------------------------------------------
    def set_positions(self, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        self.n = xy.size // 2
        self._check(self._l.cg_set_positions(self._ctx, fptr(xy), self.n))

    def set_links(self, links, strength=None, rand_in=None, rand_out=None) -> None:
        links = np.ascontiguousarray(links, dtype=np.float32).reshape(-1)
        m = links.size // 2
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        strength, rand_in, rand_out = conv(strength), conv(rand_in), conv(rand_out)
        self._check(self._l.cg_set_links(self._ctx, fptr(links), m, fptr(strength),
                                         fptr(rand_in), fptr(rand_out)))

    def set_jitter(self, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        self._check(self._l.cg_set_jitter(self._ctx, fptr(xy), xy.size // 2))

    def set_point_sizes(self, sizes) -> None:
        if sizes is None:
            self._check(self._l.cg_set_point_sizes(self._ctx, None, 0))
            return
        s = np.ascontiguousarray(sizes, dtype=np.float32)
        self._check(self._l.cg_set_point_sizes(self._ctx, fptr(s), s.size))

    def set_pinned(self, indices) -> None:
        idx = np.ascontiguousarray(indices if indices is not None else [], dtype=np.uint32)
        self._check(self._l.cg_set_pinned(self._ctx, uptr(idx) if idx.size else None, idx.size))

    def set_alpha(self, alpha: float) -> None:
        self._check(self._l.cg_set_alpha(self._ctx, float(alpha)))

    def set_point_clusters(self, clusters, gid: int | None = None) -> None:
        a = None if clusters is None else np.ascontiguousarray(clusters, dtype=np.int32)
        n = 0 if a is None else a.size
        if gid is None:
            self._check(self._l.cg_set_point_clusters(self._ctx, iptr(a), n))
        else:
            self._check(self._l.cg_graph_set_point_clusters(self._ctx, gid, iptr(a), n))

    def set_cluster_positions(self, xy, gid: int | None = None) -> None:
        a = None if xy is None else np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        n = 0 if a is None else a.size // 2
        if gid is None:
            self._check(self._l.cg_set_cluster_positions(self._ctx, fptr(a), n))
        else:
            self._check(self._l.cg_graph_set_cluster_positions(self._ctx, gid, fptr(a), n))

    def set_point_cluster_strength(self, strength, gid: int | None = None) -> None:
        a = None if strength is None else np.ascontiguousarray(strength, dtype=np.float32)
        n = 0 if a is None else a.size
        if gid is None:
            self._check(self._l.cg_set_point_cluster_strength(self._ctx, fptr(a), n))
        else:
            self._check(self._l.cg_graph_set_point_cluster_strength(self._ctx, gid, fptr(a), n))

    def get_alpha(self) -> float:
        a = C.c_double()
        self._check(self._l.cg_get_alpha(self._ctx, C.byref(a)))
        return a.value

    # --- tick -----------------------------------------------------------------
    def step(self, params: CgSimParams, nticks: int = 1, seeds=None, trace: bool = False):
        s = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.float32)
        tr = np.zeros(nticks, dtype=np.float64) if trace else None
        self._check(self._l.cg_step(self._ctx, C.byref(params), nticks, fptr(s), dptr(tr)))
        return tr

    def set_tuning(self, tuning: dict) -> None:
        """Free implementation parameters (cg_tuning); keys missing from `tuning` keep their defaults."""
        t = CgTuning()
        self._l.cg_default_tuning(C.byref(t))
        for k, v in tuning.items():
            setattr(t, k, int(v))
        self._check(self._l.cg_set_tuning(self._ctx, C.byref(t)))

    def get_tuning(self) -> dict:
        t = CgTuning()
        self._check(self._l.cg_get_tuning(self._ctx, C.byref(t)))
        return t.as_dict()

    def prepare(self, params: CgSimParams) -> None:
        """Packs graph 0 and captures the tick graphs cg_step replays, running no tick."""
        self._check(self._l.cg_prepare(self._ctx, C.byref(params)))

    def get_positions(self) -> np.ndarray:
        out = np.empty(2 * self.n, dtype=np.float32)
        self._check(self._l.cg_get_positions(self._ctx, fptr(out), self.n))
        return out

    def drag_points(self, indices, xy, gid: int | None = None) -> None:
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        p = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        assert p.size == 2 * idx.size
        if gid is None:
            self._check(self._l.cg_drag_points(self._ctx, uptr(idx), fptr(p), idx.size))
        else:
            self._check(self._l.cg_graph_drag_points(self._ctx, gid, uptr(idx), fptr(p), idx.size))

    def get_positions_at(self, indices, gid: int | None = None) -> np.ndarray:
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        out = np.empty(2 * idx.size, dtype=np.float32)
        if gid is None:
            self._check(self._l.cg_get_positions_at(self._ctx, uptr(idx), idx.size, fptr(out)))
        else:
            self._check(self._l.cg_graph_get_positions_at(self._ctx, gid, uptr(idx), idx.size, fptr(out)))
        return out

    # --- inspection -----------------------------------------------------------
    def eval_force(self, params: CgSimParams, force: int, alpha: float, seed: float = 0.0) -> np.ndarray:
        out = np.zeros(2 * self.n, dtype=np.float32)
        self._check(self._l.cg_eval_force(self._ctx, C.byref(params), force, alpha, seed, fptr(out)))
        return out

    def graph_eval_force(self, gid: int, force: int, alpha: float, seed: float = 0.0) -> np.ndarray:
        """One force pass of graph `gid` inside the packed set, with its stored parameters."""
        out = np.zeros(2 * self._sizes.get(gid, 0), dtype=np.float32)
        self._check(self._l.cg_graph_eval_force(self._ctx, gid, force, alpha, seed, fptr(out)))
        return out

    def eval_manybody_noalloc(self, params: CgSimParams, alpha: float, seed: float, xy, jitter=None):
        """Many-body velocity of the given positions through the allocation-free
        entry point (the static device scratch; no allocation)."""
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        jit = None if jitter is None else np.ascontiguousarray(jitter, dtype=np.float32).reshape(-1)
        n = xy.size // 2
        out = np.zeros(2 * n, dtype=np.float32)
        self._check(self._l.cg_eval_manybody_noalloc(self._ctx, C.byref(params), alpha, seed, fptr(xy),
                                                     fptr(jit), n, fptr(out)))
        return out

    def debug_slots(self):
        gf, L = C.c_uint32(), C.c_uint32()
        self._check(self._l.cg_debug_slots(self._ctx, None, 0, C.byref(gf), C.byref(L)))
        out = np.empty(gf.value * gf.value * 8, dtype=np.int32)
        self._check(self._l.cg_debug_slots(self._ctx, iptr(out), out.size, C.byref(gf), C.byref(L)))
        return out, gf.value, L.value

    def debug_cells(self, level: int, members: bool = False):
        """(counts[g*g], centroids[g*g, 2], members or None) of pyramid level
        `level` from the last many-body pass; members (finest level only) are
        the present points grouped by row-major cell."""
        g = 4 << level
        counts = np.empty(g * g, dtype=np.uint32)
        cent = np.empty(2 * g * g, dtype=np.float32)
        mem = np.empty(max(self.n, 1), dtype=np.uint32) if members else None
        self._check(self._l.cg_debug_cells(self._ctx, level, uptr(counts), fptr(cent), uptr(mem),
                                           0 if mem is None else mem.size))
        if mem is not None:
            mem = mem[: int(counts.sum(dtype=np.int64))]
        return counts, cent.reshape(-1, 2), mem

    def debug_collision_cells(self):
        """(keys, points, cells, counts, G) of graph 0's last collision pass: the
        sorted (base cell, point) pairs, every point's cell in each of the four
        offset grids (cells[i, o]; UINT32_MAX for absent points) and the point
        count of every cell aggregate (counts[o, cell])."""
        cnt, G = C.c_uint32(), C.c_uint32()
        self._check(self._l.cg_debug_collision_cells(self._ctx, None, None, None, None, 0, C.byref(cnt),
                                                     C.byref(G)))
        k = np.empty(max(cnt.value, 1), dtype=np.uint32)
        p = np.empty(max(cnt.value, 1), dtype=np.uint32)
        cells = np.empty(4 * max(self.n, 1), dtype=np.uint32)
        counts = np.empty(4 * G.value * G.value, dtype=np.uint32)
        self._check(self._l.cg_debug_collision_cells(self._ctx, uptr(k), uptr(p), uptr(cells), uptr(counts),
                                                     k.size, C.byref(cnt), C.byref(G)))
        return (k[: cnt.value], p[: cnt.value], cells[: 4 * self.n].reshape(-1, 4),
                counts.reshape(4, -1), G.value)

    def debug_link_csr(self, direction: int):
        E = C.c_uint32()
        self._check(self._l.cg_debug_link_csr(self._ctx, direction, None, None, None, None, None, 0,
                                              C.byref(E)))
        n = self.n
        off = np.empty(n + 1, dtype=np.uint32)
        nbr = np.empty(max(E.value, 1), dtype=np.uint32)
        bias = np.empty(max(E.value, 1), dtype=np.float32)
        strength = np.empty(max(E.value, 1), dtype=np.float32)
        deg = np.empty(max(n, 1), dtype=np.uint32)
        self._check(self._l.cg_debug_link_csr(self._ctx, direction, uptr(off), uptr(nbr), fptr(bias),
                                              fptr(strength), uptr(deg), nbr.size, C.byref(E)))
        e = E.value
        return off, nbr[:e], bias[:e], strength[:e], deg[:n]

    def set_profiling(self, stages, link_stats: bool = False) -> None:
        mask = (1 << 31) if link_stats else 0
        for s in stages:
            mask |= 1 << (STAGES.index(s) if isinstance(s, str) else int(s))
        self._check(self._l.cg_set_profiling(self._ctx, mask))

    def get_profile(self) -> dict:
        ms = np.zeros(len(STAGES), np.float64)
        cnt = np.zeros(len(STAGES), np.uint32)
        self._check(self._l.cg_get_profile(self._ctx, dptr(ms), uptr(cnt), len(STAGES)))
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(STAGES) if cnt[i]}

    def debug_link_stats(self) -> dict:
        out = np.zeros(16, np.uint32)
        self._check(self._l.cg_debug_link_stats(self._ctx, uptr(out)))
        t = {}
        for name, o in (("t64", 0), ("t256", 4)):
            s, n, mx, cap = (int(v) for v in out[o:o + 4])
            t[name] = {"mean_sweeps": s / n if n else 0.0, "units": n, "max_sweeps": mx, "at_cap": cap}
        s, n, mx, cap = (int(v) for v in out[8:12])
        t["split"] = {"mean_rounds": s / n if n else 0.0, "points": n, "max_rounds": mx, "at_cap": cap}
        return t

    def device_bytes(self) -> int:
        return int(self._l.cg_device_bytes(self._ctx))

    # --- graph set ------------------------------------------------------------
    def graph_add(self) -> int:
        gid = C.c_uint32()
        self._check(self._l.cg_graph_add(self._ctx, C.byref(gid)))
        self._sizes[gid.value] = 0
        return gid.value

    def graph_remove(self, gid: int) -> None:
        self._check(self._l.cg_graph_remove(self._ctx, gid))
        self._sizes.pop(gid, None)

    def graph_set_params(self, gid: int, params: CgSimParams) -> None:
        self._check(self._l.cg_graph_set_params(self._ctx, gid, C.byref(params)))

    def graph_set_seed(self, gid: int, seed: int) -> None:
        self._check(self._l.cg_graph_set_seed(self._ctx, gid, int(seed)))

    def graph_set_positions(self, gid: int, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        self._sizes[gid] = xy.size // 2
        self._check(self._l.cg_graph_set_positions(self._ctx, gid, fptr(xy), xy.size // 2))

    def graph_set_links(self, gid: int, links, strength=None, rand_in=None, rand_out=None) -> None:
        links = np.ascontiguousarray(links, dtype=np.float32).reshape(-1)
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        strength, rand_in, rand_out = conv(strength), conv(rand_in), conv(rand_out)
        self._check(self._l.cg_graph_set_links(self._ctx, gid, fptr(links), links.size // 2,
                                               fptr(strength), fptr(rand_in), fptr(rand_out)))

    def graph_set_jitter(self, gid: int, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        self._check(self._l.cg_graph_set_jitter(self._ctx, gid, fptr(xy), xy.size // 2))

    def graph_set_point_sizes(self, gid: int, sizes) -> None:
        if sizes is None:
            self._check(self._l.cg_graph_set_point_sizes(self._ctx, gid, None, 0))
            return
        s = np.ascontiguousarray(sizes, dtype=np.float32)
        self._check(self._l.cg_graph_set_point_sizes(self._ctx, gid, fptr(s), s.size))

    def graph_set_pinned(self, gid: int, indices) -> None:
        idx = np.ascontiguousarray(indices if indices is not None else [], dtype=np.uint32)
        self._check(self._l.cg_graph_set_pinned(self._ctx, gid, uptr(idx) if idx.size else None, idx.size))

    def graph_set_alpha(self, gid: int, alpha: float) -> None:
        self._check(self._l.cg_graph_set_alpha(self._ctx, gid, float(alpha)))

    def graph_get_alpha(self, gid: int) -> float:
        a = C.c_double()
        self._check(self._l.cg_graph_get_alpha(self._ctx, gid, C.byref(a)))
        return a.value

    def graph_get_positions(self, gid: int) -> np.ndarray:
        n = self._sizes.get(gid, 0)
        out = np.empty(2 * n, dtype=np.float32)
        self._check(self._l.cg_graph_get_positions(self._ctx, gid, fptr(out), n))
        return out

    def step_graphs(self, gids, nticks: int) -> None:
        ids = np.ascontiguousarray(list(gids), dtype=np.uint32)
        self._check(self._l.cg_step_graphs(self._ctx, uptr(ids) if ids.size else None, ids.size, nticks))

    def sync(self) -> None:
        self._check(self._l.cg_sync(self._ctx))

    def set_capture(self, enable: bool) -> None:
        self._check(self._l.cg_set_capture(self._ctx, 1 if enable else 0))
