#!/usr/bin/env python3
"""Benchmark of the force-simulation tick on one MI355X.

A step is one tick (gravity -> many-body -> link IN -> link OUT, each with its
integrate pass, then alpha decay) over the resident 1M-point / 5M-link graph.
Prints one JSON line: whole-job ticks/s with inputs already in HBM, the
roofline of the kernel that dominates the tick (timed with HIP events on the
library's stream over the timed region), and the oracle's CPU tick timed on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ticks/s at 1M pts/5M links; batched layouts/s + p50/p99 latency, 1 GPU"
HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 157.3

# The kernels each stage's HIP events bracket (one bracketed launch = one pass;
# the link passes fork a helper stream and join it before their closing event).
STAGE_KERNELS = {
    "gravity": "k_pointwise<GRAVITY>",
    "mb_sort": "k_mb_clear + k_bin_count + k_count_sums/k_count_scan/k_count_bounds + k_bin_place",
    "mb_cells": "k_cells_seg + k_cells_join",
    "mb_pyramid": "k_pyramid x2",
    "mb_force": "k_mb_force",
    "link_in": "link pass IN: k_link_lane + k_link_chunk<64,4> + k_link_chunk<256,6> + k_hub_final (two streams)",
    "link_out": "link pass OUT: k_link_lane + k_link_chunk<64,4> + k_link_chunk<256,6> + k_hub_final (two streams)",
    "collision": "k_col_keys + radix passes + k_fill32 + k_col_bounds + k_col_cells + k_col_force",
}


def stage_work(stage: str, info: dict) -> tuple[str, float]:
    """Algorithmic work of one pass of a stage: the compulsory HBM traffic of the
    reference's dataflow in SoA f32/int32 (each array counted once per pass that
    consumes or produces it; cache-served re-reads not counted), or the FLOPs of
    the far/near-field pass.  np present points, E directed entries of the pass's
    direction, C_f finest cells, C_tot level cells."""
    np_, L, Cf, Ct = info["np"], info["L"], info["gf"] ** 2, info["C_tot"]
    if stage in ("gravity", "center", "mouse", "cluster"):
        return "hbm", 16.0 * np_                      # position in 8, position out 8
    if stage == "mb_sort":
        return "hbm", 16.0 * np_                      # position in 8; cell key 4 + point index 4 out
    if stage == "mb_cells":
        return "hbm", 12.0 * np_ + 80.0 * Cf          # index 4 + position 8 in; finest sums 16 + 8 slots x 8 out
    if stage == "mb_pyramid":
        return "hbm", 16.0 * Cf + 16.0 * Ct           # finest sums in; every level's (x, y, N) out
    if stage == "mb_force":
        # flops per far-field cell: d (2), |d|^2 (3), 1/l (1), N/l (1), accumulate (4) = 11,
        # over 9.75 level-0 cells (mean over query cells of the 4x4 grid) and 27 per finer
        # level; per near-field pair 13 (the same plus min/sqrt of the clamp), at most 72 pairs
        far = 9.75 + 27 * (L - 1)
        return "valu", (11.0 * far + 13.0 * info.get("near_pairs", 72)) * np_
    if stage in ("link_in", "link_out"):
        E = info["E_live"][0 if stage == "link_in" else 1]
        return "hbm", 24.0 * E + 16.0 * np_          # per entry: neighbour, draw, bias, strength 16 + gathered position 8
    if stage == "collision":
        return "hbm", 92.0 * np_                      # four offset grids: keys, cell sums, sizes, positions
    return "hbm", 0.0


def level_cells(L: int) -> int:
    """Cells of a pyramid of L levels (4^2, 8^2, ..., (4 << (L-1))^2)."""
    return sum((4 << lv) ** 2 for lv in range(L))


def stage_rooflines(breakdown: dict, info: dict) -> dict:
    """Per stage: ms per pass, algorithmic bytes (or flops) per pass, achieved rate
    and its fraction of the MI355X peak (HBM 8 TB/s; FP32 VALU 157.3 TFLOP/s)."""
    out = {}
    for st, (ms, cnt) in breakdown.items():
        if cnt == 0:
            continue
        kind, work = stage_work(st, info)
        sec = ms / cnt / 1e3
        if work <= 0 or sec <= 0:
            continue
        if kind == "hbm":
            rate = work / sec / 1e9
            out[st] = {"ms_per_pass": sec * 1e3, "bound": "hbm", "bytes_per_pass": work,
                       "achieved": rate, "unit": "GB/s", "frac": rate / HBM_PEAK_GBS}
        else:
            rate = work / sec / 1e12
            out[st] = {"ms_per_pass": sec * 1e3, "bound": "valu", "flops_per_pass": work,
                       "achieved": rate, "unit": "TFLOP/s", "frac": rate / VALU_FP32_PEAK_TFLOPS}
    return out


def link_tiers(links: np.ndarray, n: int, tuning: dict) -> dict:
    """Points and entries of each link tier per direction, from the link list and
    the tier bounds (every point present): a point's entries are its valid links
    as the owner of the direction (IN: the source, OUT: the target,
    ForceLink/index.ts:49); the lane tier also holds the points without any.
    Algorithmic bytes as for the whole pass: 24 per entry and 16 per point."""
    lk = np.asarray(links, dtype=np.float32).reshape(-1, 2)
    ok = ((lk >= 0) & (lk < n) & (lk == np.floor(lk))).all(axis=1)
    lane, group = tuning["link_lane_max"], tuning["link_group_max"]
    hub = 256 * tuning["link_hub_entries"]
    out = {}
    for name, col in (("in", 0), ("out", 1)):
        deg = np.bincount(lk[ok, col].astype(np.int64), minlength=n)
        tiers = {"lane": deg <= lane, "lanes64": (deg > lane) & (deg <= group),
                 "lanes256": (deg > group) & (deg <= hub), "split": deg > hub}
        out[name] = {}
        for k, m in tiers.items():
            pts, ent = int(m.sum()), int(deg[m].sum())
            out[name][k] = {"points": pts, "entries": ent, "algorithmic_bytes": 24 * ent + 16 * pts}
    return out


def profiled_traffic(kernel: str) -> tuple[float | None, str | None]:
    """HBM bytes per pass of a stage (or launch of a kernel) from the committed
    rocprofv3 --pmc summary (read requests by size and WRITE_SIZE over
    `bench.py --probe cfg1`, checked against calibration kernels of known
    traffic there), or None when it holds no figure for it."""
    path = ROOT / "profiles" / "traffic_cfg1.md"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
            return float(parts[2]), ("profiles/traffic_cfg1.md (rocprofv3 --pmc: 128/64/32-byte read requests "
                                     "(= FETCH_SIZE x 2.0) + WRITE_SIZE x 1.0, per pass)")
    return None, None


def oracle_ticks_per_s(wl: dict, seconds: float, threads: int) -> tuple[float, int, float, int]:
    """Full FP64 oracle ticks of a workload for about `seconds` on `threads`
    threads (0 = the OpenMP default): (ticks/s, ticks, seconds, threads)."""
    from oracle import oracle as O

    O.set_threads(threads)
    used = O.get_threads()
    p = wl["params"]
    pd = dict(space_size=wl["S"], gravity=p.gravity, center=0.0, repulsion=p.repulsion,
              link_spring=p.link_spring, link_distance=p.link_distance, lo=1.0, hi=1.2,
              friction=p.friction, decay=p.decay)
    Lk = O.Links(wl["n"], wl["links"], wl["r_in"], wl["r_out"])
    pos, a = wl["xy"].astype(np.float64), 1.0
    ticks = 0
    t0 = time.perf_counter()
    while True:
        pos, a = O.tick(pos, a, 0.5 + 0.01 * ticks, pd, links=Lk, jitter=wl["jitter"])
        ticks += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    O.set_threads(0)
    return ticks / el, ticks, el, used


def cpu_baseline(wl: dict, wl0: dict, seconds: float) -> dict:
    """The oracle's FP64 tick on the host: the 1M/5M workload on every host
    thread OpenMP is given (the reported baseline), and beside it the same on
    one thread and the 10k workload on every thread."""
    v, t, el, used = oracle_ticks_per_s(wl, seconds, 0)
    v1, t1, el1, _ = oracle_ticks_per_s(wl, seconds * 0.5, 1)
    v0, t0, el0, used0 = oracle_ticks_per_s(wl0, seconds * 0.5, 0)
    return {"value": v, "unit": "ticks/s", "cores": used, "kind": "port",
            "sample": f"{t} full ticks of the same 1M-point/5M-link workload through the FP64 oracle "
                      f"(oracle/cosmos_oracle.c, OpenMP over {used} threads) in {el:.1f} s",
            "one_thread": {"value": v1, "unit": "ticks/s", "cores": 1,
                           "sample": f"{t1} ticks of the 1M/5M workload on 1 thread in {el1:.1f} s"},
            "cfg0": {"value": v0, "unit": "ticks/s", "cores": used0,
                     "sample": f"{t0} ticks of the 10k-point BA workload on {used0} threads in {el0:.1f} s"},
            "reference_cpu_path": "unavailable: the reference's shaders need Node >= 22, its npm packages "
                                  "(luma.gl 9.3.6) and a headless WebGL2 (llvmpipe), none of which is "
                                  "present here or on the GPU host"}


def bench_cfg0(ticks: int, warmup: int) -> dict:
    """The 10k-point BA workload on the GPU: `ticks` ticks (the run's length)
    after `warmup`, timed between synchronisations."""
    import torch

    from graph_amd import _lib, autotune, datagen, workloads

    wl = workloads.cfg0()
    ctx = _lib.Context(device=0)
    tuning = autotune.apply(ctx, wl["n"], wl["links"].size / wl["n"])
    ctx.set_positions(wl["xy"])
    ctx.set_jitter(wl["jitter"])
    ctx.set_links(wl["links"], None, wl["r_in"], wl["r_out"])
    ctx.set_alpha(1.0)
    p = wl["params"]
    seeds = datagen.uniform(warmup + ticks, 0.0, 1.0, 95)
    ctx.prepare(p)
    ctx.step(p, warmup, seeds=seeds[:warmup])
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(p, ticks, seeds=seeds[warmup:])
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ctx.close()
    return {"workload": wl["name"], "points": wl["n"], "links": int(wl["links"].size // 2), "ticks": ticks,
            "warmup": warmup, "ticks_per_s": ticks / el, "ms_per_tick": el / ticks * 1e3, "tuning": tuning,
            "point_ticks_per_s": ticks * wl["n"] / el}


def force_error_log(wl: dict, ticks: int, every: int, sample: int, log) -> dict:
    """The 1M/5M run with its per-pass force error: every `every` ticks, each
    pass (gravity, many-body, link IN, link OUT) is evaluated on the GPU from the
    live state and by the FP64 oracle from the same state on a fixed sample of
    points; the error is max |v_gpu - v_oracle| / spaceSize over the sample."""
    from graph_amd import _lib
    from oracle import oracle as O

    n, S, p = wl["n"], wl["S"], wl["params"]
    ctx = _lib.Context(device=0)
    ctx.set_positions(wl["xy"])
    ctx.set_jitter(wl["jitter"])
    ctx.set_links(wl["links"], None, wl["r_in"], wl["r_out"])
    ctx.set_alpha(1.0)
    Lk = O.Links(n, wl["links"], wl["r_in"], wl["r_out"])
    # a fixed random sample plus the 64 longest link chains of each direction
    # (the hubs, whose chains are solved by segments)
    lk = wl["links"].reshape(-1, 2).astype(np.int64)
    hubs = [np.argsort(np.bincount(lk[:, c], minlength=n))[-64:] for c in (0, 1)]
    idx = np.unique(np.concatenate([np.random.default_rng(5).choice(n, sample, replace=False)] + hubs))
    idx = idx.astype(np.uint32)
    lo, hi = p.link_dist_random_variation_range[0], p.link_dist_random_variation_range[1]
    rows, worst = [], 0.0
    for t in range(0, ticks + 1, every):
        if t:
            ctx.step(p, every)
        xy = ctx.get_positions()
        a = float(np.float32(ctx.get_alpha()))
        seed = 0.25 + (t % 97) / 200.0

        def err(got, want):
            d = got.reshape(-1, 2)[idx].astype(np.float64) - want.reshape(-1, 2)[idx]
            return float(np.hypot(d[:, 0], d[:, 1]).max() / S)

        e = {"gravity": err(ctx.eval_force(p, _lib.FORCE_GRAVITY, a), O.gravity(xy, S, a, p.gravity)),
             "manybody": err(ctx.eval_force(p, _lib.FORCE_MANYBODY, a, seed),
                             O.manybody(xy, S, a, p.repulsion, seed, jitter=wl["jitter"], eval_idx=idx))}
        for d, name in ((0, "link_in"), (1, "link_out")):
            e[name] = err(ctx.eval_force(p, _lib.FORCE_LINK_IN + d, a),
                          Lk.force(xy, d, a, p.link_spring, p.link_distance, lo, hi))
        worst = max(worst, max(e.values()))
        rows.append({"tick": t, "alpha": a, **{k: float(f"{v:.3e}") for k, v in e.items()}})
        if t % 250 == 0:
            log(f"force error at tick {t}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
    ctx.close()
    return {"ticks": ticks, "every": every, "sample_points": int(idx.size),
            "sample": f"{sample} random points and the 64 longest link chains per direction",
            "tolerance": 1e-5, "max_error_over_S": worst, "within_tolerance": worst <= 1e-5, "log": rows}


def densest_cells_sample(xy: np.ndarray, S: float, g: int, cells: int, cap: int) -> np.ndarray:
    """Points of the `cells` most populated cells of a g x g grid over [0, S)^2
    (at most `cap` points), by the oracle's cell rule."""
    from oracle import oracle as O

    c = O.cells(xy, O.cell_size(S, g), g, O.absent_mask(xy))
    ok = c[:, 0] >= 0
    key = c[:, 1].astype(np.int64) * g + c[:, 0]
    key[~ok] = -1
    cnt = np.bincount(key[ok], minlength=g * g)
    top = np.argsort(cnt)[-cells:]
    return np.flatnonzero(np.isin(key, top))[:cap]


def bench_cfg2(ticks: int, every: int, sample: int, log) -> dict:
    """The 4M-point SBM graph with collision: `ticks` ticks from alpha 1, timed in
    chunks of `every` ticks (each bracketed by synchronisations); between chunks,
    outside the timing, every pass (gravity, many-body, link IN, link OUT,
    collision) is evaluated on the GPU from the live state and by the FP64 oracle
    from the same state on a sample: random points, the points of the most
    crowded finest and collision cells and the longest link chains."""
    import torch

    from graph_amd import _lib, autotune, datagen, workloads
    from oracle import oracle as O

    t_gen = time.perf_counter()
    wl = workloads.cfg2()
    n, S, p = wl["n"], wl["S"], wl["params"]
    m = int(wl["links"].size // 2)
    log(f"cfg2: {n} points, {m} links generated in {time.perf_counter() - t_gen:.1f}s")
    ctx = _lib.Context(device=0)
    tuning = autotune.apply(ctx, n, 2.0 * m / n)
    ctx.set_positions(wl["xy"])
    ctx.set_jitter(wl["jitter"])
    ctx.set_links(wl["links"], None, wl["r_in"], wl["r_out"])
    ctx.set_point_sizes(wl["sizes"])
    ctx.set_alpha(1.0)
    ctx.prepare(p)
    sizes = wl["sizes"]
    max_size = max(float(p.point_default_size), float(sizes.max()))
    G, h = O.collision_grid(S, float(p.collision_radius), float(p.collision_padding), max_size)
    Lk = O.Links(n, wl["links"], wl["r_in"], wl["r_out"])
    lk = wl["links"].reshape(-1, 2).astype(np.int64)
    hubs = np.concatenate([np.argsort(np.bincount(lk[:, c], minlength=n))[-64:] for c in (0, 1)])
    rnd = np.random.default_rng(6).choice(n, sample, replace=False)
    lo, hi = p.link_dist_random_variation_range[0], p.link_dist_random_variation_range[1]
    seeds = datagen.uniform(ticks, 0.0, 1.0, 111)
    rows, worst, chunk_s = [], {}, []
    for t in range(0, ticks + 1, every):
        if t:
            ctx.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.step(p, every, seeds=seeds[t - every: t])
            ctx.sync()
            torch.cuda.synchronize()
            chunk_s.append(time.perf_counter() - t0)
        xy = ctx.get_positions()
        a = float(np.float32(ctx.get_alpha()))
        seed = 0.25 + (t % 97) / 200.0
        idx = np.unique(np.concatenate([rnd, hubs, densest_cells_sample(xy, S, 512, 8, 512),
                                        densest_cells_sample(xy, S, G, 8, 512)])).astype(np.uint32)

        def err(got, want):
            d = got.reshape(-1, 2)[idx].astype(np.float64) - want.reshape(-1, 2)[idx]
            return float(np.hypot(d[:, 0], d[:, 1]).max() / S)

        e = {"gravity": err(ctx.eval_force(p, _lib.FORCE_GRAVITY, a), O.gravity(xy, S, a, p.gravity)),
             "manybody": err(ctx.eval_force(p, _lib.FORCE_MANYBODY, a, seed),
                             O.manybody(xy, S, a, p.repulsion, seed, jitter=wl["jitter"], eval_idx=idx))}
        for d, name in ((0, "link_in"), (1, "link_out")):
            e[name] = err(ctx.eval_force(p, _lib.FORCE_LINK_IN + d, a),
                          Lk.force(xy, d, a, p.link_spring, p.link_distance, lo, hi))
        e["collision"] = err(ctx.eval_force(p, _lib.FORCE_COLLISION, a),
                             O.collision(xy, sizes, S, a, p.collision, p.collision_radius, p.collision_padding,
                                         max_size=max_size))
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
        rows.append({"tick": t, "alpha": a, "sample_points": int(idx.size),
                     **{k: float(f"{v:.3e}") for k, v in e.items()}})
        if t % 100 == 0:
            log(f"cfg2 force error at tick {t}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
    # stage breakdown over 20 direct-launch ticks after the run
    ctx.set_profiling(_lib.STAGES)
    ctx.step(p, 20, seeds=datagen.uniform(20, 0.0, 1.0, 112))
    breakdown = ctx.get_profile()
    ctx.set_profiling([])
    dev_gib = ctx.device_bytes() / 2**30
    ctx.close()
    info = dict(n=n, np=n, L=8, gf=512, C_tot=level_cells(8), E_live=[m, m])
    timed = chunk_s[1:] or chunk_s  # the first chunk also runs the first ticks' one-off work
    tps = len(timed) * every / sum(timed)
    return {
        "workload": wl["name"] + f" ({m} links, log-normal radii max size {max_size:.1f}, "
                    f"collision strength {p.collision:g})",
        "points": n, "links": m, "space_size": S,
        "space_size_note": "the config's 16384; Store.adjustSpaceSize (Store/index.ts:209-233) would halve "
                           "it to 8192 on a WebGL device whose maxTextureDimension2D is 16384 (none here)",
        "collision_grid": G, "collision_cell": h, "finest_grid": 512, "levels": 8, "tuning": tuning,
        "ticks": ticks, "ticks_per_s": tps, "ms_per_tick": 1e3 / tps, "point_ticks_per_s": tps * n,
        "timed_window": f"ticks {ticks - len(timed) * every}-{ticks}: {len(timed)} chunks of {every} ticks, each between "
                        f"synchronisations (the error evaluations between chunks are outside the timing)",
        "stage_ms_per_tick": {st: v[0] / 20.0 for st, v in breakdown.items()},
        "stages": stage_rooflines(breakdown, info),
        "device_gib": dev_gib,
        "force_error": {"every": every, "tolerance": 1e-5, "max_error_over_S": worst,
                        "within_tolerance": max(worst.values()) <= 1e-5,
                        "sample": f"{sample} random points, the 64 longest link chains per direction, the "
                                  f"points of the 8 most crowded finest cells and of the 8 most crowded "
                                  f"collision cells (up to 512 each)",
                        "log": rows},
    }


def bench_batched(count: int, ticks: int, sample: int, log) -> dict:
    """Batched layouts: `count` independent BA graphs packed into one graph set,
    `ticks` ticks each, against the plain one-graph-at-a-time loop (the
    single-graph entry points, one context per graph) timed on an evenly
    strided sample of the same graphs and scaled to the whole batch."""
    from graph_amd import _lib, autotune, serving, workloads

    ns, ms = workloads.cfg3_sizes(count)
    graphs = [workloads.cfg3_graph(i, int(ns[i]), int(ms[i])) for i in range(count)]
    links_total = sum(g["links"].size // 2 for g in graphs)
    ctx = _lib.Context(device=0)
    tuning = autotune.apply(ctx, int(ns.sum()), 2.0 * links_total / float(ns.sum()))
    ids = []
    for g in graphs:
        gid = ctx.graph_add()
        ids.append(gid)
        ctx.graph_set_params(gid, g["params"])
        ctx.graph_set_positions(gid, g["xy"])
        ctx.graph_set_jitter(gid, g["jitter"])
        ctx.graph_set_links(gid, g["links"], None, g["r_in"], g["r_out"])
    t0 = time.perf_counter()
    ctx.step_graphs(ids, 0)  # packs the set (no tick)
    ctx.sync()
    pack_s = time.perf_counter() - t0
    dev_gib = ctx.device_bytes() / 2**30
    log(f"batched: {count} graphs, {int(ns.sum())} points packed in {pack_s:.2f}s ({dev_gib:.1f} GiB)")
    t0 = time.perf_counter()
    ctx.step_graphs(ids, ticks)
    ctx.sync()
    batch_s = time.perf_counter() - t0
    picks = workloads.cfg3_plain_sample(count, sample)
    xb = {i: ctx.graph_get_positions(ids[i]) for i in picks}
    ctx.close()
    log(f"batched: {ticks} ticks in {batch_s:.2f}s")
    plain_s, same = [], True
    for i in picks:
        g = graphs[i]
        with _lib.Context(device=0) as c:
            autotune.apply(c, g["n"], g["links"].size / g["n"])
            c.set_positions(g["xy"])
            c.set_jitter(g["jitter"])
            c.set_links(g["links"], None, g["r_in"], g["r_out"])
            c.step(g["params"], 0)
            c.get_positions()
            t0 = time.perf_counter()
            c.step(g["params"], ticks)
            x = c.get_positions()
            plain_s.append(time.perf_counter() - t0)
        same &= bool(np.array_equal(x.view(np.uint32), xb[i].view(np.uint32)))
    plain_total = float(np.mean(plain_s)) * count
    log(f"batched: plain loop {np.mean(plain_s) * 1e3:.1f} ms per graph on {len(picks)} graphs")
    # completion latencies: the same graphs submitted together at t = 0 to the
    # batch server with a resident cap of a quarter of the points, so graphs
    # queue and are admitted in order as others finish
    jobs = [serving.Job(jid=i, arrival=0.0, xy=g["xy"], links=g["links"], params=g["params"], r_in=g["r_in"],
                        r_out=g["r_out"], jitter=g["jitter"], ticks=ticks) for i, g in enumerate(graphs)]
    cap = int(ns.sum()) // 4
    srv = serving.BatchServer(budget=200 * 2**30, segment=100, tuning=tuning, max_points=cap)
    res, st = srv.run(jobs)
    srv.close()
    stream = serving.latency_summary(res)
    stream.update(resident_point_cap=cap, repacks=st["repacks"], pack_s=st["pack_s"], tick_s=st["tick_s"])
    log(f"batched: streamed with a {cap}-point cap: p50 {stream['p50_s']:.2f}s p99 {stream['p99_s']:.2f}s, "
        f"{stream['layouts_per_s']:.1f} layouts/s")
    same_stream = all(np.array_equal(r.xy.view(np.uint32), xb[r.jid].view(np.uint32)) for r in res if r.jid in xb)
    return {
        "workload": f"{count} Barabasi-Albert graphs, n log-uniform in [1000, 200000] (seed 23), "
                    f"m uniform in 1..4, {ticks} ticks each",
        "graphs": count, "points": int(ns.sum()), "ticks": ticks,
        "layouts_per_s": count / batch_s, "ms_per_tick": batch_s / ticks * 1e3,
        "point_ticks_per_s": float(ns.sum()) * ticks / batch_s,
        "latency_all_at_once_s": batch_s,
        "p50_latency_s": stream["p50_s"], "p99_latency_s": stream["p99_s"],
        "streamed": stream, "streamed_equals_batched": same_stream, "tuning": tuning,
        "pack_s": pack_s, "device_gib": dev_gib,
        "plain_layouts_per_s": count / plain_total,
        "plain_sample": f"{len(picks)} evenly strided graphs of the batch run one at a time through "
                        f"cg_step, mean {np.mean(plain_s) * 1e3:.1f} ms per {ticks}-tick layout, "
                        f"scaled to {count} graphs",
        "speedup_vs_plain": plain_total / batch_s,
        "plain_equals_batched": same,
    }


def bench_serving(window_s: float, budget_gib: float, segment: int, stride: int, log,
                  max_points: int | None = None) -> dict:
    """Continuous batching: Poisson arrivals at 40/s over `window_s` seconds,
    each job laid out until alpha < 0.001, all of them through the batch
    server; the plain first-come first-served path runs every `stride`-th job
    (an evenly strided sample of i.i.d. requests), whose mean service time
    gives its layouts/s (the sequential path is saturated: requests arrive
    far faster than it serves them)."""
    from graph_amd import autotune, serving, workloads

    jobs = workloads.cfg4_jobs(duration=window_s)
    pts, lnk = sum(j.n for j in jobs), sum(j.m for j in jobs)
    log(f"serving: {len(jobs)} jobs, {pts} points, {lnk} links")
    tuning = autotune.lookup(pts, 2.0 * lnk / pts)
    srv = serving.BatchServer(budget=int(budget_gib * 2**30), segment=segment, tuning=tuning,
                              max_points=max_points)
    res, st = srv.run(jobs)
    srv.close()
    log(f"serving: batched done in {st['wall_s']:.1f}s")
    sample = jobs[::stride]
    seq, seq_st = serving.serve_sequential(sample, tuned=True)
    service = float(np.mean([r.service for r in seq]))
    log(f"serving: sequential sample of {len(sample)} jobs, mean service {service:.3f}s")
    by_id = {r.jid: r for r in res}
    same = all(np.array_equal(a.xy.view(np.uint32), by_id[a.jid].xy.view(np.uint32)) for a in seq)
    b = serving.latency_summary(res)
    seq_rate = 1.0 / service
    return {
        "workload": f"Poisson arrivals 40/s over {window_s:g}s (seed 31), n log-uniform in [500, 500000], "
                    f"Chung-Lu links 2-8x points, 10% with collision, until alpha < 0.001 "
                    f"({serving.alpha_ticks(5000.0)} ticks)",
        "jobs": len(jobs), "points": pts, "links": lnk, "hbm_budget_gib": budget_gib, "segment_ticks": segment,
        "resident_point_cap": max_points,
        "batched": b,
        "batched_breakdown_s": {k: st[k] for k in ("ingest_s", "overlap_s", "pack_s", "tick_s", "retire_s",
                                                   "idle_s", "wall_s")},
        "repacks": st["repacks"], "max_active": st["max_active"], "segments": st["segments"],
        "point_ticks_per_s": st["point_ticks"] / st["wall_s"],
        "sequential": {"sample_jobs": len(sample), "stride": stride, "mean_service_s": service,
                       "layouts_per_s": seq_rate,
                       "note": "plain path, one context per job, context creation to readback, on every "
                               f"{stride}-th request; layouts/s = 1 / mean service time"},
        "speedup_layouts_per_s": b["layouts_per_s"] / seq_rate,
        "sequential_equals_batched": bool(same), "tuning": tuning,
    }


def gather_probe() -> dict:
    """Time per pass of 5M random gathers of 8-byte positions from the 1M-point
    array (torch index_select with int32 indices, HIP events over 20
    repetitions): indices uniform, then the 1M/5M graph's neighbour indices
    in the CSR (owner) order the link passes read them.  Per gather: index 4 B
    read, position 8 B gathered, 8 B written."""
    import torch

    from graph_amd import workloads

    wl = workloads.cfg1()
    n = wl["n"]
    lk = wl["links"].reshape(-1, 2).astype(np.int64)
    pos = torch.rand(n, 2, device="cuda", dtype=torch.float32)
    cases = {"uniform": np.random.default_rng(7).integers(0, n, lk.shape[0])}
    for d, name in ((0, "csr_in"), (1, "csr_out")):
        cases[name] = lk[np.argsort(lk[:, d], kind="stable"), 1 - d]
    out = {"points": n, "gathers": int(lk.shape[0]), "bytes_per_gather": 20}
    for name, idx in cases.items():
        t = torch.from_numpy(idx.astype(np.int32)).cuda()
        for _ in range(3):
            pos.index_select(0, t)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            pos.index_select(0, t)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / 20 * 1e3
        out[name] = {"us": us, "gathers_per_s": idx.size / us * 1e6, "algorithmic_gb_s": idx.size * 20 / us / 1e3}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=2048)
    ap.add_argument("--batch-ticks", type=int, default=300)
    ap.add_argument("--batch-sample", type=int, default=32)
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-stride", type=int, default=5, help="sequential path on every k-th request")
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    ap.add_argument("--serve-segment", type=int, default=100)
    ap.add_argument("--serve-max-points", type=int, default=0,
                    help="resident points cap of the batch server (0: the HBM budget alone)")
    ap.add_argument("--error-ticks", type=int, default=1000)
    ap.add_argument("--cfg2-ticks", type=int, default=500)
    ap.add_argument("--cfg2-every", type=int, default=50, help="ticks between force-error checkpoints")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc: calibration kernels of known traffic (256 MiB), then 5 "
                         "direct-launch ticks of the 1M/5M workload; prints nothing else (= --probe cfg1)")
    ap.add_argument("--probe", choices=["cfg1", "cfg2", "serving"], default=None,
                    help="for rocprofv3: the calibration kernels, then --probe-ticks direct-launch ticks of "
                         "one workload (cfg1: 1M/5M; cfg2: 4M SBM with collision; serving: the 202 requests "
                         "of the serving leg's 5 s window packed in one set, after --probe-warm ticks); "
                         "prints nothing else")
    ap.add_argument("--probe-warm", type=int, default=2000,
                    help="serving probe: ticks run before the traced ones, so the layouts have collapsed")
    ap.add_argument("--probe-ticks", type=int, default=5)
    ap.add_argument("--gather-probe", action="store_true",
                    help="the random-gather ceiling of the link passes: 5M gathers of 8-byte positions from "
                         "the 1M-point array (torch index_select), uniform and in the 1M/5M graph's CSR "
                         "order; prints one JSON line")
    ap.add_argument("--skip", default="", help="comma list of legs to skip: errors,cfg2,batched,serving,cfg0")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of it "
              "(each graph or request would be assigned to one GPU)", file=sys.stderr)
        return 2

    import torch

    from graph_amd import _lib, autotune, datagen, workloads

    if args.gather_probe:
        print(json.dumps(gather_probe()), flush=True)
        return 0
    if args.traffic_probe or args.probe:
        # calibration kernels of known traffic: a 256 MiB fill (writes 256 MiB), then an
        # elementwise scale of it into a second buffer (reads 256 MiB, writes 256 MiB)
        src = torch.ones(64 * 2**20, dtype=torch.float32, device="cuda")
        dst = src * 2.0
        torch.cuda.synchronize()
        del dst, src
        if args.probe == "serving":
            from graph_amd import serving

            jobs = workloads.cfg4_jobs(duration=5.0)
            pts, lnk = sum(j.n for j in jobs), sum(j.m for j in jobs)
            ctx = _lib.Context(device=0, hbm_budget=200 << 30)
            ctx.set_tuning(autotune.lookup(pts, 2.0 * lnk / pts))
            ids = []
            for j in jobs:
                g = ctx.graph_add()
                serving._load(ctx, g, j)
                ids.append(g)
            ctx.step_graphs(ids, args.probe_warm)
            ctx.set_capture(False)
            ctx.step_graphs(ids, args.probe_ticks)
            ctx.sync()
            ctx.close()
            return 0
        pw = workloads.cfg2() if args.probe == "cfg2" else workloads.cfg1()
        ctx = _lib.Context(device=0)
        autotune.apply(ctx, pw["n"], pw["links"].size / pw["n"])
        ctx.set_positions(pw["xy"])
        ctx.set_jitter(pw["jitter"])
        ctx.set_links(pw["links"], None, pw["r_in"], pw["r_out"])
        if pw.get("sizes") is not None:
            ctx.set_point_sizes(pw["sizes"])
        ctx.set_alpha(1.0)
        ctx.set_capture(False)
        ctx.step(pw["params"], args.probe_ticks if args.probe else 5)
        ctx.sync()
        ctx.close()
        return 0
    wl = workloads.cfg1()
    n = wl["n"]
    ctx = _lib.Context(device=0)
    tuning = autotune.apply(ctx, n, wl["links"].size / n)  # the table's bit-neutral parameters
    ctx.set_positions(wl["xy"])
    ctx.set_jitter(wl["jitter"])
    ctx.set_links(wl["links"], None, wl["r_in"], wl["r_out"])
    ctx.set_alpha(1.0)
    p = wl["params"]
    seeds = datagen.uniform(args.warmup + args.steps + 1, 0.0, 1.0, 99)

    # capture and instantiate the tick graphs the timed region replays (runs no
    # tick), then the warm-up ticks on the production path
    ctx.prepare(p)
    ctx.step(p, args.warmup, seeds=seeds[: args.warmup])

    # timed region: the production path (tick sequences replayed from captured graphs)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(p, args.steps, seeds=seeds[args.warmup: args.warmup + args.steps])
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    el = t1 - t0
    ticks_per_s = args.steps / el
    # after the timed region: 50 ticks with every stage bracketed (the per-stage
    # breakdown and rooflines; picks the dominant stage by time per tick)
    ctx.set_profiling(_lib.STAGES)
    ctx.step(p, 50, seeds=datagen.uniform(50, 0.0, 1.0, 97))
    breakdown = ctx.get_profile()
    ctx.set_profiling([])
    gf, L = 512, 8
    m = int(wl["links"].size // 2)
    info = dict(n=n, np=n, L=L, gf=gf, C_tot=level_cells(L), E_live=[m, m])
    modelled = {st: v for st, v in breakdown.items() if stage_work(st, info)[1] > 0}
    dominant = max(modelled, key=lambda st: modelled[st][0])
    # the dominant stage's duration: HIP events on the library stream around each
    # of its passes (a link pass joins its helper stream before the closing event)
    # over an equal run of ticks right after (events force direct launches)
    prof_seeds = datagen.uniform(args.steps, 0.0, 1.0, 98)
    ctx.set_profiling([dominant])
    t2 = time.perf_counter()
    ctx.step(p, args.steps, seeds=prof_seeds)
    timed = ctx.get_profile()
    direct_ms = (time.perf_counter() - t2) / args.steps * 1e3
    ctx.set_profiling([])

    kind, work = stage_work(dominant, info)
    k_ms, k_cnt = timed.get(dominant, (0.0, 0))
    k_avg_s = (k_ms / max(k_cnt, 1)) / 1000.0
    if kind == "hbm":
        achieved = work / k_avg_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": work}
    else:
        achieved = work / k_avg_s / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / VALU_FP32_PEAK_TFLOPS, "traffic": None, "algorithmic_flops": work}
    roof["kernel"] = STAGE_KERNELS.get(dominant, dominant)
    roof["stage"] = dominant
    roof["traffic"], src = profiled_traffic(dominant)
    if src:
        roof["traffic_unit"] = "bytes per pass"
        roof["traffic_source"] = src
    roof["avg_ms"] = k_avg_s * 1e3
    roof["timing"] = (f"HIP events around each of {k_cnt} passes over {args.steps} direct-launch ticks "
                      f"run after the timed region ({direct_ms:.3f} ms/tick there); dominant = the stage "
                      f"with the most time per tick over 50 stage-timed ticks")

    out = {
        "metric": METRIC,
        "value": ticks_per_s,
        "unit": "ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "points": n, "links": int(wl["links"].size // 2),
                   "space_size": wl["S"], "finest_grid": gf, "levels": L, "ticks_per_step": 1,
                   "tuning": tuning},
        "point_ticks_per_s": ticks_per_s * n,
        "stage_ms_per_tick": {st: v[0] / 50.0 for st, v in breakdown.items()},
        "stages": stage_rooflines(breakdown, info),
        "roofline": roof,
        "link_tiers": link_tiers(wl["links"], n, tuning),
    }
    ctx.close()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    skip = set(filter(None, args.skip.split(",")))
    if "errors" not in skip:
        out["force_error"] = force_error_log(wl, args.error_ticks, 50, 4096, log)
    if "cfg2" not in skip:
        out["cfg2"] = bench_cfg2(args.cfg2_ticks, args.cfg2_every, 4096, log)
    if "batched" not in skip:
        out["batched"] = bench_batched(args.batch_graphs, args.batch_ticks, args.batch_sample, log)
    if "serving" not in skip:
        out["serving"] = bench_serving(args.serve_seconds, args.serve_budget_gib, args.serve_segment,
                                       args.serve_stride, log, args.serve_max_points or None)
    if "cfg0" not in skip:
        out["cfg0"] = bench_cfg0(300, 20)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, workloads.cfg0(), args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
