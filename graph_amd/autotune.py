"""Autotuner of the tick's free implementation parameters (cg_tuning).

The parameters change how the work is laid out on the device, never the
discretisation: the pyramid depth, the cell formula, the slot rule and every
force formula are fixed by the reference (ForceManyBody/index.ts:646-656 fixes
the pyramid depth; nothing here touches it).  Two classes:

- bit-neutral (BIT_NEUTRAL): ticks per captured HIP graph, the crowded-cell
  tier bounds of the cell pass (integer sums and a set-valued slot rule: the
  same bits in any tier), and the force kernel's LDS window (the same values
  summed in the same order, staged or not);
- link tiers (LINK): the degree bounds of the one-lane and 64-lane link tiers
  and the entries per lane of 256-lane units.  They move chains between the
  serial loop and the multiple-shooting solver, so results change within the
  force tolerance but not bit for bit.

``tune`` measures ms per tick of a workload for candidate values, one
parameter at a time (coordinate descent from the defaults), each candidate
timed back to back with the current choice so that the layout's slow drift
does not bias the comparison.  The table it writes (``tuning.json`` beside
this file) is keyed on the workload's present points and mean degree (2 x
links / points); ``lookup`` returns the entry nearest in (log2 points, log2
degree).  ``apply`` sets a context's tuning from the table; by default only
the bit-neutral part, so a tuned run keeps the bits of an untuned one.

CLI (on the GPU):  python -m graph_amd.autotune --out tuning.json
"""
from __future__ import annotations

import argparse
import json
import math
import time
from pathlib import Path

import numpy as np

from . import _lib

TABLE = Path(__file__).resolve().parent / "tuning.json"

BIT_NEUTRAL = {
    "graph_ticks": [4, 8, 16, 32],
    "cell_items": [4, 8],
    "force_tile": [512, 1024],
}
LINK = {
    "link_lane_max": [32, 48, 64, 96],
    "link_group_max": [128, 192, 256],
    "link_hub_entries": [4, 6, 8],
}


def valid(t: dict) -> bool:
    """The constraints cg_set_tuning enforces."""
    return (t["cell_items"] in (4, 8) and t["force_tile"] in (512, 1024)
            and 2 <= t["link_lane_max"] < t["link_group_max"] <= 256)


def key(points: int, mean_degree: float) -> tuple[float, float]:
    return math.log2(max(points, 1)), math.log2(max(mean_degree, 0.5))


def load_table(path: Path = TABLE) -> dict:
    if not path.exists():
        return {"entries": []}
    return json.loads(path.read_text())


def lookup(points: int, mean_degree: float, bit_exact: bool = True, table: dict | None = None) -> dict:
    """Tuning of the table entry nearest to (points, mean degree); with
    bit_exact only its bit-neutral parameters (the rest at their defaults)."""
    table = load_table() if table is None else table
    t = _lib.default_tuning()
    if not table.get("entries"):
        return t
    kx, ky = key(points, mean_degree)
    best = min(table["entries"], key=lambda e: (key(e["points"], e["mean_degree"])[0] - kx) ** 2
               + (key(e["points"], e["mean_degree"])[1] - ky) ** 2)
    for k, v in best["tuning"].items():
        if k in BIT_NEUTRAL or not bit_exact:
            t[k] = int(v)
    return t


def apply(ctx: _lib.Context, points: int, mean_degree: float, bit_exact: bool = True) -> dict:
    t = lookup(points, mean_degree, bit_exact)
    ctx.set_tuning(t)
    return t


class Workload:
    """A context holding one workload: `step(k)` advances it k ticks."""

    def __init__(self, name: str, graphs: list[dict], single: bool):
        self.name = name
        self.ctx = _lib.Context(device=0)
        self.single = single
        self.points = sum(g["n"] for g in graphs)
        self.links = sum(g["links"].size // 2 for g in graphs)
        if single:
            g = graphs[0]
            self.params = g["params"]
            self.ctx.set_positions(g["xy"])
            self.ctx.set_jitter(g["jitter"])
            self.ctx.set_links(g["links"], None, g["r_in"], g["r_out"])
            if g.get("sizes") is not None:
                self.ctx.set_point_sizes(g["sizes"])
            self.ctx.set_alpha(1.0)
            self.ids = None
        else:
            self.ids = []
            for g in graphs:
                gid = self.ctx.graph_add()
                self.ctx.graph_set_params(gid, g["params"])
                self.ctx.graph_set_positions(gid, g["xy"])
                self.ctx.graph_set_jitter(gid, g["jitter"])
                self.ctx.graph_set_links(gid, g["links"], None, g["r_in"], g["r_out"])
                if g.get("sizes") is not None:
                    self.ctx.graph_set_point_sizes(gid, g["sizes"])
                self.ids.append(gid)

    @property
    def mean_degree(self) -> float:
        return 2.0 * self.links / max(self.points, 1)

    def step(self, k: int) -> None:
        if self.single:
            self.ctx.step(self.params, k)
        else:
            self.ctx.step_graphs(self.ids, k)

    def ms_per_tick(self, tuning: dict, ticks: int) -> float:
        self.ctx.set_tuning(tuning)
        self.step(0)          # repack for the new layout parameters (not timed)
        self.step(min(ticks, tuning["graph_ticks"] + 1))  # capture the tick graphs (not timed)
        self.ctx.sync()
        t0 = time.perf_counter()
        self.step(ticks)
        self.ctx.sync()
        return (time.perf_counter() - t0) / ticks * 1e3

    def close(self) -> None:
        self.ctx.close()


def tune(w: Workload, ticks: int, space: dict, rounds: int = 1, log=print) -> dict:
    """Coordinate descent over `space` from the defaults; a candidate replaces
    the current value only when faster than it measured just before and just
    after (by 2%)."""
    cur = _lib.default_tuning()
    base = w.ms_per_tick(cur, ticks)
    log(f"{w.name}: default {base:.4f} ms/tick")
    for _ in range(rounds):
        for name, values in space.items():
            for v in values:
                if v == cur[name]:
                    continue
                cand = dict(cur, **{name: v})
                if not valid(cand):
                    continue
                a = w.ms_per_tick(cur, ticks)
                b = w.ms_per_tick(cand, ticks)
                a2 = w.ms_per_tick(cur, ticks)
                if b < 0.98 * min(a, a2):
                    log(f"{w.name}: {name} {cur[name]} -> {v}: {min(a, a2):.4f} -> {b:.4f} ms/tick")
                    cur = cand
                else:
                    log(f"{w.name}: {name} {v} kept {cur[name]}: {min(a, a2):.4f} vs {b:.4f} ms/tick")
    return cur


def measure(w: Workload, tuning: dict, ticks: int, reps: int = 3) -> float:
    return float(np.median([w.ms_per_tick(tuning, ticks) for _ in range(reps)]))


def workloads(names: list[str]):
    """The benchmark workloads as (name, graphs, single-graph?, ticks per measurement)."""
    from . import datagen, workloads as W

    for name in names:
        if name == "cfg0":
            yield name, [W.cfg0()], True, 60
        elif name == "cfg1":
            yield name, [W.cfg1()], True, 40
        elif name == "cfg2":
            yield name, [W.cfg2()], True, 20
        elif name == "cfg3":
            ns, ms = W.cfg3_sizes(2048)
            yield name, [W.cfg3_graph(i, int(ns[i]), int(ms[i])) for i in range(2048)], False, 8
        elif name == "cfg4":
            jobs = W.cfg4_jobs(duration=5.0)
            graphs = [dict(n=j.n, xy=j.xy, links=j.links, r_in=j.r_in, r_out=j.r_out, jitter=j.jitter,
                           sizes=j.sizes, params=j.params) for j in jobs]
            yield name, graphs, False, 10
        else:
            raise ValueError(name)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--workloads", default="cfg0,cfg1,cfg2,cfg3,cfg4")
    ap.add_argument("--link", action="store_true", help="also tune the link tiers (not bit-neutral)")
    args = ap.parse_args()
    log = lambda *a: print(*a, flush=True)
    entries = []
    for name, graphs, single, ticks in workloads(args.workloads.split(",")):
        w = Workload(name, graphs, single)
        space = dict(BIT_NEUTRAL, **(LINK if args.link else {}))
        best = tune(w, ticks, space, log=log)
        bits = {k: v for k, v in best.items() if k in BIT_NEUTRAL}
        t_def = measure(w, _lib.default_tuning(), ticks)
        t_bits = measure(w, dict(_lib.default_tuning(), **bits), ticks)
        t_all = measure(w, best, ticks) if args.link else t_bits
        log(f"{name}: default {t_def:.4f}, tuned (bit-neutral) {t_bits:.4f}, tuned (all) {t_all:.4f} ms/tick")
        entries.append({"workload": name, "points": w.points, "mean_degree": w.mean_degree,
                        "tuning": best, "ms_per_tick": {"default": t_def, "bit_neutral": t_bits, "all": t_all},
                        "ticks_per_measurement": ticks})
        w.close()
    Path(args.out).write_text(json.dumps({"device": "MI355X", "entries": entries}, indent=1) + "\n")
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
