"""Layout serving: continuous batching of layout jobs over one device.

Every job is one graph laid out from alpha = 1 until alpha < ALPHA_MIN, the
point where the reference's frame loop ends the simulation
(src/index.ts:2164-2169, Store/index.ts:8).  ``BatchServer`` keeps every admitted
job resident in one context's graph set (positions, velocities and alpha stay
on the device between ticks), advances all of them together, one launch per
force pass per tick, ingests arrivals while the HBM budget allows (also while
a segment runs: their link builds go on the library's ingest stream), adds
them to the set at the next segment boundary, and retires each job on the tick
its alpha crosses the cutoff.
``serve_sequential`` is the plain path: one job at a time through the
single-graph entry points, in arrival order.

A job computes the same bits in either path (the graph set packs graphs
without changing any graph's arithmetic), so the two differ only in time.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

ALPHA_MIN = 0.001


@dataclass
class Job:
    """One layout request: the graph, its parameters and its arrival time (s)."""
    jid: int
    arrival: float
    xy: np.ndarray
    links: np.ndarray | None
    params: _lib.CgSimParams
    r_in: np.ndarray | None = None
    r_out: np.ndarray | None = None
    jitter: np.ndarray | None = None
    sizes: np.ndarray | None = None
    ticks: int | None = None   # ticks to run; None = until alpha < ALPHA_MIN

    @property
    def n(self) -> int:
        return self.xy.size // 2

    @property
    def m(self) -> int:
        return 0 if self.links is None else self.links.size // 2


@dataclass
class Result:
    jid: int
    arrival: float
    start: float     # admitted (batched) or started (sequential)
    finish: float
    ticks: int
    xy: np.ndarray = field(repr=False)

    @property
    def latency(self) -> float:
        return self.finish - self.arrival

    @property
    def queue_wait(self) -> float:
        return self.start - self.arrival

    @property
    def service(self) -> float:
        return self.finish - self.start


def finest_grid(n: int) -> int:
    """ForceManyBody/index.ts:646-656: smallest power of two >= 2 sqrt(n), in [8, 512]."""
    p = 1
    while (p // 2) * (p // 2) < n:
        p <<= 1
    return min(512, max(8, p))


def footprint_bytes(n: int, m: int, collision: bool) -> int:
    """Device bytes one job holds in a packed set, with the transient copy a
    repack makes (positions and link schedules exist twice while it runs)."""
    gf = finest_grid(n)
    cells = gf * gf
    per_point = 16 + 1 + 4 + 8 + 4 + 4 + 16 + 8 + 32 + 13 + 6  # positions x2, flags, gid, jitter, size,
    #        present, keys/values x2, sorted, link schedule, staged flags/jitter/present, ranks and
    #        the cell pass's partials
    per_cell = 8 + 96 + 4 + 54                          # bounds, 8 slots, counts, levels
    per_link = 2 * (16 + 16 + 16 + 16) + 24             # CSR + schedule + unit/warm state + hub terms, build
    col = (n * 40 + 4 * 512 * 512 * 24) if collision else 0   # sort keys/values/positions x2, sizes x2, cells
    return int(1.3 * (per_point * n + per_cell * cells + per_link * m + col)) + (1 << 20)


def collides(p: _lib.CgSimParams) -> bool:
    """Whether the collision pass runs for these parameters: simulationCollision
    is truthy (src/index.ts:2066), as the library decides it."""
    return p.collision != 0 and not math.isnan(p.collision)


def alpha_ticks(decay: float, alpha0: float = 1.0) -> int:
    """Ticks until alpha < ALPHA_MIN under Store.addAlpha (Store/index.ts:328-332)."""
    f = 1.0 - math.pow(ALPHA_MIN, 1.0 / decay)
    a, t = alpha0, 0
    while a >= ALPHA_MIN and t < 10_000_000:
        a += (0.0 - a) * f
        t += 1
    return t


def _load(ctx: _lib.Context, gid: int, job: Job) -> None:
    ctx.graph_set_params(gid, job.params)
    ctx.graph_set_positions(gid, job.xy)
    if job.jitter is not None:
        ctx.graph_set_jitter(gid, job.jitter)
    if job.links is not None:
        ctx.graph_set_links(gid, job.links, None, job.r_in, job.r_out)
    if job.sizes is not None:
        ctx.graph_set_point_sizes(gid, job.sizes)
    ctx.graph_set_alpha(gid, 1.0)


class BatchServer:
    """Continuous batching over one context.

    ``segment`` caps the ticks between two admission points; a segment also
    ends on the tick the first active job reaches its cutoff, so no job runs
    past it.  ``budget`` is the HBM budget admission works against."""

    def __init__(self, budget: int, segment: int = 100, device: int = 0, tuning: dict | None = None,
                 max_points: int | None = None):
        self.budget = int(budget)
        self.segment = int(segment)
        self.max_points = max_points  # resident points cap (None: the HBM budget alone)
        self.ctx = _lib.Context(device=device, hbm_budget=self.budget)
        if tuning:
            self.ctx.set_tuning(tuning)
        self._cancel: set[int] = set()

    def close(self) -> None:
        self.ctx.close()

    def cancel(self, jid: int) -> None:
        """Withdraws job `jid`: a waiting job is never admitted, a resident one
        leaves the device at the next segment boundary.  Either way it has no
        result and is listed in the run's stats["cancelled"]."""
        self._cancel.add(int(jid))

    def run(self, jobs: list[Job], log=None, on_segment=None) -> tuple[list[Result], dict]:
        """Serves `jobs` (arrival times relative to the call).  on_segment(ctx,
        {graph id: job}) runs after every segment with the jobs it advanced,
        before retirements."""
        ctx = self.ctx
        pending = sorted(jobs, key=lambda j: j.arrival)
        active: dict[int, tuple[Job, int, int, float, int]] = {}  # gid -> job, left, total, start, bytes
        results: list[Result] = []
        stats = dict(segments=0, ticks=0, max_active=0, point_ticks=0, rejected=[], cancelled=[],
                     ingest_s=0.0, pack_s=0.0, tick_s=0.0, retire_s=0.0, idle_s=0.0, repacks=0,
                     overlap_s=0.0)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        held = 0  # bytes of the admitted jobs' footprints

        def admit(t: float) -> bool:
            """Ingests every job arrived by `t` that fits, in arrival order (no
            overtaking); it joins the set at the next segment boundary."""
            nonlocal held
            got = False
            while pending and pending[0].arrival <= t:
                job = pending[0]
                need = footprint_bytes(job.n, job.m, collides(job.params))
                if need > self.budget:
                    # could never fit, even alone: refused, not run
                    pending.pop(0)
                    stats["rejected"].append(job.jid)
                    if log:
                        log(f"job {job.jid} n={job.n} refused: footprint {need} B exceeds the budget")
                    continue
                if held + need > self.budget:
                    break
                if self.max_points and active and sum(v[0].n for v in active.values()) + job.n > self.max_points:
                    break
                pending.pop(0)
                ti = time.perf_counter()
                gid = ctx.graph_add()
                _load(ctx, gid, job)
                stats["ingest_s"] += time.perf_counter() - ti
                total = job.ticks if job.ticks is not None else alpha_ticks(job.params.decay)
                active[gid] = (job, total, total, t, need)
                held += need
                got = True
            return got

        changed = True
        while pending or active:
            t = now()
            if self._cancel:
                for gid in [g for g, v in active.items() if v[0].jid in self._cancel]:
                    ctx.graph_remove(gid)
                    held -= active[gid][4]
                    stats["cancelled"].append(active.pop(gid)[0].jid)
                    changed = True
                for j in [j for j in pending if j.jid in self._cancel]:
                    pending.remove(j)
                    stats["cancelled"].append(j.jid)
            changed |= admit(t)
            if not active:
                if not pending:
                    break
                # idle until the next arrival
                ti = time.perf_counter()
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                stats["idle_s"] += time.perf_counter() - ti
                continue
            stats["max_active"] = max(stats["max_active"], len(active))
            k = min(self.segment, min(v[1] for v in active.values()))
            ids = list(active)
            if changed:
                # the set changed: repack now (no tick), so the time is its own
                ti = time.perf_counter()
                ctx.step_graphs(ids, 0)
                ctx.sync()
                stats["pack_s"] += time.perf_counter() - ti
                stats["repacks"] += 1
                changed = False
            ti = time.perf_counter()
            ctx.step_graphs(ids, k)
            # while the segment runs: ingest arrivals (their link builds run on
            # the library's ingest stream beside the ticks)
            tb = time.perf_counter()
            changed |= admit(now())
            stats["overlap_s"] += time.perf_counter() - tb
            ctx.sync()
            stats["tick_s"] += time.perf_counter() - ti
            stats["segments"] += 1
            stats["ticks"] += k
            stats["point_ticks"] += k * sum(active[g][0].n for g in ids)
            if on_segment:
                on_segment(ctx, {gid: active[gid][0] for gid in ids})
            ti = time.perf_counter()
            for gid in ids:
                job, left, total, start, need = active[gid]
                left -= k
                if left > 0:
                    active[gid] = (job, left, total, start, need)
                    continue
                xy = ctx.graph_get_positions(gid)
                fin = now()
                ctx.graph_remove(gid)
                del active[gid]
                held -= need
                changed = True
                results.append(Result(job.jid, job.arrival, start, fin, total, xy))
                if log:
                    log(f"job {job.jid} n={job.n} done at {fin:.2f}s latency {fin - job.arrival:.2f}s")
            stats["retire_s"] += time.perf_counter() - ti
        stats["wall_s"] = now()
        return sorted(results, key=lambda r: r.jid), stats


def serve_sequential(jobs: list[Job], device: int = 0, log=None, tuned: bool = False) -> tuple[list[Result], dict]:
    """The plain path: each job alone through the single-graph entry points,
    first come first served (tuned: each job's context takes the bit-neutral
    tuning of its own size from the autotuner's table).  A job's service time is measured from context
    creation (ingest and link build included, as the batched server's loop
    includes them) to its readback; a job starts when it has arrived and the
    previous one has finished."""
    results = []
    clock = 0.0
    busy = 0.0
    for job in sorted(jobs, key=lambda j: j.arrival):
        s0 = time.perf_counter()
        with _lib.Context(device=device) as ctx:
            if tuned:
                from . import autotune

                autotune.apply(ctx, job.n, 2.0 * job.m / max(job.n, 1))
            ctx.set_positions(job.xy)
            if job.jitter is not None:
                ctx.set_jitter(job.jitter)
            if job.links is not None:
                ctx.set_links(job.links, None, job.r_in, job.r_out)
            if job.sizes is not None:
                ctx.set_point_sizes(job.sizes)
            ctx.set_alpha(1.0)
            total = job.ticks if job.ticks is not None else alpha_ticks(job.params.decay)
            ctx.step(job.params, total)
            xy = ctx.get_positions()
        service = time.perf_counter() - s0
        start = max(clock, job.arrival)
        clock = start + service
        busy += service
        results.append(Result(job.jid, job.arrival, start, clock, total, xy))
        if log:
            log(f"sequential job {job.jid} n={job.n} service {service:.3f}s")
    return sorted(results, key=lambda r: r.jid), dict(busy_s=busy, makespan_s=clock)


def latency_summary(results: list[Result]) -> dict:
    """Completion latency (arrival to result), split into queue wait (arrival to
    admission) and service (admission to result); layouts/s over the makespan."""
    lat = np.array([r.latency for r in results], dtype=np.float64)
    wait = np.array([r.queue_wait for r in results], dtype=np.float64)
    serv = np.array([r.service for r in results], dtype=np.float64)
    fin = max(r.finish for r in results) - min(r.arrival for r in results)
    pct = lambda a, q: float(np.percentile(a, q))
    return dict(jobs=len(results), p50_s=pct(lat, 50), p99_s=pct(lat, 99), mean_s=float(lat.mean()),
                queue_wait_p50_s=pct(wait, 50), queue_wait_p99_s=pct(wait, 99),
                service_p50_s=pct(serv, 50), service_p99_s=pct(serv, 99),
                makespan_s=float(fin), layouts_per_s=len(results) / fin)
