"""The benchmark configurations' shapes under test, through the C ABI:

- the 4M-point SBM graph with collision (spaceSize 16384): every pass against
  the oracle from the initial state and from a live state 20 ticks in;
- a packed set of 64 graphs drawn from the 2048-graph batch's size law (1k to
  200k points, BA m in 1..4, pyramids of 5 to 8 levels): every pass of three
  graphs at different places of the packing against the oracle, evaluated in
  place, before and after 20 batched ticks; a graph's 20 ticks alone and in
  the set give the same bits;
- the serving loop over a window of the request stream (collision requests
  included): batched results equal the sequential path's bit for bit, and one
  resident job's passes match the oracle mid-run.

Tolerance (per point, vector norm): 1e-5 * spaceSize.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, serving, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(got, want, S, what, idx=None):
    d = got.reshape(-1, 2).astype(np.float64) - want.reshape(-1, 2)
    e = np.hypot(d[:, 0], d[:, 1])
    if idx is not None:
        e = e[idx]
    worst = float(np.nanmax(e)) if e.size else 0.0
    say(f"{what}: {worst / S:.2e} S")
    assert np.isfinite(e).all(), f"{what}: non-finite"
    assert worst <= TOL * S, f"{what}: {worst / S:.3e} S"


def passes_vs_oracle(eval_fn, xy, S, p, alpha, seed, Lk, jitter, sizes=None, idx=None, what=""):
    """eval_fn(force, alpha, seed) -> GPU velocity; every pass the graph runs."""
    absent = O.absent_mask(xy)
    live = absent == 0
    lo, hi = p.link_dist_random_variation_range[0], p.link_dist_random_variation_range[1]
    check(eval_fn(_lib.FORCE_GRAVITY, alpha, 0.0), O.gravity(np.nan_to_num(xy), S, alpha, p.gravity), S,
          f"{what} gravity", live)
    ev = idx if idx is not None else np.flatnonzero(live).astype(np.uint32)
    want = O.manybody(xy, S, alpha, p.repulsion, seed, jitter=jitter, absent=absent, eval_idx=ev)
    check(eval_fn(_lib.FORCE_MANYBODY, alpha, seed), want, S, f"{what} many-body", ev)
    for d in (0, 1):
        check(eval_fn(_lib.FORCE_LINK_IN + d, alpha, 0.0),
              Lk.force(xy, d, alpha, p.link_spring, p.link_distance, lo, hi, absent), S,
              f"{what} link {'IN' if d == 0 else 'OUT'}", live)
    if sizes is not None:
        ms = max(float(p.point_default_size), float(np.max(sizes)))
        check(eval_fn(_lib.FORCE_COLLISION, alpha, 0.0),
              O.collision(xy, sizes, S, alpha, p.collision, p.collision_radius, p.collision_padding,
                          max_size=ms, absent=absent), S, f"{what} collision", live)


def test_cfg2_every_pass_vs_oracle():
    wl = workloads.cfg2()
    n, S, p = wl["n"], wl["S"], wl["params"]
    say(f"cfg2: n={n} m={wl['links'].size // 2} S={S}")
    c = _lib.Context()
    c.set_positions(wl["xy"])
    c.set_jitter(wl["jitter"])
    c.set_links(wl["links"], None, wl["r_in"], wl["r_out"])
    c.set_point_sizes(wl["sizes"])
    Lk = O.Links(n, wl["links"], wl["r_in"], wl["r_out"])
    rng = np.random.default_rng(2)
    for state in ("initial", "after 20 ticks"):
        if state != "initial":
            c.set_alpha(1.0)
            c.step(p, 20)
        xy = c.get_positions()
        a = 1.0 if state == "initial" else float(np.float32(c.get_alpha()))
        # many-body on a sample: 20k random points plus the most crowded finest cells
        g = 512
        cell = O.cells(xy, O.cell_size(S, g), g)
        key = cell[:, 1].astype(np.int64) * g + cell[:, 0]
        crowded = np.argsort(np.bincount(key, minlength=g * g))[-8:]
        idx = np.unique(np.concatenate([rng.choice(n, 20_000, replace=False),
                                        np.flatnonzero(np.isin(key, crowded))])).astype(np.uint32)
        passes_vs_oracle(lambda f, al, sd: c.eval_force(p, f, al, sd), xy, S, p, a, 0.4375, Lk, wl["jitter"],
                         sizes=wl["sizes"], idx=idx, what=f"cfg2 {state}")
    c.close()


def cfg3_pick(count=64):
    """`count` graphs of the 2048-graph batch evenly spread over its sorted sizes."""
    ns, ms = workloads.cfg3_sizes(2048)
    order = np.argsort(ns, kind="stable")
    pick = order[np.linspace(0, 2047, count).round().astype(int)]
    return [(int(i), int(ns[i]), int(ms[i])) for i in pick]


def test_cfg3_packed_set_every_pass_and_alone_equals_mix():
    pick = cfg3_pick(64)
    levels = [O.geometry(n)[1] for _, n, _ in pick]
    assert min(levels) == 5 and max(levels) == 8
    graphs = [workloads.cfg3_graph(i, n, mba) for i, n, mba in pick]
    say(f"cfg3 set: {len(graphs)} graphs, {sum(g['n'] for g in graphs)} points, levels {min(levels)}..{max(levels)}")
    c = _lib.Context()
    ids = []
    for g in graphs:
        gid = c.graph_add()
        ids.append(gid)
        c.graph_set_params(gid, g["params"])
        c.graph_set_positions(gid, g["xy"])
        c.graph_set_jitter(gid, g["jitter"])
        c.graph_set_links(gid, g["links"], None, g["r_in"], g["r_out"])
    checked = (0, 31, 63)
    links = {k: O.Links(graphs[k]["n"], graphs[k]["links"], graphs[k]["r_in"], graphs[k]["r_out"]) for k in checked}
    for state in ("initial", "after 20 ticks"):
        if state != "initial":
            c.step_graphs(ids, 20)
        for k in checked:
            g, gid = graphs[k], ids[k]
            xy = c.graph_get_positions(gid)
            a = float(np.float32(c.graph_get_alpha(gid)))
            passes_vs_oracle(lambda f, al, sd: c.graph_eval_force(gid, f, al, sd), xy, g["S"], g["params"], a,
                             0.5625, links[k], g["jitter"], what=f"cfg3 slot {k} (n={g['n']}) {state}")
    mixed = c.graph_get_positions(ids[31])
    c.close()
    # the same graph's 20 ticks alone
    g = graphs[31]
    with _lib.Context() as solo:
        gid = solo.graph_add()
        solo.graph_set_params(gid, g["params"])
        solo.graph_set_positions(gid, g["xy"])
        solo.graph_set_jitter(gid, g["jitter"])
        solo.graph_set_links(gid, g["links"], None, g["r_in"], g["r_out"])
        solo.step_graphs([gid], 20)
        alone = solo.graph_get_positions(gid)
    assert np.array_equal(alone.view(np.uint32), mixed.view(np.uint32))


def test_cfg4_serving_equals_sequential_with_oracle_mid_run():
    """A window of the request stream with the point counts capped at 40k so the
    test stays short (the full sizes run in bench.py) and a 30% collision share."""
    jobs = workloads.cfg4_jobs(duration=0.5, n_hi=40_000, collision_frac=0.3)
    assert any(j.sizes is not None for j in jobs)
    say(f"cfg4 window: {len(jobs)} jobs, {sum(j.n for j in jobs)} points, "
        f"{sum(j.sizes is not None for j in jobs)} with collision")
    checked = []

    def mid_run(ctx, active):
        if checked:
            return
        col = [gid for gid, j in active.items() if j.sizes is not None]
        gid = col[0] if col else next(iter(active))
        job = active[gid]
        xy = ctx.graph_get_positions(gid)
        a = float(np.float32(ctx.graph_get_alpha(gid)))
        Lk = O.Links(job.n, job.links, job.r_in, job.r_out)
        passes_vs_oracle(lambda f, al, sd: ctx.graph_eval_force(gid, f, al, sd), xy, job.params.space_size,
                         job.params, a, 0.8125, Lk, job.jitter, sizes=job.sizes,
                         what=f"cfg4 job {job.jid} (n={job.n}) mid-run")
        checked.append(job.jid)

    srv = serving.BatchServer(budget=200 * 2**30, segment=100)
    res, st = srv.run(jobs, on_segment=mid_run)
    srv.close()
    seq, _ = serving.serve_sequential(jobs)
    assert checked
    assert [r.jid for r in res] == [r.jid for r in seq]
    for a, b in zip(res, seq):
        assert np.array_equal(a.xy.view(np.uint32), b.xy.view(np.uint32)), f"job {a.jid} differs"
    say(f"cfg4: {len(res)} jobs equal, max active {st['max_active']}, {st['segments']} segments")
