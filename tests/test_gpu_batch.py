"""GPU tests of the packed graph set (batched layouts, continuous batching).

A graph's result must not depend on what else shares the device: every kernel
looks its graph up and works in graph-local order, so the check is bit-exact
(positions compared as raw f32 bits, NaNs included), not a tolerance.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def spec(kind: str, seed: int) -> dict:
    """Small graphs covering every pass and the link tiers."""
    if kind == "ba":
        n, S = 3000, 2048.0
        links = datagen.barabasi_albert(n, 2, seed)
        p = workloads.sim_params(S)
        sizes = None
    elif kind == "hubs":  # chung-lu: lane, 64-lane and 256-lane link tiers
        n, S = 20000, 4096.0
        links = datagen.chung_lu(n, 2.1, 1.0, 5 * n, seed)
        p = workloads.sim_params(S, center=0.3)
        sizes = None
    elif kind == "collide":
        n, S = 2500, 1024.0
        links = datagen.barabasi_albert(n, 1, seed)
        p = workloads.sim_params(S, collision=1.0)
        sizes = (2.0 * datagen.lognormal(n, np.log(2.0), 0.5, seed)).astype(np.float32)
    elif kind == "nolinks":  # and clusters (centroid and custom targets)
        n, S = 1500, 1024.0
        links = None
        p = workloads.sim_params(S, gravity=0.0, cluster=0.2)
        sizes = None
    else:
        raise ValueError(kind)
    xy = datagen.uniform_square(n, S, seed)
    if kind == "ba":
        xy = xy.reshape(-1, 2).copy()
        xy[::97] = np.nan  # absent points
        xy[5::97, 1] = np.nan  # half-NaN: absent too (Points/position-utils.ts:17-24)
    r_in, r_out = (None, None) if links is None else workloads._draws(links.size // 2, seed)
    clusters = None
    if kind == "nolinks":
        clusters = dict(index=(np.arange(n) % 7 - 1).astype(np.int32),
                        positions=np.array([100.0, 900.0, -1.0, 0.0], np.float32))
    return dict(kind=kind, n=n, xy=xy, links=links, r_in=r_in, r_out=r_out, params=p, sizes=sizes,
                pinned=np.arange(0, n, 211, dtype=np.uint32), clusters=clusters)


def load(ctx: _lib.Context, s: dict) -> int:
    gid = ctx.graph_add()
    ctx.graph_set_params(gid, s["params"])
    ctx.graph_set_positions(gid, s["xy"])
    if s["links"] is not None:
        ctx.graph_set_links(gid, s["links"], None, s["r_in"], s["r_out"])
    if s["sizes"] is not None:
        ctx.graph_set_point_sizes(gid, s["sizes"])
    ctx.graph_set_pinned(gid, s["pinned"])
    if s["clusters"] is not None:
        ctx.set_point_clusters(s["clusters"]["index"], gid=gid)
        ctx.set_cluster_positions(s["clusters"]["positions"], gid=gid)
    return gid


def alone(s: dict, ticks: int) -> np.ndarray:
    with _lib.Context() as ctx:
        gid = load(ctx, s)
        ctx.step_graphs([gid], ticks)
        return ctx.graph_get_positions(gid)


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_alone_equals_continuous_mix():
    kinds = ["ba", "hubs", "collide", "nolinks", "ba"]
    specs = [spec(k, 100 + i) for i, k in enumerate(kinds)]
    # schedule: (tick block, graphs admitted before it, graphs removed after it)
    with _lib.Context() as ctx:
        ids, done, ticks_run = {}, {}, {}
        say("admit 0,1")
        for i in (0, 1):
            ids[i] = load(ctx, specs[i])
            ticks_run[i] = 0
        for block, admit, leave in ((20, (2,), ()), (15, (3,), (0,)), (25, (4,), ()), (10, (), (1, 2, 3, 4))):
            live = [i for i in ids if i not in done]
            ctx.step_graphs([ids[i] for i in live], block)
            for i in live:
                ticks_run[i] += block
            for i in leave:
                done[i] = ctx.graph_get_positions(ids[i])
                ctx.graph_remove(ids[i])
            for i in admit:
                say(f"admit {i} after a block of {block}")
                ids[i] = load(ctx, specs[i])
                ticks_run[i] = 0
    for i, s in enumerate(specs):
        ref = alone(s, ticks_run[i])
        same = np.array_equal(bits(ref), bits(done[i]))
        say(f"graph {i} ({s['kind']}, n={s['n']}, {ticks_run[i]} ticks): bit-identical={same}")
        assert same, f"graph {i}: mixed result differs from the graph run alone"
        present = ~np.isnan(s["xy"].reshape(-1, 2)).any(axis=1)
        assert np.isfinite(ref.reshape(-1, 2)[present]).all()
        assert np.isnan(ref.reshape(-1, 2)[~present]).all()


def test_graph_without_present_points_in_a_mix():
    """A graph whose points are all absent (no sort segment, no link units)
    between two others: they still match their runs alone bit for bit, and it
    stays all NaN."""
    a, b = spec("ba", 11), spec("hubs", 12)
    empty = spec("ba", 13)
    empty["xy"] = np.full_like(empty["xy"], np.nan)
    with _lib.Context() as ctx:
        ga, ge, gb = load(ctx, a), load(ctx, empty), load(ctx, b)
        ctx.step_graphs([ga, ge, gb], 20)
        xa, xe, xb = (ctx.graph_get_positions(g) for g in (ga, ge, gb))
    assert np.isnan(xe).all()
    for s, x in ((a, xa), (b, xb)):
        same = np.array_equal(bits(alone(s, 20)), bits(x))
        say(f"{s['kind']}: bit-identical={same}")
        assert same


def tiny(n: int, seed: int) -> dict:
    """A graph of n points (1 .. 12): a chain of links from 2 points on, every
    third one with collision, one with a pinned point."""
    S = 64.0
    links = None
    if n >= 2:
        links = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.float32).reshape(-1)
    p = workloads.sim_params(S, collision=1.0 if seed % 3 == 0 else 0.0)
    r_in, r_out = (None, None) if links is None else workloads._draws(links.size // 2, seed)
    xy = datagen.uniform_square(n, S, seed)
    return dict(kind=f"tiny{n}", n=n, xy=xy, links=links, r_in=r_in, r_out=r_out, params=p, sizes=None,
                pinned=np.array([0], np.uint32) if seed % 5 == 0 else np.zeros(0, np.uint32), clusters=None)


def test_many_tiny_graphs_in_a_mix():
    """Forty graphs of 1 to 12 points packed around two larger ones: lanes of
    the cell pass then hold points of three or more graphs, cell runs cross
    wave ends next to graph ends, and the collision cells of tiny graphs come
    from their few occupied base cells.  Every graph matches its run alone
    bit for bit."""
    specs = [spec("ba", 31)] + [tiny(1 + i % 12, 40 + i) for i in range(40)] + [spec("collide", 32)]
    with _lib.Context() as ctx:
        ids = [load(ctx, s) for s in specs]
        ctx.step_graphs(ids, 15)
        got = [ctx.graph_get_positions(g) for g in ids]
    for s, x in zip(specs, got):
        same = np.array_equal(bits(alone(s, 15)), bits(x))
        if not same:
            say(f"{s['kind']} (n={s['n']}): differs from its run alone")
        assert same, f"{s['kind']}: mixed result differs from the graph run alone"
    say(f"{len(specs)} graphs bit-identical to their runs alone")


def test_unstepped_graphs_keep_state_and_alpha():
    a, b = spec("ba", 7), spec("nolinks", 8)
    with _lib.Context() as ctx:
        ga, gb = load(ctx, a), load(ctx, b)
        ctx.step_graphs([ga, gb], 5)
        xb = ctx.graph_get_positions(gb)
        alpha_b = ctx.graph_get_alpha(gb)
        ctx.step_graphs([ga], 12)
        assert np.array_equal(bits(xb), bits(ctx.graph_get_positions(gb)))
        assert ctx.graph_get_alpha(gb) == alpha_b
        assert ctx.graph_get_alpha(ga) < alpha_b


def test_default_graph_api_matches_graph_set_api():
    s = spec("hubs", 3)
    with _lib.Context() as ctx:
        ctx.set_positions(s["xy"])
        ctx.set_links(s["links"], None, s["r_in"], s["r_out"])
        ctx.set_pinned(s["pinned"])
        ctx.step(s["params"], 30)
        x0 = ctx.get_positions()
    x1 = alone(s, 30)
    assert np.array_equal(bits(x0), bits(x1))


def test_captured_ticks_match_direct_launches():
    """Tick sequences replayed from captured HIP graphs give the bits of the
    same kernels launched one by one; checked after each of several calls,
    with eager work (readback, a force evaluation) between the replays."""
    a, b = spec("hubs", 11), spec("collide", 12)
    runs = {}
    for capture in (True, False):
        with _lib.Context() as ctx:
            ctx.set_capture(capture)
            ga, gb = load(ctx, a), load(ctx, b)
            out = []
            for ticks in (13, 8, 21, 1):
                ctx.step_graphs([ga, gb], ticks)
                out.append((ctx.graph_get_positions(ga), ctx.graph_get_positions(gb)))
            runs[capture] = out
    for k, ((xa, xb), (ya, yb)) in enumerate(zip(runs[True], runs[False])):
        same = np.array_equal(bits(xa), bits(ya)) and np.array_equal(bits(xb), bits(yb))
        say(f"call {k}: captured == direct: {same}")
        assert same


def test_drag_and_tracked_positions():
    s = spec("ba", 21)  # has absent points (every 97th)
    with _lib.Context() as ctx:
        ctx.set_positions(s["xy"])
        ctx.set_links(s["links"], None, s["r_in"], s["r_out"])
        ctx.step(s["params"], 3)
        # Points.drag: a dragged point takes the mouse position, unclamped; absent stays absent
        idx = np.array([6, 97, 1234], np.uint32)  # 97 is absent
        new = np.array([[-10.0, 3000.0], [1.0, 2.0], [500.5, 600.25]], np.float32)
        ctx.drag_points(idx, new)
        got = ctx.get_positions_at(np.array([6, 97, 1234, s["n"] + 3], np.uint32)).reshape(-1, 2)
        assert np.array_equal(got[0], new[0]) and np.array_equal(got[2], new[2])
        assert np.isnan(got[1]).all() and np.isnan(got[3]).all()
        # the next tick starts from the dragged positions; tracking agrees with the full readback
        ctx.step(s["params"], 2)
        full = ctx.get_positions().reshape(-1, 2)
        track = ctx.get_positions_at(np.array([6, 1234, 7], np.uint32)).reshape(-1, 2)
        assert np.array_equal(bits(track), bits(full[[6, 1234, 7]]))
        assert 0.0 <= full[6, 0] <= s["params"].space_size  # integrated and clamped after the drag
