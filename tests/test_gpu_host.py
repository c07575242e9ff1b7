"""The host side above the C ABI on the GPU: the Python mirror of the
reference's Graph simulation API (graph_amd/simulation.py) against the oracle,
its lifecycle and callbacks (src/index.ts:1457-1540, 2006-2099, 2152-2177,
2335-2341), and the failure paths a caller can force from the host: empty
graphs, no links, odd-length inputs, a point-count change between steps, an
HBM budget refusal and a cancelled serving job.
"""
import math
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen, serving, simulation, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg0_config(**over):
    c = {"spaceSize": 4096, "simulationRepulsion": 0.5, "simulationGravity": 0.25, "simulationFriction": 0.85,
         "simulationLinkSpring": 1.0, "simulationLinkDistance": 10, "simulationDecay": 5000, "randomSeed": 42}
    c.update(over)
    return c


def mirror_draws(n, links, seed=42):
    """The mirror's random draws in its documented order (jitter over the
    texture's texels, IN then OUT rest lengths per CSR entry, then one seed per
    tick), replayed from the same generator."""
    rng = np.random.default_rng(seed)
    tex = math.ceil(math.sqrt(n))
    jit = (rng.uniform(-1, 1, size=2 * tex * tex) * 0.00001).astype(np.float32)[: 2 * n]
    m = links.size // 2
    s, t = links[0::2], links[1::2]
    valid = (s >= 0) & (t >= 0) & (np.floor(s) == s) & (np.floor(t) == t) & (s < n) & (t < n)
    idx = np.nonzero(valid)[0]
    r_in, r_out = np.zeros(m, np.float32), np.zeros(m, np.float32)
    r_in[idx[np.argsort(s[idx], kind="stable")]] = rng.uniform(0, 1, size=idx.size)
    r_out[idx[np.argsort(t[idx], kind="stable")]] = rng.uniform(0, 1, size=idx.size)
    return jit, r_in, r_out, rng


def test_graph_mirror_ticks_vs_oracle():
    """render -> start -> step x4 through the mirror at cfg0 (gravity off, so
    each tick's first pass starts from the state read back): every tick within
    1e-5 S of the oracle's composed passes from the same state, with the
    mirror's own draws, and onSimulationTick reporting the oracle's alpha."""
    wl = workloads.cfg0()
    n, S = wl["n"], wl["S"]
    ticks = []
    g = simulation.Graph(cfg0_config(simulationGravity=0, onSimulationTick=lambda a, i, p: ticks.append(a)))
    g.setPointPositions(wl["xy"])
    g.setLinks(wl["links"])
    g.render()
    g.start()
    jit, r_in, r_out, rng = mirror_draws(n, wl["links"])
    Lk = O.Links(n, wl["links"], r_in, r_out)
    f32 = lambda q: q.astype(np.float32).astype(np.float64)
    a = 1.0
    for t in range(4):
        before = g.get_point_positions_array().astype(np.float64)
        seed = float(rng.uniform(0, 1, size=1).astype(np.float32)[0])
        g.step()
        got = g.get_point_positions_array().reshape(-1, 2).astype(np.float64)
        af = float(np.float32(a))
        pos = f32(O.integrate(before, O.manybody(before, S, af, 0.5, seed, jitter=jit), 0.85, S))
        for d in (0, 1):
            pos = f32(O.integrate(pos, Lk.force(pos, d, af, 1.0, 10.0, 1.0, 1.2), 0.85, S))
        e = np.hypot(*(got - pos.reshape(-1, 2)).T).max()
        say(f"mirror tick {t}: {e / S:.2e} S")
        assert e <= TOL * S
        a = O.alpha_step(a, 5000)
        assert ticks[-1] == a
    assert len(ticks) == 4
    assert g.progress == pytest.approx(math.sqrt(min(1.0, 0.001 / a)))
    g.destroy()


def test_graph_mirror_layout_quality_cfg0():
    """300 frames of cfg0 through the mirror's frame loop against the oracle's
    300 ticks with the same draws: edge-length percentiles and stress within 2%
    (trajectories diverge chaotically, the layouts' statistics do not)."""
    wl = workloads.cfg0()
    n, S = wl["n"], wl["S"]
    g = simulation.Graph(cfg0_config())
    g.setPointPositions(wl["xy"])
    g.setLinks(wl["links"])
    g.render()
    g.start()
    for _ in range(300):
        g.frame()
    gpu = g.get_point_positions_array()
    g.destroy()
    jit, r_in, r_out, rng = mirror_draws(n, wl["links"])
    seeds = [float(rng.uniform(0, 1, size=1).astype(np.float32)[0]) for _ in range(300)]
    pd = dict(space_size=S, gravity=0.25, center=0, repulsion=0.5, link_spring=1.0, link_distance=10.0,
              lo=1.0, hi=1.2, friction=0.85, decay=5000.0)
    Lk = O.Links(n, wl["links"], r_in, r_out)
    pos, a = wl["xy"].astype(np.float64), 1.0
    for t in range(300):
        pos, a = O.tick(pos, a, seeds[t], pd, links=Lk, jitter=jit)

    def stats(xy):
        P = xy.reshape(-1, 2).astype(np.float64)
        s, t = wl["links"][0::2].astype(np.int64), wl["links"][1::2].astype(np.int64)
        L = np.hypot(*(P[s] - P[t]).T)
        D = (r_in.astype(np.float64) * 0.2 + 1.0) * 10.0
        return np.percentile(L, [10, 50, 90]), float(np.sum((L - D) ** 2) / np.sum(D**2))

    q_g, s_g = stats(gpu)
    q_o, s_o = stats(pos)
    say(f"mirror layout: p10/50/90 {q_g} vs {q_o}; stress {s_g:.4f} vs {s_o:.4f}")
    np.testing.assert_allclose(q_g, q_o, rtol=0.02)
    assert abs(s_g - s_o) <= 0.02 * s_o


def test_graph_mirror_lifecycle_and_callbacks():
    events = []
    cb = {k: (lambda k=k: events.append(k)) for k in ("onSimulationStart", "onSimulationEnd",
                                                          "onSimulationPause", "onSimulationUnpause")}
    g = simulation.Graph(cfg0_config(**cb))
    g.start()                      # no points yet: a no-op
    xy = datagen.uniform_square(2000, 4096.0, 3)
    xy[10:12] = np.nan             # point 5 absent
    g.setPointPositions(np.concatenate([xy, [7.0]]))  # odd length: the trailing value is ignored
    g.setLinks(np.concatenate([datagen.barabasi_albert(2000, 2, 3), [1.0]]))
    g.render()
    assert g.points_number == 2000 and g.links_number == datagen.barabasi_albert(2000, 2, 3).size // 2
    g.start()
    g.start(0.5)                   # already running: no second onSimulationStart
    assert g.alpha == 0.5
    g.frame()
    a1 = g.alpha
    assert a1 < 0.5
    g.pause()
    g.pause()                      # already paused: nothing
    g.frame()                      # paused: no forces, alpha unchanged
    assert g.alpha == a1
    g.step()                       # a manual step runs even when paused
    assert g.alpha < a1
    g.unpause()
    g.start(0.0009)                # below ALPHA_MIN: the next frame ends the run first
    g.frame()
    assert not g.is_simulation_running and g.progress == 1.0
    g.stop()                       # alpha > 0: still counts as active
    assert g.alpha == 0
    g.stop()                       # nothing active: no second onSimulationEnd
    assert events == ["onSimulationStart", "onSimulationPause", "onSimulationUnpause",
                      "onSimulationEnd", "onSimulationEnd"]
    pos = np.array(g.getPointPositions()).reshape(-1, 2)
    assert np.isnan(pos[5]).all() and np.isfinite(np.delete(pos, 5, axis=0)).all()
    g.trackPointPositionsByIndices([0, 5, 5000])
    tr = np.array(g.getTrackedPointPositionsArray()).reshape(-1, 2)
    assert np.array_equal(tr[0], pos[0]) and np.isnan(tr[1:]).all()
    assert list(g.getTrackedPointPositionsMap()) == [0]
    g.destroy()


def test_empty_graph_and_no_links():
    p = _lib.default_params()
    with _lib.Context() as c:
        c.set_positions(np.zeros(0, np.float32))       # n = 0
        c.set_links(np.zeros(0, np.float32))           # m = 0
        c.set_alpha(1.0)
        c.step(p, 3)
        assert c.get_positions().size == 0
        assert c.eval_force(p, _lib.FORCE_MANYBODY, 1.0, 0.5).size == 0
    xy = datagen.uniform_square(3000, 4096.0, 9)
    with _lib.Context() as c:
        c.set_positions(xy)
        c.set_links(np.zeros(0, np.float32))           # points, no links: the link passes are skipped
        c.set_alpha(1.0)
        c.step(p, 3)
        out = c.get_positions()
        assert np.isfinite(out).all()
        assert not c.eval_force(p, _lib.FORCE_LINK_IN, 1.0).any()
    g = simulation.Graph(cfg0_config())
    g.setPointPositions([])
    g.render()
    g.start()
    g.step()
    assert g.getPointPositions() == [] and not g.is_simulation_running
    g.destroy()


def test_point_count_change_between_steps():
    """Positions of a new length between two steps: the link CSR is rebuilt
    from the stored links against the new count (endpoints past it become
    invalid and are skipped), and the run continues with the same bits as a
    context that starts from the new positions."""
    S = 4096.0
    links = datagen.barabasi_albert(8000, 2, 5)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 6), datagen.uniform(m, 0, 1, 7)
    p = _lib.default_params()
    p.space_size = S
    p.repulsion = 0.5
    seeds = datagen.uniform(6, 0, 1, 8)
    xy8 = datagen.uniform_square(8000, S, 9)
    xy5 = datagen.uniform_square(5000, S, 10)
    jit5 = datagen.jitter(5000, 11)
    with _lib.Context() as c:
        c.set_positions(xy8)
        c.set_links(links, None, r_in, r_out)
        c.set_alpha(1.0)
        c.step(p, 3, seeds=seeds[:3])
        c.set_positions(xy5)
        c.set_jitter(jit5)
        c.set_alpha(0.9)
        c.step(p, 3, seeds=seeds[3:])
        got = c.get_positions()
        _, nbr, _, _, _ = c.debug_link_csr(0)
    with _lib.Context() as f:
        f.set_positions(xy5)
        f.set_jitter(jit5)
        f.set_links(links, None, r_in, r_out)
        f.set_alpha(0.9)
        f.step(p, 3, seeds=seeds[3:])
        want = f.get_positions()
    assert (nbr < 5000).all()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_budget_refusal_of_a_single_graph():
    """A context whose budget cannot hold the graph refuses the step with
    CG_ERR_BUDGET, keeps the positions as set, and the error names the budget."""
    xy = datagen.uniform_square(200_000, 8192.0, 12)
    with _lib.Context(hbm_budget=1 << 20) as c:
        c.set_positions(xy)
        with pytest.raises(_lib.CosmosGLError) as e:
            c.step(_lib.default_params(), 1)
        assert e.value.code == _lib.CG_ERR_BUDGET and "budget" in str(e.value)
        assert np.array_equal(c.get_positions().view(np.uint32), xy.view(np.uint32))


def test_cancelled_serving_job():
    """A job cancelled after the first segment leaves the device with no
    result; the others finish with the same bits as the sequential path, and a
    job refused for its footprint is listed, not run."""
    jobs = workloads.cfg4_jobs(duration=0.3, n_hi=20_000, collision_frac=0.2)
    victim = jobs[1].jid
    srv = serving.BatchServer(budget=200 * 2**30, segment=100)
    srv_calls = []

    def hook(ctx, active):
        if not srv_calls:
            srv.cancel(victim)
        srv_calls.append(len(active))

    res, st = srv.run(jobs, on_segment=hook)
    srv.close()
    assert st["cancelled"] == [victim]
    assert victim not in [r.jid for r in res]
    seq, _ = serving.serve_sequential([j for j in jobs if j.jid != victim])
    for a, b in zip(res, seq):
        assert a.jid == b.jid and np.array_equal(a.xy.view(np.uint32), b.xy.view(np.uint32))
    # a budget no single job fits: every job refused
    tiny = serving.BatchServer(budget=1 << 20, segment=100)
    res2, st2 = tiny.run(jobs[:3])
    tiny.close()
    assert res2 == [] and st2["rejected"] == [j.jid for j in jobs[:3]]
