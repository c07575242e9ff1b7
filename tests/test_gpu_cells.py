"""GPU parity of the integer structures of the tick, bit for bit, through the C ABI:

- many-body binning (calculate-level.vert:47-48): per-point finest cell,
  per-cell membership as sets (the order inside a cell follows the previous
  tick's order by design), per-level cell counts of the pyramid, and the
  level centroids within 1e-5 * spaceSize of the f64 means;
- spatial-hash collision cells (build-grid.vert:46-55): every point's cell in
  each of the four offset grids, and the device's sorted (base cell, point)
  sequence equals the lexicographic sort of the oracle's keys;

plus the context's behaviour around them: a repack refused by the HBM budget
leaves the previous packing intact, repeated repacks hold the device memory
constant, positions far outside [0, spaceSize]
bin and sum without overflow, and whole ticks with the many-body pass on
match the oracle's composed passes from each tick's input state.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative to spaceSize
OFFSETS = ((0.0, 0.0), (0.5, 0.0), (0.0, 0.5), (0.5, 0.5))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(S, **kw):
    p = _lib.default_params()
    p.space_size = S
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def positions(case):
    rng = np.random.default_rng(11)
    if case == "cfg0":
        return 4096.0, datagen.uniform_square(10_000, 4096.0, 42)
    if case == "cfg1":
        return 8192.0, datagen.uniform_disk(1_000_000, 4096, 4096, 2048, 7)
    if case == "clustered":
        n = 20_000
        xy = np.clip(rng.normal(2048, 40, 2 * n), 0, 4096).astype(np.float32)
        xy[: 2 * 500] = rng.uniform(0, 4096, 1000)
        xy[2 * 7: 2 * 9] = np.nan
        return 4096.0, xy
    if case == "crowded":
        n = 12_000
        xy = rng.uniform(0, 4096, 2 * n).astype(np.float32)
        xy[: 2 * 9000] = rng.uniform(1001.0, 1003.0, 2 * 9000).astype(np.float32)
        return 4096.0, xy
    if case == "boundary":
        # exact multiples of the finest cell (16), the space edge, one ulp either side
        n, S = 10_000, 4096.0
        xy = datagen.uniform_square(n, S, 3)
        k = rng.integers(0, 257, 2 * 3000)
        xy[: 2 * 3000] = (k * 16.0).astype(np.float32)
        xy[2 * 3000: 2 * 3500] = S
        xy[2 * 3500: 2 * 4000] = np.nextafter(np.float32(k[:1000] * 16.0), np.float32(-1))
        xy[2 * 4000: 2 * 4500] = np.nextafter(np.float32(k[1000:2000] * 16.0), np.float32(1e9))
        xy[2 * 4500: 2 * 4600] = 0.0
        return S, xy.astype(np.float32)
    if case == "outside":
        # points outside [0, S] (as set, before any integrate): clamped into the edge cells
        n, S = 10_000, 4096.0
        xy = datagen.uniform_square(n, S, 4)
        xy[: 2 * 300] = rng.uniform(-3 * S, -0.5, 2 * 300).astype(np.float32)
        xy[2 * 300: 2 * 600] = rng.uniform(S + 0.5, 5 * S, 2 * 300).astype(np.float32)
        return S, xy
    if case == "odd_space":
        # a space size that is not a power of two: the cell divisions are inexact
        S = 5000.0
        return S, datagen.uniform_square(30_000, S, 5)
    raise ValueError(case)


def check_cells(c, xy, S, tag, gid=None):
    """Per-level counts and centroids, and finest-cell membership as sets, of
    the context's last many-body pass against the oracle's binning of xy."""
    n = xy.size // 2
    absent = O.absent_mask(xy)
    pres = np.flatnonzero(absent == 0).astype(np.int64)
    gf, L = O.geometry(n)
    P = xy.reshape(-1, 2)[pres].astype(np.float64)
    for lev in range(L):
        g = 4 << lev
        finest = lev == L - 1
        counts, cent, members = c.debug_cells(lev, members=finest)
        cell = O.cells(xy, O.cell_size(S, g), g, absent)[pres]
        key = cell[:, 1].astype(np.int64) * g + cell[:, 0]
        want = np.bincount(key, minlength=g * g)
        assert np.array_equal(counts.astype(np.int64), want), \
            f"{tag} level {lev}: counts differ in {int(np.sum(counts != want))} cells"
        # centroids: the f64 mean of each cell's points
        occ = want > 0
        mean = np.zeros((g * g, 2))
        np.add.at(mean, key, P)
        mean[occ] /= want[occ, None]
        err = float(np.abs(cent[occ] - mean[occ]).max()) if occ.any() else 0.0
        assert err <= TOL * max(S, float(np.abs(P).max())), f"{tag} level {lev}: centroid off by {err}"
        if finest:
            dev_key = np.repeat(np.arange(g * g, dtype=np.int64), counts.astype(np.int64))
            dev = members.astype(np.int64)[np.lexsort((members, dev_key))]
            ora = pres[np.lexsort((pres, key))]
            assert np.array_equal(dev, ora), f"{tag}: finest-cell membership differs for {int(np.sum(dev != ora))} points"
    return L, pres.size


@pytest.mark.parametrize("case", ["cfg0", "cfg1", "clustered", "crowded", "boundary", "outside",
                                  "odd_space"])
def test_cell_assignment_and_levels_bit_exact(case):
    S, xy = positions(case)
    n = xy.size // 2
    say(f"cells {case}: n={n} S={S}")
    c = _lib.Context()
    c.set_positions(xy)
    c.set_jitter(datagen.jitter(n, 6))
    c.eval_force(params(S, repulsion=0.5), _lib.FORCE_MANYBODY, 0.9, 0.375)
    L, npres = check_cells(c, xy, S, f"cells {case}")
    say(f"cells {case}: {L} levels, {npres} present points bit-exact")
    c.close()


@pytest.mark.parametrize("n", [200, 600, 12_000])
def test_cells_after_collapsing_ticks_bit_exact(n):
    """A graph's layout collapses from a uniform square into a few clumps within
    its first ticks, so most pyramid tiles empty out: every pass rewrites only
    the tiles that held points at that pass or the one before, and the others
    keep their empty centroids.  After the ticks, one more pass bins the final
    positions; its cells, centroids and membership must equal the oracle's at
    every level (finest grids of 32, 64 and 256: tiles below, at and above one
    block of 256 finest cells)."""
    S = float(1 << int(np.ceil(np.log2(32 * np.sqrt(n)))))
    xy = datagen.uniform_square(n, S, 31)
    links = datagen.barabasi_albert(n, 2, 31)
    p = params(S, repulsion=0.5, gravity=0.25, link_spring=1.0, link_distance=10.0)
    c = _lib.Context()
    c.set_positions(xy)
    c.set_jitter(datagen.jitter(n, 32))
    c.set_links(links)
    c.set_alpha(1.0)
    say(f"collapse n={n} S={S}")
    for t in (5, 15, 40):
        c.step(p, t)
        cur = c.get_positions()
        c.eval_force(p, _lib.FORCE_MANYBODY, 0.5, 0.25)
        L, npres = check_cells(c, cur, S, f"collapse n={n} after +{t} ticks")
        gf, _ = O.geometry(n)
        occ = np.unique(O.cells(cur, O.cell_size(S, gf), gf, O.absent_mask(cur)), axis=0).shape[0]
        say(f"collapse n={n}: +{t} ticks, {occ} of {gf * gf} finest cells occupied, {L} levels bit-exact")
    c.close()


@pytest.mark.parametrize("case", ["clump", "boundary", "odd_space", "outside"])
def test_collision_keys_and_sort_order_bit_exact(case):
    rng = np.random.default_rng(21)
    if case == "clump":
        S, n = 4096.0, 20_000
        xy = np.clip(rng.normal(2048, 300, 2 * n), 0, S).astype(np.float32)
    elif case == "boundary":
        S, n = 4096.0, 8000
        xy = datagen.uniform_square(n, S, 7)
    elif case == "odd_space":
        S, n = 3000.0, 15_000
        xy = datagen.uniform_square(n, S, 8)
    else:
        S, n = 2048.0, 6000
        xy = datagen.uniform_square(n, S, 9)
        xy[: 2 * 200] = rng.uniform(-S, 2 * S, 400).astype(np.float32)
    sizes = (2 * datagen.lognormal(n, np.log(2), 0.5, 10)).astype(np.float32)
    G, h = O.collision_grid(S, 0.0, 0.0, max(4.0, float(sizes.max())))
    if case == "boundary":
        # exact cell edges and half-cell edges (the offset grids' edges), the space edge
        k = rng.integers(0, G + 1, 2 * 2000).astype(np.float32)
        xy[: 2 * 1000] = (k[: 2 * 1000] * np.float32(h)).astype(np.float32)
        xy[2 * 1000: 2 * 2000] = ((k[2 * 1000:] + np.float32(0.5)) * np.float32(h)).astype(np.float32)
        xy[2 * 2000: 2 * 2100] = S
    xy[2 * 11: 2 * 13] = np.nan
    absent = O.absent_mask(xy)
    pres = np.flatnonzero(absent == 0).astype(np.int64)
    say(f"collision keys {case}: n={n} S={S} G={G} h={h}")
    c = _lib.Context()
    c.set_positions(xy)
    c.set_point_sizes(sizes)
    c.eval_force(params(S, collision=0.7), _lib.FORCE_COLLISION, 0.8)
    keys, pts, cells, counts, Gd = c.debug_collision_cells()
    assert Gd == G
    # the sorted (base cell, point) sequence
    base = O.cells(xy, h, G, absent)[pres]
    wk = base[:, 1].astype(np.int64) * G + base[:, 0]
    order = np.lexsort((pres, wk))
    assert keys.size == pres.size
    assert np.array_equal(keys.astype(np.int64), wk[order]), "base-grid keys differ"
    assert np.array_equal(pts.astype(np.int64), pres[order]), "collision sort order differs"
    # every point's cell in each offset grid
    for o, off in enumerate(OFFSETS):
        cell = O.cells(xy, h, G, absent, off)[pres]
        want = cell[:, 1].astype(np.int64) * G + cell[:, 0]
        got = cells[pres, o].astype(np.int64)
        assert np.array_equal(got, want), f"grid {o}: {int(np.sum(got != want))} points in other cells"
        # the cell aggregates hold exactly those points
        assert np.array_equal(counts[o].astype(np.int64), np.bincount(want, minlength=G * G)), \
            f"grid {o}: aggregate counts differ"
    assert (cells[absent == 1] == 0xFFFFFFFF).all()
    c.close()


def test_repack_refused_by_budget_keeps_previous_packing():
    """A graph added beyond the HBM budget makes the next repack fail with
    CG_ERR_BUDGET; the previous packing stays: positions read back unchanged,
    a drag still lands, and once the graph is removed the layout continues
    with the same bits as a context that never saw the refusal."""
    S, n = 4096.0, 5000
    xy = datagen.uniform_square(n, S, 12)
    links = datagen.barabasi_albert(n, 2, 12)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 13), datagen.uniform(m, 0, 1, 14)
    jit = datagen.jitter(n, 15)
    p = params(S, repulsion=0.5)
    seeds = datagen.uniform(10, 0, 1, 16)

    def load(c):
        c.set_positions(xy)
        c.set_jitter(jit)
        c.set_links(links, None, r_in, r_out)
        c.set_alpha(1.0)

    with _lib.Context() as probe:
        load(probe)
        probe.step(p, 1, seeds=seeds[:1])
        held = probe.device_bytes()
    say(f"budget test: one graph holds {held} bytes")
    ref = _lib.Context()
    load(ref)
    ref.step(p, 5, seeds=seeds[:5])
    ref.drag_points([3], [100.0, 200.0])
    ref.step(p, 5, seeds=seeds[5:])
    want = ref.get_positions()
    ref.close()

    c = _lib.Context(hbm_budget=int(2.5 * held))
    load(c)
    c.step(p, 5, seeds=seeds[:5])
    before = c.get_positions()
    big = c.graph_add()
    c.graph_set_positions(big, datagen.uniform_square(1_000_000, S, 17))
    with pytest.raises(_lib.CosmosGLError) as e:
        c.step(p, 3, seeds=seeds[:3])
    assert e.value.code == _lib.CG_ERR_BUDGET
    assert np.array_equal(c.get_positions().view(np.uint32), before.view(np.uint32))
    c.drag_points([3], [100.0, 200.0])
    c.graph_remove(big)
    c.step(p, 5, seeds=seeds[5:])
    got = c.get_positions()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    c.close()


def test_repeated_repacks_hold_constant_device_memory():
    """Every repack frees all packed arrays it allocates (binning counts,
    ranks and scan sums included): after the first repack has filled the
    allocator's cache, further repacks of the same graphs keep the held
    device bytes constant, and the layout matches a context packed once."""
    S, n = 4096.0, 20000
    xy = datagen.uniform_square(n, S, 21)
    links = datagen.barabasi_albert(n, 2, 21)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 22), datagen.uniform(m, 0, 1, 23)
    jit = datagen.jitter(n, 24)
    p = params(S, repulsion=0.5, collision=0.5)
    seeds = datagen.uniform(8, 0, 1, 25)

    def load(c):
        c.set_positions(xy)
        c.set_jitter(jit)
        c.set_links(links, None, r_in, r_out)
        c.set_alpha(1.0)

    ref = _lib.Context()
    load(ref)
    ref.step(p, 8, seeds=seeds)
    want = ref.get_positions()
    ref.close()

    c = _lib.Context()
    load(c)
    c.step(p, 1, seeds=seeds[:1])
    held = []
    for t in range(1, 8):
        c.set_point_cluster_strength(None)  # marks the packing stale: the next step repacks
        c.step(p, 1, seeds=seeds[t:t + 1])
        held.append(c.device_bytes())
    say(f"repacks: held device bytes {held}")
    assert held[2:] == [held[1]] * (len(held) - 2)
    assert np.array_equal(c.get_positions().view(np.uint32), want.view(np.uint32))
    c.close()


def _device_free_bytes() -> int:
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_serving_repacks_return_device_memory():
    """A repack destroys the set's tick graphs and frees the old packing; the
    device must get that memory back.  The serving mix (202 requests, 11.5M
    points, Chung-Lu links with split hubs, 10% with collision): six rounds of
    16 ticks, one request removed and a repack.  Device memory outside the
    context's own accounting (free + held bytes) may not shrink by more than
    512 MiB (it lost ~7 GB per repack while such sets replayed captured
    graphs)."""
    from graph_amd import serving, workloads

    jobs = workloads.cfg4_jobs(duration=5.0)
    with _lib.Context() as c:
        ids = []
        for j in jobs:
            g = c.graph_add()
            serving._load(c, g, j)
            ids.append(g)
        c.step_graphs(ids, 16)
        c.sync()
        base = _device_free_bytes() + c.device_bytes()
        lost = []
        for _ in range(6):
            c.step_graphs(ids, 16)
            c.sync()
            c.graph_remove(ids.pop())
            c.step_graphs(ids, 0)  # repack
            c.sync()
            lost.append(base - (_device_free_bytes() + c.device_bytes()))
        say(f"device bytes outside the context after each repack: {[x >> 20 for x in lost]} MiB")
        assert max(lost) < 512 << 20


def test_manybody_with_positions_far_outside_space():
    """Positions as set are not clamped (only integrate clamps), so with
    gravity off the many-body pass can see points far outside [0, S]: its
    exact cell sums must not overflow.  300 points near (1000 S, 1000 S) and
    300 near (-1000 S, 0) against the oracle."""
    S, n = 4096.0, 5000
    rng = np.random.default_rng(18)
    xy = datagen.uniform_square(n, S, 19)
    xy[: 2 * 300] = (1000 * S + rng.uniform(0, 50, 600)).astype(np.float32)
    xy[2 * 300: 2 * 600: 2] = (-1000 * S - rng.uniform(0, 50, 300)).astype(np.float32)
    xy[2 * 300 + 1: 2 * 600: 2] = rng.uniform(0, S, 300).astype(np.float32)
    jit = datagen.jitter(n, 20)
    c = _lib.Context()
    c.set_positions(xy)
    c.set_jitter(jit)
    p = params(S, repulsion=0.5, gravity=0.0)
    got = c.eval_force(p, _lib.FORCE_MANYBODY, 1.0, 0.6)
    want = O.manybody(xy, S, 1.0, 0.5, 0.6, jitter=jit)
    e = np.hypot(*(got.reshape(-1, 2).astype(np.float64) - want.reshape(-1, 2)).T)
    say(f"far outside: max |dv| = {e.max():.3e} ({e.max() / S:.2e} S)")
    assert np.isfinite(got).all()
    assert e.max() <= TOL * S
    # and a tick from there: the out-of-range points are clamped into [0, S]
    c.set_alpha(1.0)
    c.step(p, 1)
    out = c.get_positions().reshape(-1, 2)
    assert (out >= 0).all() and (out <= S).all()
    c.close()


@pytest.mark.parametrize("case", ["cfg0", "pinned_absent"])
def test_ticks_with_manybody_vs_oracle_passes(case):
    """Whole ticks with the many-body pass on (gravity off, so the tick's first
    pass starts from the identical state): many-body, link IN and link OUT,
    each integrated, against the oracle's passes composed with f32 positions
    between passes.  Every tick starts the oracle from the GPU's state after
    the previous tick (forces from identical input states; the trajectories
    themselves diverge chaotically).  Per point |dp| <= 1e-5 * spaceSize."""
    S, n, ticks = 4096.0, 10_000, 4
    xy = datagen.uniform_square(n, S, 42)
    links = datagen.barabasi_albert(n, 2, 42)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 1), datagen.uniform(m, 0, 1, 2)
    jit = datagen.jitter(n, 3)
    seeds = datagen.uniform(ticks, 0, 1, 4)
    pinned = []
    if case == "pinned_absent":
        xy[2 * 20: 2 * 25] = np.nan
        pinned = [0, 1, 30, 31]
    c = _lib.Context()
    c.set_positions(xy)
    c.set_jitter(jit)
    c.set_links(links, None, r_in, r_out)
    if pinned:
        c.set_pinned(pinned)
    c.set_alpha(1.0)
    p = params(S, repulsion=0.5, gravity=0.0, friction=0.85, decay=5000, link_spring=1.0,
               link_distance=10.0)
    absent = O.absent_mask(xy)
    pin = np.zeros(n, np.uint8)
    pin[pinned] = 1
    Lk = O.Links(n, links, r_in, r_out)
    f32 = lambda q: q.astype(np.float32).astype(np.float64)
    live = absent == 0
    a = 1.0
    for t in range(ticks):
        start = c.get_positions().astype(np.float64)
        af = float(np.float32(a))
        c.step(p, 1, seeds=seeds[t: t + 1])
        gpu = c.get_positions().reshape(-1, 2).astype(np.float64)
        pos = f32(O.integrate(start, O.manybody(start, S, af, 0.5, float(seeds[t]), jitter=jit, absent=absent),
                              0.85, S, pin, absent))
        for d in (0, 1):
            pos = f32(O.integrate(pos, Lk.force(pos, d, af, 1.0, 10.0, 1.0, 1.2, absent), 0.85, S, pin, absent))
        e = np.hypot(*(gpu - pos.reshape(-1, 2)).T)
        worst = float(e[live].max())
        say(f"tick {t} ({case}): max |dp| = {worst:.3e} ({worst / S:.2e} S)")
        assert np.isnan(gpu[~live]).all()
        assert worst <= TOL * S
        if pinned:
            assert np.array_equal(gpu[pinned], xy.reshape(-1, 2)[pinned].astype(np.float64))
        a = O.alpha_step(a, 5000)
    c.close()
