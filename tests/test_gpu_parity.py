"""GPU parity: every force pass of the HIP path against the oracle from an
identical input state, integer structures bit-exact, through the C ABI.

Tolerance (per point, vector norm): |v_gpu - v_oracle| <= 1e-5 * spaceSize
for forces evaluated in f32 on the GPU vs f64 in the oracle.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative to spaceSize


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(S, **kw):
    p = _lib.default_params()
    p.space_size = S
    for k, v in kw.items():
        if k == "range":
            p.link_dist_random_variation_range[0], p.link_dist_random_variation_range[1] = v
        else:
            setattr(p, k, v)
    return p


def ctx_for(xy, links=None, jitter=None, r_in=None, r_out=None, strength=None):
    c = _lib.Context()
    c.set_positions(xy)
    if jitter is not None:
        c.set_jitter(jitter)
    if links is not None:
        c.set_links(links, strength, r_in, r_out)
    return c


def err_norm(v_gpu, v_or):
    d = v_gpu.reshape(-1, 2).astype(np.float64) - v_or.reshape(-1, 2)
    return np.hypot(d[:, 0], d[:, 1])


def check(v_gpu, v_or, S, what, idx=None):
    e = err_norm(v_gpu, v_or)
    if idx is not None:
        e = e[idx]
    worst = float(np.nanmax(e)) if e.size else 0.0
    say(f"{what}: max |dv| = {worst:.3e} ({worst / S:.2e} S)")
    assert np.isfinite(e).all()
    assert worst <= TOL * S, f"{what}: {worst / S:.3e} S > {TOL} S"


def test_pointwise_forces():
    S, n = 4096.0, 5000
    xy = datagen.uniform_square(n, S, 1)
    xy[10:14] = np.nan
    absent = O.absent_mask(xy)
    c = ctx_for(xy)
    p = params(S, gravity=0.25, center=0.3, repulsion_from_mouse=2.0)
    p.mouse_position[0], p.mouse_position[1] = 1000.0, 2000.0
    pres = absent == 0
    xy0 = np.where(np.isnan(xy), 0, xy)
    for force, want in [
        (_lib.FORCE_GRAVITY, O.gravity(xy0, S, 0.7, 0.25)),
        (_lib.FORCE_CENTER, O.center(xy, 0.7, 0.3, absent)),
        (_lib.FORCE_MOUSE, O.mouse(xy0, 1000.0, 2000.0, 2.0)),
    ]:
        got = c.eval_force(p, force, 0.7)
        check(got, want, S, f"force {force}", idx=pres)
        assert (got.reshape(-1, 2)[~pres] == 0).all()


@pytest.mark.parametrize("case", ["cfg0", "clustered", "crowded", "tiny"])
def test_manybody_slots_bit_exact_and_forces(case):
    rng = np.random.default_rng(0)
    if case == "cfg0":
        S, n = 4096.0, 10_000
        xy = datagen.uniform_square(n, S, 42)
    elif case == "clustered":
        # a dense Gaussian clump: finest cells far above 64 points (crowded-cell path)
        S, n = 4096.0, 20_000
        xy = np.clip(rng.normal(2048, 40, 2 * n), 0, S).astype(np.float32)
        xy[: 2 * 500] = rng.uniform(0, S, 1000)
        xy[2 * 7: 2 * 9] = np.nan
    elif case == "crowded":
        # 9000 points inside one finest cell: several 2048-point chunks per cell
        S, n = 4096.0, 12_000
        xy = rng.uniform(0, S, 2 * n).astype(np.float32)
        xy[: 2 * 9000] = rng.uniform(1001.0, 1003.0, 2 * 9000).astype(np.float32)
    else:
        S, n = 1024.0, 7
        xy = rng.uniform(0, S, 2 * n).astype(np.float32)
    jit = datagen.jitter(n, 5)
    absent = O.absent_mask(xy)
    c = ctx_for(xy, jitter=jit)
    p = params(S, repulsion=0.5)
    for seed in (0.125, 0.8765):
        got = c.eval_force(p, _lib.FORCE_MANYBODY, 0.9, seed)
        slots, gf, L = c.debug_slots()
        want_slots = O.slots(xy, S, seed, absent)
        assert (gf, L) == O.geometry(n)
        assert np.array_equal(slots, want_slots), f"slots differ in {np.sum(slots != want_slots)} entries"
        want = O.manybody(xy, S, 0.9, 0.5, seed, jitter=jit, absent=absent)
        check(got, want, S, f"manybody {case} seed {seed}", idx=absent == 0)


def test_manybody_cfg1_shape():
    # 1M points uniform in a disk of radius 2048 centred in S = 8192: L = 8, g_f = 512
    S, n = 8192.0, 1_000_000
    xy = datagen.uniform_disk(n, 4096, 4096, 2048, 7)
    jit = datagen.jitter(n, 8)
    c = ctx_for(xy, jitter=jit)
    p = params(S, repulsion=0.5)
    seed = 0.4242
    got = c.eval_force(p, _lib.FORCE_MANYBODY, 1.0, seed)
    slots, gf, L = c.debug_slots()
    assert (gf, L) == (512, 8)
    want_slots = O.slots(xy, S, seed)
    assert np.array_equal(slots, want_slots)
    idx = np.random.default_rng(1).choice(n, 20_000, replace=False).astype(np.uint32)
    want = O.manybody(xy, S, 1.0, 0.5, seed, jitter=jit, eval_idx=idx)
    check(got, want, S, "manybody cfg1 (20k-point sample)", idx=idx)


def test_link_csr_bit_exact():
    n = 10_000
    links = datagen.barabasi_albert(n, 2, 42)
    bad = np.array([3, 3, 1.5, 2, -1, 0, n, 1, 5, 5, 17, 4], np.float32)
    links = np.concatenate([links, bad])
    m = links.size // 2
    user = datagen.uniform(m, -0.5, 2.0, 3)
    user[::7] = np.nan
    xy = datagen.uniform_square(n, 4096.0, 1)
    for strength in (None, user):
        c = ctx_for(xy, links=links, r_in=datagen.uniform(m, 0, 1, 1), r_out=datagen.uniform(m, 0, 1, 2),
                    strength=strength)
        for d in (0, 1):
            off, nbr, bias, st, deg = c.debug_link_csr(d)
            o_off, o_nbr, o_lof, o_deg = O.link_csr(n, links, d)
            o_bias, o_st = O.link_props(n, o_off, o_nbr, o_lof, o_deg, strength)
            assert np.array_equal(off, o_off)
            assert np.array_equal(nbr, o_nbr)
            assert np.array_equal(deg, o_deg)
            assert np.array_equal(bias.view(np.uint32), o_bias.view(np.uint32))
            assert np.array_equal(st.view(np.uint32), o_st.view(np.uint32))


@pytest.mark.parametrize("case", ["ba_cfg0", "chung_lu_hubs", "absent_neighbours"])
def test_link_forces_vs_serial_oracle(case):
    if case == "chung_lu_hubs":
        S, n = 8192.0, 200_000
        links = datagen.chung_lu(n, 2.1, 1.0, 1_000_000, 7)
        xy = datagen.uniform_disk(n, 4096, 4096, 2048, 3)
    else:
        S, n = 4096.0, 10_000
        links = datagen.barabasi_albert(n, 2, 42)
        xy = datagen.uniform_square(n, S, 42)
        if case == "absent_neighbours":
            xy[2 * 5: 2 * 40] = np.nan
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 11), datagen.uniform(m, 0, 1, 12)
    absent = O.absent_mask(xy)
    c = ctx_for(xy, links=links, r_in=r_in, r_out=r_out)
    p = params(S, link_spring=1.0, link_distance=10.0, range=(1.0, 1.2))
    L = O.Links(n, links, r_in, r_out)
    for alpha in (1.0, 0.3):
        for d in (0, 1):
            got = c.eval_force(p, _lib.FORCE_LINK_IN + d, alpha)
            want = L.force(xy, d, alpha, 1.0, 10.0, 1.0, 1.2, absent)
            check(got, want, S, f"link {case} dir {d} alpha {alpha}", idx=absent == 0)


def test_link_forces_after_ticks_vs_oracle():
    """Link forces from a live state, 40 ticks into a layout (warm starts in
    play; the points above 1536 entries per direction, up to 41,950, are split
    into segments and fixed up), against the serial oracle for every point."""
    S, n = 8192.0, 200_000
    links = datagen.chung_lu(n, 2.1, 1.0, 1_000_000, 7)
    xy = datagen.uniform_disk(n, 4096, 4096, 2048, 3)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 11), datagen.uniform(m, 0, 1, 12)
    c = ctx_for(xy, links=links, r_in=r_in, r_out=r_out, jitter=datagen.jitter(n, 4))
    p = params(S, link_spring=1.0, link_distance=10.0, range=(1.0, 1.2), gravity=0.25, repulsion=0.5)
    c.set_alpha(1.0)
    c.step(p, 40)
    live = c.get_positions()
    a = float(np.float32(c.get_alpha()))
    L = O.Links(n, links, r_in, r_out)
    absent = O.absent_mask(live)
    for d in (0, 1):
        got = c.eval_force(p, _lib.FORCE_LINK_IN + d, a)
        want = L.force(live, d, a, 1.0, 10.0, 1.0, 1.2, absent)
        check(got, want, S, f"link dir {d} after 40 ticks", idx=absent == 0)


@pytest.mark.parametrize("S,radius,path", [(4096.0, 5.0, "sparse"), (16384.0, 0.0, "sparse"),
                                          (4096.0, 0.0, "dense4"), (4096.0, 80.0, "dense16")])
def test_collision_vs_oracle(S, radius, path):
    """Each cell pass of the collision force: from the occupied base cells
    (more base cells than points), or over every cell with 4 or 16 lanes;
    radius 0 takes each point's radius from its size."""
    n = 20_000
    rng = np.random.default_rng(4)
    xy = np.clip(rng.normal(2048, 300, 2 * n), 0, S).astype(np.float32)
    sizes = 2 * datagen.lognormal(n, np.log(2), 0.5, 9)
    mx = max(4.0, float(sizes.max()))
    G, _ = O.collision_grid(S, radius, 0.0, mx)
    assert path == ("sparse" if G * G > n else ("dense16" if n >= 8 * G * G else "dense4"))
    c = ctx_for(xy)
    c.set_point_sizes(sizes)
    p = params(S, collision=0.7, collision_radius=radius)
    got = c.eval_force(p, _lib.FORCE_COLLISION, 0.8)
    want = O.collision(xy, sizes, S, 0.8, 0.7, radius=radius, max_size=mx)
    check(got, want, S, f"collision ({path}, grid {G})")


def test_ticks_deterministic_alpha_and_masks():
    S, n = 4096.0, 10_000
    xy = datagen.uniform_square(n, S, 42)
    xy[2 * 3: 2 * 6] = np.nan
    links = datagen.barabasi_albert(n, 2, 42)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 1), datagen.uniform(m, 0, 1, 2)
    jit = datagen.jitter(n, 3)
    seeds = datagen.uniform(40, 0, 1, 4)
    outs = []
    for _ in range(2):
        c = ctx_for(xy, links=links, jitter=jit, r_in=r_in, r_out=r_out)
        c.set_pinned([0, 1, 2, 10])
        c.set_alpha(1.0)
        p = params(S, repulsion=0.5, gravity=0.25, friction=0.85, decay=5000)
        tr = c.step(p, 40, seeds=seeds, trace=True)
        outs.append(c.get_positions())
        a = 1.0
        for t in range(40):
            a = O.alpha_step(a, 5000)
            assert tr[t] == a
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    out = outs[0].reshape(-1, 2)
    assert np.isnan(out[3:6]).all()
    assert np.array_equal(out[[0, 1, 2, 10]], xy.reshape(-1, 2)[[0, 1, 2, 10]])
    live = np.ones(n, bool)
    live[3:6] = False
    assert (out[live] >= 0).all() and (out[live] <= S).all()
    assert not np.array_equal(out[live], xy.reshape(-1, 2)[live])


def edge_stats(xy, links, r, lo=1.0, hi=1.2, dist=10.0):
    P = xy.reshape(-1, 2).astype(np.float64)
    s, t = links[0::2].astype(np.int64), links[1::2].astype(np.int64)
    L = np.hypot(*(P[s] - P[t]).T)
    D = (r.astype(np.float64) * (hi - lo) + lo) * dist
    stress = float(np.sum((L - D) ** 2) / np.sum(D**2))
    return np.percentile(L, [10, 50, 90]), stress


def test_layout_quality_vs_oracle_cfg0():
    """cfg0 for 300 ticks on the GPU and in the FP64 oracle from the same inputs:
    trajectories diverge chaotically, so the layouts are compared by their
    edge-length distribution and edge stress (within 2%)."""
    S, n, ticks = 4096.0, 10_000, 300
    xy = datagen.uniform_square(n, S, 42)
    links = datagen.barabasi_albert(n, 2, 42)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 1), datagen.uniform(m, 0, 1, 2)
    jit = datagen.jitter(n, 3)
    seeds = datagen.uniform(ticks, 0, 1, 4)
    c = ctx_for(xy, links=links, jitter=jit, r_in=r_in, r_out=r_out)
    c.set_alpha(1.0)
    p = params(S, repulsion=0.5, gravity=0.25, friction=0.85, decay=5000, link_spring=1.0,
               link_distance=10.0)
    c.step(p, ticks, seeds=seeds)
    gpu = c.get_positions()
    pd = dict(space_size=S, gravity=0.25, center=0, repulsion=0.5, link_spring=1.0, link_distance=10.0,
              lo=1.0, hi=1.2, friction=0.85, decay=5000.0)
    Lk = O.Links(n, links, r_in, r_out)
    pos, a = xy.astype(np.float64), 1.0
    for t in range(ticks):
        pos, a = O.tick(pos, a, float(seeds[t]), pd, links=Lk, jitter=jit)
    q_g, s_g = edge_stats(gpu, links, r_in)
    q_o, s_o = edge_stats(pos, links, r_in)
    say(f"edge length p10/p50/p90 gpu {q_g} oracle {q_o}; stress gpu {s_g:.4f} oracle {s_o:.4f}")
    np.testing.assert_allclose(q_g, q_o, rtol=0.02)
    assert abs(s_g - s_o) <= 0.02 * s_o


def test_napi_addon_matches_ctypes(tmp_path):
    """The Node-API addon drives the same ABI: identical inputs give identical bits."""
    import json
    import shutil
    import subprocess

    node = shutil.which("node")
    addon = _lib.PKG / "napi" / "cosmosgl.node"
    if node is None or not addon.exists():
        pytest.skip("node or the addon is not available")
    S, n = 4096.0, 10_000
    xy = datagen.uniform_square(n, S, 42)
    links = datagen.barabasi_albert(n, 2, 42)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 1), datagen.uniform(m, 0, 1, 2)
    jit = datagen.jitter(n, 3)
    seeds = datagen.uniform(25, 0, 1, 4)
    for name, arr in (("xy", xy), ("links", links), ("rin", r_in), ("rout", r_out), ("jitter", jit),
                      ("seeds", seeds)):
        arr.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    (tmp_path / "params.json").write_text(json.dumps({"spaceSize": S, "simulationRepulsion": 0.5}))
    subprocess.run([node, str(_lib.PKG / "napi" / "run_ticks.js"), str(tmp_path)], check=True, timeout=120)
    got = np.fromfile(tmp_path / "out.f32", np.float32)
    c = ctx_for(xy, links=links, jitter=jit, r_in=r_in, r_out=r_out)
    c.set_alpha(1.0)
    tr = c.step(params(S, repulsion=0.5), 25, seeds=seeds, trace=True)
    want = c.get_positions()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(np.fromfile(tmp_path / "alpha.f64", np.float64), tr)


def test_napi_addon_forces_vs_oracle(tmp_path):
    """Node drives the addon (graph_amd/napi/run_force.js): every force pass it
    evaluates through evalForce matches the oracle from the same inputs."""
    import json
    import shutil
    import subprocess

    node = shutil.which("node")
    addon = _lib.PKG / "napi" / "cosmosgl.node"
    if node is None or not addon.exists():
        pytest.skip("node or the addon is not available")
    S, n = 4096.0, 10_000
    xy = datagen.uniform_square(n, S, 42)
    links = datagen.barabasi_albert(n, 2, 42)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 1), datagen.uniform(m, 0, 1, 2)
    jit = datagen.jitter(n, 3)
    sizes = (2 * datagen.lognormal(n, np.log(2), 0.5, 4)).astype(np.float32)
    for name, arr in (("xy", xy), ("links", links), ("rin", r_in), ("rout", r_out), ("jitter", jit),
                      ("sizes", sizes)):
        arr.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    forces = [_lib.FORCE_GRAVITY, _lib.FORCE_MANYBODY, _lib.FORCE_LINK_IN, _lib.FORCE_LINK_OUT,
              _lib.FORCE_COLLISION]
    alpha, seed = 0.75, 0.40625
    (tmp_path / "params.json").write_text(json.dumps({
        "params": {"spaceSize": S, "simulationRepulsion": 0.5, "simulationGravity": 0.25,
                   "simulationCollision": 0.6},
        "alpha": alpha, "seed": seed, "forces": forces}))
    subprocess.run([node, str(_lib.PKG / "napi" / "run_force.js"), str(tmp_path)], check=True, timeout=120)
    got = {f: np.fromfile(tmp_path / f"v{f}.f32", np.float32) for f in forces}
    Lk = O.Links(n, links, r_in, r_out)
    check(got[_lib.FORCE_GRAVITY], O.gravity(xy, S, alpha, 0.25), S, "addon gravity")
    check(got[_lib.FORCE_MANYBODY], O.manybody(xy, S, alpha, 0.5, seed, jitter=jit), S, "addon many-body")
    for d in (0, 1):
        check(got[_lib.FORCE_LINK_IN + d], Lk.force(xy, d, alpha, 1.0, 10.0, 1.0, 1.2), S, f"addon link {d}")
    check(got[_lib.FORCE_COLLISION],
          O.collision(xy, sizes, S, alpha, 0.6, max_size=max(4.0, float(sizes.max()))), S, "addon collision")


def test_cluster_force_vs_oracle():
    # Clusters/force-cluster.frag: custom targets, centroid targets (unset or a negative
    # coordinate), unclustered points, absent points out of the centroids, per-point strength
    rng = np.random.default_rng(17)
    S, n = 4096.0, 6000
    xy = rng.uniform(0, S, 2 * n).astype(np.float32)
    xy[2 * 50: 2 * 60] = np.nan
    cl = rng.integers(-1, 10, n).astype(np.int32)
    cpos = np.array([1000.0, 3000.0, -1.0, 200.0], np.float32)  # cluster 0 custom, 1 centroid
    strength = rng.uniform(0.2, 2.0, n).astype(np.float32)
    absent = O.absent_mask(xy)
    c = ctx_for(xy)
    c.set_point_clusters(cl)
    c.set_cluster_positions(cpos)
    c.set_point_cluster_strength(strength)
    p = params(S, cluster=0.3)
    got = c.eval_force(p, _lib.FORCE_CLUSTER, 0.8)
    want = O.cluster(xy, cl, 0.8, 0.3, cluster_pos=cpos, strength=strength, absent=absent)
    check(got, want, S, "cluster", idx=absent == 0)
    # unclustered points get no pull
    assert np.all(got.reshape(-1, 2)[cl < 0] == 0)


@pytest.mark.parametrize("center", [0.0, 0.5])
def test_ticks_vs_oracle_pass_order(center):
    """Whole ticks (gravity, centre, many-body, both link directions, each
    integrated) against the oracle's passes composed with positions stored in
    f32 between passes, as the reference's position textures hold them.
    Repulsion is 0 so the many-body pass integrates a zero velocity: over
    several ticks the check then covers the pass order and the buffer
    hand-offs between passes without the cell-boundary discontinuities of the
    pyramid (a point one rounding away from a cell edge may bin differently
    in f32 and f64).  The many-body pass inside whole ticks is checked per
    tick from identical states in test_gpu_cells.py.  Per point
    |dp| <= 1e-5 * spaceSize per tick."""
    S, n, ticks = 4096.0, 10_000, 3
    xy = datagen.uniform_square(n, S, 5)
    xy[2 * 7: 2 * 9] = np.nan
    links = datagen.barabasi_albert(n, 2, 5)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 6), datagen.uniform(m, 0, 1, 7)
    pinned = [0, 11, 12]
    c = ctx_for(xy, links=links, r_in=r_in, r_out=r_out)
    c.set_pinned(pinned)
    c.set_alpha(1.0)
    p = params(S, repulsion=0.0, gravity=0.25, center=center, friction=0.85, decay=5000, link_spring=1.0,
               link_distance=10.0, range=(1.0, 1.2))
    c.step(p, ticks)
    gpu = c.get_positions().reshape(-1, 2).astype(np.float64)
    absent = O.absent_mask(xy)
    pin = np.zeros(n, np.uint8)
    pin[pinned] = 1
    Lk = O.Links(n, links, r_in, r_out)
    pos, a = xy.astype(np.float32).astype(np.float64), 1.0
    f32 = lambda q: q.astype(np.float32).astype(np.float64)
    for _ in range(ticks):
        pos = f32(O.integrate(pos, O.gravity(pos, S, a, 0.25), 0.85, S, pin, absent))
        if center:
            pos = f32(O.integrate(pos, O.center(pos, a, center, absent), 0.85, S, pin, absent))
        for d in (0, 1):
            v = Lk.force(pos, d, a, 1.0, 10.0, 1.0, 1.2, absent)
            pos = f32(O.integrate(pos, v, 0.85, S, pin, absent))
        a = O.alpha_step(a, 5000)
    e = np.hypot(*(gpu - pos.reshape(-1, 2)).T)
    live = absent == 0
    assert np.isnan(gpu[~live]).all()
    assert np.array_equal(gpu[pinned], xy.reshape(-1, 2)[pinned].astype(np.float64))
    worst = float(e[live].max())
    say(f"{ticks} ticks, centre {center}: max |dp| = {worst:.3e} ({worst / S:.2e} S)")
    assert worst <= TOL * S * ticks
