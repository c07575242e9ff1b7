"""CPU-side checks of the native library and the host logic (no GPU calls)."""
import ctypes
import math
import subprocess

import numpy as np
import pytest

from graph_amd import _lib, datagen


def test_header_declares_the_boundary():
    syms = _lib.header_symbols()
    for name in ["cg_create", "cg_set_positions", "cg_set_links", "cg_step", "cg_get_positions",
                 "cg_destroy", "cg_last_error", "cg_set_jitter", "cg_set_point_sizes",
                 "cg_set_pinned", "cg_get_alpha", "cg_set_alpha"]:
        assert name in syms


def test_library_exports_every_header_symbol():
    l = _lib.lib()
    missing = [s for s in _lib.header_symbols() if not hasattr(l, s)]
    assert not missing, missing


def test_default_params_mirror_reference_defaults():
    # src/variables.ts:10-19, 68-83
    p = _lib.default_params()
    assert p.space_size == 4096 and p.decay == 5000
    assert p.gravity == pytest.approx(0.25) and p.center == 0
    assert p.repulsion == 1.0 and p.link_spring == 1 and p.link_distance == 10
    assert list(p.link_dist_random_variation_range) == pytest.approx([1, 1.2])
    assert p.repulsion_from_mouse == 2 and p.friction == pytest.approx(0.85)
    assert p.collision == 0 and p.collision_padding == 0 and p.point_default_size == 4
    assert p.enable_simulation == 1


def test_struct_layout_matches_header(tmp_path):
    # the ctypes mirror must agree with the C compiler's layout of the header structs
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "cosmosgl.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu\", sizeof(cg_sim_params),"
        " offsetof(cg_sim_params, friction), offsetof(cg_sim_params, enable_simulation),"
        " offsetof(cg_sim_params, mouse_position), sizeof(cg_create_params));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(_lib.HEADER.parent), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    P = _lib.CgSimParams
    assert got == [ctypes.sizeof(P), P.friction.offset, P.enable_simulation.offset,
                   P.mouse_position.offset, ctypes.sizeof(_lib.CgCreateParams)]


def test_barabasi_albert():
    links = datagen.barabasi_albert(10_000, 2, 42)
    m = links.size // 2
    assert m == sum(min(2, t) for t in range(1, 10_000))  # 19,997
    s, t = links[0::2].astype(np.int64), links[1::2].astype(np.int64)
    assert (s > t).all()                      # new node -> earlier node
    assert np.array_equal(links, datagen.barabasi_albert(10_000, 2, 42))
    pairs = set(zip(s.tolist(), t.tolist()))
    assert len(pairs) == m                    # distinct targets per new node
    deg = np.bincount(np.concatenate([s, t]), minlength=10_000)
    assert deg.max() > 50                     # preferential attachment grows hubs


def test_chung_lu_power_law():
    n, m = 200_000, 1_000_000
    links = datagen.chung_lu(n, 2.1, 1.0, m, 7)
    assert links.size == 2 * m
    s, t = links[0::2].astype(np.int64), links[1::2].astype(np.int64)
    assert (s != t).all() and s.min() >= 0 and max(s.max(), t.max()) < n
    deg = np.bincount(np.concatenate([s, t]), minlength=n)
    assert deg.max() > 5_000                  # hubs for the chunked link path


def test_sbm_cross_fraction():
    n = 100_000
    links = datagen.sbm(n, 100, 1.1, 5.0, 0.1, 11)
    m = links.size // 2
    assert 0.9 * n * 2.5 / 0.9 < m < 1.1 * n * 2.5 / 0.9


def test_positions_generators():
    xy = datagen.uniform_disk(100_000, 4096, 4096, 2048, 7).reshape(-1, 2)
    r = np.hypot(xy[:, 0] - 4096, xy[:, 1] - 4096)
    assert r.max() <= 2048.01
    assert abs(np.median(r) - 2048 / math.sqrt(2)) < 20
    sq = datagen.uniform_square(10_000, 4096, 42)
    assert sq.min() >= 0 and sq.max() < 4096
    j = datagen.jitter(1000, 3)
    assert np.abs(j).max() <= 1e-5


def test_space_size_adjustment():
    from graph_amd.simulation import adjust_space_size

    # Store/index.ts:209-233
    assert adjust_space_size(16384, 16384) == 8192
    assert adjust_space_size(8192, 16384) == 8192
    assert adjust_space_size(-5, 16384) == 4096
    assert adjust_space_size(1, None) == 2


def test_napi_addon_exports():
    import shutil

    node = shutil.which("node")
    addon = _lib.PKG / "napi" / "cosmosgl.node"
    if node is None or not addon.exists():
        pytest.skip("node or the addon is not available")
    code = ("const a=require(process.argv[1]);"
            "console.log(Object.keys(a).sort().join(','));"
            "console.log(JSON.stringify(a.defaultParams()))")
    out = subprocess.run([node, "-e", code, str(addon)], capture_output=True, text=True, check=True).stdout
    names, params = out.strip().split("\n")
    assert names.split(",") == sorted(["create", "destroy", "setPositions", "setLinks", "setJitter",
                                        "setPointSizes", "setPinned", "setPointClusters",
                                        "setClusterPositions", "setPointClusterStrength", "setAlpha",
                                        "getAlpha", "step", "getPositions", "evalForce", "defaultParams"])
    import json

    d = json.loads(params)
    assert d["spaceSize"] == 4096 and d["simulationDecay"] == 5000
    assert abs(d["simulationCluster"] - 0.1) < 1e-7


def test_typescript_dropin_matches_addon_exports():
    """graph_amd/napi/ForceSimulationHip.ts declares exactly the addon's exports
    (the property table of cosmosgl_napi.c) and calls nothing else."""
    import re

    napi = _lib.PKG / "napi"
    exports = set(re.findall(r'\{"(\w+)", 0, js_\w+', (napi / "cosmosgl_napi.c").read_text()))
    ts = (napi / "ForceSimulationHip.ts").read_text()
    iface = ts[ts.index("interface CosmosGLAddon"):ts.index("}\n", ts.index("interface CosmosGLAddon"))]
    declared = set(re.findall(r"^\s+(\w+) \(", iface, re.M))
    called = set(re.findall(r"\baddon\.(\w+)\(", ts))
    assert declared == exports - {"defaultParams"}, (declared ^ exports)
    assert called <= exports, called - exports
    # every key the drop-in passes to step() is one the addon reads
    read = set(re.findall(r'read_\w+\(env, obj, "(\w+)"', (napi / "cosmosgl_napi.c").read_text()))
    assert {"spaceSize", "runForces", "mouseRepulsion", "mousePosition"} <= read


def test_tuning_struct_and_table_lookup(tmp_path):
    from graph_amd import autotune

    d = _lib.default_tuning()
    assert d == {"graph_ticks": 8, "cell_items": 8, "force_tile": 1024,
                 "link_lane_max": 64, "link_group_max": 256, "link_hub_entries": 6}
    assert ctypes.sizeof(_lib.CgTuning) == 6 * 4
    table = {"entries": [
        {"points": 10_000, "mean_degree": 4.0, "tuning": dict(d, graph_ticks=32, link_lane_max=32)},
        {"points": 1_000_000, "mean_degree": 10.0, "tuning": dict(d, force_tile=512, link_hub_entries=8)}]}
    # nearest in (log2 points, log2 degree); bit-exact lookups keep the link tiers at their defaults
    t = autotune.lookup(20_000, 4.0, table=table)
    assert t["graph_ticks"] == 32 and t["link_lane_max"] == 64
    t = autotune.lookup(2_000_000, 9.0, bit_exact=False, table=table)
    assert t["force_tile"] == 512 and t["link_hub_entries"] == 8
    assert autotune.lookup(5, 1.0, table={"entries": []}) == d
    assert all(autotune.valid(dict(d, **{k: v})) for k, vs in autotune.BIT_NEUTRAL.items() for v in vs)
    assert not autotune.valid(dict(d, cell_items=16))
