"""The free implementation parameters (cg_tuning) through the C ABI: the
bit-neutral ones (ticks per captured graph, points per lane of the cell pass, the force
kernel's LDS window) leave every result bit unchanged over whole ticks; the
link tiers keep every link force within 1e-5 * spaceSize of the oracle
(hubs split into segments included); invalid values are refused.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(S, **kw):
    p = _lib.default_params()
    p.space_size = S
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_bit_neutral_tuning_keeps_bits():
    rng = np.random.default_rng(3)
    S, n = 4096.0, 30_000
    xy = datagen.uniform_square(n, S, 3)
    xy[: 2 * 8000] = np.clip(rng.normal(1500, 25, 2 * 8000), 0, S).astype(np.float32)  # crowded cells
    links = datagen.barabasi_albert(n, 3, 3)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 4), datagen.uniform(m, 0, 1, 5)
    jit = datagen.jitter(n, 6)
    seeds = datagen.uniform(40, 0, 1, 7)
    p = params(S, repulsion=0.5)
    outs = []
    for t in ({}, {"graph_ticks": 1}, {"graph_ticks": 32, "force_tile": 512},
              {"cell_items": 4}, {"cell_items": 4, "graph_ticks": 16, "force_tile": 512}):
        with _lib.Context() as c:
            c.set_tuning(t)
            c.set_positions(xy)
            c.set_jitter(jit)
            c.set_links(links, None, r_in, r_out)
            c.set_alpha(1.0)
            c.step(p, 40, seeds=seeds)
            outs.append(c.get_positions())
            assert c.get_tuning() == dict(_lib.default_tuning(), **t)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))


@pytest.mark.parametrize("tiers", [{"link_lane_max": 32, "link_group_max": 128, "link_hub_entries": 4},
                                   {"link_lane_max": 96, "link_group_max": 256, "link_hub_entries": 8},
                                   {"link_lane_max": 2, "link_group_max": 64, "link_hub_entries": 6}])
def test_link_tiers_stay_within_tolerance(tiers):
    S, n = 8192.0, 200_000
    links = datagen.chung_lu(n, 2.1, 1.0, 1_000_000, 7)
    xy = datagen.uniform_disk(n, 4096, 4096, 2048, 3)
    m = links.size // 2
    r_in, r_out = datagen.uniform(m, 0, 1, 11), datagen.uniform(m, 0, 1, 12)
    c = _lib.Context()
    c.set_tuning(tiers)
    c.set_positions(xy)
    c.set_links(links, None, r_in, r_out)
    p = params(S, gravity=0.25, repulsion=0.5)
    c.set_alpha(1.0)
    c.step(p, 20)
    live = c.get_positions()
    a = float(np.float32(c.get_alpha()))
    Lk = O.Links(n, links, r_in, r_out)
    for d in (0, 1):
        got = c.eval_force(p, _lib.FORCE_LINK_IN + d, a)
        want = Lk.force(live, d, a, 1.0, 10.0, 1.0, 1.2)
        e = np.hypot(*(got.reshape(-1, 2).astype(np.float64) - want.reshape(-1, 2)).T).max()
        say(f"tiers {tiers} dir {d}: {e / S:.2e} S")
        assert e <= 1e-5 * S
    c.close()


def test_invalid_tuning_refused():
    with _lib.Context() as c:
        for bad in ({"graph_ticks": 0}, {"force_tile": 700}, {"cell_items": 2},
                    {"cell_items": 16}, {"link_lane_max": 256},
                    {"link_group_max": 512}, {"link_hub_entries": 5}):
            with pytest.raises(_lib.CosmosGLError) as e:
                c.set_tuning(bad)
            assert e.value.code == _lib.CG_ERR_INVALID
        assert c.get_tuning() == _lib.default_tuning()
