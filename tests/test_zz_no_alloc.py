"""The allocation-free entry point (cg_eval_manybody_noalloc) against the entry
point it varies, cg_eval_force(CG_FORCE_MANYBODY), through the C ABI: bit for
bit on two calls in a row, the second with a smaller input (the static scratch
keeps the first call's contents), the size limit refused cleanly, the
context's own memory untouched, and the oracle's force within 1e-5 * spaceSize.
"""
import sys

import numpy as np
import pytest

from graph_amd import _lib, datagen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_POINTS = 262144


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reference(xy, jit, p, alpha, seed):
    with _lib.Context() as c:
        c.set_positions(xy)
        c.set_jitter(jit)
        return c.eval_force(p, _lib.FORCE_MANYBODY, alpha, seed)


def test_noalloc_matches_eval_force_on_two_calls():
    S = 8192.0
    p = _lib.default_params()
    p.space_size = S
    p.repulsion = 0.5
    ctx = _lib.Context()
    held = ctx.device_bytes()
    cases = []
    # first: 200k points in a disk with absent points; second: 3000 points, crowded corner
    xy1 = datagen.uniform_disk(200_000, 4096, 4096, 2048, 31)
    xy1[2 * 100: 2 * 140] = np.nan
    cases.append((xy1, datagen.jitter(200_000, 32), 0.9, 0.3125))
    rng = np.random.default_rng(33)
    xy2 = rng.uniform(0, S, 6000).astype(np.float32)
    xy2[: 2 * 1500] = rng.uniform(10.0, 12.0, 3000).astype(np.float32)
    cases.append((xy2, datagen.jitter(3000, 34), 0.4, 0.75))
    for k, (xy, jit, alpha, seed) in enumerate(cases):
        say(f"no-alloc call {k}: n={xy.size // 2}")
        got = ctx.eval_manybody_noalloc(p, alpha, seed, xy, jit)
        want = reference(xy, jit, p, alpha, seed)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            f"call {k}: {int(np.sum(got != want))} values differ from cg_eval_force"
        assert ctx.device_bytes() == held
    # the smaller call also against the oracle
    xy, jit, alpha, seed = cases[1]
    ora = O.manybody(xy, S, alpha, 0.5, seed, jitter=jit)
    e = np.hypot(*(got.reshape(-1, 2).astype(np.float64) - ora.reshape(-1, 2)).T)
    say(f"no-alloc vs oracle: {e.max() / S:.2e} S")
    assert e.max() <= 1e-5 * S
    ctx.close()


def test_noalloc_size_limit():
    S = 4096.0
    p = _lib.default_params()
    p.space_size = S
    ctx = _lib.Context()
    xy = datagen.uniform_square(MAX_POINTS + 1, S, 35)
    with pytest.raises(_lib.CosmosGLError) as e:
        ctx.eval_manybody_noalloc(p, 1.0, 0.5, xy)
    assert e.value.code == _lib.CG_ERR_INVALID
    say(f"refused: {e.value}")
    # the largest input fits the scratch and matches the context path
    xy = xy[: 2 * MAX_POINTS].copy()
    got = ctx.eval_manybody_noalloc(p, 1.0, 0.5, xy)
    want = reference(xy, np.zeros(2 * MAX_POINTS, np.float32), p, 1.0, 0.5)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    ctx.close()
