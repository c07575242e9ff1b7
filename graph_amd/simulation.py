"""Host-side mirror of the simulation half of cosmos.gl's ``Graph``.

The reference's public simulation API (src/index.ts) — ``setPointPositions``,
``setLinks``, ``setLinkStrength``, ``setPinnedPoints``, ``setPointSizes``,
``setPointClusters``, ``setClusterPositions``, ``setPointClusterStrength``,
``render``, ``start``, ``stop``, ``pause``, ``unpause``, ``step``,
``getPointPositions``, ``trackPointPositionsByIndices`` with
``getTrackedPointPositionsArray`` / ``Map``, point dragging and the
``simulation*`` config keys with their callbacks — over the C ABI of
``include/cosmosgl.h``.  Rendering, zoom and
picking are not part of this path.

Random draws follow the reference's order: the many-body jitter (two draws per
texel over ceil(sqrt(n))^2 texels, ForceManyBody/index.ts:133-138), then one
rest-length draw per INCOMING CSR entry, then per OUTGOING entry
(ForceLink/index.ts:76, instances created IN then OUT, src/index.ts:1663-1666),
then one near-field seed per tick (ForceManyBody/index.ts:521).  The stream is
numpy's PCG64 seeded by ``randomSeed``: the reference's npm ``random`` /
``seedrandom`` stream cannot be reproduced here.
"""
from __future__ import annotations

import math
from typing import Any, Callable

import numpy as np

from . import _lib

ALPHA_MIN = 0.001  # Store/index.ts:8

# src/variables.ts:10-19, 68-90 (simulation subset)
DEFAULT_CONFIG: dict[str, Any] = {
    "enableSimulation": True,
    "spaceSize": 4096,
    "pointDefaultSize": 4,
    "simulationDecay": 5000,
    "simulationGravity": 0.25,
    "simulationCenter": 0,
    "simulationRepulsion": 1.0,
    "simulationLinkSpring": 1,
    "simulationLinkDistance": 10,
    "simulationLinkDistRandomVariationRange": [1, 1.2],
    "simulationRepulsionFromMouse": 2,
    "simulationFriction": 0.85,
    "simulationCollision": 0,
    "simulationCollisionRadius": None,
    "simulationCollisionPadding": 0,
    "simulationCluster": 0.1,
    "enableRightClickRepulsion": False,
    "enableSimulationDuringZoom": False,
    "randomSeed": None,
    "onSimulationStart": None,
    "onSimulationTick": None,
    "onSimulationEnd": None,
    "onSimulationPause": None,
    "onSimulationUnpause": None,
}


def adjust_space_size(config_space_size: float, max_texture_size: float | None) -> float:
    """Store.adjustSpaceSize (Store/index.ts:209-233)."""
    s = config_space_size
    if s is None or not (s > 0) or not math.isfinite(s):
        s = DEFAULT_CONFIG["spaceSize"]
    s = max(s, 2)
    if max_texture_size is None or not math.isfinite(max_texture_size) or max_texture_size < 2:
        return float(s)
    if s >= max_texture_size:
        return float(max(max_texture_size / 2, 2))
    return float(s)


class Graph:
    """The simulation part of ``Graph`` driven through the MI355X C ABI."""

    def __init__(self, config: dict | None = None, device: int = 0, max_texture_size: float | None = None,
                 hbm_budget: int = 0):
        self.config = dict(DEFAULT_CONFIG)
        self.config.update(config or {})
        self._ctx = _lib.Context(device=device, hbm_budget=hbm_budget)
        self._max_texture_size = max_texture_size
        seed = self.config.get("randomSeed")
        self._rng = np.random.default_rng(None if seed is None else abs(hash(seed)) % (2**63))
        self._input_positions: np.ndarray | None = None
        self._input_links: np.ndarray | None = None
        self._input_strength: np.ndarray | None = None
        self._input_sizes: np.ndarray | None = None
        self._pinned: list[int] | None = None
        self._clusters: np.ndarray | None = None
        self._cluster_positions: np.ndarray | None = None
        self._cluster_strength: np.ndarray | None = None
        self._tracked: np.ndarray = np.zeros(0, np.uint32)
        self._dirty = {"positions": False, "links": False, "sizes": False, "pinned": False,
                       "clusters": False}
        self.alpha = 1.0
        self.is_simulation_running = False
        self.simulation_progress = 0.0
        self.is_right_click_mouse = False
        self.mouse_position = (0.0, 0.0)
        self.is_zooming = False
        self.points_number: int | None = None
        self.links_number: int | None = None

    # ---- config ----------------------------------------------------------------
    def setConfig(self, config: dict) -> None:
        keep = {k: self.config.get(k) for k in ("randomSeed",)}
        self.config = dict(DEFAULT_CONFIG)
        self.config.update(config or {})
        self.config.update(keep)

    def setConfigPartial(self, config: dict) -> None:
        for k, v in (config or {}).items():
            if k == "randomSeed":
                continue
            self.config[k] = DEFAULT_CONFIG.get(k) if v is None else v

    @property
    def space_size(self) -> float:
        return adjust_space_size(self.config["spaceSize"], self._max_texture_size)

    def _params(self, run_forces: bool) -> _lib.CgSimParams:
        c = self.config
        p = _lib.default_params()
        p.space_size = self.space_size
        p.decay = c["simulationDecay"]
        p.gravity = c["simulationGravity"] or 0.0
        p.center = c["simulationCenter"] or 0.0
        p.repulsion = c["simulationRepulsion"]
        p.link_spring = c["simulationLinkSpring"]
        p.link_distance = c["simulationLinkDistance"]
        rng = c["simulationLinkDistRandomVariationRange"]
        rng = rng if isinstance(rng, (list, tuple)) and len(rng) == 2 else [0, 0]  # ensureVec2
        p.link_dist_random_variation_range[0] = rng[0]
        p.link_dist_random_variation_range[1] = rng[1]
        p.repulsion_from_mouse = c["simulationRepulsionFromMouse"]
        p.friction = c["simulationFriction"]
        p.collision = c["simulationCollision"] or 0.0
        p.collision_radius = c["simulationCollisionRadius"] or 0.0
        p.collision_padding = c["simulationCollisionPadding"] or 0.0
        p.point_default_size = c["pointDefaultSize"]
        p.enable_simulation = 1 if c["enableSimulation"] else 0
        p.run_forces = 1 if run_forces else 0
        p.mouse_repulsion = 1 if (self.is_right_click_mouse and c["enableRightClickRepulsion"]) else 0
        p.mouse_position[0], p.mouse_position[1] = self.mouse_position
        p.cluster = c["simulationCluster"] if c["simulationCluster"] is not None else 0.1
        return p

    # ---- data setters (src/index.ts:548, 627, 723, 836, 912) -------------------------
    def setPointPositions(self, positions, dontRescale: bool | None = None) -> None:
        a = np.asarray(positions, dtype=np.float32).reshape(-1)
        if a.size % 2:
            a = a[:-1]  # GraphData/index.ts:117-120
        self._input_positions = a
        self._dirty["positions"] = True

    def setLinks(self, links) -> None:
        a = np.asarray(links, dtype=np.float32).reshape(-1)
        if a.size % 2:
            a = a[:-1]  # GraphData/index.ts:296-299
        self._input_links = a
        self._dirty["links"] = True

    def setLinkStrength(self, strength) -> None:
        self._input_strength = None if strength is None else np.asarray(strength, dtype=np.float32)
        self._dirty["links"] = True

    def setPointSizes(self, sizes) -> None:
        self._input_sizes = None if sizes is None else np.asarray(sizes, dtype=np.float32)
        self._dirty["sizes"] = True

    def setPinnedPoints(self, indices) -> None:
        self._pinned = None if indices is None else [int(i) for i in indices if float(i).is_integer()]
        self._dirty["pinned"] = True

    def setPointClusters(self, clusters) -> None:
        """src/index.ts:854: one cluster index per point; None (undefined) = no cluster."""
        if clusters is None:
            self._clusters = None
        else:
            self._clusters = np.array([-1 if c is None or not (c == c) else int(c) for c in clusters],
                                      dtype=np.int32)
        self._dirty["clusters"] = True

    def setClusterPositions(self, positions) -> None:
        """src/index.ts:871: [x, y] per cluster; None entries pull toward the centroid."""
        if positions is None:
            self._cluster_positions = None
        else:
            self._cluster_positions = np.array([-1.0 if v is None else float(v) for v in positions],
                                               dtype=np.float32)
        self._dirty["clusters"] = True

    def setPointClusterStrength(self, strength) -> None:
        """src/index.ts:888: per-point cluster force coefficient (default 1)."""
        self._cluster_strength = None if strength is None else np.asarray(strength, dtype=np.float32)
        self._dirty["clusters"] = True

    # ---- lifecycle ---------------------------------------------------------------
    def render(self, simulationAlpha: float | None = None) -> None:
        """GraphData.update + create (src/index.ts:936-1003, 1635-1688)."""
        if self._dirty["positions"] or self._dirty["links"]:
            self._create()
        if self._dirty["sizes"] and self.points_number is not None:
            sizes = self._input_sizes
            if sizes is not None and sizes.size != self.points_number:
                sizes = None  # GraphData.updatePointSize: mismatched length -> defaults
            self._ctx.set_point_sizes(sizes)
        if self._dirty["pinned"] or self._dirty["positions"]:
            n = self.points_number or 0
            self._ctx.set_pinned([i for i in (self._pinned or []) if 0 <= i < n])
        if self._dirty["clusters"]:
            self._ctx.set_point_clusters(self._clusters)
            self._ctx.set_cluster_positions(self._cluster_positions)
            self._ctx.set_point_cluster_strength(self._cluster_strength)
        for k in self._dirty:
            self._dirty[k] = False
        if simulationAlpha is not None:
            self.alpha = float(simulationAlpha)
            self._ctx.set_alpha(self.alpha)

    def _create(self) -> None:
        pos = self._input_positions if self._input_positions is not None else np.zeros(0, np.float32)
        n = pos.size // 2
        self.points_number = n
        self._ctx.set_positions(pos)
        tex = math.ceil(math.sqrt(n))
        # ForceManyBody.create: two draws per texel, the first n pairs belong to points
        jit = (self._rng.uniform(-1, 1, size=2 * tex * tex) * 0.00001).astype(np.float32)
        if n:
            self._ctx.set_jitter(jit[: 2 * n])
        links = self._input_links
        if links is None:
            self.links_number = None
            return
        m = links.size // 2
        self.links_number = m
        strength = self._input_strength
        if strength is not None and strength.size != m:
            strength = None  # GraphData.updateLinkStrength
        r_in = np.zeros(m, np.float32)
        r_out = np.zeros(m, np.float32)
        s, t = links[0::2], links[1::2]
        valid = (s >= 0) & (t >= 0) & (np.floor(s) == s) & (np.floor(t) == t) & (s < n) & (t < n)
        idx = np.nonzero(valid)[0]
        order_in = idx[np.argsort(s[idx], kind="stable")]
        order_out = idx[np.argsort(t[idx], kind="stable")]
        r_in[order_in] = self._rng.uniform(0, 1, size=order_in.size)
        r_out[order_out] = self._rng.uniform(0, 1, size=order_out.size)
        self._ctx.set_links(links, strength, r_in, r_out)

    def start(self, alpha: float = 1.0) -> None:
        """src/index.ts:1457-1476: (re)starts at `alpha`; onSimulationStart only
        when the simulation was not already running."""
        if not self.config["enableSimulation"] or not self.points_number:
            return
        was_running = self.is_simulation_running
        self.is_simulation_running = True
        self.simulation_progress = 0.0
        self.alpha = float(alpha)
        self._ctx.set_alpha(self.alpha)
        cb = self.config.get("onSimulationStart")
        if cb and not was_running:
            cb()

    def stop(self) -> None:
        """src/index.ts:1482-1490: stops and resets alpha and progress;
        onSimulationEnd only if the simulation was active."""
        was_active = self.is_simulation_running or self.alpha > 0 or self.simulation_progress > 0
        self.is_simulation_running = False
        self.simulation_progress = 0.0
        self.alpha = 0.0
        self._ctx.set_alpha(0.0)
        cb = self.config.get("onSimulationEnd")
        if cb and was_active:
            cb()

    def pause(self) -> None:
        """src/index.ts:1496-1501: keeps alpha and progress."""
        if not self.is_simulation_running:
            return
        self.is_simulation_running = False
        cb = self.config.get("onSimulationPause")
        if cb:
            cb()

    def unpause(self) -> None:
        """src/index.ts:1510-1521."""
        if not self.config["enableSimulation"] or self.is_simulation_running:
            return
        self.is_simulation_running = True
        cb = self.config.get("onSimulationUnpause")
        if cb:
            cb()

    def step(self) -> None:
        """Graph.step (src/index.ts:1528-1540): one tick even when paused."""
        if not self.config["enableSimulation"] or not self.points_number:
            return  # no points: pointsTextureSize is 0
        self._run_simulation_step(True)

    def frame(self) -> None:
        """One rAF frame of the simulation (src/index.ts:2152-2177): the run ends
        (end(), src/index.ts:2335-2341) before the frame's step once alpha has
        fallen below ALPHA_MIN; the step then runs no forces."""
        if not self.points_number:
            return
        if self.is_simulation_running and self.alpha < ALPHA_MIN:
            self.is_simulation_running = False
            self.simulation_progress = 1.0
            cb = self.config.get("onSimulationEnd")
            if cb:
                cb()
        self._run_simulation_step(False)

    def run_ticks(self, nticks: int) -> np.ndarray:
        """nticks forced ticks in one call (fewer host round trips); returns the alpha trace."""
        return self._run_simulation_step(True, nticks)

    def _run_simulation_step(self, force_execution: bool, nticks: int = 1):
        if not self.config["enableSimulation"]:
            return None
        zoom_gate = self.is_zooming and not self.config["enableSimulationDuringZoom"]
        run = force_execution or (self.is_simulation_running and not zoom_gate)
        seeds = self._rng.uniform(0, 1, size=nticks).astype(np.float32) if run else None
        trace = self._ctx.step(self._params(run), nticks, seeds=seeds, trace=True)
        if run:
            self.alpha = float(trace[-1])
            self.simulation_progress = math.sqrt(min(1.0, ALPHA_MIN / self.alpha))
            cb = self.config.get("onSimulationTick")
            if cb:
                for a in trace:
                    cb(float(a), None, None)
        return trace

    # ---- tracking and dragging (src/index.ts:1346-1388; Points/index.ts:1961-1988) ----
    def trackPointPositionsByIndices(self, indices) -> None:
        self._tracked = np.asarray(list(indices), dtype=np.uint32)

    def getTrackedPointPositionsArray(self) -> list[float]:
        """Tracked positions in tracking order; absent (or past the count) read NaN."""
        if self._tracked.size == 0:
            return []
        return self._ctx.get_positions_at(self._tracked).astype(np.float64).tolist()

    def getTrackedPointPositionsMap(self) -> dict[int, tuple[float, float]]:
        """Tracked positions by index; absent points and indices past the count are omitted."""
        if self._tracked.size == 0:
            return {}
        xy = self._ctx.get_positions_at(self._tracked).reshape(-1, 2)
        return {int(i): (float(x), float(y)) for i, (x, y) in zip(self._tracked, xy)
                if x == x and y == y}

    def drag_point(self, index: int, position) -> None:
        """Points.drag: the dragged (hovered) point takes the mouse position."""
        self._ctx.drag_points([index], np.asarray(position, dtype=np.float32).reshape(1, 2))

    def getPointPositions(self) -> list[float]:
        """src/index.ts:1106-1130 — absent points read back as NaN."""
        if not self.points_number:
            return []
        return self._ctx.get_positions().astype(np.float64).tolist()

    def get_point_positions_array(self) -> np.ndarray:
        return self._ctx.get_positions()

    @property
    def progress(self) -> float:
        return self.simulation_progress

    def destroy(self) -> None:
        self._ctx.close()
