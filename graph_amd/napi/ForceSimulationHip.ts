// ForceSimulationHip — the drop-in for cosmos.gl's force passes on an MI355X.
//
// A maintainer adds this file as src/modules/ForceSimulationHip/index.ts of the
// cosmos.gl 3.4.1 tree.  It replaces the Force* modules and the
// Points.updatePosition calls inside Graph.runSimulationStep
// (src/index.ts:2031-2078: gravity, center, many-body, link IN/OUT, clusters,
// collision, each with its integrate pass, and the alpha decay of :2084-2088)
// by one call into the Node-API addon built from graph_amd/napi/cosmosgl_napi.c,
// which drives the C ABI of include/cosmosgl.h.  The public Graph API, the
// simulation* config keys, the callbacks and the luma.gl rendering stay as they
// are: Points uploads readPositions() into currentPositionTexture
// (Points/index.ts:3025-3033) on frames that draw.
//
// Random draws keep the reference's stream and order: the many-body jitter
// (ForceManyBody/index.ts:132-163: two draws per texel of the points texture),
// one rest-length draw per CSR entry of the INCOMING then the OUTGOING instance
// (ForceLink/index.ts:76, created in that order by src/index.ts:1663-1666) and
// one near-field seed per tick (ForceManyBody/index.ts:521), all from
// store.getRandomFloat.

import type { GraphConfigInterface } from '@/graph/config'
import { CoreModule } from '@/graph/modules/core-module'

/** The addon's exports (graph_amd/napi/cosmosgl_napi.c). */
interface CosmosGLAddon {
  create (device: number, hbmBudget: number): unknown;
  destroy (h: unknown): void;
  setPositions (h: unknown, xy: Float32Array): void;
  setLinks (h: unknown, links: Float32Array, strength: Float32Array | null,
    randIn: Float32Array, randOut: Float32Array): void;
  setJitter (h: unknown, xy: Float32Array): void;
  setPointSizes (h: unknown, sizes: Float32Array | null): void;
  setPinned (h: unknown, indices: Uint32Array): void;
  setPointClusters (h: unknown, clusters: Int32Array | null): void;
  setClusterPositions (h: unknown, xy: Float32Array | null): void;
  setPointClusterStrength (h: unknown, strength: Float32Array | null): void;
  setAlpha (h: unknown, alpha: number): void;
  getAlpha (h: unknown): number;
  step (h: unknown, params: SimParams, nticks: number, seeds: Float32Array | null): Float64Array;
  getPositions (h: unknown, out: Float32Array): Float32Array;
  evalForce (h: unknown, params: SimParams, force: number, alpha: number, seed: number,
    out: Float32Array): Float32Array;
}

/** The simulation* keys plus the gates of runSimulationStep (cg_sim_params). */
type SimParams = Partial<GraphConfigInterface> & {
  spaceSize: number;
  runForces: boolean;
  mouseRepulsion: boolean;
  mousePosition?: number[];
}

// eslint-disable-next-line @typescript-eslint/no-require-imports
const addon: CosmosGLAddon = require('@cosmos.gl/mi355x') // the .node built from cosmosgl_napi.c

export class ForceSimulationHip extends CoreModule {
  private readonly h = addon.create(0, 0)
  private xy = new Float32Array(0)

  /** On every data update (src/index.ts:1655-1666), where the reference calls
   * ForceManyBody.create and ForceLink.create (INCOMING, then OUTGOING). */
  public create (): void {
    const { data, store } = this
    const n = data.pointsNumber ?? 0
    addon.setPositions(this.h, data.pointPositions ?? new Float32Array(0))
    // ForceManyBody.create: two draws per texel; the first n texels are points
    const texels = store.pointsTextureSize * store.pointsTextureSize
    const jitter = new Float32Array(2 * n)
    for (let i = 0; i < texels; ++i) {
      const rx = store.getRandomFloat(-1, 1) * 0.00001
      const ry = store.getRandomFloat(-1, 1) * 0.00001
      if (i < n) { jitter[2 * i] = rx; jitter[2 * i + 1] = ry }
    }
    addon.setJitter(this.h, jitter)
    if (data.links) {
      // ForceLink.create x2: one draw per entry, entries in CSR order, IN then OUT
      const m = data.links.length / 2
      const randIn = new Float32Array(m)
      const randOut = new Float32Array(m)
      data.sourceIndexToTargetIndices?.forEach(entries => entries?.forEach(([, link]) => {
        randIn[link] = store.getRandomFloat(0, 1)
      }))
      data.targetIndexToSourceIndices?.forEach(entries => entries?.forEach(([, link]) => {
        randOut[link] = store.getRandomFloat(0, 1)
      }))
      addon.setLinks(this.h, data.links, data.linkStrength ?? null, randIn, randOut)
    }
    addon.setPointSizes(this.h, data.pointSizes ?? null)
    addon.setPinned(this.h, Uint32Array.from(data.inputPinnedPoints ?? []))
    if (data.pointClusters || data.clusterPositions) {
      addon.setPointClusters(this.h, Int32Array.from(data.pointClusters ?? [], c => c ?? -1))
      addon.setClusterPositions(this.h, Float32Array.from(data.clusterPositions ?? [], v => v ?? -1))
      addon.setPointClusterStrength(this.h, data.clusterStrength ?? null)
    } else {
      addon.setPointClusters(this.h, null)
      addon.setClusterPositions(this.h, null)
    }
    this.xy = new Float32Array(2 * n)
  }

  /** One tick: the block of runSimulationStep from the mouse pass to the alpha
   * decay.  Returns alpha after the tick, which the caller stores in
   * store.alpha before onSimulationTick (src/index.ts:2084-2094). */
  public run (forceExecution: boolean, isSimulationRunning: boolean, zoomGate: boolean,
    rightClickRepulsion: boolean): number {
    const { config, store } = this
    const runForces = forceExecution || (isSimulationRunning && !zoomGate)
    // the per-tick near-field seed is drawn only when the many-body pass runs
    const seeds = runForces ? Float32Array.of(store.getRandomFloat(0, 1)) : null
    addon.setAlpha(this.h, store.alpha)
    const trace = addon.step(this.h, {
      ...config,
      spaceSize: store.adjustedSpaceSize,
      runForces,
      mouseRepulsion: rightClickRepulsion,
      mousePosition: store.mousePosition,
    }, 1, seeds)
    return trace[0] ?? store.alpha
  }

  /** Current positions [x0, y0, ...] (absent points NaN), for Points to upload. */
  public readPositions (): Float32Array {
    addon.getPositions(this.h, this.xy)
    return this.xy
  }

  public destroy (): void {
    addon.destroy(this.h)
  }
}
