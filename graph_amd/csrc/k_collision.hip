// k_collision.hip — spatial-hash collision over four offset grids.
//
// build-grid.vert bins every point into four grids of the same cell size h,
// offset by o in {(0,0), (.5,0), (0,.5), (.5,.5)} cells: cell = clamp(floor((p +
// o h) / h), 0, G-1) per axis.  Here each point is sorted once, by its cell of
// the unshifted (base) grid, its position riding the sort; a point's cell in a
// shifted grid is its base cell or the next one along each shifted axis, so a
// cell of grid o gathers the points of at most 4 base cells and keeps those
// whose own grid-o cell (the exact formula above) is that cell.  Per cell, in
// a fixed order (base cells ascending, points by index within each), the sums
// of (x - x0, y - y0, size) in f64 and the count; relative to the cell's
// origin so the force pass can remove the point's own contribution without
// f32 cancellation.  Where base cells outnumber points, the cells are found from
// the occupied base cells instead of visiting every cell (k_col_cells_sparse),
// after clearing the last pass's ones (k_col_clear).  The force pass visits the
// points in base-cell order (a wave reads the same few cells), sweeps the 3x3
// cell aggregates of all four grids and integrates.
//
// Reference: ForceCollision/index.ts:61-164 (grid geometry, sizes),
// :264-345 (run); build-grid.vert:28-62; force-collision-spatial.frag:41-168.
#include <algorithm>
#include <cmath>

#include "cg_internal.h"

namespace cg {
namespace {

__constant__ float c_off[4][2] = {{0.f, 0.f}, {0.5f, 0.f}, {0.f, 0.5f}, {0.5f, 0.5f}};

__global__ void k_fill32(uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// build-grid.vert:46-55: cell = clamp(floor((p + o*h) / h), 0, G-1)
__device__ inline int col_cell(float p, float o, float h, int G) {
    return cell_of(__fadd_rn(p, __fmul_rn(o, h)), h, G);
}

// One base-grid key per present collision point (the graph's block of base
// cells, col_c0 / 4, then the row-major cell), its index and position.
__global__ void k_col_keys(const float2* __restrict__ pos0, const float2* __restrict__ pos1,
                           const uint32_t* __restrict__ cpresent, uint32_t npc,
                           const uint32_t* __restrict__ gid, const GDesc* __restrict__ desc,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                           float2* __restrict__ pays, uint32_t one_graph) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    const uint32_t i = cpresent[k];
    const GDesc& d = desc[one_graph ? 0u : gid[i]];
    const float2 p = ((d.inbuf >> P_COLLISION) & 1u) ? pos1[i] : pos0[i];
    const int G = (int)d.col_G;
    const int cx = col_cell(p.x, 0.f, d.col_h, G), cy = col_cell(p.y, 0.f, d.col_h, G);
    keys[k] = d.col_c0 / 4u + (uint32_t)(cy * G + cx);
    vals[k] = i;
    pays[k] = p;
}

// Base-cell bounds from the sorted keys (end cleared first; end 0 = empty) and
// the sizes in sorted order.
__global__ void k_col_bounds(const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv, uint32_t nk,
                             const float* __restrict__ size, uint32_t* __restrict__ begin,
                             uint32_t* __restrict__ end, float* __restrict__ ssize) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const uint32_t key = sk[k];
    if (k == 0 || sk[k - 1] != key) begin[key] = k;
    if (k == nk - 1 || sk[k + 1] != key) end[key] = k + 1;
    ssize[k] = size[sv[k]];
}

// x / d for x < 2^24 and d >= 1: a float reciprocal estimate corrected to
// the exact quotient.
__device__ inline uint32_t udiv24(uint32_t x, uint32_t d) {
    uint32_t q = (uint32_t)((float)x * __builtin_amdgcn_rcpf((float)d));
    int32_t r = (int32_t)(x - q * d);
    while (r < 0) {
        --q;
        r += (int32_t)d;
    }
    while (r >= (int32_t)d) {
        ++q;
        r -= (int32_t)d;
    }
    return q;
}

// CELL_LANES lanes per cell of every graph's four grids: (sum(x - x0),
// sum(y - y0), sum(size), count) in f64, stored as f32, with (x0, y0) = cell
// index * h.  The lanes stride over the points of the cell's base cells
// (batches of COL_BATCH loads in flight), keep those whose own cell of this
// grid is this cell, and reduce in a fixed tree, so the result depends only on
// the points' order in the sort (by index within a base cell) and the lanes
// per cell.  Every term is the exact f64 difference of two f32 values (or an
// f32 size), so a cell's sums are exact, and independent of the lanes, while
// n terms whose magnitudes span less than 2^30 / n (always away from the
// grid's first row and column, where coordinates near 0 can be tiny).  A wave
// finds its graph once when all its cells belong to one graph.
// A cell whose base cells hold more than crowd_limit<CELL_LANES> points is
// deferred to k_col_crowd (one wave each: a hub's neighbours can pile
// thousands of points into one cell, which a few lanes would sum serially);
// crowd[0] counts them.  Measured: at 4 lanes per cell, deferring above 128
// points cut the serving mix's cell pass 590 -> ~100 us; at 16 lanes, cells of
// 512-2048 points (the 4M SBM graph's contracting communities) cost less
// in place than deferred.
constexpr int COL_BATCH = 4;
constexpr uint32_t COL_CROWD_MIN = 128;
template <int CELL_LANES>
constexpr uint32_t crowd_limit() { return CELL_LANES >= 16 ? 2048u : COL_CROWD_MIN; }
template <int CELL_LANES>
__global__ void __launch_bounds__(256) k_col_cells(const uint32_t* __restrict__ begin,
                                                   const uint32_t* __restrict__ end,
                                                   const float2* __restrict__ spos,
                                                   const float* __restrict__ ssize,
                                                   const uint32_t* __restrict__ colc0s,
                                                   const GDesc* __restrict__ desc, uint32_t G,
                                                   uint32_t CG, float4* __restrict__ cells,
                                                   uint32_t* __restrict__ crowd, uint32_t crowd_cap) {
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = gt / CELL_LANES, sub = gt % CELL_LANES;
    const bool valid = c < CG;
    // the graph of the wave's first and last cells; per lane only if they differ
    const uint32_t cw0 = __builtin_amdgcn_readfirstlane(valid ? c : CG - 1);
    const uint32_t cw1 = __builtin_amdgcn_readlane(valid ? c : CG - 1, 63);
    const uint32_t gw0 = find_graph(colc0s, G, cw0);
    const uint32_t gw1 = find_graph(colc0s, G, cw1);
    uint32_t gs = gw0;
    if (gw0 != gw1) gs = find_graph(colc0s, G, valid ? c : CG - 1);
    double ax = 0.0, ay = 0.0, as = 0.0;
    uint32_t n = 0;
    bool crowded = false;
    if (valid) {
        const GDesc& d = desc[gs];
        const int Gc = (int)d.col_G;
        const uint32_t GG = d.col_G * d.col_G;
        const uint32_t local = c - d.col_c0;
        const int o = (int)udiv24(local, GG);
        const uint32_t cell = local - (uint32_t)o * GG;
        const int cy = (int)udiv24(cell, d.col_G), cx = (int)cell - cy * Gc;
        const float h = d.col_h;
        const float x0 = (float)cx * h, y0 = (float)cy * h;
        const float ox = c_off[o][0], oy = c_off[o][1];
        const int sx = ox != 0.f ? 1 : 0, sy = oy != 0.f ? 1 : 0;  // base cells cx - sx .. cx
        const uint32_t b0 = d.col_c0 / 4u;
        // the base cells' point ranges, in order (empty ranges for cells off the grid)
        uint32_t rb[4], re[4], total = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int by = cy - sy + (q >> 1), bx = cx - sx + (q & 1);
            const bool use = (q >> 1) <= sy && (q & 1) <= sx && by >= 0 && bx >= 0;
            // q = 1 and 3 exist only with sx, q = 2 and 3 only with sy
            const bool in = use && ((q & 1) == 0 || sx) && ((q >> 1) == 0 || sy);
            const uint32_t bc = b0 + (uint32_t)(max(by, 0) * Gc + max(bx, 0));
            const uint32_t e = in ? end[bc] : 0u;
            const uint32_t b = e ? begin[bc] : 0u;
            rb[q] = b;
            re[q] = e;
            total += e - b;
        }
        crowded = total > crowd_limit<CELL_LANES>();
#pragma unroll 1
        for (int q = 0; q < 4 && !crowded; ++q) {
            const uint32_t e = re[q];
            if (e == 0) continue;
            {
                for (uint32_t k = rb[q] + sub; k < e; k += COL_BATCH * CELL_LANES) {
                    float2 pp[COL_BATCH];
                    float sz[COL_BATCH];
#pragma unroll
                    for (int u = 0; u < COL_BATCH; ++u) {
                        const uint32_t kk = min(k + u * CELL_LANES, e - 1);
                        pp[u] = spos[kk];
                        sz[u] = ssize[kk];
                    }
#pragma unroll
                    for (int u = 0; u < COL_BATCH; ++u) {
                        if (k + u * CELL_LANES >= e) break;
                        const float2 p = pp[u];
                        if ((sx && col_cell(p.x, ox, h, Gc) != cx) || (sy && col_cell(p.y, oy, h, Gc) != cy))
                            continue;
                        ax += (double)p.x - (double)x0;
                        ay += (double)p.y - (double)y0;
                        as += (double)sz[u];
                        ++n;
                    }
                }
            }
        }
    }
    // crowded cells to the list: one atomic per wave
    const bool push = crowded && sub == 0;
    const uint64_t pm = __ballot(push);
    if (pm) {
        const int first = __ffsll((long long)pm) - 1;
        const int lane = (int)(threadIdx.x & 63u);
        uint32_t base = 0;
        if (lane == first) base = atomicAdd(crowd, (uint32_t)__popcll(pm));
        base = __shfl(base, first, 64);
        if (push) {
            const uint32_t idx = base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
            if (idx < crowd_cap) crowd[1 + idx] = c;
        }
    }
    // fixed-tree reduction over the cell's lanes (aligned groups of CELL_LANES)
#pragma unroll
    for (int w = CELL_LANES / 2; w > 0; w >>= 1) {
        ax += __shfl_down(ax, w, CELL_LANES);
        ay += __shfl_down(ay, w, CELL_LANES);
        as += __shfl_down(as, w, CELL_LANES);
        n += __shfl_down(n, w, CELL_LANES);
    }
    if (valid && sub == 0 && !crowded) cells[c] = make_float4((float)ax, (float)ay, (float)as, (float)n);
}

// The deferred crowded cells, one wave each (grid-stride over the list): the
// same sums as k_col_cells, each lane striding over the points of the cell's
// base cells, then a fixed tree over the wave.
__global__ void __launch_bounds__(256) k_col_crowd(const uint32_t* __restrict__ crowd, uint32_t crowd_cap,
                                                   const uint32_t* __restrict__ begin,
                                                   const uint32_t* __restrict__ end,
                                                   const float2* __restrict__ spos,
                                                   const float* __restrict__ ssize,
                                                   const uint32_t* __restrict__ colc0s,
                                                   const GDesc* __restrict__ desc, uint32_t G,
                                                   float4* __restrict__ cells) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t count = min(crowd[0], crowd_cap);
    for (uint32_t j = wave; j < count; j += waves) {
        const uint32_t c = crowd[1 + j];
        const GDesc& d = desc[find_graph(colc0s, G, c)];
        const int Gc = (int)d.col_G;
        const uint32_t GG = d.col_G * d.col_G;
        const uint32_t local = c - d.col_c0;
        const int o = (int)udiv24(local, GG);
        const uint32_t cell = local - (uint32_t)o * GG;
        const int cy = (int)udiv24(cell, d.col_G), cx = (int)cell - cy * Gc;
        const float h = d.col_h;
        const float x0 = (float)cx * h, y0 = (float)cy * h;
        const float ox = c_off[o][0], oy = c_off[o][1];
        const int sx = ox != 0.f ? 1 : 0, sy = oy != 0.f ? 1 : 0;
        const uint32_t b0 = d.col_c0 / 4u;
        double ax = 0.0, ay = 0.0, as = 0.0;
        uint32_t n = 0;
        for (int by = max(cy - sy, 0); by <= cy; ++by) {
            for (int bx = max(cx - sx, 0); bx <= cx; ++bx) {
                const uint32_t bc = b0 + (uint32_t)(by * Gc + bx);
                const uint32_t e = end[bc];
                if (e == 0) continue;
                for (uint32_t k = begin[bc] + lane; k < e; k += COL_BATCH * 64) {
                    float2 pp[COL_BATCH];
                    float sz[COL_BATCH];
#pragma unroll
                    for (int u = 0; u < COL_BATCH; ++u) {
                        const uint32_t kk = min(k + u * 64, e - 1);
                        pp[u] = spos[kk];
                        sz[u] = ssize[kk];
                    }
#pragma unroll
                    for (int u = 0; u < COL_BATCH; ++u) {
                        if (k + u * 64 >= e) break;
                        const float2 p = pp[u];
                        if ((sx && col_cell(p.x, ox, h, Gc) != cx) || (sy && col_cell(p.y, oy, h, Gc) != cy))
                            continue;
                        ax += (double)p.x - (double)x0;
                        ay += (double)p.y - (double)y0;
                        as += (double)sz[u];
                        ++n;
                    }
                }
            }
        }
#pragma unroll
        for (int w = 32; w > 0; w >>= 1) {
            ax += __shfl_down(ax, w, 64);
            ay += __shfl_down(ay, w, 64);
            as += __shfl_down(as, w, 64);
            n += __shfl_down(n, w, 64);
        }
        if (lane == 0) cells[c] = make_float4((float)ax, (float)ay, (float)as, (float)n);
    }
}

// Sets whose base cells outnumber their points (many small graphs, each
// with a grid of up to 512^2 cells) leave most cells empty: there the cells
// are found from the points.  A cell of grid o gathers base cells
// (cx - sx .. cx, cy - sy .. cy); an occupied base cell (a run of the sorted
// keys) is a source of the cells (x .. x + sx, y .. y + sy) of each grid, and
// each cell is computed once, by its first occupied source in the order
// k_col_cells reads them, with the same sums.  The cells a pass writes are
// those of its occupied base cells, so the next pass first clears exactly
// them (and their base cells' ends), from the keys this pass sorted.

// Clears the cells (and base-cell ends) written by the pass whose sorted
// keys and points are `old` and `ov`: every cell of every grid that an
// occupied base cell is a source of.
__global__ void __launch_bounds__(256) k_col_clear(const uint32_t* __restrict__ old,
                                                   const uint32_t* __restrict__ ov, uint32_t npc,
                                                   const uint32_t* __restrict__ gid,
                                                   const GDesc* __restrict__ desc, uint32_t CB,
                                                   uint32_t one_graph, uint32_t* __restrict__ end,
                                                   float4* __restrict__ cells) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    const uint32_t key = old[k];
    if ((k > 0 && old[k - 1] == key) || key >= CB) return;
    end[key] = 0u;
    const GDesc& d = desc[one_graph ? 0u : gid[ov[k]]];
    const int Gc = (int)d.col_G;
    const uint32_t local = key - d.col_c0 / 4u, GG = d.col_G * d.col_G;
    const int y = (int)udiv24(local, d.col_G), x = (int)local - y * Gc;
    for (int o = 0; o < 4; ++o) {
        const int sx = c_off[o][0] != 0.f ? 1 : 0, sy = c_off[o][1] != 0.f ? 1 : 0;
        for (int dy = 0; dy <= sy; ++dy)
            for (int dx = 0; dx <= sx; ++dx) {
                const int cx = x + dx, cy = y + dy;
                if (cx < Gc && cy < Gc)
                    cells[d.col_c0 + (uint32_t)o * GG + (uint32_t)(cy * Gc + cx)] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
    }
}

// Nine lanes per sorted point; at the head of a run (an occupied base cell)
// lane j takes the j-th cell of the four grids the base cell is a source of
// (1 + 2 + 2 + 4: grid o's cells x .. x + sx by y .. y + sy) and computes it
// if the base cell is its first occupied source, summing it in one lane over
// its sources' points in order, COL_BATCH loads at a time (a cell whose
// sources hold more than COL_CROWD_MIN points goes to k_col_crowd, as in
// k_col_cells).
__constant__ int8_t c_cand[9][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {2, 0, 0}, {2, 0, 1},
                                    {3, 0, 0}, {3, 1, 0}, {3, 0, 1}, {3, 1, 1}};  // {grid, dx, dy}
__global__ void __launch_bounds__(256) k_col_cells_sparse(const uint32_t* __restrict__ sk,
                                                          const uint32_t* __restrict__ sv, uint32_t npc,
                                                          const uint32_t* __restrict__ begin,
                                                          const uint32_t* __restrict__ end,
                                                          const float2* __restrict__ spos,
                                                          const float* __restrict__ ssize,
                                                          const uint32_t* __restrict__ gid,
                                                          const GDesc* __restrict__ desc,
                                                          float4* __restrict__ cells,
                                                          uint32_t* __restrict__ crowd, uint32_t crowd_cap,
                                                          uint32_t one_graph) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = t / 9u, j = t - k * 9u;
    if (k >= npc) return;
    const uint32_t key = sk[k];
    if (k > 0 && sk[k - 1] == key) return;
    const GDesc& d = desc[one_graph ? 0u : gid[sv[k]]];
    const int Gc = (int)d.col_G;
    const uint32_t b0 = d.col_c0 / 4u, local = key - b0;
    const int y = (int)udiv24(local, d.col_G), x = (int)local - y * Gc;
    const int o = c_cand[j][0], dx = c_cand[j][1], dy = c_cand[j][2];
    const int cx = x + dx, cy = y + dy;
    if (cx >= Gc || cy >= Gc) return;
    const float ox = c_off[o][0], oy = c_off[o][1];
    const int sx = ox != 0.f ? 1 : 0, sy = oy != 0.f ? 1 : 0;
    // the cell's sources in k_col_cells' order; this base cell is source
    // (sy - dy, sx - dx) of them
    const int mine = (sy - dy) * 2 + (sx - dx);
    uint32_t rb[4], re[4], total = 0;
    bool first = true;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int by = cy - sy + (q >> 1), bx = cx - sx + (q & 1);
        const bool in = (q >> 1) <= sy && (q & 1) <= sx && by >= 0 && bx >= 0;
        const uint32_t bc = b0 + (uint32_t)(max(by, 0) * Gc + max(bx, 0));
        const uint32_t e = in ? end[bc] : 0u;
        const uint32_t bb = in ? begin[bc] : 0u;
        first &= !(q < mine && e != 0u);
        rb[q] = e ? bb : 0u;
        re[q] = e;
        total += e ? e - bb : 0u;
    }
    if (!first) return;
    const uint32_t c = d.col_c0 + (uint32_t)o * (uint32_t)(Gc * Gc) + (uint32_t)(cy * Gc + cx);
    if (total > COL_CROWD_MIN) {
        const uint32_t idx = atomicAdd(crowd, 1u);
        if (idx < crowd_cap) crowd[1 + idx] = c;
        return;
    }
    const float h = d.col_h;
    const float x0 = (float)cx * h, y0 = (float)cy * h;
    double ax = 0.0, ay = 0.0, as = 0.0;
    uint32_t n = 0;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const uint32_t e = re[q];
        for (uint32_t kk = rb[q]; kk < e; kk += COL_BATCH) {
            float2 pp[COL_BATCH];
            float sz[COL_BATCH];
#pragma unroll
            for (int u = 0; u < COL_BATCH; ++u) {
                const uint32_t ku = min(kk + u, e - 1);
                pp[u] = spos[ku];
                sz[u] = ssize[ku];
            }
#pragma unroll
            for (int u = 0; u < COL_BATCH; ++u) {
                if (kk + u >= e) break;
                const float2 p = pp[u];
                if ((sx && col_cell(p.x, ox, h, Gc) != cx) || (sy && col_cell(p.y, oy, h, Gc) != cy)) continue;
                ax += (double)p.x - (double)x0;
                ay += (double)p.y - (double)y0;
                as += (double)sz[u];
                ++n;
            }
        }
    }
    cells[c] = make_float4((float)ax, (float)ay, (float)as, (float)n);
}

// One lane per present collision point, in base-cell order.
__global__ void __launch_bounds__(256) k_col_force(const float4* __restrict__ cells,
                                                   float2* __restrict__ pos0,
                                                   float2* __restrict__ pos1,
                                                   const uint32_t* __restrict__ sv,
                                                   const float2* __restrict__ spos,
                                                   const float* __restrict__ ssize,
                                                   uint32_t npc, const uint32_t* __restrict__ gid,
                                                   const GDesc* __restrict__ desc,
                                                   const uint8_t* __restrict__ flags,
                                                   float2* __restrict__ vout, uint32_t one_graph) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    const uint32_t i = sv[k];
    const GDesc& a = desc[one_graph ? 0u : gid[i]];  // a set of one graph skips the slot load
    if (!(a.passes & (1u << P_COLLISION))) return;
    const bool bsel = (a.inbuf >> P_COLLISION) & 1u;
    const float2 p = spos[k];
    const float si = ssize[k];
    const float ri = (a.col_radius > 0.f ? a.col_radius : si * 0.5f) + a.col_padding;
    const int G = (int)a.col_G;
    const size_t GG = (size_t)G * G;
    float vx = 0.f, vy = 0.f;
    for (int o = 0; o < 4; ++o) {
        const float4* __restrict__ cg_ = cells + a.col_c0 + o * GG;
        const int mx = col_cell(p.x, c_off[o][0], a.col_h, G);
        const int my = col_cell(p.y, c_off[o][1], a.col_h, G);
        float ux = 0.f, uy = 0.f, total = 0.f;
        // the 3x3 cells' aggregates, all loads in flight before any use (cells
        // outside the grid read as empty)
        float4 nb[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int cx = mx + q / 3 - 1, cy = my + q % 3 - 1;
            const bool in = cx >= 0 && cx < G && cy >= 0 && cy < G;
            nb[q] = in ? cg_[cy * G + cx] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int dx = q / 3 - 1, dy = q % 3 - 1;
            const int cx = mx + dx, cy = my + dy;
            {
                float4 cd = nb[q];
                float cnt = cd.w;
                if (cnt < 0.5f) continue;
                const float x0 = (float)cx * a.col_h, y0 = (float)cy * a.col_h;
                float rx = cd.x, ry = cd.y, ss = cd.z;
                if (dx == 0 && dy == 0) {
                    cnt -= 1.f;
                    rx -= (p.x - x0);
                    ry -= (p.y - y0);
                    ss -= si;
                }
                if (cnt < 0.5f) continue;
                total += cnt;
                // average of the other points in the cell, relative to p (the
                // hardware reciprocal and square root, ~1 ulp: the divisions
                // and lengths of the shader, well inside the force tolerance)
                const float ic = __builtin_amdgcn_rcpf(cnt);
                float ddx = (p.x - x0) - rx * ic;
                float ddy = (p.y - y0) - ry * ic;
                float as = ss * ic;
                float ro = (a.col_radius > 0.f ? a.col_radius : as * 0.5f) + a.col_padding;
                float rho = ri + ro;
                float dd = __builtin_amdgcn_sqrtf(ddx * ddx + ddy * ddy);
                if (dd < rho && dd > 0.001f) {
                    const float id = __builtin_amdgcn_rcpf(dd);
                    float ov = (rho - dd) * __builtin_amdgcn_rcpf(rho);
                    float soft = __builtin_amdgcn_sqrtf(ov) * rho * 0.5f;
                    float f = a.alpha * a.col_strength * soft * 0.25f * cnt;
                    f = fminf(f, rho * 0.5f);
                    ux += f * (ddx * id);
                    uy += f * (ddy * id);
                } else if (dd <= 0.001f && cnt > 0.f) {
                    // golden-angle kick by the point's index in its graph
                    float ang = (float)(i - a.p0) * 0.618033988749895f;
                    float f = fminf(a.alpha * a.col_strength * rho * 0.1f, rho * 0.3f);
                    float sn, cs;
                    sincosf(ang, &sn, &cs);
                    ux += f * cnt * cs;
                    uy += f * cnt * sn;
                }
            }
        }
        if (total > 2.f) {
            float damp = 2.f * __builtin_amdgcn_rcpf(total);
            ux *= damp;
            uy *= damp;
        }
        float maxc = ri * 0.1f;
        float corr = __builtin_amdgcn_sqrtf(ux * ux + uy * uy);
        if (corr > maxc) {
            float s = maxc * __builtin_amdgcn_rcpf(corr);
            ux *= s;
            uy *= s;
        }
        vx += ux;
        vy += uy;
    }
    const float2 v = make_float2(vx, vy);
    if (vout) {
        vout[i] = v;
        return;
    }
    float2* __restrict__ pout = bsel ? pos0 : pos1;
    pout[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, a.friction, a.S);
}

// The four grids' cells of every present collision point (inspection), by
// the same cell function the cell and force kernels use.
__global__ void k_col_point_cells(const float2* __restrict__ pos0, const float2* __restrict__ pos1,
                                  const uint32_t* __restrict__ cpresent, uint32_t npc,
                                  const uint32_t* __restrict__ gid, const GDesc* __restrict__ desc,
                                  uint32_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    const uint32_t i = cpresent[k];
    const GDesc& d = desc[gid[i]];
    const float2 p = ((d.inbuf >> P_COLLISION) & 1u) ? pos1[i] : pos0[i];
    const int G = (int)d.col_G;
    for (int o = 0; o < 4; ++o)
        out[4 * (size_t)k + o] = (uint32_t)(col_cell(p.y, c_off[o][1], d.col_h, G) * G +
                                            col_cell(p.x, c_off[o][0], d.col_h, G));
}

}  // namespace

// Grid geometry and resolved sizes of one graph — ForceCollision/index.ts:65-95,
// 121-125; GraphData.getResolvedPointSize: NaN -> 0 for an absent point, else
// pointDefaultSize.
int col_prepare_graph(Ctx* c, Graph& g) {
    (void)c;
    const cg_sim_params* p = &g.params;
    const uint32_t n = g.n;
    g.col_size.assign(n, 0.f);
    double maxs = p->point_default_size;
    for (uint32_t i = 0; i < n; ++i) {
        float raw = g.sizes.empty() ? NAN : g.sizes[i];
        float v = !std::isnan(raw) ? raw : ((g.flags[i] & F_ABSENT) ? 0.f : p->point_default_size);
        g.col_size[i] = v;
        maxs = std::max(maxs, (double)v);
    }
    double R = p->collision_radius, pad = p->collision_padding;
    double eff = (R > 0.0 ? R : maxs * 0.5) + pad;
    double cs = std::max(eff * 2.0, 8.0);
    double gg = std::floor((double)p->space_size / cs);
    gg = std::min(512.0, std::max(1.0, gg));
    g.col_G = (uint32_t)gg;
    g.col_h = (float)((double)p->space_size / gg);
    g.col_dirty = false;
    return CG_OK;
}

int col_run(Ctx* c, float2* vout) {
    const uint32_t npc = c->NPC, CG = c->CG, CB = c->CG / 4;
    if (npc == 0 || CG == 0) return CG_OK;
    const int fin = c->col_rs.passes & 1;
    // more base cells than points: cells from the occupied base cells, after
    // clearing those of the last pass (whose sorted keys the key kernel is
    // about to overwrite)
    const bool sparse = CB > npc;
    if (sparse)
        hipLaunchKernelGGL(k_col_clear, dim3(blocks_for(npc, 256)), dim3(256), 0, c->stream, c->col_keys[fin],
                           c->col_vals[fin], npc, c->gid, c->desc, CB, c->G == 1 ? 1u : 0u, c->col_end,
                           c->col_cell);
    hipLaunchKernelGGL(k_col_keys, dim3(blocks_for(npc, 256)), dim3(256), 0, c->stream, c->pos[0], c->pos[1],
                       c->cpresent, npc, c->gid, c->desc, c->col_keys[0], c->col_vals[0], c->col_pos[0],
                       c->G == 1 ? 1u : 0u);
    radix_run(c, c->col_rs, c->col_keys, c->col_vals, false, c->col_pos);
    if (!sparse)
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(CB, 256)), dim3(256), 0, c->stream, c->col_end, CB, 0u);
    hipLaunchKernelGGL(k_fill32, dim3(1), dim3(64), 0, c->stream, c->col_crowd, 1u, 0u);
    hipLaunchKernelGGL(k_col_bounds, dim3(blocks_for(npc, 256)), dim3(256), 0, c->stream, c->col_keys[fin],
                       c->col_vals[fin], npc, c->size, c->col_begin, c->col_end, c->col_ssize);
    if (sparse)
        hipLaunchKernelGGL(k_col_cells_sparse, dim3(blocks_for((uint64_t)npc * 9, 256)), dim3(256), 0, c->stream,
                           c->col_keys[fin], c->col_vals[fin], npc, c->col_begin, c->col_end, c->col_pos[fin],
                           c->col_ssize, c->gid, c->desc, c->col_cell, c->col_crowd, c->col_crowd_cap,
                           c->G == 1 ? 1u : 0u);
    // lanes per cell from the points per base cell: 16 where cells are crowded
    // (each lane then strides over a few points), 4 where most are sparse;
    // crowded cells then take a wave each
    else if (npc >= 8u * CB)
        hipLaunchKernelGGL(k_col_cells<16>, dim3(blocks_for((uint64_t)CG * 16, 256)), dim3(256), 0, c->stream,
                           c->col_begin, c->col_end, c->col_pos[fin], c->col_ssize, c->colc0s, c->desc, c->G, CG,
                           c->col_cell, c->col_crowd, c->col_crowd_cap);
    else
        hipLaunchKernelGGL(k_col_cells<4>, dim3(blocks_for((uint64_t)CG * 4, 256)), dim3(256), 0, c->stream,
                           c->col_begin, c->col_end, c->col_pos[fin], c->col_ssize, c->colc0s, c->desc, c->G, CG,
                           c->col_cell, c->col_crowd, c->col_crowd_cap);
    hipLaunchKernelGGL(k_col_crowd, dim3(blocks_for(std::min<uint32_t>(c->col_crowd_cap, 16384u), 4)), dim3(256), 0, c->stream,
                       c->col_crowd, c->col_crowd_cap, c->col_begin, c->col_end, c->col_pos[fin], c->col_ssize,
                       c->colc0s, c->desc, c->G, c->col_cell);
    hipLaunchKernelGGL(k_col_force, dim3(blocks_for(npc, 256)), dim3(256), 0, c->stream, c->col_cell,
                       c->pos[0], c->pos[1], c->col_vals[fin], c->col_pos[fin], c->col_ssize, npc, c->gid, c->desc,
                       c->flags, vout, c->G == 1 ? 1u : 0u);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

int col_point_cells(Ctx* c, uint32_t* out) {
    if (c->NPC == 0) return CG_OK;
    hipLaunchKernelGGL(k_col_point_cells, dim3(blocks_for(c->NPC, 256)), dim3(256), 0, c->stream, c->pos[0],
                       c->pos[1], c->cpresent, c->NPC, c->gid, c->desc, out);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Packed collision arrays for the graphs with collision on (called by the
// packer): the base-grid sort (segmented per graph: the present list of the
// collision graphs is graph-major), cell bounds and the four grids' cells.
int col_alloc(Ctx* c) {
    int rc = CG_OK;
    const uint32_t npc = c->NPC;
    if (npc == 0 || c->CG == 0) return CG_OK;
    for (int b = 0; b < 2; ++b) {
        c->col_keys[b] = (uint32_t*)dalloc(c, (size_t)npc * 4, &rc); if (rc) return rc;
        c->col_vals[b] = (uint32_t*)dalloc(c, (size_t)npc * 4, &rc); if (rc) return rc;
        c->col_pos[b] = (float2*)dalloc(c, (size_t)npc * 8, &rc); if (rc) return rc;
    }
    c->col_ssize = (float*)dalloc(c, (size_t)npc * 4, &rc); if (rc) return rc;
    c->col_begin = (uint32_t*)dalloc(c, (size_t)(c->CG / 4) * 4, &rc); if (rc) return rc;
    c->col_end = (uint32_t*)dalloc(c, (size_t)(c->CG / 4) * 4, &rc); if (rc) return rc;
    c->col_cell = (float4*)dalloc(c, (size_t)c->CG * 16, &rc); if (rc) return rc;
    // each point's base cell is a base cell of at most 1 + 2 + 2 + 4 cells of
    // the four grids, so at most 9 npc / (COL_CROWD_MIN + 1) cells can be
    // deferred
    c->col_crowd_cap = (uint32_t)(9ull * npc / (COL_CROWD_MIN + 1) + 1);
    c->col_crowd = (uint32_t*)dalloc(c, ((size_t)c->col_crowd_cap + 1) * 4, &rc); if (rc) return rc;
    std::vector<uint4> blk;
    std::vector<uint32_t> key0;
    uint64_t range = 1;
    uint32_t k0 = 0;
    for (uint32_t s = 0; s < c->G; ++s) {
        const GDesc& d = c->h_desc[s];
        if (d.col_G == 0) continue;  // collision off for this graph
        range = std::max<uint64_t>(range, (uint64_t)d.col_G * d.col_G);
        const uint32_t fb = (uint32_t)blk.size(), nbk = (d.np + RADIX_TILE - 1) / RADIX_TILE;
        for (uint32_t q = 0; q < nbk; ++q) {
            const uint32_t b0 = k0 + q * RADIX_TILE;
            blk.push_back(make_uint4(b0, std::min(k0 + d.np, b0 + RADIX_TILE), fb, fb + nbk));
            key0.push_back(d.col_c0 / 4u);
        }
        k0 += d.np;
    }
    const bool one = blk.size() && blk.front().z == 0 && blk.front().w == blk.size() && key0.front() == 0;
    rc = one ? radix_alloc(c, c->col_rs, npc, range) : radix_alloc(c, c->col_rs, npc, range, &blk, &key0);
    if (rc) return rc;
    // every cell and base-cell end starts empty; the first pass's clear reads
    // sorted keys of 0 (base cell 0, cleared again: always safe) for points
    // of the set's first collision graph
    CG_HIP(c, hipMemsetAsync(c->col_cell, 0, (size_t)c->CG * 16, c->stream));
    CG_HIP(c, hipMemsetAsync(c->col_end, 0, (size_t)(c->CG / 4) * 4, c->stream));
    CG_HIP(c, hipMemsetAsync(c->col_keys[c->col_rs.passes & 1], 0, (size_t)npc * 4, c->stream));
    CG_HIP(c, hipMemcpyAsync(c->col_vals[c->col_rs.passes & 1], c->cpresent, (size_t)npc * 4,
                             hipMemcpyDeviceToDevice, c->stream));
    return CG_OK;
}

}  // namespace cg
