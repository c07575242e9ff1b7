// k_collision.hip — spatial-hash collision over four offset grids.
//
// Per offset o in {(0,0), (.5,0), (0,.5), (.5,.5)}: integer row-major cell key
// per present point, one radix sort of (key, index) over all four grids and all
// graphs, segment bounds, and a fixed-order per-cell reduction of
// (x, y, size, count).  Sums are kept
// relative to the cell's origin so the force pass can remove the point's own
// contribution without f32 cancellation.  The force pass sweeps the 3x3 cell
// aggregates of all four grids per point and integrates.
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

// One key per (point, offset grid): the graph's collision block, then the grid,
// then the row-major cell.  Work item t = o * npc + k.
__global__ void k_col_keys(const float2* __restrict__ pos0, const float2* __restrict__ pos1,
                           const uint32_t* __restrict__ cpresent, uint32_t npc,
                           const uint32_t* __restrict__ gid, const GDesc* __restrict__ desc,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * npc) return;
    const uint32_t o = t / npc, k = t - o * npc;
    const uint32_t i = cpresent[k];
    const GDesc& d = desc[gid[i]];
    const float2 p = ((d.inbuf >> P_COLLISION) & 1u) ? pos1[i] : pos0[i];
    const int G = (int)d.col_G;
    int cx = col_cell(p.x, c_off[o][0], d.col_h, G), cy = col_cell(p.y, c_off[o][1], d.col_h, G);
    keys[t] = d.col_c0 + o * (uint32_t)(G * G) + (uint32_t)(cy * G + cx);
    vals[t] = i;
}

__global__ void k_col_bounds(const uint32_t* __restrict__ sk, uint32_t nk,
                             uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    uint32_t key = sk[k];
    if (k == 0 || sk[k - 1] != key) begin[key] = k;
    if (k == nk - 1 || sk[k + 1] != key) end[key] = k + 1;
}

// One lane per cell of every graph's four grids: (sum(x - x0), sum(y - y0),
// sum(size), count) in f64, stored as f32, with (x0, y0) = cell index * h.
__global__ void k_col_cells(const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end,
                            const uint32_t* __restrict__ sv, const float2* __restrict__ pos0,
                            const float2* __restrict__ pos1, const float* __restrict__ size,
                            const uint32_t* __restrict__ colc0s, const GDesc* __restrict__ desc,
                            uint32_t G, uint32_t CG, float4* __restrict__ cells) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= CG) return;
    const GDesc& d = desc[find_graph(colc0s, G, c)];
    const float2* __restrict__ pin = ((d.inbuf >> P_COLLISION) & 1u) ? pos1 : pos0;
    const uint32_t GG = d.col_G * d.col_G;
    const uint32_t local = (c - d.col_c0) % GG;
    const uint32_t e = end[c], b = e ? begin[c] : 0u;  // end 0: empty (begin not written)
    float x0 = (float)(local % d.col_G) * d.col_h, y0 = (float)(local / d.col_G) * d.col_h;
    double sx = 0.0, sy = 0.0, ss = 0.0;
    for (uint32_t k = b; k < e; ++k) {
        uint32_t i = sv[k];
        float2 p = pin[i];
        sx += (double)p.x - (double)x0;
        sy += (double)p.y - (double)y0;
        ss += (double)size[i];
    }
    cells[c] = make_float4((float)sx, (float)sy, (float)ss, (float)(e - b));
}

__global__ void __launch_bounds__(256) k_col_force(const float4* __restrict__ cells,
                                                   float2* __restrict__ pos0,
                                                   float2* __restrict__ pos1,
                                                   const uint32_t* __restrict__ cpresent,
                                                   uint32_t npc, const uint32_t* __restrict__ gid,
                                                   const GDesc* __restrict__ desc,
                                                   const float* __restrict__ size,
                                                   const uint8_t* __restrict__ flags,
                                                   float2* __restrict__ vout) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npc) return;
    const uint32_t i = cpresent[k];
    const uint32_t gslot = gid[i];
    const GDesc& a = desc[gslot];
    if (!(a.passes & (1u << P_COLLISION))) return;
    const bool bsel = (a.inbuf >> P_COLLISION) & 1u;
    const float2 p = (bsel ? pos1 : pos0)[i];
    const float si = size[i];
    const float ri = (a.col_radius > 0.f ? a.col_radius : si * 0.5f) + a.col_padding;
    const int G = (int)a.col_G;
    const size_t GG = (size_t)G * G;
    float vx = 0.f, vy = 0.f;
    for (int o = 0; o < 4; ++o) {
        const float4* __restrict__ cg_ = cells + a.col_c0 + o * GG;
        const int mx = col_cell(p.x, c_off[o][0], a.col_h, G);
        const int my = col_cell(p.y, c_off[o][1], a.col_h, G);
        float ux = 0.f, uy = 0.f, total = 0.f;
        for (int dx = -1; dx <= 1; ++dx) {
            const int cx = mx + dx;
            if (cx < 0 || cx >= G) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int cy = my + dy;
                if (cy < 0 || cy >= G) continue;
                float4 cd = cg_[cy * G + cx];
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
                // average of the other points in the cell, relative to p
                float ddx = (p.x - x0) - rx / cnt;
                float ddy = (p.y - y0) - ry / cnt;
                float as = ss / cnt;
                float ro = (a.col_radius > 0.f ? a.col_radius : as * 0.5f) + a.col_padding;
                float rho = ri + ro;
                float dd = sqrtf(ddx * ddx + ddy * ddy);
                if (dd < rho && dd > 0.001f) {
                    float ov = (rho - dd) / rho;
                    float soft = sqrtf(ov) * rho * 0.5f;
                    float f = a.alpha * a.col_strength * soft * 0.25f * cnt;
                    f = fminf(f, rho * 0.5f);
                    ux += f * (ddx / dd);
                    uy += f * (ddy / dd);
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
            float damp = 2.f / total;
            ux *= damp;
            uy *= damp;
        }
        float maxc = ri * 0.1f;
        float corr = sqrtf(ux * ux + uy * uy);
        if (corr > maxc) {
            float s = maxc / corr;
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
    const uint32_t npc = c->NPC, CG = c->CG;
    if (npc == 0 || CG == 0) return CG_OK;
    const uint32_t nk = 4 * npc;
    hipLaunchKernelGGL(k_col_keys, dim3(blocks_for(nk, 256)), dim3(256), 0, c->stream, c->pos[0],
                       c->pos[1], c->cpresent, npc, c->gid, c->desc, c->col_keys[0], c->col_vals[0]);
    radix_run(c, c->col_rs, c->col_keys, c->col_vals, false);
    const int fin = c->col_rs.passes & 1;
    hipLaunchKernelGGL(k_fill32, dim3(blocks_for(CG, 256)), dim3(256), 0, c->stream, c->col_end, CG, 0u);
    hipLaunchKernelGGL(k_col_bounds, dim3(blocks_for(nk, 256)), dim3(256), 0, c->stream, c->col_keys[fin],
                       nk, c->col_begin, c->col_end);
    hipLaunchKernelGGL(k_col_cells, dim3(blocks_for(CG, 256)), dim3(256), 0, c->stream, c->col_begin,
                       c->col_end, c->col_vals[fin], c->pos[0], c->pos[1], c->size, c->colc0s, c->desc,
                       c->G, CG, c->col_cell);
    hipLaunchKernelGGL(k_col_force, dim3(blocks_for(npc, 256)), dim3(256), 0, c->stream, c->col_cell,
                       c->pos[0], c->pos[1], c->cpresent, npc, c->gid, c->desc, c->size, c->flags, vout);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Packed collision arrays for the graphs with collision on (called by the packer).
int col_alloc(Ctx* c) {
    int rc = CG_OK;
    const uint32_t nk = 4 * c->NPC;
    if (nk == 0 || c->CG == 0) return CG_OK;
    for (int b = 0; b < 2; ++b) {
        c->col_keys[b] = (uint32_t*)dalloc(c, (size_t)nk * 4, &rc); if (rc) return rc;
        c->col_vals[b] = (uint32_t*)dalloc(c, (size_t)nk * 4, &rc); if (rc) return rc;
    }
    c->col_begin = (uint32_t*)dalloc(c, (size_t)c->CG * 4, &rc); if (rc) return rc;
    c->col_end = (uint32_t*)dalloc(c, (size_t)c->CG * 4, &rc); if (rc) return rc;
    c->col_cell = (float4*)dalloc(c, (size_t)c->CG * 16, &rc); if (rc) return rc;
    return radix_alloc(c, c->col_rs, nk, c->CG);
}

}  // namespace cg
