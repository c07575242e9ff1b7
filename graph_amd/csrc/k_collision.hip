// k_collision.hip — spatial-hash collision over four offset grids.
//
// Per offset o in {(0,0), (.5,0), (0,.5), (.5,.5)}: integer row-major cell key
// per present point, stable radix sort of (key, index), segment bounds, and a
// fixed-order per-cell reduction of (x, y, size, count).  Sums are kept
// relative to the cell's origin so the force pass can remove the point's own
// contribution without f32 cancellation.  The force pass sweeps the 3x3 cell
// aggregates of all four grids per point and integrates.
//
// Reference: ForceCollision/index.ts:61-164 (grid geometry, sizes),
// :264-345 (run); build-grid.vert:28-62; force-collision-spatial.frag:103-230.
#include <hipcub/hipcub.hpp>

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

__global__ void k_col_keys(const float2* __restrict__ pin, const uint32_t* __restrict__ present,
                           uint32_t np, int o, float h, int G, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ vals) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t i = present[k];
    float2 p = pin[i];
    int cx = col_cell(p.x, c_off[o][0], h, G), cy = col_cell(p.y, c_off[o][1], h, G);
    keys[k] = (uint32_t)(cy * G + cx);
    vals[k] = i;
}

__global__ void k_col_bounds(const uint32_t* __restrict__ sk, uint32_t np,
                             uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t key = sk[k];
    if (k == 0 || sk[k - 1] != key) begin[key] = k;
    if (k == np - 1 || sk[k + 1] != key) end[key] = k + 1;
}

// One lane per cell: (sum(x - x0), sum(y - y0), sum(size), count) in f64,
// stored as f32, with (x0, y0) = cell index * h.
__global__ void k_col_cells(const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end,
                            const uint32_t* __restrict__ sv, const float2* __restrict__ pin,
                            const float* __restrict__ size, float h, int G,
                            float4* __restrict__ cells) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (uint32_t)(G * G)) return;
    uint32_t b = begin[c], e = end[c];
    float x0 = (float)(c % G) * h, y0 = (float)(c / G) * h;
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

struct ColArgs {
    float alpha, strength, radius, padding;
    float h;
    int G;
    float friction, S;
};

__global__ void __launch_bounds__(256) k_col_force(ColArgs a, const float4* __restrict__ cells,
                                                   const float2* __restrict__ pin,
                                                   const uint32_t* __restrict__ present, uint32_t np,
                                                   const float* __restrict__ size,
                                                   const uint8_t* __restrict__ flags,
                                                   float2* __restrict__ pout,
                                                   float2* __restrict__ vout) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t i = present[k];
    const float2 p = pin[i];
    const float si = size[i];
    const float ri = (a.radius > 0.f ? a.radius : si * 0.5f) + a.padding;
    const int G = a.G;
    const size_t GG = (size_t)G * G;
    float vx = 0.f, vy = 0.f;
    for (int o = 0; o < 4; ++o) {
        const float4* __restrict__ cg_ = cells + o * GG;
        const int mx = col_cell(p.x, c_off[o][0], a.h, G);
        const int my = col_cell(p.y, c_off[o][1], a.h, G);
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
                const float x0 = (float)cx * a.h, y0 = (float)cy * a.h;
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
                float ro = (a.radius > 0.f ? a.radius : as * 0.5f) + a.padding;
                float rho = ri + ro;
                float d = sqrtf(ddx * ddx + ddy * ddy);
                if (d < rho && d > 0.001f) {
                    float ov = (rho - d) / rho;
                    float soft = sqrtf(ov) * rho * 0.5f;
                    float f = a.alpha * a.strength * soft * 0.25f * cnt;
                    f = fminf(f, rho * 0.5f);
                    ux += f * (ddx / d);
                    uy += f * (ddy / d);
                } else if (d <= 0.001f && cnt > 0.f) {
                    float ang = (float)i * 0.618033988749895f;
                    float f = fminf(a.alpha * a.strength * rho * 0.1f, rho * 0.3f);
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
    if (vout) vout[i] = v;
    if (pout) pout[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, a.friction, a.S);
}

}  // namespace

// Grid geometry and resolved sizes — ForceCollision/index.ts:65-95, 121-125;
// GraphData.getResolvedPointSize: NaN -> 0 for an absent point, else pointDefaultSize.
int col_prepare(Ctx* c, const cg_sim_params* p) {
    const uint32_t n = c->n;
    int rc = CG_OK;
    std::vector<float> rs(n ? n : 1, 0.f);
    double maxs = p->point_default_size;
    for (uint32_t i = 0; i < n; ++i) {
        float raw = c->h_sizes.empty() ? NAN : c->h_sizes[i];
        float v = !std::isnan(raw) ? raw : ((c->h_flags[i] & F_ABSENT) ? 0.f : p->point_default_size);
        rs[i] = v;
        maxs = std::max(maxs, (double)v);
    }
    double R = p->collision_radius, pad = p->collision_padding;
    double eff = (R > 0.0 ? R : maxs * 0.5) + pad;
    double cs = std::max(eff * 2.0, 8.0);
    double gg = std::floor((double)p->space_size / cs);
    gg = std::min(512.0, std::max(1.0, gg));
    uint32_t G = (uint32_t)gg;
    if (G != c->col_G || !c->col_cell) {
        dfree(c, c->col_begin); dfree(c, c->col_end); dfree(c, c->col_cell);
        c->col_begin = (uint32_t*)dalloc(c, (size_t)G * G * 4, &rc); if (rc) return rc;
        c->col_end = (uint32_t*)dalloc(c, (size_t)G * G * 4, &rc); if (rc) return rc;
        c->col_cell = (float4*)dalloc(c, (size_t)4 * G * G * 16, &rc); if (rc) return rc;
        c->col_G = G;
    }
    if (!c->col_size) {
        c->col_size = (float*)dalloc(c, (size_t)(n ? n : 1) * 4, &rc);
        if (rc) return rc;
    }
    if (n) {
        CG_HIP(c, hipMemcpyAsync(c->col_size, rs.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    c->col_max_size = (float)maxs;
    c->col_last_default = p->point_default_size;
    return CG_OK;
}

int col_run(Ctx* c, const cg_sim_params* p, float alpha, const float2* pin, float2* pout,
            float2* vout) {
    if (c->np == 0) return CG_OK;
    int rc = CG_OK;
    // re-derive grid and sizes when the inputs that shape them change
    {
        double R = p->collision_radius, pad = p->collision_padding;
        double eff = (R > 0.0 ? R : (double)std::max(c->col_max_size, p->point_default_size) * 0.5) + pad;
        double cs = std::max(eff * 2.0, 8.0);
        uint32_t G = (uint32_t)std::min(512.0, std::max(1.0, std::floor((double)p->space_size / cs)));
        if (!c->col_cell || c->col_last_default != p->point_default_size || G != c->col_G) {
            rc = col_prepare(c, p);
            if (rc) return rc;
        }
    }
    const uint32_t G = c->col_G, cells = G * G, np = c->np;
    const float h = (float)((double)p->space_size / (double)G);
    int bits = 1;
    while ((1u << bits) < cells) ++bits;
    for (int o = 0; o < 4; ++o) {
        hipLaunchKernelGGL(k_col_keys, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, pin,
                           c->present, np, o, h, (int)G, c->keys[0], c->vals[0]);
        size_t tmp = c->sort_tmp_bytes;
        CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, tmp, c->keys[0], c->keys[1],
                                                     c->vals[0], c->vals[1], (int)np, 0, bits,
                                                     c->stream));
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                           c->col_begin, cells, 0u);
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                           c->col_end, cells, 0u);
        hipLaunchKernelGGL(k_col_bounds, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream,
                           c->keys[1], np, c->col_begin, c->col_end);
        hipLaunchKernelGGL(k_col_cells, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                           c->col_begin, c->col_end, c->vals[1], pin, c->col_size, h, (int)G,
                           c->col_cell + (size_t)o * cells);
    }
    ColArgs a;
    a.alpha = alpha;
    a.strength = p->collision;
    a.radius = p->collision_radius;
    a.padding = p->collision_padding;
    a.h = h;
    a.G = (int)G;
    a.friction = p->friction;
    a.S = p->space_size;
    hipLaunchKernelGGL(k_col_force, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, a,
                       c->col_cell, pin, c->present, np, c->col_size, c->flags, pout, vout);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

}  // namespace cg
