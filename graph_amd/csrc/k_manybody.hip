// k_manybody.hip — grid-pyramid many-body repulsion with the Monte-Carlo
// near field, as one deterministic pipeline per tick:
//
//   1. finest-cell Morton key per present point (k_pointwise.hip)
//   2. stable radix sort of (key, point index)  -> points grouped by finest
//      cell, ascending index inside a cell; every pyramid cell at every level
//      is then a contiguous run of the sorted order
//   3. segment bounds per finest cell
//   4. per finest cell: f64 sums of (x, y) and count, and the top-8 near-field
//      slots by (ticket, index) — one lane per cell, one wave per crowded cell
//   5. f64 2x2 merges up the pyramid, f32 centroids (cx, cy, N) per level
//   6. per point (in sorted order): far field over every level + near field,
//      fused with the integrate pass
//
// No float atomics anywhere: every sum has a fixed order, so a graph gives the
// same bits alone or packed with others.
//
// Reference: ForceManyBody/index.ts:399-417 (run), :464-502 (drawLevels),
// :511-559 (drawNearFieldSlots), :565-638 (drawForces), :646-694 (geometry);
// calculate-level.vert:28-53; build-nearfield-slots.vert:46-103;
// force-level.frag:43-98; force-nearfield.frag:64-166.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "cg_internal.h"

namespace cg {
namespace {

constexpr uint32_t TMAX = 0xffffffffu;

// ---- 3. segment bounds -----------------------------------------------------

__global__ void k_fill_u32(uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_mark_bounds(const uint32_t* __restrict__ sk, uint32_t np,
                              uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t key = sk[k];
    if (k == 0 || sk[k - 1] != key) begin[key] = k;
    if (k == np - 1 || sk[k + 1] != key) end[key] = k + 1;
}

// ---- 4. finest cells: sums + slots ------------------------------------------

// Sorted insert of (t, i) into a top-8 list of distinct tickets.  Points arrive
// in ascending index, so an equal ticket already in the list has the smaller
// index and wins (the depth test's first-drawn rule, build-nearfield-slots.vert:85-101).
__device__ inline void top8_insert(uint32_t (&tb)[SLOTS], uint32_t (&ti)[SLOTS], uint32_t t,
                                   uint32_t i) {
    if (t >= tb[SLOTS - 1]) return;
    bool dup = false;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) dup |= (tb[s] == t);
    if (dup) return;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        bool sw = t < tb[s];
        uint32_t a = tb[s], b = ti[s];
        tb[s] = sw ? t : a;
        ti[s] = sw ? i : b;
        t = sw ? a : t;
        i = sw ? b : i;
    }
}

__device__ inline void write_cell(const GDesc& d, uint32_t rm, double sx, double sy, uint32_t cnt,
                                  const uint32_t (&ti)[SLOTS], const float2* __restrict__ pin,
                                  double* __restrict__ sums, int32_t* __restrict__ slot_idx,
                                  float2* __restrict__ slot_pos) {
    const size_t c = (size_t)d.c0 + level_off(d.L - 1) + rm;
    sums[3 * c + 0] = sx;
    sums[3 * c + 1] = sy;
    sums[3 * c + 2] = (double)cnt;
    // slots of an empty cell are never read (its count is 0)
    if (cnt == 0) return;
    const size_t sb = ((size_t)d.f0 + rm) * SLOTS;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        bool ok = ti[s] != TMAX;
        slot_idx[sb + s] = ok ? (int32_t)ti[s] : -1;
        slot_pos[sb + s] = ok ? pin[ti[s]] : make_float2(0.f, 0.f);
    }
}

// One lane per finest cell of every graph.  Sums in index order, top-8 slots by
// (ticket of the graph-local index, index).
__global__ void __launch_bounds__(256) k_cells(const uint32_t* __restrict__ begin,
                                               const uint32_t* __restrict__ end,
                                               const uint32_t* __restrict__ sv,
                                               const float2* __restrict__ pos0,
                                               const float2* __restrict__ pos1,
                                               float2* __restrict__ sorted_pos,
                                               double* __restrict__ sums,
                                               int32_t* __restrict__ slot_idx,
                                               float2* __restrict__ slot_pos,
                                               uint32_t* __restrict__ big_list,
                                               uint32_t* __restrict__ big_count,
                                               const uint32_t* __restrict__ f0s,
                                               const GDesc* __restrict__ desc, uint32_t G,
                                               uint32_t F) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const GDesc& d = desc[find_graph(f0s, G, f)];
    if (!(d.passes & (1u << P_MANYBODY))) return;
    const float2* __restrict__ pin = ((d.inbuf >> P_MANYBODY) & 1u) ? pos1 : pos0;
    const uint32_t b = begin[f], e = end[f];
    if (e - b > BIG_CELL) {
        big_list[atomicAdd(big_count, 1u)] = f;
        return;
    }
    uint32_t tb[SLOTS], ti[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) { tb[s] = TMAX; ti[s] = TMAX; }
    double sx = 0.0, sy = 0.0;
    for (uint32_t k = b; k < e; ++k) {
        uint32_t i = sv[k];
        float2 p = pin[i];
        sorted_pos[k] = p;
        sx += (double)p.x;
        sy += (double)p.y;
        top8_insert(tb, ti, ticket_bits(i - d.p0, d.seed), i);
    }
    const uint32_t mc = f - d.f0;
    const uint32_t rm = compact1by1(mc >> 1) * d.gf + compact1by1(mc);
    write_cell(d, rm, sx, sy, e - b, ti, pin, sums, slot_idx, slot_pos);
}

// One wave per crowded cell: lanes stride the segment keeping private top-8
// lists and f64 partial sums, then a fixed shuffle tree reduces the sums and
// eight wave-wide argmin rounds merge the lists.
__global__ void __launch_bounds__(256) k_cells_big(const uint32_t* __restrict__ begin,
                                                   const uint32_t* __restrict__ end,
                                                   const uint32_t* __restrict__ sv,
                                                   const float2* __restrict__ pos0,
                                                   const float2* __restrict__ pos1,
                                                   float2* __restrict__ sorted_pos,
                                                   double* __restrict__ sums,
                                                   int32_t* __restrict__ slot_idx,
                                                   float2* __restrict__ slot_pos,
                                                   const uint32_t* __restrict__ big_list,
                                                   const uint32_t* __restrict__ big_count,
                                                   const uint32_t* __restrict__ f0s,
                                                   const GDesc* __restrict__ desc, uint32_t G) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t nbig = *big_count;
    for (uint32_t w = wave; w < nbig; w += nwaves) {
        const uint32_t f = big_list[w];
        const GDesc& d = desc[find_graph(f0s, G, f)];
        const float2* __restrict__ pin = ((d.inbuf >> P_MANYBODY) & 1u) ? pos1 : pos0;
        const uint32_t b = begin[f], e = end[f];
        uint32_t tb[SLOTS], ti[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) { tb[s] = TMAX; ti[s] = TMAX; }
        double sx = 0.0, sy = 0.0;
        for (uint32_t k = b + lane; k < e; k += 64) {
            uint32_t i = sv[k];
            float2 p = pin[i];
            sorted_pos[k] = p;
            sx += (double)p.x;
            sy += (double)p.y;
            top8_insert(tb, ti, ticket_bits(i - d.p0, d.seed), i);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
        }
        uint32_t out[SLOTS];
#pragma unroll
        for (int r = 0; r < SLOTS; ++r) {
            unsigned long long key = ((unsigned long long)tb[0] << 32) | ti[0];
            unsigned long long mn = key;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                unsigned long long other = __shfl_xor(mn, o, 64);
                mn = other < mn ? other : mn;
            }
            uint32_t wt = (uint32_t)(mn >> 32), wi = (uint32_t)mn;
            out[r] = (wt == TMAX) ? TMAX : wi;
            if (wt != TMAX && tb[0] == wt) {
#pragma unroll
                for (int s = 0; s < SLOTS - 1; ++s) { tb[s] = tb[s + 1]; ti[s] = ti[s + 1]; }
                tb[SLOTS - 1] = TMAX;
                ti[SLOTS - 1] = TMAX;
            }
        }
        if (lane == 0) {
            const uint32_t mc = f - d.f0;
            const uint32_t rm = compact1by1(mc >> 1) * d.gf + compact1by1(mc);
            write_cell(d, rm, sx, sy, e - b, out, pin, sums, slot_idx, slot_pos);
        }
    }
}

// ---- 5. pyramid ---------------------------------------------------------------

__device__ inline float4 centroid(double sx, double sy, double n) {
    if (n <= 0.0) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4((float)(sx / n), (float)(sy / n), (float)n, 0.f);
}

// One block per tile of one graph's level `lin` (work item {graph slot, lin,
// T | nout << 16, tile}); produces the `nout` coarser levels of that tile down
// to one cell per tile.  Children are summed in a fixed order.
__global__ void __launch_bounds__(256) k_pyramid(const uint4* __restrict__ work,
                                                 const GDesc* __restrict__ desc,
                                                 double* __restrict__ sums,
                                                 float4* __restrict__ cent, int write_in) {
    __shared__ double s0[32 * 32 * 3];
    __shared__ double s1[16 * 16 * 3];
    const uint4 wk = work[blockIdx.x];
    const GDesc& d = desc[wk.x];
    if (!(d.passes & (1u << P_MANYBODY))) return;
    const int lin = (int)wk.y, T = (int)(wk.z & 0xffffu), nout = (int)(wk.z >> 16);
    const int gin = 4 << lin;
    const int tiles_x = gin / T;
    const int tx = (int)wk.w % tiles_x, ty = (int)wk.w / tiles_x;
    const size_t in0 = (size_t)d.c0 + level_off(lin);
    for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
        int x = c % T, y = c / T;
        size_t rm = in0 + (size_t)(ty * T + y) * gin + (tx * T + x);
        double sx = sums[3 * rm], sy = sums[3 * rm + 1], n = sums[3 * rm + 2];
        s0[3 * c] = sx;
        s0[3 * c + 1] = sy;
        s0[3 * c + 2] = n;
        if (write_in) cent[rm] = centroid(sx, sy, n);
    }
    __syncthreads();
    double* cur = s0;
    double* nxt = s1;
    int size = T;
    for (int lev = 1; lev <= nout; ++lev) {
        int half = size / 2;
        int lout = lin - lev;
        int gout = gin >> lev;
        const size_t out0 = (size_t)d.c0 + level_off(lout);
        for (int c = threadIdx.x; c < half * half; c += blockDim.x) {
            int x = c % half, y = c / half;
            int c00 = (2 * y) * size + 2 * x, c10 = c00 + 1, c01 = c00 + size, c11 = c01 + 1;
            double v[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                v[q] = (cur[3 * c00 + q] + cur[3 * c10 + q]) + (cur[3 * c01 + q] + cur[3 * c11 + q]);
            nxt[3 * c] = v[0];
            nxt[3 * c + 1] = v[1];
            nxt[3 * c + 2] = v[2];
            size_t rm = out0 + (size_t)(ty * half + y) * gout + (tx * half + x);
            sums[3 * rm] = v[0];
            sums[3 * rm + 1] = v[1];
            sums[3 * rm + 2] = v[2];
            cent[rm] = centroid(v[0], v[1], v[2]);
        }
        __syncthreads();
        double* t = cur;
        cur = nxt;
        nxt = t;
        size = half;
    }
}

// ---- 6. forces ------------------------------------------------------------------

// Clamped inverse-distance falloff of force-level.frag:43-56 /
// force-nearfield.frag:64-80 without the alpha*k*N factor:
// returns c with contribution = c * d.
__device__ inline float falloff_coef(float l) {
    float inv = __builtin_amdgcn_rsqf(l);            // 1/|d|
    return (l < 1.f) ? inv * __builtin_sqrtf(inv) : inv * inv;
}

__global__ void __launch_bounds__(256) k_mb_force(const uint32_t* __restrict__ sv,
                                                  const float2* __restrict__ sorted_pos,
                                                  const float4* __restrict__ cent,
                                                  const int32_t* __restrict__ slot_idx,
                                                  const float2* __restrict__ slot_pos,
                                                  const float2* __restrict__ jitter,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ gid,
                                                  const GDesc* __restrict__ desc, uint32_t np,
                                                  float2* __restrict__ pos0,
                                                  float2* __restrict__ pos1,
                                                  float2* __restrict__ vout) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t i = sv[k];
    const GDesc& d = desc[gid[i]];
    if (!(d.passes & (1u << P_MANYBODY))) return;
    const float2 p = sorted_pos[k];
    const int L = (int)d.L;
    float fx = 0.f, fy = 0.f;

    // far field, level 0: every cell outside the 3x3 neighbourhood (force-level.frag:70-79)
    {
        const float4* __restrict__ lv = cent + d.c0;
        int qx = cell_of(p.x, d.h[0], 4), qy = cell_of(p.y, d.h[0], 4);
#pragma unroll
        for (int cy = 0; cy < 4; ++cy) {
#pragma unroll
            for (int cx = 0; cx < 4; ++cx) {
                if (abs(cx - qx) <= 1 && abs(cy - qy) <= 1) continue;
                float4 c = lv[cy * 4 + cx];
                if (c.z <= 0.f) continue;
                float dx = p.x - c.x, dy = p.y - c.y;
                float l = dx * dx + dy * dy;
                if (l <= 0.f) continue;
                float w = c.z * falloff_coef(l);
                fx += w * dx;
                fy += w * dy;
            }
        }
    }
    // far field, levels 1..L-1: the aligned 6x6 child block minus the 3x3 (force-level.frag:80-95)
    int qx = 0, qy = 0;
    for (int lvl = 1; lvl < L; ++lvl) {
        const int g = 4 << lvl;
        const float4* __restrict__ lv = cent + d.c0 + level_off(lvl);
        qx = cell_of(p.x, d.h[lvl], g);
        qy = cell_of(p.y, d.h[lvl], g);
        const int bx = (qx >> 1) * 2 - 2, by = (qy >> 1) * 2 - 2;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int cy = by + j;
            if (cy < 0 || cy >= g) continue;
#pragma unroll
            for (int ii = 0; ii < 6; ++ii) {
                const int cx = bx + ii;
                if (cx < 0 || cx >= g) continue;
                if (abs(cx - qx) <= 1 && abs(cy - qy) <= 1) continue;
                float4 c = lv[cy * g + cx];
                if (c.z <= 0.f) continue;
                float dx = p.x - c.x, dy = p.y - c.y;
                float l = dx * dx + dy * dy;
                if (l <= 0.f) continue;
                float w = c.z * falloff_coef(l);
                fx += w * dx;
                fy += w * dy;
            }
        }
    }
    // near field at the finest level (force-nearfield.frag:101-163)
    const int gf = (int)d.gf;
    const float4* __restrict__ fin = cent + d.c0 + level_off(L - 1);
    const float2 r = jitter[i];
    float wx = 0.f, wy = 0.f;
    for (int dj = -1; dj <= 1; ++dj) {
        const int cy = qy + dj;
        if (cy < 0 || cy >= gf) continue;
        for (int di = -1; di <= 1; ++di) {
            const int cx = qx + di;
            if (cx < 0 || cx >= gf) continue;
            const int cell = cy * gf + cx;
            float others = fin[cell].z - ((di == 0 && dj == 0) ? 1.f : 0.f);
            if (others <= 0.f) continue;
            const size_t sb = ((size_t)d.f0 + cell) * SLOTS;
            float sxp = 0.f, syp = 0.f, sampled = 0.f;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) {
                const int j = slot_idx[sb + s];
                if (j < 0) break;
                if ((uint32_t)j == i) continue;
                sampled += 1.f;
                const float2 q = slot_pos[sb + s];
                float dx = p.x - q.x, dy = p.y - q.y;
                float l = dx * dx + dy * dy;
                if (l <= 0.f) {
                    dx = r.x;
                    dy = r.y;
                    l = dx * dx + dy * dy;
                    if (l <= 0.f) continue;
                }
                float w = falloff_coef(l);
                sxp += w * dx;
                syp += w * dy;
            }
            if (sampled > 0.f) {
                float hw = others / sampled;
                wx += hw * sxp;
                wy += hw * syp;
            }
        }
    }
    wx *= d.ak;
    wy *= d.ak;
    wx += wx * r.x;
    wy += wy * r.y;
    {
        const float maxstep = 2.0f * d.h[L - 1];  // force-nearfield.frag:161
        float sp = sqrtf(wx * wx + wy * wy);
        if (sp > maxstep) {
            float s = maxstep / sp;
            wx *= s;
            wy *= s;
        }
    }
    const float2 v = make_float2(d.ak * fx + wx, d.ak * fy + wy);
    if (vout) {
        vout[i] = v;
        return;
    }
    float2* __restrict__ pout = ((d.inbuf >> P_MANYBODY) & 1u) ? pos0 : pos1;
    pout[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, d.friction, d.S);
}

}  // namespace

int mb_alloc(Ctx* c) {
    int rc = CG_OK;
    const size_t np = c->NP ? c->NP : 1;
    const size_t F = c->F ? c->F : 1, C = c->C ? c->C : 1;
    for (int b = 0; b < 2; ++b) {
        c->keys[b] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
        c->vals[b] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
    }
    c->sorted_pos = (float2*)dalloc(c, np * 8, &rc); if (rc) return rc;
    c->cell_begin = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    c->cell_end = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    c->slot_idx = (int32_t*)dalloc(c, F * SLOTS * 4, &rc); if (rc) return rc;
    c->slot_pos = (float2*)dalloc(c, F * SLOTS * 8, &rc); if (rc) return rc;
    c->big_list = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    c->big_count = (uint32_t*)dalloc(c, 64, &rc); if (rc) return rc;
    c->sums = (double*)dalloc(c, C * 3 * sizeof(double), &rc); if (rc) return rc;
    c->cent = (float4*)dalloc(c, C * sizeof(float4), &rc); if (rc) return rc;
    c->key_bits = 1;
    while (c->key_bits < 32 && (1ull << c->key_bits) < (unsigned long long)F) ++c->key_bits;
    // sized for full 32-bit keys: the collision pass sorts its cell keys with it too
    size_t tmp = 0;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, c->keys[0], c->keys[1], c->vals[0],
                                                 c->vals[1], (int)np, 0, 32, c->stream));
    c->sort_tmp_bytes = tmp;
    c->sort_tmp = dalloc(c, tmp ? tmp : 16, &rc);
    return rc;
}

int mb_run(Ctx* c, bool keys_ready, float2* vout) {
    if (c->NP == 0) return CG_OK;
    int rc = CG_OK;
    prof_begin(c, CG_STAGE_MB_SORT);
    if (!keys_ready) {
        rc = keys_run(c);
        if (rc) return rc;
    }
    const uint32_t np = c->NP, F = c->F;
    size_t tmp = c->sort_tmp_bytes;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, tmp, c->keys[0], c->keys[1],
                                                 c->vals[0], c->vals[1], (int)np, 0, c->key_bits,
                                                 c->stream));
    prof_end(c, CG_STAGE_MB_SORT);
    prof_begin(c, CG_STAGE_MB_CELLS);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream,
                       c->cell_begin, F, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream,
                       c->cell_end, F, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->big_count, 1u, 0u);
    hipLaunchKernelGGL(k_mark_bounds, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream,
                       c->keys[1], np, c->cell_begin, c->cell_end);
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream,
                       c->cell_begin, c->cell_end, c->vals[1], c->pos[0], c->pos[1], c->sorted_pos,
                       c->sums, c->slot_idx, c->slot_pos, c->big_list, c->big_count, c->f0s, c->desc,
                       c->G, F);
    hipLaunchKernelGGL(k_cells_big, dim3(256), dim3(256), 0, c->stream, c->cell_begin, c->cell_end,
                       c->vals[1], c->pos[0], c->pos[1], c->sorted_pos, c->sums, c->slot_idx,
                       c->slot_pos, c->big_list, c->big_count, c->f0s, c->desc, c->G);
    prof_end(c, CG_STAGE_MB_CELLS);
    prof_begin(c, CG_STAGE_MB_PYRAMID);
    // pyramid: finest -> tiles of up to 32x32 -> remaining top levels
    if (c->n_pyr[0])
        hipLaunchKernelGGL(k_pyramid, dim3(c->n_pyr[0]), dim3(256), 0, c->stream, c->pyr[0], c->desc,
                           c->sums, c->cent, 1);
    if (c->n_pyr[1])
        hipLaunchKernelGGL(k_pyramid, dim3(c->n_pyr[1]), dim3(256), 0, c->stream, c->pyr[1], c->desc,
                           c->sums, c->cent, 0);
    prof_end(c, CG_STAGE_MB_PYRAMID);
    prof_begin(c, CG_STAGE_MB_FORCE);
    hipLaunchKernelGGL(k_mb_force, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, c->vals[1],
                       c->sorted_pos, c->cent, c->slot_idx, c->slot_pos, c->jitter, c->flags, c->gid,
                       c->desc, np, c->pos[0], c->pos[1], vout);
    prof_end(c, CG_STAGE_MB_FORCE);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Pyramid work lists for one graph slot (stage 0: finest level in tiles of up
// to 32x32; stage 1: the remaining top levels of graphs finer than 128^2).
void pyramid_work_host(uint32_t slot, uint32_t gf, uint32_t L, std::vector<uint4>& w0,
                       std::vector<uint4>& w1) {
    const uint32_t lf = L - 1;
    const uint32_t T1 = std::min<uint32_t>(32u, gf / 4);
    uint32_t nout1 = 0;
    while ((1u << nout1) < T1) ++nout1;
    const uint32_t tiles = (gf / T1) * (gf / T1);
    for (uint32_t t = 0; t < tiles; ++t) w0.push_back(make_uint4(slot, lf, T1 | (nout1 << 16), t));
    const uint32_t g2 = gf / T1;
    if (g2 > 4) {
        const uint32_t lin2 = lf - nout1, T2 = g2 / 4;
        uint32_t nout2 = 0;
        while ((1u << nout2) < T2) ++nout2;
        for (uint32_t t = 0; t < 16; ++t) w1.push_back(make_uint4(slot, lin2, T2 | (nout2 << 16), t));
    }
}

}  // namespace cg
