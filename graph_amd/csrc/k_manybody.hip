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

__device__ inline void write_cell(uint32_t rm, double sx, double sy, uint32_t cnt,
                                  const uint32_t (&ti)[SLOTS], const float2* __restrict__ pin,
                                  double* __restrict__ sums, int32_t* __restrict__ slot_idx,
                                  float2* __restrict__ slot_pos) {
    sums[3 * (size_t)rm + 0] = sx;
    sums[3 * (size_t)rm + 1] = sy;
    sums[3 * (size_t)rm + 2] = (double)cnt;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        bool ok = ti[s] != TMAX;
        slot_idx[(size_t)rm * SLOTS + s] = ok ? (int32_t)ti[s] : -1;
        slot_pos[(size_t)rm * SLOTS + s] = ok ? pin[ti[s]] : make_float2(0.f, 0.f);
    }
}

__global__ void __launch_bounds__(256) k_cells(const uint32_t* __restrict__ begin,
                                               const uint32_t* __restrict__ end,
                                               const uint32_t* __restrict__ sv,
                                               const float2* __restrict__ pin,
                                               float2* __restrict__ sorted_pos,
                                               double* __restrict__ sums,
                                               int32_t* __restrict__ slot_idx,
                                               float2* __restrict__ slot_pos,
                                               uint32_t* __restrict__ big_list,
                                               uint32_t* __restrict__ big_count, uint32_t gf,
                                               float seed) {
    uint32_t mc = blockIdx.x * blockDim.x + threadIdx.x;
    if (mc >= gf * gf) return;
    uint32_t b = begin[mc], e = end[mc];
    if (e - b > BIG_CELL) {
        big_list[atomicAdd(big_count, 1u)] = mc;
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
        top8_insert(tb, ti, ticket_bits(i, seed), i);
    }
    uint32_t rm = compact1by1(mc >> 1) * gf + compact1by1(mc);
    write_cell(rm, sx, sy, e - b, ti, pin, sums, slot_idx, slot_pos);
}

// One wave per crowded cell: lanes stride the segment keeping private top-8
// lists and f64 partial sums, then a fixed shuffle tree reduces the sums and
// eight wave-wide argmin rounds merge the lists.
__global__ void __launch_bounds__(256) k_cells_big(const uint32_t* __restrict__ begin,
                                                   const uint32_t* __restrict__ end,
                                                   const uint32_t* __restrict__ sv,
                                                   const float2* __restrict__ pin,
                                                   float2* __restrict__ sorted_pos,
                                                   double* __restrict__ sums,
                                                   int32_t* __restrict__ slot_idx,
                                                   float2* __restrict__ slot_pos,
                                                   const uint32_t* __restrict__ big_list,
                                                   const uint32_t* __restrict__ big_count,
                                                   uint32_t gf, float seed) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t nbig = *big_count;
    for (uint32_t w = wave; w < nbig; w += nwaves) {
        uint32_t mc = big_list[w];
        uint32_t b = begin[mc], e = end[mc];
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
            top8_insert(tb, ti, ticket_bits(i, seed), i);
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
            uint32_t rm = compact1by1(mc >> 1) * gf + compact1by1(mc);
            write_cell(rm, sx, sy, e - b, out, pin, sums, slot_idx, slot_pos);
        }
    }
}

// ---- 5. pyramid ---------------------------------------------------------------

struct PyrArgs {
    double* sums[MAX_LEVELS];
    float4* cent[MAX_LEVELS];
};

__device__ inline float4 centroid(double sx, double sy, double n) {
    if (n <= 0.0) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4((float)(sx / n), (float)(sy / n), (float)n, 0.f);
}

// One block per T x T tile of level `lin`; produces the `nout` coarser levels
// of that tile (down to one cell per tile).  Children are summed in a fixed order.
__global__ void __launch_bounds__(256) k_pyramid(PyrArgs a, int lin, int T, int nout,
                                                 int write_in) {
    __shared__ double s0[32 * 32 * 3];
    __shared__ double s1[16 * 16 * 3];
    const int gin = 4 << lin;
    const int tiles_x = gin / T;
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
        int x = c % T, y = c / T;
        size_t rm = (size_t)(ty * T + y) * gin + (tx * T + x);
        double sx = a.sums[lin][3 * rm], sy = a.sums[lin][3 * rm + 1], n = a.sums[lin][3 * rm + 2];
        s0[3 * c] = sx;
        s0[3 * c + 1] = sy;
        s0[3 * c + 2] = n;
        if (write_in) a.cent[lin][rm] = centroid(sx, sy, n);
    }
    __syncthreads();
    double* cur = s0;
    double* nxt = s1;
    int size = T;
    for (int lev = 1; lev <= nout; ++lev) {
        int half = size / 2;
        int lout = lin - lev;
        int gout = gin >> lev;
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
            size_t rm = (size_t)(ty * half + y) * gout + (tx * half + x);
            a.sums[lout][3 * rm] = v[0];
            a.sums[lout][3 * rm + 1] = v[1];
            a.sums[lout][3 * rm + 2] = v[2];
            a.cent[lout][rm] = centroid(v[0], v[1], v[2]);
        }
        __syncthreads();
        double* t = cur;
        cur = nxt;
        nxt = t;
        size = half;
    }
}

// ---- 6. forces ------------------------------------------------------------------

struct MbArgs {
    const float4* cent[MAX_LEVELS];
    float h[MAX_LEVELS];
    int L;
    int gf;
    float ak;        // alpha * repulsion
    float friction;
    float S;
    float maxstep;   // 2 * finest cell size (force-nearfield.frag:161)
};

// Clamped inverse-distance falloff of force-level.frag:43-56 /
// force-nearfield.frag:64-80 without the alpha*k*N factor:
// returns c with contribution = c * d.
__device__ inline float falloff_coef(float l) {
    float inv = __builtin_amdgcn_rsqf(l);            // 1/|d|
    return (l < 1.f) ? inv * __builtin_sqrtf(inv) : inv * inv;
}

__global__ void __launch_bounds__(256) k_mb_force(MbArgs a, const uint32_t* __restrict__ sv,
                                                  const float2* __restrict__ sorted_pos,
                                                  const int32_t* __restrict__ slot_idx,
                                                  const float2* __restrict__ slot_pos,
                                                  const float2* __restrict__ jitter,
                                                  const uint8_t* __restrict__ flags,
                                                  uint32_t np, float2* __restrict__ pout,
                                                  float2* __restrict__ vout) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t i = sv[k];
    const float2 p = sorted_pos[k];
    float fx = 0.f, fy = 0.f;

    // far field, level 0: every cell outside the 3x3 neighbourhood (force-level.frag:70-79)
    {
        const float4* __restrict__ lv = a.cent[0];
        int qx = cell_of(p.x, a.h[0], 4), qy = cell_of(p.y, a.h[0], 4);
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
    for (int lvl = 1; lvl < a.L; ++lvl) {
        const int g = 4 << lvl;
        const float4* __restrict__ lv = a.cent[lvl];
        qx = cell_of(p.x, a.h[lvl], g);
        qy = cell_of(p.y, a.h[lvl], g);
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
    const int gf = a.gf;
    if (a.L == 1) {
        qx = cell_of(p.x, a.h[0], gf);
        qy = cell_of(p.y, a.h[0], gf);
    }
    const float4* __restrict__ fin = a.cent[a.L - 1];
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
            float sxp = 0.f, syp = 0.f, sampled = 0.f;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) {
                const int j = slot_idx[cell * SLOTS + s];
                if (j < 0) break;
                if ((uint32_t)j == i) continue;
                sampled += 1.f;
                const float2 q = slot_pos[cell * SLOTS + s];
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
    wx *= a.ak;
    wy *= a.ak;
    wx += wx * r.x;
    wy += wy * r.y;
    {
        float sp = sqrtf(wx * wx + wy * wy);
        if (sp > a.maxstep) {
            float s = a.maxstep / sp;
            wx *= s;
            wy *= s;
        }
    }
    const float2 v = make_float2(a.ak * fx + wx, a.ak * fy + wy);
    if (vout) vout[i] = v;
    if (pout) pout[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, a.friction, a.S);
}

}  // namespace

int mb_alloc(Ctx* c) {
    int rc = CG_OK;
    const size_t np = c->np ? c->np : 1;
    const size_t cells = (size_t)c->gf * c->gf;
    for (int b = 0; b < 2; ++b) {
        c->keys[b] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
        c->vals[b] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
    }
    c->sorted_pos = (float2*)dalloc(c, np * 8, &rc); if (rc) return rc;
    c->cell_begin = (uint32_t*)dalloc(c, cells * 4, &rc); if (rc) return rc;
    c->cell_end = (uint32_t*)dalloc(c, cells * 4, &rc); if (rc) return rc;
    c->slot_idx = (int32_t*)dalloc(c, cells * SLOTS * 4, &rc); if (rc) return rc;
    c->slot_pos = (float2*)dalloc(c, cells * SLOTS * 8, &rc); if (rc) return rc;
    c->big_list = (uint32_t*)dalloc(c, cells * 4, &rc); if (rc) return rc;
    c->big_count = (uint32_t*)dalloc(c, 64, &rc); if (rc) return rc;
    for (uint32_t l = 0; l < c->L; ++l) {
        size_t g = 4u << l;
        c->sums[l] = (double*)dalloc(c, g * g * 3 * sizeof(double), &rc); if (rc) return rc;
        c->cent[l] = (float4*)dalloc(c, g * g * sizeof(float4), &rc); if (rc) return rc;
    }
    // sized for full 32-bit keys: the collision pass sorts its cell keys with it too
    size_t tmp = 0;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, c->keys[0], c->keys[1], c->vals[0],
                                                 c->vals[1], (int)np, 0, 32, c->stream));
    c->sort_tmp_bytes = tmp;
    c->sort_tmp = dalloc(c, tmp ? tmp : 16, &rc);
    return rc;
}

int mb_run(Ctx* c, const cg_sim_params* p, float alpha, float seed, const float2* pin,
           float2* pout, float2* vout, bool keys_ready) {
    if (c->np == 0) return CG_OK;
    int rc = CG_OK;
    prof_begin(c, CG_STAGE_MB_SORT);
    if (!keys_ready) {
        rc = keys_only(c, pin, p->space_size);
        if (rc) return rc;
    }
    const uint32_t np = c->np, gf = c->gf, cells = gf * gf;
    int bits = 0;
    while ((1u << bits) < gf) ++bits;
    size_t tmp = c->sort_tmp_bytes;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, tmp, c->keys[0], c->keys[1],
                                                 c->vals[0], c->vals[1], (int)np, 0, 2 * bits,
                                                 c->stream));
    prof_end(c, CG_STAGE_MB_SORT);
    prof_begin(c, CG_STAGE_MB_CELLS);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                       c->cell_begin, cells, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                       c->cell_end, cells, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, c->big_count, 1u, 0u);
    hipLaunchKernelGGL(k_mark_bounds, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream,
                       c->keys[1], np, c->cell_begin, c->cell_end);
    const uint32_t lf = c->L - 1;
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(cells, 256)), dim3(256), 0, c->stream,
                       c->cell_begin, c->cell_end, c->vals[1], pin, c->sorted_pos, c->sums[lf],
                       c->slot_idx, c->slot_pos, c->big_list, c->big_count, gf, seed);
    hipLaunchKernelGGL(k_cells_big, dim3(256), dim3(256), 0, c->stream, c->cell_begin, c->cell_end,
                       c->vals[1], pin, c->sorted_pos, c->sums[lf], c->slot_idx, c->slot_pos,
                       c->big_list, c->big_count, gf, seed);
    prof_end(c, CG_STAGE_MB_CELLS);
    prof_begin(c, CG_STAGE_MB_PYRAMID);
    // pyramid: finest -> tiles of up to 32x32 -> remaining top levels
    PyrArgs pa;
    for (int l = 0; l < MAX_LEVELS; ++l) {
        pa.sums[l] = c->sums[l];
        pa.cent[l] = c->cent[l];
    }
    {
        int T1 = (int)std::min<uint32_t>(32u, gf / 4);
        int nout1 = 0;
        while ((1 << nout1) < T1) ++nout1;
        int tiles = (int)(gf / T1) * (int)(gf / T1);
        hipLaunchKernelGGL(k_pyramid, dim3(tiles), dim3(256), 0, c->stream, pa, (int)lf, T1, nout1,
                           1);
        uint32_t g2 = gf / T1;
        if (g2 > 4) {
            int lin2 = (int)lf - nout1;
            int T2 = (int)(g2 / 4);
            int nout2 = 0;
            while ((1 << nout2) < T2) ++nout2;
            hipLaunchKernelGGL(k_pyramid, dim3(16), dim3(256), 0, c->stream, pa, lin2, T2, nout2, 0);
        }
    }
    prof_end(c, CG_STAGE_MB_PYRAMID);
    MbArgs ma;
    for (int l = 0; l < MAX_LEVELS; ++l) {
        ma.cent[l] = c->cent[l];
        ma.h[l] = (l < (int)c->L) ? (float)((double)p->space_size / (double)(4u << l)) : 1.f;
    }
    ma.L = (int)c->L;
    ma.gf = (int)gf;
    ma.ak = alpha * p->repulsion;
    ma.friction = p->friction;
    ma.S = p->space_size;
    ma.maxstep = 2.0f * ma.h[c->L - 1];
    prof_begin(c, CG_STAGE_MB_FORCE);
    hipLaunchKernelGGL(k_mb_force, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, ma,
                       c->vals[1], c->sorted_pos, c->slot_idx, c->slot_pos, c->jitter, c->flags, np,
                       pout, vout);
    prof_end(c, CG_STAGE_MB_FORCE);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

}  // namespace cg
