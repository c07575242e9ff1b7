// k_link.hip — link springs: CSR build on device and the per-direction
// Gauss-Seidel spring pass fused with the integrate pass.
//
// Reference: GraphData/index.ts:286-289 (isPointIndex), :525-568 (adjacency,
// degrees); ForceLink/index.ts:40-171 (create: CSR, bias, strength, random
// rest length), :221-263 (run); ForceLink/force-spring.ts:38-96 (the loop).
//
// The reference evaluates each point's links serially, every term seeing the
// velocity accumulated so far.  Points are scheduled by live degree:
//   - up to LANE_MAX entries: one lane per point, the serial loop as written;
//   - more: one workgroup per point.  The chain is cut into one chunk per lane;
//     each sweep runs every chunk serially from a guessed start velocity,
//     records the chunk's end velocity and its slope prod(1 - t_k), and an
//     affine prefix scan over the chunks yields the next guesses (multiple
//     shooting).  Sweeps repeat until the guesses stop moving; a last sweep
//     from the converged starts gives the point's velocity.
#include <hipcub/hipcub.hpp>

#include "cg_internal.h"

namespace cg {
namespace {

// ---- CSR build ------------------------------------------------------------------

// isPointIndex (GraphData/index.ts:286-289): an integer in [0, n).
__device__ inline bool valid_index(float v, uint32_t n) {
    return v >= 0.f && floorf(v) == v && (double)v < (double)n;
}

__global__ void k_validate(const float2* __restrict__ links, uint32_t m, uint32_t n,
                           uint32_t* __restrict__ src, uint32_t* __restrict__ tgt,
                           uint32_t* __restrict__ outdeg, uint32_t* __restrict__ indeg) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= m) return;
    float2 st = links[l];
    bool ok = valid_index(st.x, n) && valid_index(st.y, n);
    uint32_t s = ok ? (uint32_t)st.x : n, t = ok ? (uint32_t)st.y : n;
    src[l] = s;
    tgt[l] = t;
    if (ok) {
        atomicAdd(&outdeg[s], 1u);
        atomicAdd(&indeg[t], 1u);
    }
}

__global__ void k_degree(const uint32_t* __restrict__ outdeg, const uint32_t* __restrict__ indeg,
                         uint32_t n, uint32_t* __restrict__ degree) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) degree[p] = outdeg[p] + indeg[p];
}

__global__ void k_iota_keys(const uint32_t* __restrict__ owner, uint32_t m,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= m) return;
    keys[l] = owner[l];
    vals[l] = l;
}

// Entry e of the point-ordered CSR (ForceLink/index.ts:58-79): neighbour,
// random draw, strength, bias — bias and strength in f64, stored as f32.
__global__ void k_entries(const uint32_t* __restrict__ sorted_owner,
                          const uint32_t* __restrict__ sorted_link, uint32_t E,
                          const uint32_t* __restrict__ other_of, const uint32_t* __restrict__ degree,
                          const float* __restrict__ strength, const float* __restrict__ rnd,
                          const uint8_t* __restrict__ flags, uint4* __restrict__ ent,
                          uint32_t* __restrict__ live) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    uint32_t l = sorted_link[e];
    uint32_t p = sorted_owner[e];
    uint32_t q = other_of[l];
    double dq = (double)degree[q], dp = (double)degree[p];
    double sum = dq + dp;
    double bias = sum != 0.0 ? dq / sum : 0.5;
    double s = -1.0;
    if (strength) {
        double u = (double)strength[l];
        if (isfinite(u) && u >= 0.0) s = u;
    }
    if (s < 0.0) {
        double mind = dq < dp ? dq : dp;
        s = 1.0 / (mind > 1.0 ? mind : 1.0);
    }
    ent[e] = make_uint4(q, __float_as_uint(rnd[l]), __float_as_uint((float)sqrt(s)),
                        __float_as_uint((float)bias));
    if (!(flags[q] & F_ABSENT)) atomicAdd(&live[p], 1u);
}

__global__ void k_sched_keys(const uint32_t* __restrict__ present, uint32_t np,
                             const uint32_t* __restrict__ live, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t p = present[k];
    keys[k] = live[p];
    vals[k] = p;
}

__global__ void k_tier_count(const uint32_t* __restrict__ skeys, uint32_t ns,
                             uint32_t* __restrict__ counts) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    uint32_t d = skeys[k];
    if (d <= LANE_MAX) atomicAdd(&counts[0], 1u);
    else if (d <= WAVE_MAX) atomicAdd(&counts[1], 1u);
}

// Copy each scheduled point's live entries, in link order, to its schedule slot.
__global__ void k_permute(const uint32_t* __restrict__ sched, uint32_t ns,
                          const uint32_t* __restrict__ soff, const uint32_t* __restrict__ off_pt,
                          const uint4* __restrict__ ent_pt, const uint8_t* __restrict__ flags,
                          uint4* __restrict__ ent) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    uint32_t p = sched[j];
    uint32_t dst = soff[j];
    for (uint32_t e = off_pt[p]; e < off_pt[p + 1]; ++e) {
        uint4 v = ent_pt[e];
        if (!(flags[v.x] & F_ABSENT)) ent[dst++] = v;
    }
}

// ---- spring pass ------------------------------------------------------------------

struct LkArgs {
    float spring_alpha;  // linkSpring * alpha
    float dist;          // linkDistance
    float lo, span;      // variation range [lo, hi): lo and hi - lo
    float friction, S;
    float tol;           // multiple-shooting convergence threshold on chunk starts
};

// One term of force-spring.ts:61-90.  Returns the step fraction t (v += x*t).
__device__ inline float spring_term(const LkArgs& a, uint4 e, float2 pp, float2 v, float2 q,
                                    float& x, float& y) {
    float r = __uint_as_float(e.y), s = __uint_as_float(e.z), b = __uint_as_float(e.w);
    float D = (r * a.span + a.lo) * a.dist;
    x = q.x - (pp.x + v.x);
    y = q.y - (pp.y + v.y);
    float l = sqrtf(x * x + y * y);
    l = fmaxf(l, D * 0.99f);
    l = (l - D) / l;
    l *= a.spring_alpha;
    l *= s;
    l *= b;
    return l;
}

__global__ void __launch_bounds__(256) k_link_lane(LkArgs a, const uint32_t* __restrict__ sched,
                                                   const uint32_t* __restrict__ soff,
                                                   const uint4* __restrict__ ent, uint32_t count,
                                                   const float2* __restrict__ pin,
                                                   const uint8_t* __restrict__ flags,
                                                   float2* __restrict__ pout,
                                                   float2* __restrict__ vout) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const uint32_t p = sched[j];
    const uint32_t b = soff[j], e = soff[j + 1];
    const float2 pp = pin[p];
    float2 v = make_float2(0.f, 0.f);
    for (uint32_t k = b; k < e; ++k) {
        uint4 en = ent[k];
        float2 q = pin[en.x];
        float x, y;
        float t = spring_term(a, en, pp, v, q, x, y);
        v.x += x * t;
        v.y += y * t;
    }
    if (vout) vout[p] = v;
    if (pout) pout[p] = (flags[p] & F_PINNED) ? pp : integrate(pp, v, a.friction, a.S);
}

constexpr int MAX_SWEEPS = 12;

template <int B>
__global__ void __launch_bounds__(B) k_link_chunk(LkArgs a, const uint32_t* __restrict__ sched,
                                                  const uint32_t* __restrict__ soff,
                                                  const uint4* __restrict__ ent, uint32_t first,
                                                  const float2* __restrict__ pin,
                                                  const uint8_t* __restrict__ flags,
                                                  float2* __restrict__ pout,
                                                  float2* __restrict__ vout) {
    __shared__ float sA[B], sX[B], sY[B];
    __shared__ float s_delta[B / 64];
    __shared__ float2 s_final;
    const int t = threadIdx.x;
    const uint32_t j = first + blockIdx.x;
    const uint32_t p = sched[j];
    const uint32_t base = soff[j];
    const uint32_t d = soff[j + 1] - base;
    const uint32_t c = (d + B - 1) / B;
    const uint32_t lo = min(d, (uint32_t)t * c), hi = min(d, lo + c);
    const int last = (int)((d - 1) / c);
    const float2 pp = pin[p];
    float2 g = make_float2(0.f, 0.f);
    for (int sweep = 0; sweep <= MAX_SWEEPS; ++sweep) {
        float2 v = g;
        float A = 1.f;
        for (uint32_t k = lo; k < hi; ++k) {
            uint4 en = ent[base + k];
            float2 q = pin[en.x];
            float x, y;
            float tk = spring_term(a, en, pp, v, q, x, y);
            v.x += x * tk;
            v.y += y * tk;
            A *= (1.f - tk);
        }
        if (sweep == MAX_SWEEPS || (sweep > 0 && s_delta[0] <= a.tol)) {
            // g is converged: v is this chunk's exact serial end from g
            if (t == last) s_final = v;
            break;
        }
        // chunk map  start -> A*start + (v - A*g); inclusive scan of the composition
        float cA = A, cx = v.x - A * g.x, cy = v.y - A * g.y;
        sA[t] = cA;
        sX[t] = cx;
        sY[t] = cy;
        __syncthreads();
        for (int off = 1; off < B; off <<= 1) {
            float pA = 1.f, px = 0.f, py = 0.f;
            if (t >= off) { pA = sA[t - off]; px = sX[t - off]; py = sY[t - off]; }
            __syncthreads();
            if (t >= off) {
                // mine after theirs: s -> cA*(pA*s + p) + c
                cx = cA * px + cx;
                cy = cA * py + cy;
                cA = cA * pA;
                sA[t] = cA;
                sX[t] = cx;
                sY[t] = cy;
            }
            __syncthreads();
        }
        float2 start = (t == 0) ? make_float2(0.f, 0.f) : make_float2(sX[t - 1], sY[t - 1]);
        float dlt = fmaxf(fabsf(start.x - g.x), fabsf(start.y - g.y));
        if (hi <= lo) dlt = 0.f;
        g = start;
        // block max of the start corrections
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dlt = fmaxf(dlt, __shfl_xor(dlt, o, 64));
        __syncthreads();
        if ((t & 63) == 0) s_delta[t >> 6] = dlt;
        __syncthreads();
        if (t == 0) {
            float mx = s_delta[0];
            for (int w = 1; w < B / 64; ++w) mx = fmaxf(mx, s_delta[w]);
            s_delta[0] = mx;
        }
        __syncthreads();
    }
    __syncthreads();
    if (t == 0) {
        float2 v = s_final;
        if (vout) vout[p] = v;
        if (pout) pout[p] = (flags[p] & F_PINNED) ? pp : integrate(pp, v, a.friction, a.S);
    }
}

}  // namespace

// Builds both directions' CSR, degrees and schedules from the host link copy.
int link_build(Ctx* c) {
    c->links_ready = false;
    for (int d = 0; d < 2; ++d) {
        LinkDir& L = c->dir[d];
        dfree(c, L.ent_pt); dfree(c, L.off_pt); dfree(c, L.ent); dfree(c, L.sched); dfree(c, L.soff);
        L = LinkDir();
    }
    dfree(c, c->degree);
    c->degree = nullptr;
    if (!c->links_set || c->n == 0) return CG_OK;
    const uint32_t n = c->n, m = c->m;
    int rc = CG_OK;
    c->degree = (uint32_t*)dalloc(c, (size_t)n * 4, &rc);
    if (rc) return rc;
    if (m == 0) {
        CG_HIP(c, hipMemsetAsync(c->degree, 0, (size_t)n * 4, c->stream));
    }
    // temporaries
    std::vector<void*> tmp;
    auto talloc = [&](size_t bytes) -> void* {
        void* p = dalloc(c, bytes ? bytes : 16, &rc);
        if (p) tmp.push_back(p);
        return p;
    };
    auto cleanup = [&]() { for (void* p : tmp) dfree(c, p); };
    const size_t mm = m ? m : 1;
    float2* d_links = (float2*)talloc(mm * 8);
    float* d_str = c->h_strength.empty() ? nullptr : (float*)talloc(mm * 4);
    float* d_rin = (float*)talloc(mm * 4);
    float* d_rout = (float*)talloc(mm * 4);
    uint32_t* src = (uint32_t*)talloc(mm * 4);
    uint32_t* tgt = (uint32_t*)talloc(mm * 4);
    uint32_t* outdeg = (uint32_t*)talloc(((size_t)n + 1) * 4);
    uint32_t* indeg = (uint32_t*)talloc(((size_t)n + 1) * 4);
    const size_t kk = std::max<size_t>(mm, (size_t)c->np) + 1;  // CSR sort and schedule sort
    uint32_t* k0 = (uint32_t*)talloc(kk * 4);
    uint32_t* k1 = (uint32_t*)talloc(kk * 4);
    uint32_t* v0 = (uint32_t*)talloc(kk * 4);
    uint32_t* v1 = (uint32_t*)talloc(kk * 4);
    uint32_t* live = (uint32_t*)talloc(((size_t)n + 1) * 4);
    uint32_t* counts = (uint32_t*)talloc(64);
    if (rc) { cleanup(); return rc; }
    if (m) {
        CG_HIP(c, hipMemcpyAsync(d_links, c->h_links.data(), (size_t)m * 8, hipMemcpyHostToDevice, c->stream));
        if (d_str)
            CG_HIP(c, hipMemcpyAsync(d_str, c->h_strength.data(), (size_t)m * 4, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipMemcpyAsync(d_rin, c->h_rin.data(), (size_t)m * 4, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipMemcpyAsync(d_rout, c->h_rout.data(), (size_t)m * 4, hipMemcpyHostToDevice, c->stream));
    }
    CG_HIP(c, hipMemsetAsync(outdeg, 0, ((size_t)n + 1) * 4, c->stream));
    CG_HIP(c, hipMemsetAsync(indeg, 0, ((size_t)n + 1) * 4, c->stream));
    if (m) {
        hipLaunchKernelGGL(k_validate, dim3(blocks_for(m, 256)), dim3(256), 0, c->stream, d_links, m,
                           n, src, tgt, outdeg, indeg);
        hipLaunchKernelGGL(k_degree, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, outdeg, indeg,
                           n, c->degree);
    }
    int nbits = 1;
    while (nbits < 32 && (1ull << nbits) <= (unsigned long long)n) ++nbits;  // n itself must fit
    // sort / scan temp storage sized for the largest use
    size_t sort_bytes = 0, scan_bytes = 0, sched_bytes = 0;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, v0, v1, (int)mm, 0,
                                                 nbits, c->stream));
    CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, outdeg, outdeg, (int)n + 1,
                                               c->stream));
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sched_bytes, k0, k1, v0, v1,
                                                 (int)(c->np ? c->np : 1), 0, 32, c->stream));
    size_t tbytes = std::max(sort_bytes, std::max(scan_bytes, sched_bytes));
    void* d_tmp = talloc(tbytes);
    if (rc) { cleanup(); return rc; }
    uint32_t E = 0;
    for (int d = 0; d < 2; ++d) {
        LinkDir& L = c->dir[d];
        uint32_t* owner = d == 0 ? src : tgt;
        uint32_t* other = d == 0 ? tgt : src;
        uint32_t* cnt = d == 0 ? outdeg : indeg;  // entries owned per point
        L.off_pt = (uint32_t*)dalloc(c, ((size_t)n + 1) * 4, &rc);
        if (rc) { cleanup(); return rc; }
        // offsets: exclusive scan of the per-point counts (cnt[n] == 0)
        size_t sb = tbytes;
        CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(d_tmp, sb, cnt, L.off_pt, (int)n + 1, c->stream));
        CG_HIP(c, hipMemcpyAsync(&E, L.off_pt + n, 4, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
        L.E = E;
        L.ent_pt = (uint4*)dalloc(c, (size_t)(E ? E : 1) * 16, &rc);
        if (rc) { cleanup(); return rc; }
        CG_HIP(c, hipMemsetAsync(live, 0, ((size_t)n + 1) * 4, c->stream));
        if (m) {
            hipLaunchKernelGGL(k_iota_keys, dim3(blocks_for(m, 256)), dim3(256), 0, c->stream, owner,
                               m, k0, v0);
            sb = tbytes;
            CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(d_tmp, sb, k0, k1, v0, v1, (int)m, 0, nbits,
                                                         c->stream));
            if (E)
                hipLaunchKernelGGL(k_entries, dim3(blocks_for(E, 256)), dim3(256), 0, c->stream, k1,
                                   v1, E, other, c->degree, d_str, d == 0 ? d_rin : d_rout,
                                   c->flags, L.ent_pt, live);
        }
        // schedule: present points by ascending live degree
        const uint32_t ns = c->np;
        L.n_sched = ns;
        L.sched = (uint32_t*)dalloc(c, (size_t)(ns ? ns : 1) * 4, &rc);
        L.soff = (uint32_t*)dalloc(c, ((size_t)ns + 1) * 4, &rc);
        if (rc) { cleanup(); return rc; }
        if (ns) {
            hipLaunchKernelGGL(k_sched_keys, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream,
                               c->present, ns, live, k0, v0);
            sb = tbytes;
            CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(d_tmp, sb, k0, k1, v0, L.sched, (int)ns, 0,
                                                         32, c->stream));
            CG_HIP(c, hipMemsetAsync(k1 + ns, 0, 4, c->stream));
            sb = tbytes;
            CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(d_tmp, sb, k1, L.soff, (int)ns + 1, c->stream));
            CG_HIP(c, hipMemsetAsync(counts, 0, 64, c->stream));
            hipLaunchKernelGGL(k_tier_count, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, k1,
                               ns, counts);
            uint32_t h[3] = {0, 0, 0};
            CG_HIP(c, hipMemcpyAsync(h, counts, 8, hipMemcpyDeviceToHost, c->stream));
            CG_HIP(c, hipMemcpyAsync(&h[2], L.soff + ns, 4, hipMemcpyDeviceToHost, c->stream));
            CG_HIP(c, hipStreamSynchronize(c->stream));
            L.n_lane = h[0];
            L.n_wave = h[1];
            L.n_block = ns - h[0] - h[1];
            L.E_live = h[2];
            L.ent = (uint4*)dalloc(c, (size_t)(L.E_live ? L.E_live : 1) * 16, &rc);
            if (rc) { cleanup(); return rc; }
            hipLaunchKernelGGL(k_permute, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream,
                               L.sched, ns, L.soff, L.off_pt, L.ent_pt, c->flags, L.ent);
        }
        CG_HIP(c, hipGetLastError());
    }
    CG_HIP(c, hipStreamSynchronize(c->stream));
    cleanup();
    if (rc) return rc;
    c->links_ready = true;
    return CG_OK;
}

int link_run(Ctx* c, const cg_sim_params* p, int dir, float alpha, const float2* pin,
             float2* pout, float2* vout) {
    if (!c->links_ready) return CG_OK;
    const LinkDir& L = c->dir[dir];
    LkArgs a;
    a.spring_alpha = p->link_spring * alpha;
    a.dist = p->link_distance;
    a.lo = p->link_dist_random_variation_range[0];
    a.span = p->link_dist_random_variation_range[1] - p->link_dist_random_variation_range[0];
    a.friction = p->friction;
    a.S = p->space_size;
    a.tol = 1e-7f * p->space_size;
    if (L.n_lane)
        hipLaunchKernelGGL(k_link_lane, dim3(blocks_for(L.n_lane, 256)), dim3(256), 0, c->stream, a,
                           L.sched, L.soff, L.ent, L.n_lane, pin, c->flags, pout, vout);
    if (L.n_wave)
        hipLaunchKernelGGL(k_link_chunk<64>, dim3(L.n_wave), dim3(64), 0, c->stream, a, L.sched,
                           L.soff, L.ent, L.n_lane, pin, c->flags, pout, vout);
    if (L.n_block)
        hipLaunchKernelGGL(k_link_chunk<1024>, dim3(L.n_block), dim3(1024), 0, c->stream, a, L.sched,
                           L.soff, L.ent, L.n_lane + L.n_wave, pin, c->flags, pout, vout);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

}  // namespace cg
