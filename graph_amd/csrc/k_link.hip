// k_link.hip — link springs: per-graph CSR build on device, the packed
// schedule over all graphs, and the per-direction Gauss-Seidel spring pass
// fused with the integrate pass.
//
// Reference: GraphData/index.ts:286-289 (isPointIndex), :525-568 (adjacency,
// degrees); ForceLink/index.ts:40-171 (create: CSR, bias, strength, random
// rest length), :221-263 (run); ForceLink/force-spring.ts:38-96 (the loop).
//
// The reference evaluates each point's links serially, every term seeing the
// velocity accumulated so far.  Points are scheduled by live degree:
//   - up to lane_max entries (Tuning): one lane per point, the serial loop
//     as written;
//   - up to group_max entries: one 64-lane workgroup per point; above: 256
//     lanes.  The chain is cut into one chunk per lane, at most 4 entries
//     (64 lanes) or hub_entries (256 lanes), whose terms the lane gathers once into
//     registers; each sweep runs every chunk from a guessed start velocity,
//     carrying the chunk's end velocity and 2x2 Jacobian, and an affine
//     prefix scan over the chunks yields the next guesses (multiple shooting
//     with Newton-accurate chunk maps).  Sweeps repeat until no start moves
//     more than FIX_ACCEPT tolerances (or the correction stalls near the
//     tolerance, or MAX_SWEEPS); the last chunk's end, moved to its corrected
//     start, is the point's velocity.  Converged starts are kept per point as
//     the next call's first guesses.
//   - a point above 256 x hub_entries entries is cut into segments, one 256-lane
//     workgroup each (one CU no longer carries a whole hub): each segment is
//     solved as above from a guessed start and reports its end and end
//     Jacobian; a fix-up composes the segments from the exact start 0 and
//     corrects the guesses (the same Newton step, one level up), for up to
//     HUB_ROUNDS rounds, until the corrections hold within FIX_ACCEPT
//     tolerances or the next one predicted from their rate is within one.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "cg_internal.h"

namespace cg {
namespace {

// ---- per-graph CSR build ------------------------------------------------------

// isPointIndex (GraphData/index.ts:286-289): an integer in [0, n).
__device__ inline bool valid_index(float v, uint32_t n) {
    return v >= 0.f && floorf(v) == v && (double)v < (double)n;
}

__global__ void k_fill_words(uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
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
// random draw, strength, bias — bias and strength in f64, stored as f32 —
// and whether it is live (owner and neighbour present).
__global__ void k_entries(const uint32_t* __restrict__ sorted_owner,
                          const uint32_t* __restrict__ sorted_link, uint32_t E,
                          const uint32_t* __restrict__ other_of, const uint32_t* __restrict__ degree,
                          const float* __restrict__ strength, const float* __restrict__ rnd,
                          const uint8_t* __restrict__ flags, uint4* __restrict__ ent,
                          uint32_t* __restrict__ live_flag) {
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
    // only entries of a present owner toward a present neighbour are scheduled
    // (force-spring.ts:69-74 skips an absent neighbour; an absent owner never moves)
    live_flag[e] = ((flags[q] | flags[p]) & F_ABSENT) ? 0u : 1u;
}

// Live entries per point, from the scan of live flags.
__global__ void k_live_count(const uint32_t* __restrict__ off_pt, const uint32_t* __restrict__ lrank,
                             uint32_t n, uint32_t* __restrict__ live) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) live[p] = lrank[off_pt[p + 1]] - lrank[off_pt[p]];
}

// ---- packed schedule ------------------------------------------------------------

struct GLinkPtr {
    const uint4* ent_pt;
    const uint32_t* off_pt;
    const uint32_t* live;
    const uint32_t* lrank;
    uint32_t n;
};

__device__ inline uint32_t live_degree(uint32_t p, const uint32_t* __restrict__ gid,
                                       const GDesc* __restrict__ desc, const GLinkPtr* __restrict__ tab) {
    const uint32_t g = gid[p];
    const GLinkPtr t = tab[g];
    return t.live ? t.live[p - desc[g].p0] : 0u;
}

// Schedule sort keys.  Chunked-tier points sort by live degree.  Lane-tier
// points sort into three classes only (at most 2 entries, at most LANE_SPLIT,
// then the rest), so within a class the stable sort keeps ascending point
// index and a wave's own-point loads and stores fall on neighbouring
// addresses, while a wave's chains stay of similar length.
__global__ void k_sched_keys(const uint32_t* __restrict__ present, uint32_t np,
                             const uint32_t* __restrict__ gid, const GDesc* __restrict__ desc,
                             const GLinkPtr* __restrict__ tab, uint32_t lane_max, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t p = present[k];
    const uint32_t deg = live_degree(p, gid, desc, tab);
    keys[k] = deg > lane_max ? deg : (deg > LANE_SPLIT ? 2u : (deg > 2 ? 1u : 0u));
    vals[k] = p;
}

// Live degree of every schedule slot, in schedule order.
__global__ void k_sched_degree(const uint32_t* __restrict__ sched, uint32_t ns,
                               const uint32_t* __restrict__ gid, const GDesc* __restrict__ desc,
                               const GLinkPtr* __restrict__ tab, uint32_t* __restrict__ deg) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) deg[j] = live_degree(sched[j], gid, desc, tab);
}

// Tier boundaries over the schedule's degrees (lane-tier degrees first, then
// the chunked tiers ascending): first position above lane_max and above
// group_max (counts pre-filled with ns).  One writer each.
__global__ void k_tier_bounds(const uint32_t* __restrict__ skeys, uint32_t ns, uint32_t lane_max,
                              uint32_t group_max, uint32_t* __restrict__ counts) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    uint32_t d = skeys[k], prev = k ? skeys[k - 1] : 0u;
    if (d > lane_max && (k == 0 || prev <= lane_max)) counts[0] = k;
    if (d > group_max && (k == 0 || prev <= group_max)) counts[1] = k;
}

// Graph slot of each lane-tier point, in schedule order.
__global__ void k_sched_graph(const uint32_t* __restrict__ sched, uint32_t n, const uint32_t* __restrict__ gid,
                              uint32_t* __restrict__ sgraph) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) sgraph[j] = gid[sched[j]];
}

__global__ void k_slot_of(const uint32_t* __restrict__ sched, uint32_t ns, uint32_t* __restrict__ slot_of) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) slot_of[sched[j]] = j;
}

// One lane per CSR entry of every graph (graph s holds entries [ebase[s],
// ebase[s+1])): a live entry (owner and neighbour present) goes to its owner's
// schedule slot at its rank among the owner's live entries, in link order, with
// a global neighbour index.
__global__ void k_assemble(const uint32_t* __restrict__ ebase, uint32_t G, uint32_t total,
                           const GLinkPtr* __restrict__ tab, const GDesc* __restrict__ desc,
                           const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ soff,
                           const uint8_t* __restrict__ flags, uint4* __restrict__ ent) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t s = find_graph(ebase, G, t);
    const GLinkPtr tb = tab[s];
    const uint32_t e = t - ebase[s];
    // owner: the last local point whose entries start at or before e
    uint32_t lo = 0, hi = tb.n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tb.off_pt[mid] <= e) lo = mid;
        else hi = mid;
    }
    const uint32_t p0 = desc[s].p0, p = p0 + lo;
    uint4 v = tb.ent_pt[e];
    v.x += p0;
    if ((flags[p] | flags[v.x]) & F_ABSENT) return;
    ent[soff[slot_of[p]] + (tb.lrank[e] - tb.lrank[tb.off_pt[lo]])] = v;
}

// ---- spring pass ------------------------------------------------------------------

// One term of force-spring.ts:61-90.  Returns the step fraction t (v += x*t).
__device__ inline float spring_term(const GDesc& d, uint4 e, float2 pp, float2 v, float2 q,
                                    float& x, float& y) {
    float r = __uint_as_float(e.y), s = __uint_as_float(e.z), b = __uint_as_float(e.w);
    float D = (r * d.span + d.lo) * d.dist;
    x = q.x - (pp.x + v.x);
    y = q.y - (pp.y + v.y);
    float l = __builtin_amdgcn_sqrtf(x * x + y * y);
    l = fmaxf(l, D * 0.99f);
    l = (l - D) * __builtin_amdgcn_rcpf(l);
    l *= d.spring_alpha;
    l *= s;
    l *= b;
    return l;
}

// One lane's chain of entries [b, e), G entries fetched together (G entry
// loads, then G neighbour gathers).
template <int G>
__device__ inline float2 lane_chain(const GDesc& d, const uint4* __restrict__ ent, uint32_t b, uint32_t e,
                                    const float2* __restrict__ pin, float2 pp) {
    float2 v = make_float2(0.f, 0.f);
    for (uint32_t k = b; k < e; k += G) {
        uint4 en[G];
        float2 q[G];
#pragma unroll
        for (int u = 0; u < G; ++u) en[u] = ent[min(k + u, e - 1)];
#pragma unroll
        for (int u = 0; u < G; ++u) q[u] = pin[en[u].x];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            float x, y;
            float t = spring_term(d, en[u], pp, v, q[u], x, y);
            t = (k + u < e) ? t : 0.f;
            v.x += x * t;
            v.y += y * t;
        }
    }
    return v;
}

constexpr uint32_t LANE_RUN = 16;  // lane-tier blocks per XCD run

// Largest value over the wave, by DPP row shifts and row broadcasts (no LDS
// round trips), read from lane 63.
__device__ inline uint32_t wave_max_u32(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// One lane per lane-tier point.  Every load that the entry loads and the
// neighbour gathers wait on comes from the schedule itself (the point, its
// entry range and, for sets of several graphs, its graph slot), so a wave
// waits on three memory latencies (schedule, entries and descriptor,
// neighbours) before its first spring step; a point of a graph without the
// pass runs the chain and stores nothing.
__global__ void __launch_bounds__(256) k_link_lane(int pass, const uint32_t* __restrict__ sched,
                                                   const uint32_t* __restrict__ soff,
                                                   const uint32_t* __restrict__ sgraph,
                                                   const uint4* __restrict__ ent, uint32_t count,
                                                   const GDesc* __restrict__ desc,
                                                   float2* __restrict__ pos0,
                                                   float2* __restrict__ pos1,
                                                   const uint8_t* __restrict__ flags,
                                                   float2* __restrict__ vout, uint32_t one_graph) {
    // the longer-chain classes start first; each XCD takes runs of LANE_RUN
    // consecutive blocks of the schedule (graph-major inside a class, so in a
    // set of many graphs an XCD's L2 holds the positions of fewer graphs)
    const uint32_t r = xcd_run_block(blockIdx.x, gridDim.x, LANE_RUN) * blockDim.x + threadIdx.x;
    const bool in = r < count;
    const uint32_t j = in ? count - 1 - r : 0u;
    const uint32_t p = sched[j];
    const uint32_t b = soff[j], e1 = soff[j + 1];
    const GDesc& d = desc[one_graph ? 0u : sgraph[j]];
    const uint32_t e = in ? e1 : b;
    // batch depth from the wave's longest chain (every lane is still here)
    const uint32_t mx = wave_max_u32(e - b);
    if (!in) return;
    const bool bsel = (d.inbuf >> pass) & 1u;
    const float2* __restrict__ pin = bsel ? pos1 : pos0;
    const float2 pp = pin[p];
    const float2 v = mx <= 2 ? lane_chain<2>(d, ent, b, e, pin, pp)
                             : (mx <= 4 ? lane_chain<4>(d, ent, b, e, pin, pp)
                                        : (mx <= 16 ? lane_chain<8>(d, ent, b, e, pin, pp) : lane_chain<16>(d, ent, b, e, pin, pp)));
    if (!(d.passes & (1u << pass))) return;
    if (vout) {
        vout[p] = v;
        return;
    }
    float2* __restrict__ pout = bsel ? pos0 : pos1;
    pout[p] = (flags[p] & F_PINNED) ? pp : integrate(pp, v, d.friction, d.S);
}

constexpr int MAX_SWEEPS = 8;
constexpr float NOISE_FLOOR = 16.f;  // stalls below this many tolerances end the iteration
constexpr float FIX_ACCEPT = 4.f;    // largest accepted start correction, in tolerances
// A split point's fix-up correction d that follows a correction p is also
// accepted when d * d / p (the next correction if the rounds kept only the
// rate d / p they just showed; Newton's own rate, d ~ p^2, is faster) is
// within one tolerance and d within RATE_ACCEPT tolerances: the output then
// moves by d through the segment Jacobians and misses the solution by about
// that next correction.  Split points only: their thousands of entries make
// every spring step weak (bias = the neighbour's degree share), so their
// rounds converge at Newton's rate; the chunk sweeps of strongly coupled
// shorter chains keep the strict rule (measured: a rate rule there let an
// all-multiple-shooting run of a 200k-point graph reach 2.7e-5 S).
constexpr float RATE_ACCEPT = 64.f;
// One serial spring step from an entry's terms (no change when !live),
// also applying the step's Jacobian to J (column-major 2x2, J <- M J):
//   u' = u + x t(u),  x = q - u,  t = K (1 - D / max(|x|, 0.99 D))
//   M = du'/du = (1 - t) I - (K D / |x|^3) x x^T   (the last term is 0 when clamped)
__device__ inline void spring_step(float4 e, float2 pp, float2& v, float (&J)[4], bool live) {
    float x = e.x - (pp.x + v.x), y = e.y - (pp.y + v.y);
    float l0 = __builtin_amdgcn_sqrtf(x * x + y * y);
    bool clamp = l0 < e.z * 0.99f;
    float l = clamp ? e.z * 0.99f : l0;
    float r = __builtin_amdgcn_rcpf(l);
    float t = (1.f - e.z * r) * e.w;
    float w = clamp ? 0.f : e.w * e.z * r * r * r;
    t = live ? t : 0.f;
    w = live ? w : 0.f;
    v.x += x * t;
    v.y += y * t;
    float a = 1.f - t;
    float m00 = a - w * x * x, m01 = -w * x * y, m11 = a - w * y * y;
    float j0 = m00 * J[0] + m01 * J[1], j1 = m01 * J[0] + m11 * J[1];
    float j2 = m00 * J[2] + m01 * J[3], j3 = m01 * J[2] + m11 * J[3];
    J[0] = j0; J[1] = j1; J[2] = j2; J[3] = j3;
}

// Affine map s -> A s + b of a velocity (A column-major 2x2).
struct Aff {
    float a0, a1, a2, a3, bx, by;
};

__device__ inline Aff aff_identity() { return Aff{1.f, 0.f, 0.f, 1.f, 0.f, 0.f}; }

// f then g:  s -> g(f(s))
__device__ inline Aff aff_then(const Aff& f, const Aff& g) {
    return Aff{g.a0 * f.a0 + g.a2 * f.a1, g.a1 * f.a0 + g.a3 * f.a1,
               g.a0 * f.a2 + g.a2 * f.a3, g.a1 * f.a2 + g.a3 * f.a3,
               g.a0 * f.bx + g.a2 * f.by + g.bx, g.a1 * f.bx + g.a3 * f.by + g.by};
}

__device__ inline Aff aff_shfl_up(const Aff& m, int o) {
    return Aff{__shfl_up(m.a0, o, 64), __shfl_up(m.a1, o, 64), __shfl_up(m.a2, o, 64),
               __shfl_up(m.a3, o, 64), __shfl_up(m.bx, o, 64), __shfl_up(m.by, o, 64)};
}

// Scan of the lanes' chunk maps in lane order over a block of B lanes: returns
// lane t's exclusive prefix (maps 0..t-1 composed; identity for lane 0) and
// sets `total` to all B maps composed.  Shuffles inside each wave, then wave 0
// scans the wave totals: two barriers per call for B = 1024, none for 64.
template <int B>
__device__ inline Aff block_scan(Aff m, Aff* s_wave, Aff* s_wave_inc, Aff& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const Aff p = aff_shfl_up(m, o);
        if (lane >= o) m = aff_then(p, m);
    }
    Aff ex = aff_shfl_up(m, 1);
    if (lane == 0) ex = aff_identity();
    if (B == 64) {
        total = Aff{__shfl(m.a0, 63, 64), __shfl(m.a1, 63, 64), __shfl(m.a2, 63, 64),
                    __shfl(m.a3, 63, 64), __shfl(m.bx, 63, 64), __shfl(m.by, 63, 64)};
        return ex;
    }
    constexpr int NW = B / 64;
    if (lane == 63) s_wave[w] = m;
    __syncthreads();
    if (w == 0) {
        Aff a = lane < NW ? s_wave[lane] : aff_identity();
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const Aff p = aff_shfl_up(a, o);
            if (lane >= o) a = aff_then(p, a);
        }
        if (lane < NW) s_wave_inc[lane] = a;
    }
    __syncthreads();
    total = s_wave_inc[NW - 1];
    return w == 0 ? ex : aff_then(s_wave_inc[w - 1], ex);
}

constexpr uint32_t NO_HUB = 0xffffffffu;

// Chunked tiers: one workgroup of B lanes per work unit.  unit_a = {first
// entry, len, poff, graph}; unit_b = {slot, warm offset, split hub or NO_HUB,
// segment}.  Each lane runs its chunk of the unit's chain from a guessed start
// (warm: the start converged on the previous call), the block scans the chunk
// maps, and sweeps repeat until no start moves more than the tolerance; the
// last chunk's end, moved to its corrected start, is the unit's end velocity.
// The chain of a unit starts from the velocity in lane 0's warm slot: 0 for a
// point's first unit, the fix-up's current guess for a later segment of a
// split hub.  A whole point writes its integrated position; a segment writes
// {end, start used} and its end's Jacobian in its start for the fix-up.
// Round r > 0 reruns only the segments after the first of split hubs whose
// fix-up has not converged.
// Composition of split point hb's segments (hb = {first unit, segments}) by
// one full wave: each segment's linearised end map x -> E_s + A_s (x - G_s)
// (seg_out: {E, G used}, {A}), composed from the exact start 0, gives the true
// start of every segment,
//   true_{s+1} = E_s + A_s (true_s - G_s),
// and the point's velocity, the last segment's end from its true start.  A
// wave scan of affine maps over 64 segments at a time.  Returns the true start
// of segment `want`, the end, and the largest correction |true_s - G_s|;
// `store` also writes every true start into its segment's lane-0 warm slot.
struct Composed {
    float2 want_start, end;
    float dmax;
};

__device__ inline Composed hub_compose(uint2 hb, uint32_t lane, uint32_t want,
                                       const float4* __restrict__ seg_out, const uint4* __restrict__ unit_b,
                                       float2* __restrict__ warm, bool store) {
    float2 x0 = make_float2(0.f, 0.f);  // true start of the chunk's first segment
    float dmax = 0.f;
    float2 wanted = make_float2(0.f, 0.f);
    for (uint32_t s0 = 0; s0 < hb.y; s0 += 64) {
        const uint32_t s = s0 + lane;
        const bool have = s < hb.y;
        const size_t u = (size_t)hb.x + (have ? s : s0);
        const float4 o0 = seg_out[2 * u], A = seg_out[2 * u + 1];
        Aff m = have ? Aff{A.x, A.y, A.z, A.w, o0.x - (A.x * o0.z + A.z * o0.w), o0.y - (A.y * o0.z + A.w * o0.w)}
                     : aff_identity();
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const Aff q = aff_shfl_up(m, o);
            if ((int)lane >= o) m = aff_then(q, m);
        }
        Aff ex = aff_shfl_up(m, 1);
        if (lane == 0) ex = aff_identity();
        const float2 tru = make_float2(ex.a0 * x0.x + ex.a2 * x0.y + ex.bx, ex.a1 * x0.x + ex.a3 * x0.y + ex.by);
        if (have && s > 0) {
            dmax = fmaxf(dmax, fmaxf(fabsf(tru.x - o0.z), fabsf(tru.y - o0.w)));
            if (store) warm[unit_b[u].y] = tru;
        }
        if (want >= s0 && want < s0 + 64) {
            wanted.x = __shfl(tru.x, (int)(want - s0), 64);
            wanted.y = __shfl(tru.y, (int)(want - s0), 64);
        }
        const Aff all{__shfl(m.a0, 63, 64), __shfl(m.a1, 63, 64), __shfl(m.a2, 63, 64),
                      __shfl(m.a3, 63, 64), __shfl(m.bx, 63, 64), __shfl(m.by, 63, 64)};
        x0 = make_float2(all.a0 * x0.x + all.a2 * x0.y + all.bx, all.a1 * x0.x + all.a3 * x0.y + all.by);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dmax = fmaxf(dmax, __shfl_xor(dmax, o, 64));
    return Composed{wanted, x0, dmax};
}

// Whether split point h's segments hold after the round just evaluated: every
// start correction within FIX_ACCEPT tolerances (the output applies the last
// correction through the segment Jacobians, so its error is second order), a
// correction whose rate-predicted successor is within one tolerance
// (RATE_ACCEPT), or a stall within NOISE_FLOOR tolerances (the f32 noise floor;
// a stall far above is a slow Newton step, a spring crossing its clamp, and
// takes another round).  hub_done holds [n_hubs] marks (the round that found the point
// holding, 0 while open), then the largest correction of the evaluation of
// round e at [e & 1][h] ([2][n_hubs] f32).
__device__ inline bool hub_holds(float dmax, int evaluated, uint32_t h, uint32_t n_hubs,
                                 const uint32_t* __restrict__ hub_done, float tol) {
    const float* __restrict__ dp = reinterpret_cast<const float*>(hub_done + n_hubs);
    const float prev = evaluated > 0 ? dp[((evaluated - 1) & 1) * n_hubs + h] : INFINITY;
    return dmax <= FIX_ACCEPT * tol || (evaluated > 0 && dmax > 0.5f * prev && dmax <= NOISE_FLOOR * tol) ||
           (evaluated > 0 && dmax <= RATE_ACCEPT * tol && dmax * dmax <= tol * prev);
}

// The split point's velocity, integrated (or into vout), and the round counters.
__device__ inline void hub_finish(uint32_t p, float2 v, const GDesc& d, int pass, float2* __restrict__ pos0,
                                  float2* __restrict__ pos1, const uint8_t* __restrict__ flags,
                                  float2* __restrict__ vout, uint32_t* __restrict__ stats, int rounds, bool ok) {
    const bool bsel = (d.inbuf >> pass) & 1u;
    if (vout) {
        vout[p] = v;
    } else {
        const float2 pp = (bsel ? pos1 : pos0)[p];
        float2* __restrict__ pout = bsel ? pos0 : pos1;
        pout[p] = (flags[p] & F_PINNED) ? pp : integrate(pp, v, d.friction, d.S);
    }
    if (stats) {
        // rounds used, split points, max rounds, points stopped at the round cap
        atomicAdd(&stats[8], (uint32_t)rounds);
        atomicAdd(&stats[9], 1u);
        atomicMax(&stats[10], (uint32_t)rounds);
        if (!ok) atomicAdd(&stats[11], 1u);
    }
}

// Final fix-up after the last round, one wave per split point still open:
// composes its segments, stores their true starts as the next call's guesses
// and writes the point's velocity.
__global__ void __launch_bounds__(64) k_hub_final(int pass, const uint2* __restrict__ hubs, uint32_t n_hubs,
                                                  const uint4* __restrict__ unit_a,
                                                  const uint4* __restrict__ unit_b,
                                                  const uint32_t* __restrict__ sched,
                                                  const GDesc* __restrict__ desc, float2* __restrict__ pos0,
                                                  float2* __restrict__ pos1, const uint8_t* __restrict__ flags,
                                                  const float4* __restrict__ seg_out, float2* __restrict__ warm,
                                                  const uint32_t* __restrict__ hub_done,
                                                  uint32_t* __restrict__ stats, float2* __restrict__ vout) {
    const uint32_t h = blockIdx.x, lane = threadIdx.x;
    if (h >= n_hubs) return;
    const uint2 hb = hubs[h];
    const GDesc& d = desc[unit_a[hb.x].w];
    if (!(d.passes & (1u << pass))) return;
    if (hub_done[h]) return;
    const Composed f = hub_compose(hb, lane, 0u, seg_out, unit_b, warm, true);
    if (lane != 0) return;
    const bool ok = hub_holds(f.dmax, HUB_ROUNDS - 1, h, n_hubs, hub_done, d.tol);
    hub_finish(sched[unit_b[hb.x].x], f.end, d, pass, pos0, pos1, flags, vout, stats, HUB_ROUNDS, ok);
}

// Chunked tiers: one workgroup of B lanes per work unit of at most E B
// entries.  unit_a = {first entry, len, -, graph}; unit_b = {slot, warm offset,
// split point or NO_HUB, segment}.  Each lane gathers its chunk's terms
// (q.x, q.y, D, K) once into registers (D the rest length, K = linkSpring *
// alpha * strength * bias), runs the chunk from a guessed start (warm: the
// start converged on the previous call), the block scans the chunk maps, and
// sweeps repeat until no start moves more than the tolerance; the last
// chunk's end, moved to its corrected start, is the unit's end velocity.  The
// chain of a unit starts from the velocity in lane 0's warm slot: 0 for a
// point's first unit, the fix-up's current guess for a later segment of a
// split point.  A whole point writes its integrated position; a segment writes
// {end, start used} and its end's Jacobian for the fix-up.  Round r > 0
// reruns only the segments after the first of split points whose fix-up has
// not converged.
template <int B, int E>
__global__ void __launch_bounds__(B) k_link_chunk(int pass, const uint4* __restrict__ unit_a,
                                                  const uint4* __restrict__ unit_b, uint32_t u_first,
                                                  int round, const uint32_t* __restrict__ sched,
                                                  const GDesc* __restrict__ desc,
                                                  float2* __restrict__ pos0,
                                                  float2* __restrict__ pos1,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint4* __restrict__ ent,
                                                  float2* __restrict__ warm,
                                                  float4* __restrict__ wjac,
                                                  float4* __restrict__ seg_out, uint32_t n_units,
                                                  const uint2* __restrict__ hubs, uint32_t n_hubs,
                                                  uint32_t* __restrict__ hub_done,
                                                  uint32_t* __restrict__ stats,
                                                  float2* __restrict__ vout,
                                                  float4* __restrict__ terms) {
    __shared__ Aff s_wave[B / 64], s_wave_inc[B / 64];
    __shared__ float s_delta[B / 64];
    __shared__ float2 s_final;
    __shared__ float4 s_fix;
    const int t = threadIdx.x;
    // units ascend by chain length: the longest start first
    const uint32_t u = u_first + (gridDim.x - 1 - blockIdx.x);
    const uint4 U = unit_a[u], V = unit_b[u];
    const GDesc& d = desc[U.w];
    if (!(d.passes & (1u << pass))) return;  // uniform over the block
    const bool split = V.z != NO_HUB;
    // a point that held in an earlier round is done (the round that finds it
    // holding marks it with its number, and every block of that round still
    // stores its segment's start for the next call)
    if (round > 0 && (!split || V.w == 0 || (hub_done[V.z] != 0u && hub_done[V.z] != (uint32_t)round))) return;
    if (split && V.w == 0 && t == 0) hub_done[V.z] = 0u;  // read by the rounds after this launch
    const uint32_t p = sched[V.x];
    if (round > 0) {
        // the fix-up of the previous round, by every open segment block for
        // itself (the same composition in each): its own true start, and
        // whether the point already holds; segment 1's block then writes it
        if (t < 64) {
            // segment results by round parity: this round reads the last round's
            const float4* __restrict__ seg_rd = seg_out + (size_t)((round - 1) & 1) * 2 * n_units;
            const Composed f = hub_compose(hubs[V.z], (uint32_t)t, V.w, seg_rd, unit_b, warm, false);
            const bool ok = hub_holds(f.dmax, round - 1, V.z, n_hubs, hub_done, d.tol);
            if (t == 0) {
                s_fix = make_float4(f.want_start.x, f.want_start.y, ok ? 1.f : 0.f, 0.f);
                if (V.w == 1) {
                    reinterpret_cast<float*>(hub_done + n_hubs)[((round - 1) & 1) * n_hubs + V.z] = f.dmax;
                    if (ok) {
                        hub_done[V.z] = (uint32_t)round;
                        hub_finish(p, f.end, d, pass, pos0, pos1, flags, vout, stats, round, true);
                    }
                }
            }
        }
        __syncthreads();
        const float4 fx = s_fix;
        if (t == 0) warm[V.y] = make_float2(fx.x, fx.y);  // the segment's start: this round's, or the next call's
        if (fx.z != 0.f) return;
    }
    const bool bsel = (d.inbuf >> pass) & 1u;
    const float2* __restrict__ pin = bsel ? pos1 : pos0;
    float2* __restrict__ wslot = warm + V.y + t;
    const uint32_t len = U.y;
    const uint32_t c = (len + B - 1) / B;  // <= E by construction
    const uint32_t lo = min(len, (uint32_t)t * c), hi = min(len, lo + c);
    const uint32_t steps = hi - lo;
    const int last = (int)((len - 1) / c);
    const float2 pp = pin[p];
    // the lane's entries [lo, hi) of the unit, gathered once per pass: a split
    // point's segments keep them in `terms` (slot-major, lanes contiguous) for
    // the later rounds
    float4 reg[E];
    float4* __restrict__ tslot = (split && terms) ? terms + (size_t)(u - u_first) * E * B + t : nullptr;
    if (tslot && round > 0) {
#pragma unroll
        for (int s = 0; s < E; ++s) reg[s] = tslot[s * B];
    } else {
        uint4 en[E];
#pragma unroll
        for (int s = 0; s < E; ++s) en[s] = ent[U.x + min(lo + s, len - 1)];
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const float2 q = pin[en[s].x];
            const float r = __uint_as_float(en[s].y), sv = __uint_as_float(en[s].z),
                        bv = __uint_as_float(en[s].w);
            reg[s] = make_float4(q.x, q.y, (r * d.span + d.lo) * d.dist, d.spring_alpha * sv * bv);
        }
        if (tslot) {
#pragma unroll
            for (int s = 0; s < E; ++s) tslot[s * B] = reg[s];
        }
    }
    float2 g = *wslot;
    // the unit's start: lane 0's slot (for a later round, the fix-up just made)
    const float2 G = round > 0 ? make_float2(s_fix.x, s_fix.y) : warm[V.y];
    if (split && t > 0) {
        // the lane's start converged from the previous start G0 (kept in lane
        // 0's slot of wjac), moved to the current G along its prefix Jacobian
        // (both zero before the first run)
        const float4 P = wjac[V.y + t];
        const float4 G0 = wjac[V.y];
        const float dx = G.x - G0.x, dy = G.y - G0.y;
        g.x += P.x * dx + P.z * dy;
        g.y += P.y * dx + P.w * dy;
    }
    float prev = INFINITY;
    float J[4];
    float2 v;
    Aff tot;
    for (int sweep = 0;; ++sweep) {
        v = g;
        J[0] = 1.f; J[1] = 0.f; J[2] = 0.f; J[3] = 1.f;
#pragma unroll
        for (int s = 0; s < E; ++s) spring_step(reg[s], pp, v, J, (uint32_t)s < steps);
        // chunk map  s -> J (s - g) + v;  exclusive prefix applied to the unit's start
        const Aff m{J[0], J[1], J[2], J[3], v.x - (J[0] * g.x + J[2] * g.y), v.y - (J[1] * g.x + J[3] * g.y)};
        const Aff ex = block_scan<B>(m, s_wave, s_wave_inc, tot);
        const float2 start = make_float2(ex.a0 * G.x + ex.a2 * G.y + ex.bx, ex.a1 * G.x + ex.a3 * G.y + ex.by);
        float dlt = fmaxf(fabsf(start.x - g.x), fabsf(start.y - g.y));
        if (steps == 0) dlt = 0.f;
        // block max of the start corrections
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dlt = fmaxf(dlt, __shfl_xor(dlt, o, 64));
        float mx = dlt;
        if (B > 64) {
            if ((t & 63) == 0) s_delta[t >> 6] = dlt;
            __syncthreads();
            mx = s_delta[0];
            for (int w = 1; w < B / 64; ++w) mx = fmaxf(mx, s_delta[w]);
        }
        // converged (the output moves the last chunk's end through its
        // Jacobian, so corrections up to FIX_ACCEPT tolerances leave a second
        // order error), at the f32 noise floor (no longer halving, near the
        // tolerance), or at the cap
        if (mx <= FIX_ACCEPT * d.tol || (sweep >= 2 && mx > 0.5f * prev && mx <= NOISE_FLOOR * d.tol) ||
            sweep == MAX_SWEEPS) {
            // the last chunk's end, moved to its corrected start
            if (t == last) {
                float ex_ = start.x - g.x, ey = start.y - g.y;
                s_final = make_float2(v.x + J[0] * ex_ + J[2] * ey, v.y + J[1] * ex_ + J[3] * ey);
            }
            *wslot = start;
            if (split)
                wjac[V.y + t] = t == 0 ? make_float4(G.x, G.y, 0.f, 0.f) : make_float4(ex.a0, ex.a1, ex.a2, ex.a3);
            if (stats && t == 0) {
                // sweeps used, units, max sweeps, units stopped above the tolerance (per tier)
                const int o = B == 64 ? 0 : 4;
                atomicAdd(&stats[o + 0], (uint32_t)(sweep + 1));
                atomicAdd(&stats[o + 1], 1u);
                atomicMax(&stats[o + 2], (uint32_t)(sweep + 1));
                if (mx > d.tol) atomicAdd(&stats[o + 3], 1u);
            }
            break;
        }
        prev = mx;
        g = start;
        if (B > 64) __syncthreads();  // s_delta is rewritten next sweep
    }
    __syncthreads();
    if (t == 0) {
        const float2 vf = s_final;
        if (split) {
            // this round's buffer; a first segment (run in round 0 only) fills both
            for (int b = 0; b < 2; ++b) {
                if (b != (round & 1) && !(round == 0 && V.w == 0)) continue;
                float4* __restrict__ so = seg_out + (size_t)b * 2 * n_units;
                so[2 * (size_t)u] = make_float4(vf.x, vf.y, G.x, G.y);
                so[2 * (size_t)u + 1] = make_float4(tot.a0, tot.a1, tot.a2, tot.a3);
            }
        } else if (vout) {
            vout[p] = vf;
        } else {
            float2* __restrict__ pout = bsel ? pos0 : pos1;
            pout[p] = (flags[p] & F_PINNED) ? pp : integrate(pp, vf, d.friction, d.S);
        }
    }
}

// Segmented copies: segment k = {source offset (bit 31: from the second
// source), destination offset, count, first block}, blocks laid out segment
// after segment; every element goes to both destinations (a null second one
// is skipped).  One launch moves every graph's slice of a packed array.
template <typename T>
__global__ void k_seg_copy(const T* __restrict__ srcA, const T* __restrict__ srcB, T* __restrict__ dst0,
                           T* __restrict__ dst1, const uint4* __restrict__ seg, uint32_t nseg) {
    uint32_t lo = 0, hi = nseg;  // the last segment whose first block is <= blockIdx.x
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid].w <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const uint4 sg = seg[lo];
    const uint32_t i = (blockIdx.x - sg.w) * blockDim.x + threadIdx.x;
    if (i >= sg.z) return;
    const T v = ((sg.x >> 31) ? srcB : srcA)[(sg.x & 0x7fffffffu) + i];
    dst0[sg.y + i] = v;
    if (dst1) dst1[sg.y + i] = v;
}

}  // namespace

template <typename T>
int seg_copy(Ctx* c, const T* srcA, const T* srcB, T* dst0, T* dst1, std::vector<uint4>& segs) {
    uint32_t blk = 0;
    std::vector<uint4> kept;
    kept.reserve(segs.size());
    for (uint4 sg : segs) {
        if (sg.z == 0) continue;
        sg.w = blk;
        blk += blocks_for(sg.z, 256);
        kept.push_back(sg);
    }
    segs.swap(kept);
    if (segs.empty()) return CG_OK;
    int rc = CG_OK;
    uint4* d = (uint4*)dalloc(c, segs.size() * 16, &rc);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(d, segs.data(), segs.size() * 16, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_seg_copy<T>, dim3(blk), dim3(256), 0, c->stream, srcA, srcB, dst0, dst1, d,
                           (uint32_t)segs.size());
        e = hipGetLastError();
    }
    dfree(c, d);  // reused only once the stream is idle
    if (e != hipSuccess) return hip_fail(c, e, "segmented copy");
    return CG_OK;
}
template int seg_copy<float2>(Ctx*, const float2*, const float2*, float2*, float2*, std::vector<uint4>&);
template int seg_copy<float4>(Ctx*, const float4*, const float4*, float4*, float4*, std::vector<uint4>&);

void link_free_graph(Ctx* c, Graph& g) {
    for (int d = 0; d < 2; ++d) {
        dfree(c, g.lk.ent_pt[d]);
        dfree(c, g.lk.off_pt[d]);
        dfree(c, g.lk.live[d]);
        dfree(c, g.lk.lrank[d]);
    }
    dfree(c, g.lk.degree);
    g.lk = GraphLinks();
    g.links_ready = false;
    g.wn[0] = g.wn[1] = 0;
}

// Builds one graph's CSR in both directions (graph-local indices) from its
// host link copy: validation, degrees, the reference's entry order, bias,
// strength, rest-length draws, live counts.
int link_build_graph(Ctx* c, Graph& gr) {
    link_free_graph(c, gr);
    if (!gr.links_set || gr.n == 0) return CG_OK;
    const uint32_t n = gr.n, m = gr.m;
    int rc = CG_OK;
    // its own stream: reads only this graph's inputs and writes only its own
    // buffers, so it may run while ticks of the packed set are in flight
    hipStream_t st = c->ingest ? c->ingest : c->stream;
    GraphLinks& lk = gr.lk;
    lk.degree = (uint32_t*)dalloc(c, (size_t)n * 4, &rc);
    if (rc) return rc;
    std::vector<void*> tmp;
    auto talloc = [&](size_t bytes) -> void* {
        void* p = dalloc(c, bytes ? bytes : 16, &rc);
        if (p) tmp.push_back(p);
        return p;
    };
    auto cleanup = [&]() { for (void* p : tmp) dfree(c, p); };
    // after the stream has run: only this build used them
    auto release_idle = [&]() { for (void* p : tmp) dfree_idle(c, p); };
    const size_t mm = m ? m : 1;
    float2* d_links = (float2*)talloc(mm * 8);
    float* d_str = gr.strength.empty() ? nullptr : (float*)talloc(mm * 4);
    float* d_rin = (float*)talloc(mm * 4);
    float* d_rout = (float*)talloc(mm * 4);
    uint8_t* d_flags = (uint8_t*)talloc(n);
    uint32_t* src = (uint32_t*)talloc(mm * 4);
    uint32_t* tgt = (uint32_t*)talloc(mm * 4);
    uint32_t* outdeg = (uint32_t*)talloc(((size_t)n + 1) * 4);
    uint32_t* indeg = (uint32_t*)talloc(((size_t)n + 1) * 4);
    uint32_t* k0 = (uint32_t*)talloc(mm * 4);
    uint32_t* k1 = (uint32_t*)talloc(mm * 4);
    uint32_t* v0 = (uint32_t*)talloc(mm * 4);
    uint32_t* v1 = (uint32_t*)talloc(mm * 4);
    uint32_t* lflag = (uint32_t*)talloc((mm + 1) * 4);
    uint32_t* lrank = (uint32_t*)talloc((mm + 1) * 4);  // scan-size probe only
    if (rc) { cleanup(); return rc; }
    CG_HIP(c, hipMemcpyAsync(d_flags, gr.flags.data(), n, hipMemcpyHostToDevice, st));
    CG_HIP(c, hipMemsetAsync(outdeg, 0, ((size_t)n + 1) * 4, st));
    CG_HIP(c, hipMemsetAsync(indeg, 0, ((size_t)n + 1) * 4, st));
    CG_HIP(c, hipMemsetAsync(lk.degree, 0, (size_t)n * 4, st));
    if (m) {
        CG_HIP(c, hipMemcpyAsync(d_links, gr.links.data(), (size_t)m * 8, hipMemcpyHostToDevice, st));
        if (d_str)
            CG_HIP(c, hipMemcpyAsync(d_str, gr.strength.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        CG_HIP(c, hipMemcpyAsync(d_rin, gr.rin.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        CG_HIP(c, hipMemcpyAsync(d_rout, gr.rout.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_validate, dim3(blocks_for(m, 256)), dim3(256), 0, st, d_links, m,
                           n, src, tgt, outdeg, indeg);
        hipLaunchKernelGGL(k_degree, dim3(blocks_for(n, 256)), dim3(256), 0, st, outdeg, indeg,
                           n, lk.degree);
    }
    int nbits = 1;
    while (nbits < 32 && (1ull << nbits) <= (unsigned long long)n) ++nbits;  // n itself must fit
    size_t sort_bytes = 0, scan_bytes = 0, scan2_bytes = 0;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, v0, v1, (int)mm, 0,
                                                 nbits, st));
    CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, outdeg, outdeg, (int)n + 1,
                                               st));
    CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan2_bytes, lflag, lrank, (int)mm + 1,
                                               st));
    size_t tbytes = std::max(sort_bytes, std::max(scan_bytes, scan2_bytes));
    void* d_tmp = talloc(tbytes);
    if (rc) { cleanup(); return rc; }
    for (int d = 0; d < 2; ++d) {
        uint32_t* owner = d == 0 ? src : tgt;
        uint32_t* other = d == 0 ? tgt : src;
        uint32_t* cnt = d == 0 ? outdeg : indeg;  // entries owned per point
        lk.off_pt[d] = (uint32_t*)dalloc(c, ((size_t)n + 1) * 4, &rc);
        lk.live[d] = (uint32_t*)dalloc(c, (size_t)n * 4, &rc);
        if (rc) { cleanup(); return rc; }
        size_t sb = tbytes;
        CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(d_tmp, sb, cnt, lk.off_pt[d], (int)n + 1, st));
        uint32_t E = 0;
        CG_HIP(c, hipMemcpyAsync(&E, lk.off_pt[d] + n, 4, hipMemcpyDeviceToHost, st));
        CG_HIP(c, hipStreamSynchronize(st));
        lk.E[d] = E;
        lk.ent_pt[d] = (uint4*)dalloc(c, (size_t)(E ? E : 1) * 16, &rc);
        lk.lrank[d] = (uint32_t*)dalloc(c, ((size_t)E + 1) * 4, &rc);
        if (rc) { cleanup(); return rc; }
        CG_HIP(c, hipMemsetAsync(lflag, 0, (mm + 1) * 4, st));
        if (m) {
            hipLaunchKernelGGL(k_iota_keys, dim3(blocks_for(m, 256)), dim3(256), 0, st, owner,
                               m, k0, v0);
            sb = tbytes;
            CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(d_tmp, sb, k0, k1, v0, v1, (int)m, 0, nbits,
                                                         st));
            if (E)
                hipLaunchKernelGGL(k_entries, dim3(blocks_for(E, 256)), dim3(256), 0, st, k1,
                                   v1, E, other, lk.degree, d_str, d == 0 ? d_rin : d_rout, d_flags,
                                   lk.ent_pt[d], lflag);
        }
        sb = tbytes;
        CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(d_tmp, sb, lflag, lk.lrank[d], (int)E + 1, st));
        hipLaunchKernelGGL(k_live_count, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           lk.off_pt[d], lk.lrank[d], n, lk.live[d]);
        CG_HIP(c, hipGetLastError());
    }
    CG_HIP(c, hipStreamSynchronize(st));
    release_idle();  // the temporaries are reusable by the next build
    if (rc) return rc;
    gr.links_ready = true;
    return CG_OK;
}

// Packed schedules over all live graphs, from their per-graph CSRs.  Warm-start
// regions are laid out graph by graph, each graph's chunked points in index
// order, so their layout depends only on the graph; a graph whose links did
// not change carries its region over from the previous packing (old).
void link_free_dir(Ctx* c, LinkDir& L) {
    dfree(c, L.ent); dfree(c, L.sched); dfree(c, L.soff); dfree(c, L.sgraph); dfree(c, L.unit_a); dfree(c, L.unit_b);
    dfree(c, L.hubs); dfree(c, L.seg_out); dfree(c, L.hub_done); dfree(c, L.warm); dfree(c, L.wjac);
    dfree(c, L.terms);
    L = LinkDir();
}

// Builds the packed link schedules into c->dir, reading the warm starts of
// `old` (the previous packing's schedules, which the caller frees once the
// whole repack has succeeded, or keeps if it fails).
int link_pack(Ctx* c, const LinkDir (&old)[2]) {
    for (int d = 0; d < 2; ++d) c->dir[d] = LinkDir();
    c->any_links = false;
    const uint32_t G = c->G, NP = c->NP;
    for (uint32_t s = 0; s < G; ++s) {
        const Graph& g = c->graphs[c->order[s]];
        if (g.links_ready && g.m > 0) c->any_links = true;
    }
    if (!c->any_links || NP == 0) return CG_OK;
    int rc = CG_OK;
    std::vector<void*> tmp;
    auto talloc = [&](size_t bytes) -> void* {
        void* p = dalloc(c, bytes ? bytes : 16, &rc);
        if (p) tmp.push_back(p);
        return p;
    };
    auto cleanup = [&]() { for (void* p : tmp) dfree(c, p); };
    uint32_t* k0 = (uint32_t*)talloc(((size_t)NP + 1) * 4);
    uint32_t* k1 = (uint32_t*)talloc(((size_t)NP + 1) * 4);
    uint32_t* v0 = (uint32_t*)talloc(((size_t)NP + 1) * 4);
    uint32_t* counts = (uint32_t*)talloc(64);
    GLinkPtr* tab = (GLinkPtr*)talloc((size_t)G * sizeof(GLinkPtr));
    uint32_t* ebase = (uint32_t*)talloc(((size_t)G + 1) * 4);
    uint32_t* slot_of = (uint32_t*)talloc((size_t)(c->N ? c->N : 1) * 4);
    if (rc) { cleanup(); return rc; }
    size_t sort_bytes = 0, scan_bytes = 0;
    CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, v0, v0, (int)NP, 0, 32,
                                                 c->stream));
    CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, k1, k0, (int)NP + 1, c->stream));
    const size_t tbytes = std::max(sort_bytes, scan_bytes);
    void* d_tmp = talloc(tbytes);
    if (rc) { cleanup(); return rc; }
    std::vector<uint32_t> h_p0(G);
    for (uint32_t s = 0; s < G; ++s) h_p0[s] = c->graphs[c->order[s]].p0;
    for (int d = 0; d < 2; ++d) {
        LinkDir& L = c->dir[d];
        std::vector<GLinkPtr> h_tab(G);
        std::vector<uint32_t> h_ebase(G + 1, 0);
        for (uint32_t s = 0; s < G; ++s) {
            Graph& g = c->graphs[c->order[s]];
            if (g.links_ready && g.m > 0)
                h_tab[s] = {g.lk.ent_pt[d], g.lk.off_pt[d], g.lk.live[d], g.lk.lrank[d], g.n};
            else
                h_tab[s] = {nullptr, nullptr, nullptr, nullptr, 0u};
            h_ebase[s + 1] = h_ebase[s] + ((g.links_ready && g.m > 0) ? g.lk.E[d] : 0u);
        }
        CG_HIP(c, hipMemcpyAsync(ebase, h_ebase.data(), (size_t)(G + 1) * 4, hipMemcpyHostToDevice,
                                 c->stream));
        CG_HIP(c, hipMemcpyAsync(tab, h_tab.data(), (size_t)G * sizeof(GLinkPtr), hipMemcpyHostToDevice,
                                 c->stream));
        const uint32_t ns = NP;
        L.n_sched = ns;
        L.sched = (uint32_t*)dalloc(c, (size_t)ns * 4, &rc);
        L.soff = (uint32_t*)dalloc(c, ((size_t)ns + 1) * 4, &rc);
        if (rc) { cleanup(); return rc; }
        hipLaunchKernelGGL(k_sched_keys, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, c->present,
                           ns, c->gid, c->desc, tab, c->tune.lane_max, k0, v0);
        size_t sb = tbytes;
        CG_HIP(c, hipcub::DeviceRadixSort::SortPairs(d_tmp, sb, k0, k1, v0, L.sched, (int)ns, 0, 32,
                                                     c->stream));
        hipLaunchKernelGGL(k_sched_degree, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, L.sched, ns,
                           c->gid, c->desc, tab, k1);
        CG_HIP(c, hipMemsetAsync(k1 + ns, 0, 4, c->stream));
        sb = tbytes;
        CG_HIP(c, hipcub::DeviceScan::ExclusiveSum(d_tmp, sb, k1, L.soff, (int)ns + 1, c->stream));
        hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, c->stream, counts, 2u, ns);
        hipLaunchKernelGGL(k_tier_bounds, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, k1, ns,
                           c->tune.lane_max, c->tune.group_max, counts);
        uint32_t h[2] = {0, 0};
        CG_HIP(c, hipMemcpyAsync(h, counts, 8, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipMemcpyAsync(&L.E_live, L.soff + ns, 4, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
        L.n_lane = h[0];
        const uint32_t n64 = h[1] - h[0];
        // the chunked tiers' schedule, read back before the entry assembly is
        // queued, so the host builds their work units while it runs
        const uint32_t nch = ns - L.n_lane;
        std::vector<uint32_t> hs(nch + 1), hsched(nch ? nch : 1);
        CG_HIP(c, hipMemcpyAsync(hs.data(), L.soff + L.n_lane, (size_t)(nch + 1) * 4,
                                 hipMemcpyDeviceToHost, c->stream));
        if (nch)
            CG_HIP(c, hipMemcpyAsync(hsched.data(), L.sched + L.n_lane, (size_t)nch * 4,
                                     hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
        if (G > 1 && L.n_lane) {
            L.sgraph = (uint32_t*)dalloc(c, (size_t)L.n_lane * 4, &rc);
            if (rc) { cleanup(); return rc; }
            hipLaunchKernelGGL(k_sched_graph, dim3(blocks_for(L.n_lane, 256)), dim3(256), 0, c->stream, L.sched,
                               L.n_lane, c->gid, L.sgraph);
        }
        L.ent = (uint4*)dalloc(c, (size_t)(L.E_live ? L.E_live : 1) * 16, &rc);
        if (rc) { cleanup(); return rc; }
        hipLaunchKernelGGL(k_slot_of, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, L.sched, ns,
                           slot_of);
        const uint32_t etot = h_ebase[G];
        if (etot)
            hipLaunchKernelGGL(k_assemble, dim3(blocks_for(etot, 256)), dim3(256), 0, c->stream, ebase, G,
                               etot, tab, c->desc, slot_of, L.soff, c->flags, L.ent);
        // chunked tiers: work units, warm-start offsets.
        // A 64-lane point is one unit; a 256-lane point of d entries is
        // ceil(d / seg) near-equal segments of at most seg = 256 x hub_entries,
        // one unit each, in chain order.  Every unit holds at most 4 (64 lanes)
        // or hub_entries (256 lanes) entries per lane.
        const uint32_t hub_e = c->tune.hub_entries, hub_seg = 256u * hub_e;
        std::vector<uint4> ua, ub;
        std::vector<uint32_t> slot_u0(nch + 1, 0), slot_graph(nch ? nch : 1, 0), slot_lanes(nch ? nch : 1, 0);
        std::vector<uint2> hh;
        L.first_split_unit = 0;
        for (uint32_t k = 0; k < nch; ++k) {
            const uint32_t dd = hs[k + 1] - hs[k];
            const uint32_t B = k < n64 ? 64u : 256u;
            const uint32_t S = B == 256u ? (dd + hub_seg - 1) / hub_seg : 1u;
            const uint32_t p = hsched[k];
            const uint32_t gs = (uint32_t)(std::upper_bound(h_p0.begin(), h_p0.end(), p) - h_p0.begin()) - 1;
            slot_graph[k] = gs;
            slot_lanes[k] = B;
            slot_u0[k] = (uint32_t)ua.size();
            uint32_t hub_id = NO_HUB;
            if (S > 1) {
                if (hh.empty()) L.first_split_unit = (uint32_t)ua.size();
                hub_id = (uint32_t)hh.size();
                hh.push_back(make_uint2((uint32_t)ua.size(), S));
            }
            uint32_t eb = hs[k];
            for (uint32_t q = 0; q < S; ++q) {
                const uint32_t len = dd / S + (q < dd % S ? 1u : 0u);
                if (len == 0 || len > (B == 256u ? hub_e : 4u) * B) {
                    cleanup();
                    return fail(c, CG_ERR_INVALID, "link schedule: a work unit outside its tier's entries per lane");
                }
                ua.push_back(make_uint4(eb, len, 0u, gs));
                ub.push_back(make_uint4(L.n_lane + k, 0u, hub_id, q));
                eb += len;
            }
        }
        slot_u0[nch] = (uint32_t)ua.size();
        L.n_units = (uint32_t)ua.size();
        L.n_u64 = n64;
        L.n_hubs = (uint32_t)hh.size();
        if (hh.empty()) L.first_split_unit = L.n_units;
        // warm-start regions: per graph, its chunked points in local index order,
        // each point's units in chain order, B lanes each
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per(G);  // (local point, chunk slot)
        for (uint32_t k = 0; k < nch; ++k) per[slot_graph[k]].push_back({hsched[k] - h_p0[slot_graph[k]], k});
        size_t wpos = 0;
        std::vector<std::pair<size_t, size_t>> region(G);  // (offset, length) per graph slot
        for (uint32_t s = 0; s < G; ++s) {
            std::sort(per[s].begin(), per[s].end());
            size_t start = wpos;
            for (auto& pk : per[s]) {
                for (uint32_t u = slot_u0[pk.second]; u < slot_u0[pk.second + 1]; ++u) {
                    ub[u].y = (uint32_t)wpos;
                    wpos += slot_lanes[pk.second];
                }
            }
            region[s] = {start, wpos - start};
        }
        L.warm_len = wpos;
        const size_t nu = ua.size() ? ua.size() : 1;
        L.unit_a = (uint4*)dalloc(c, nu * 16, &rc);
        L.unit_b = (uint4*)dalloc(c, nu * 16, &rc);
        L.seg_out = (float4*)dalloc(c, nu * 64, &rc);  // two buffers by round parity
        L.hubs = (uint2*)dalloc(c, (hh.size() ? hh.size() : 1) * 8, &rc);
        L.hub_done = (uint32_t*)dalloc(c, (hh.size() ? hh.size() : 1) * 12, &rc);
        L.warm = (float2*)dalloc(c, (wpos ? wpos : 1) * 8, &rc);
        L.wjac = (float4*)dalloc(c, (wpos ? wpos : 1) * 16, &rc);
        if (!hh.empty())
            L.terms = (float4*)dalloc(c, (size_t)(L.n_units - L.first_split_unit) * hub_e * 256 * 16, &rc);
        if (rc) { cleanup(); return rc; }
        if (!ua.empty()) {
            CG_HIP(c, hipMemcpyAsync(L.unit_a, ua.data(), ua.size() * 16, hipMemcpyHostToDevice, c->stream));
            CG_HIP(c, hipMemcpyAsync(L.unit_b, ub.data(), ub.size() * 16, hipMemcpyHostToDevice, c->stream));
        }
        if (!hh.empty())
            CG_HIP(c, hipMemcpyAsync(L.hubs, hh.data(), hh.size() * 8, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipMemsetAsync(L.hub_done, 0, (hh.size() ? hh.size() : 1) * 12, c->stream));
        CG_HIP(c, hipMemsetAsync(L.warm, 0, (wpos ? wpos : 1) * 8, c->stream));
        CG_HIP(c, hipMemsetAsync(L.wjac, 0, (wpos ? wpos : 1) * 16, c->stream));
        CG_HIP(c, hipMemsetAsync(L.seg_out, 0, nu * 64, c->stream));
        // warm starts of graphs whose regions kept their size, in one launch per array
        std::vector<uint4> wsegs;
        for (uint32_t s = 0; s < G; ++s) {
            Graph& g = c->graphs[c->order[s]];
            if (old[d].warm && g.wn[d] && g.wn[d] == region[s].second)
                wsegs.push_back(make_uint4(g.w0[d], (uint32_t)region[s].first, g.wn[d], 0u));
            g.w0[d] = (uint32_t)region[s].first;
            g.wn[d] = (uint32_t)region[s].second;
        }
        std::vector<uint4> jsegs = wsegs;
        rc = seg_copy<float2>(c, old[d].warm, nullptr, L.warm, nullptr, wsegs);
        if (!rc) rc = seg_copy<float4>(c, old[d].wjac, nullptr, L.wjac, nullptr, jsegs);
        if (rc) { cleanup(); return rc; }
        CG_HIP(c, hipGetLastError());
        // keep the staging vectors alive until the copies above have run
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (!c->link_stats && !rc) {
        c->link_stats = (uint32_t*)dalloc(c, 64, &rc);
        if (!rc) CG_HIP(c, hipMemsetAsync(c->link_stats, 0, 64, c->stream));
    }
    cleanup();
    return rc;
}

template <int HE>
int link_run_e(Ctx* c, int dir, float2* vout) {
    if (!c->any_links) return CG_OK;
    const LinkDir& L = c->dir[dir];
    const int pass = P_LINK_IN + dir;
    // The split points' rounds (each after the first opening with the fix-up
    // of the one before) and the final fix-up run on the side stream while the
    // whole 256-lane points, the 64-lane and the lane tiers fill the GPU on
    // the main stream, which joins the side stream at the end.  Rounds after a
    // point holds return at once.
    const uint32_t u256 = L.n_u64, us = L.first_split_unit;
    // sweep counters only while requested (same-address atomics per block)
    uint32_t* stats = (c->prof_mask & CG_PROFILE_LINK_STATS) ? c->link_stats : nullptr;
    const bool fork = L.n_hubs && c->side;
    if (L.n_hubs) {
        hipStream_t s = c->stream;
        if (fork) {
            CG_HIP(c, hipEventRecord(c->ev_fork, c->stream));
            CG_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
            s = c->side;
        }
        hipLaunchKernelGGL((k_link_chunk<256, HE>), dim3(L.n_units - us), dim3(256), 0, s, pass, L.unit_a,
                           L.unit_b, us, 0, L.sched, c->desc, c->pos[0], c->pos[1], c->flags, L.ent, L.warm,
                           L.wjac, L.seg_out, L.n_units, L.hubs, L.n_hubs, L.hub_done, stats, vout, L.terms);
        for (int r = 1; r < HUB_ROUNDS; ++r)
            hipLaunchKernelGGL((k_link_chunk<256, HE>), dim3(L.n_units - us), dim3(256), 0, s, pass, L.unit_a, L.unit_b,
                               us, r, L.sched, c->desc, c->pos[0], c->pos[1], c->flags, L.ent, L.warm, L.wjac,
                               L.seg_out, L.n_units, L.hubs, L.n_hubs, L.hub_done, stats, vout, L.terms);
        hipLaunchKernelGGL(k_hub_final, dim3(L.n_hubs), dim3(64), 0, s, pass, L.hubs, L.n_hubs, L.unit_a,
                           L.unit_b, L.sched, c->desc, c->pos[0], c->pos[1], c->flags,
                           L.seg_out + (size_t)((HUB_ROUNDS - 1) & 1) * 2 * L.n_units, L.warm, L.hub_done, stats,
                           vout);
        if (fork) CG_HIP(c, hipEventRecord(c->ev_join, c->side));
    }
    if (us > u256)
        hipLaunchKernelGGL((k_link_chunk<256, HE>), dim3(us - u256), dim3(256), 0, c->stream, pass, L.unit_a,
                           L.unit_b, u256, 0, L.sched, c->desc, c->pos[0], c->pos[1], c->flags, L.ent, L.warm,
                           L.wjac, L.seg_out, L.n_units, L.hubs, L.n_hubs, L.hub_done, stats, vout, nullptr);
    if (L.n_u64)
        hipLaunchKernelGGL((k_link_chunk<64, 4>), dim3(L.n_u64), dim3(64), 0, c->stream, pass, L.unit_a, L.unit_b,
                           0u, 0, L.sched, c->desc, c->pos[0], c->pos[1], c->flags, L.ent, L.warm, L.wjac,
                           L.seg_out, L.n_units, L.hubs, L.n_hubs, L.hub_done, stats, vout, nullptr);
    if (L.n_lane)
        hipLaunchKernelGGL(k_link_lane, dim3(blocks_for(L.n_lane, 256)), dim3(256), 0, c->stream, pass,
                           L.sched, L.soff, L.sgraph, L.ent, L.n_lane, c->desc, c->pos[0], c->pos[1],
                           c->flags, vout, c->G == 1 ? 1u : 0u);
    if (fork) CG_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

int link_run(Ctx* c, int dir, float2* vout) {
    switch (c->tune.hub_entries) {
        case 4: return link_run_e<4>(c, dir, vout);
        case 8: return link_run_e<8>(c, dir, vout);
        default: return link_run_e<6>(c, dir, vout);
    }
}

}  // namespace cg
