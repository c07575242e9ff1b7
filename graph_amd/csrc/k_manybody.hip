// k_manybody.hip — grid-pyramid many-body repulsion with the Monte-Carlo
// near field, as one deterministic pipeline per tick:
//
//   1. counting sort of the present points by finest cell, visiting them in
//      the previous pass's cell order, positions riding along; cell bounds
//      from the scan of the per-cell counts
//   2. per finest cell (a run of equal keys in the sorted points), over the
//      points rather than the cells: fixed-point sums of (x, y), count, and
//      the top-8 near-field slots by (ticket, index) — one wave per 64 N
//      sorted points, runs inside a lane finished by it, runs across lanes by
//      segmented scans, runs across waves by a merge of the waves' partials
//   3. 2x2 merges up the pyramid in the same fixed point, f32 centroids
//      (cx, cy, N) per level
//   4. per point (in cell order): far field over every level + near field,
//      fused with the integrate pass
//
// Within a cell, the points' order follows the order in which the counting
// sort's block atomics land; nothing depends on it.  Sums are integer (positions as multiples of 2^-k, k chosen per graph
// so no sum overflows), so they are exact and order-free, and the slot choice
// is a set function (the 8 smallest distinct tickets, the smallest index per
// ticket).  No float atomics anywhere: a graph gives the same bits alone or
// packed with others.
//
// Reference: ForceManyBody/index.ts:399-417 (run), :464-502 (drawLevels),
// :511-559 (drawNearFieldSlots), :565-638 (drawForces), :646-694 (geometry);
// calculate-level.vert:28-53; build-nearfield-slots.vert:46-103;
// force-level.frag:43-98; force-nearfield.frag:64-166.
#include <algorithm>

#include "cg_internal.h"

namespace cg {
namespace {

// ---- 1. counting sort by finest cell ---------------------------------------------

// Finest cell of position q: calculate-level.vert:47-48 at the finest level,
// as a Morton index inside the graph's cell range.
__device__ inline uint32_t finest_key(const GDesc& d, float2 q) {
    const float h = d.h[d.L - 1];
    uint32_t cx = (uint32_t)cell_of(q.x, h, (int)d.gf);
    uint32_t cy = (uint32_t)cell_of(q.y, h, (int)d.gf);
    return d.f0 + morton2(cx, cy);
}

// LSD radix sort of (cell key, point index) — the collision pass's sort, whose
// cell sums need points in index order inside a cell — RB_ITEMS keys per lane, digits of
// `width` <= RB_MAXW bits.  Per pass: block digit histograms (LDS atomics only),
// an exclusive scan of the [digit][block] table, then a stable scatter whose
// in-block ranks come from ballot matching in position order.  No global
// atomics: on this part they serialise at a few ns each.
constexpr int RB_ITEMS = 8;
constexpr int RB_TILE = 256 * RB_ITEMS;
static_assert(RB_TILE == (int)RADIX_TILE, "segment tables are built with RADIX_TILE");
constexpr int RB_MAXW = 9;               // digits of up to 9 bits: 18-bit cell keys in two passes
constexpr int RB_RADIX = 1 << RB_MAXW;

// Block b's positions [*bb, *be) and key offset (see RadixSort::blk).
__device__ inline void radix_block(const uint4* __restrict__ blk, const uint32_t* __restrict__ key0,
                                   uint32_t np, uint32_t& bb, uint32_t& be, uint32_t& k0) {
    if (blk) {
        const uint4 q = blk[blockIdx.x];
        bb = q.x;
        be = q.y;
        k0 = key0[blockIdx.x];
    } else {
        bb = blockIdx.x * RB_TILE;
        be = min(np, bb + RB_TILE);
        k0 = 0u;
    }
}

// The cells occupied at the previous pass (the run heads of its sorted keys,
// still in `old`): each one's end is cleared, so a cell left empty reads 0,
// and its block of CELL_BLOCK cells is marked dirty, so the pyramid rewrites
// its tile even when the tile is empty now.  Only these cells can hold a stale
// end: every other one has been 0 since the packing was built.
__global__ void __launch_bounds__(256) k_mb_clear(const uint32_t* __restrict__ old, uint32_t np,
                                                  uint32_t* __restrict__ cell_end,
                                                  uint8_t* __restrict__ cell_dirty,
                                                  uint32_t* __restrict__ cell_count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t key = old[k];
    if (k == 0 || old[k - 1] != key) {
        cell_end[key] = 0u;
        cell_count[key] = 0u;
        cell_dirty[key / CELL_BLOCK] = 1u;
    }
}

// ---- counting sort by finest cell -----------------------------------------------
//
// The keys are finest cells (at most 2^18 per graph) and repeat heavily (a
// layout collapses into clumps of up to tens of thousands of points per
// cell), so the points are placed by a counting sort: (1) each point's key
// and its rank among its cell's points; (2) an exclusive scan of the per-cell
// counts gives every cell's first position (its begin; end = begin + count);
// (3) each point goes to begin + rank.  A block of CB_TILE keys aggregates
// them per cell in an LDS hash table (ranks inside the block from LDS
// atomics) and adds each cell's block count to the cell's global count with
// one atomic, whose old value is the block's base rank.  So the order of the
// points inside a cell follows the order in which those atomics land and
// varies from run to run; nothing computed from the cells depends on it
// (integer sums, the set-valued slot rule, one lane per point in the force).
constexpr int CB_ITEMS = 8;                  // keys per lane
constexpr uint32_t CB_TILE = 256 * CB_ITEMS; // keys per block
constexpr uint32_t CB_HASH = 4096;           // LDS hash slots (at most half of them used)
constexpr uint32_t CB_EMPTY = 0xffffffffu;
constexpr uint32_t SCAN_CPT_MIN = 4;        // cells per thread of the count scan: 4 ..
constexpr uint32_t SCAN_CPT_MAX = 256;      // .. 256, so that at most ~4096 blocks cover the cells
constexpr uint32_t COUNT_OWN_PREFIX = 1024; // ranges up to which the bounds blocks sum their own prefix

// Keys of the points in the previous pass's cell order (positions as payload)
// and their ranks in their cells.
__global__ void __launch_bounds__(256) k_bin_count(const uint32_t* __restrict__ visit,
                                                   const uint32_t* __restrict__ old_keys,
                                                   const float2* __restrict__ pos0,
                                                   const float2* __restrict__ pos1,
                                                   const uint32_t* __restrict__ blk_graph,
                                                   const GDesc* __restrict__ desc, uint32_t np,
                                                   uint32_t* __restrict__ cell_count,
                                                   uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                   float2* __restrict__ pout, uint32_t* __restrict__ rank,
                                                   uint8_t* __restrict__ cell_occ, uint8_t* __restrict__ cell_dirty,
                                                   uint32_t one_graph) {
    __shared__ uint32_t hk[CB_HASH], hc[CB_HASH];
    const int t = threadIdx.x;
    const uint32_t bb = blockIdx.x * CB_TILE;
    for (uint32_t q = t; q < CB_HASH; q += 256) {
        hk[q] = CB_EMPTY;
        hc[q] = 0u;
    }
    // the lane's items go through each dependent load together (index and
    // graph, position); a point's graph is that of its previous key (read in
    // the same order as the visit), a set of one graph skips it
    uint32_t ii[CB_ITEMS], gg[CB_ITEMS];
    float2 qq[CB_ITEMS];
#pragma unroll
    for (int j = 0; j < CB_ITEMS; ++j) ii[j] = visit[min(bb + j * 256 + t, np - 1)];
#pragma unroll
    for (int j = 0; j < CB_ITEMS; ++j)
        gg[j] = one_graph ? 0u : blk_graph[old_keys[min(bb + j * 256 + t, np - 1)] / CELL_BLOCK];
#pragma unroll
    for (int j = 0; j < CB_ITEMS; ++j) {
        const GDesc& d = desc[gg[j]];
        qq[j] = ((d.inbuf >> P_MANYBODY) & 1u) ? pos1[ii[j]] : pos0[ii[j]];
    }
    __syncthreads();
    const int lane = t & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t key[CB_ITEMS], slot[CB_ITEMS], r[CB_ITEMS];
#pragma unroll
    for (int j = 0; j < CB_ITEMS; ++j) {
        key[j] = finest_key(desc[gg[j]], qq[j]);
        slot[j] = 0u;
        r[j] = 0u;
        bool pending = bb + j * 256 + t < np;
        // lanes holding the wave's first pending key go together: one of them
        // finds the slot and adds their count (points visited in the previous
        // cell order share few keys per wave); after two such rounds the rest
        // go one by one (1, 2, 4, 8 and 16 rounds measured: 1-2 fastest)
        for (int round = 0; round < 2; ++round) {
            const unsigned long long pm = __ballot(pending);
            if (pm == 0ull) break;
            const int first = (int)__ffsll((long long)pm) - 1;
            const uint32_t lead = __shfl(key[j], first, 64);
            const bool mine = pending && key[j] == lead;
            const unsigned long long m = __ballot(mine);
            uint32_t h = 0u, base = 0u;
            if (lane == first) {
                // open addressing: at most CB_TILE keys in CB_HASH slots, so a
                // free slot or the key's own is always found
                h = (lead * 2654435761u) >> 20;
                for (;;) {
                    const uint32_t o = atomicCAS(&hk[h], CB_EMPTY, lead);
                    if (o == CB_EMPTY || o == lead) break;
                    h = (h + 1u) & (CB_HASH - 1u);
                }
                base = atomicAdd(&hc[h], (uint32_t)__popcll(m));
            }
            h = __shfl(h, first, 64);
            base = __shfl(base, first, 64);
            if (mine) {
                slot[j] = h;
                r[j] = base + (uint32_t)__popcll(m & below);
                pending = false;
            }
        }
        if (pending) {
            uint32_t h = (key[j] * 2654435761u) >> 20;
            for (;;) {
                const uint32_t o = atomicCAS(&hk[h], CB_EMPTY, key[j]);
                if (o == CB_EMPTY || o == key[j]) break;
                h = (h + 1u) & (CB_HASH - 1u);
            }
            slot[j] = h;
            r[j] = atomicAdd(&hc[h], 1u);
        }
    }
    __syncthreads();
    // one global atomic per cell of the block: its old value is the block's
    // base; the cell's block of CELL_BLOCK cells is marked occupied and dirty.
    // A lane's slots go out together (its atomics in flight at once).
    constexpr int HS = (int)(CB_HASH / 256);
    uint32_t hkk[HS], hcc[HS];
#pragma unroll
    for (int u = 0; u < HS; ++u) {
        hkk[u] = hk[t + u * 256];
        hcc[u] = hc[t + u * 256];
    }
#pragma unroll
    for (int u = 0; u < HS; ++u)
        if (hkk[u] != CB_EMPTY) hcc[u] = atomicAdd(&cell_count[hkk[u]], hcc[u]);
#pragma unroll
    for (int u = 0; u < HS; ++u)
        if (hkk[u] != CB_EMPTY) {
            hc[t + u * 256] = hcc[u];
            cell_occ[hkk[u] / CELL_BLOCK] = 1u;
            cell_dirty[hkk[u] / CELL_BLOCK] = 1u;
        }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CB_ITEMS; ++j) {
        const uint32_t k = bb + j * 256 + t;
        if (k < np) {
            kout[k] = key[j];
            vout[k] = ii[j];
            pout[k] = qq[j];
            rank[k] = hc[slot[j]] + r[j];
        }
    }
}

// Keys standing in for a pass before the first: each point's graph's first
// finest cell, in the first pass's visiting order (the present list).
__global__ void k_first_keys(const uint32_t* __restrict__ present, uint32_t np, const uint32_t* __restrict__ gid,
                             const uint32_t* __restrict__ f0s, uint32_t* __restrict__ keys) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < np) keys[k] = f0s[gid[present[k]]];
}

// Block-exclusive scan of one value per thread over 256 threads (4 waves).
__device__ inline uint32_t block_exscan256(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    const uint32_t before = (w > 0 ? s_w[0] : 0u) + (w > 1 ? s_w[1] : 0u) + (w > 2 ? s_w[2] : 0u);
    total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    return before + inc - v;
}

// Count scan, step 1: the points of each range of 256 cpt cells (cpt cells
// per thread, cpt a power of two in [4, 256]; F is a multiple of CELL_BLOCK).
// A thread's range lies inside one block of CELL_BLOCK cells, or covers whole
// ones, and unoccupied blocks count 0 without being read.
__global__ void __launch_bounds__(256) k_count_sums(const uint32_t* __restrict__ cell_count, uint32_t F,
                                                    uint32_t cpt, const uint8_t* __restrict__ cell_occ,
                                                    uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[4];
    const uint32_t c0 = (blockIdx.x * 256u + threadIdx.x) * cpt;
    uint32_t v = 0u;
    for (uint32_t cb = c0; cb < min(F, c0 + cpt); cb += min(cpt, CELL_BLOCK)) {
        if (!cell_occ[cb / CELL_BLOCK]) continue;
        const uint4* __restrict__ p = reinterpret_cast<const uint4*>(cell_count + cb);
        const uint32_t nq = min(cpt, CELL_BLOCK) / 4u;
        for (uint32_t q = 0; q < nq; ++q) {
            const uint4 x = p[q];
            v += x.x + x.y + x.z + x.w;
        }
    }
    uint32_t total;
    (void)block_exscan256(v, s_w, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Count scan, step 2 (one block): exclusive prefixes of the range sums, in place.
__global__ void __launch_bounds__(256) k_count_scan(uint32_t* __restrict__ sums, uint32_t nr) {
    __shared__ uint32_t s_w[4];
    uint32_t carry = 0u;
    for (uint32_t base = 0; base < nr; base += 256 * 4) {
        const uint32_t i0 = base + 4u * threadIdx.x;
        uint32_t v[4], sum = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = i0 + q < nr ? sums[i0 + q] : 0u;
            sum += v[q];
        }
        uint32_t total;
        uint32_t run = carry + block_exscan256(sum, s_w, total);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (i0 + q < nr) sums[i0 + q] = run;
            run += v[q];
        }
        carry += total;
        __syncthreads();  // s_w is rewritten next round
    }
}

// Count scan, step 3: every occupied cell's bounds (begin = the points of the
// cells before it, end = begin + count).  Empty cells keep end 0 (k_mb_clear
// cleared the last pass's cells); unoccupied blocks of cells are not read.
// With `own_prefix` (few ranges) each block adds up the range sums before its
// own, and step 2 is not launched.
__global__ void __launch_bounds__(256) k_count_bounds(const uint32_t* __restrict__ cell_count, uint32_t F,
                                                      uint32_t cpt, const uint8_t* __restrict__ cell_occ,
                                                      const uint32_t* __restrict__ sums, uint32_t own_prefix,
                                                      uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
    __shared__ uint32_t s_w[4];
    uint32_t base;
    if (own_prefix) {
        uint32_t part = 0u;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += 256u) part += sums[b];
        (void)block_exscan256(part, s_w, base);
        __syncthreads();  // s_w is rewritten below
    } else {
        base = sums[blockIdx.x];
    }
    const uint32_t c0 = (blockIdx.x * 256u + threadIdx.x) * cpt;
    const uint32_t c1 = min(F, c0 + cpt), step = min(cpt, CELL_BLOCK);
    uint32_t v = 0u;
    for (uint32_t cb = c0; cb < c1; cb += step) {
        if (!cell_occ[cb / CELL_BLOCK]) continue;
        const uint4* __restrict__ p = reinterpret_cast<const uint4*>(cell_count + cb);
        for (uint32_t q = 0; q < step / 4u; ++q) {
            const uint4 x = p[q];
            v += x.x + x.y + x.z + x.w;
        }
    }
    uint32_t total;
    uint32_t run = base + block_exscan256(v, s_w, total);
    for (uint32_t cb = c0; cb < c1; cb += step) {
        if (!cell_occ[cb / CELL_BLOCK]) continue;
        const uint4* __restrict__ p = reinterpret_cast<const uint4*>(cell_count + cb);
        for (uint32_t q = 0; q < step / 4u; ++q) {
            const uint4 x = p[q];
            const uint32_t cs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (cs[u]) {
                    const uint32_t f = cb + 4u * q + u;
                    begin[f] = run;
                    end[f] = run + cs[u];
                }
                run += cs[u];
            }
        }
    }
}

// Counting sort, step 4: each point to its cell's begin plus its rank, with
// its key, index and position.
__global__ void __launch_bounds__(256) k_bin_place(const uint32_t* __restrict__ kin,
                                                   const uint32_t* __restrict__ vin,
                                                   const float2* __restrict__ pin,
                                                   const uint32_t* __restrict__ rank, uint32_t np,
                                                   const uint32_t* __restrict__ begin,
                                                   uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                   float2* __restrict__ pout) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t key = kin[k];
    const uint32_t at = begin[key] + rank[k];
    kout[at] = key;
    vout[at] = vin[k];
    pout[at] = pin[k];
}

__global__ void __launch_bounds__(256) k_radix_up(const uint32_t* __restrict__ kin, uint32_t np, int shift,
                                                  int width, uint32_t nb, uint32_t* __restrict__ hist,
                                                  const uint4* __restrict__ blk,
                                                  const uint32_t* __restrict__ key0) {
    __shared__ uint32_t h[RB_RADIX];
    const int t = threadIdx.x;
    const uint32_t radix = 1u << width, mask = radix - 1;
    uint32_t bb, be, k0;
    radix_block(blk, key0, np, bb, be, k0);
    for (uint32_t q = t; q < radix; q += 256) h[q] = 0;
    __syncthreads();
    uint32_t keys[RB_ITEMS];
#pragma unroll
    for (int j = 0; j < RB_ITEMS; ++j) keys[j] = kin[min(bb + j * 256 + t, np - 1)];
#pragma unroll
    for (int j = 0; j < RB_ITEMS; ++j) {
        const uint32_t k = bb + j * 256 + t;
        if (k < be) atomicAdd(&h[((keys[j] - k0) >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t q = t; q < radix; q += 256) hist[q * nb + blockIdx.x] = h[q];
}

__global__ void __launch_bounds__(256) k_radix_down(const uint32_t* __restrict__ kin,
                                                    const uint32_t* __restrict__ vin, uint32_t np,
                                                    int shift, int width, uint32_t nb,
                                                    const uint32_t* __restrict__ hist,
                                                    const uint32_t* __restrict__ totals,
                                                    uint32_t* __restrict__ kout,
                                                    uint32_t* __restrict__ vout,
                                                    const float2* __restrict__ pin,
                                                    float2* __restrict__ pout,
                                                    const uint4* __restrict__ blk,
                                                    const uint32_t* __restrict__ key0) {
    __shared__ uint32_t base[RB_RADIX];
    __shared__ uint32_t wc[4][RB_RADIX];
    __shared__ uint32_t sw[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t radix = 1u << width, mask = radix - 1;
    const unsigned long long lt = (1ull << lane) - 1ull;
    uint32_t bb, be, k0;
    radix_block(blk, key0, np, bb, be, k0);
    // the block's segment: blocks [fb, lb) from output position `start`
    const uint4 seg = blk ? blk[blockIdx.x] : make_uint4(0u, 0u, 0u, nb);
    const uint32_t fb = seg.z, lb = seg.w, start = blk ? blk[fb].x : 0u;
    {
        // the digits' offsets: exclusive scan of the segment's digit totals
        // (2 digits per lane; the rows hold exclusive prefixes over all blocks,
        // so a segment's count is a difference), plus this block's offset
        // within its segment's row
        const uint32_t d0 = 2 * t, d1 = 2 * t + 1;
        uint32_t v0 = 0u, v1 = 0u, o0 = 0u, o1 = 0u;
        if (d0 < radix) {
            const uint32_t f = blk ? hist[d0 * nb + fb] : 0u;
            v0 = (lb < nb ? hist[d0 * nb + lb] : totals[d0]) - f;
            o0 = hist[d0 * nb + blockIdx.x] - f;
        }
        if (d1 < radix) {
            const uint32_t f = blk ? hist[d1 * nb + fb] : 0u;
            v1 = (lb < nb ? hist[d1 * nb + lb] : totals[d1]) - f;
            o1 = hist[d1 * nb + blockIdx.x] - f;
        }
        uint32_t inc = v0 + v1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o, 64);
            if (lane >= o) inc += u;
        }
        if (lane == 63) sw[wave] = inc;
        __syncthreads();
        const uint32_t before = (wave > 0 ? sw[0] : 0u) + (wave > 1 ? sw[1] : 0u) + (wave > 2 ? sw[2] : 0u);
        const uint32_t ex = start + before + inc - v0 - v1;
        if (d0 < radix) base[d0] = ex + o0;
        if (d1 < radix) base[d1] = ex + v0 + o1;
    }
    // all of the lane's items in flight before the ranking rounds
    uint32_t keys[RB_ITEMS], vals[RB_ITEMS];
    float2 pays[RB_ITEMS];
#pragma unroll
    for (int j = 0; j < RB_ITEMS; ++j) {
        const uint32_t k = min(bb + j * 256 + t, np - 1);
        keys[j] = kin[k];
        vals[j] = vin[k];
        pays[j] = pin ? pin[k] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < RB_ITEMS; ++j) {
        const uint32_t k = bb + j * 256 + t;
        const bool valid = k < be;
        const uint32_t key = valid ? keys[j] : 0u;
        const uint32_t val = vals[j];
        const float2 pay = pays[j];
        const uint32_t dg = ((key - k0) >> shift) & mask;
        for (uint32_t q = t; q < radix; q += 256) {
#pragma unroll
            for (int w = 0; w < 4; ++w) wc[w][q] = 0u;
        }
        __syncthreads();
        // lanes of this wave holding the same digit, in position order
        unsigned long long m = __ballot(valid);
        for (int bit = 0; bit < width; ++bit) {
            const bool on_ = (dg >> bit) & 1u;
            const unsigned long long bb = __ballot(on_);
            m &= on_ ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(m & lt);
        if (valid && below == 0) wc[wave][dg] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t off = base[dg] + below;
            for (int w = 0; w < wave; ++w) off += wc[w][dg];
            kout[off] = key;
            vout[off] = val;
            if (pout) pout[off] = pay;
        }
        __syncthreads();
        for (uint32_t q = t; q < radix; q += 256) base[q] += wc[0][q] + wc[1][q] + wc[2][q] + wc[3][q];
        __syncthreads();
    }
}

// Exclusive scan of each digit's row of the [digit][block] histogram table in
// place, one block per digit (rows are scanned 1024 counts at a time with the
// carry), and the digit's total.  The digits' own offsets (an exclusive scan
// of the totals) are added by the scatter kernel, so one launch serves a pass.
__global__ void __launch_bounds__(256) k_scan_rows(uint32_t* __restrict__ hist, uint32_t nb,
                                                   uint32_t* __restrict__ totals) {
    __shared__ uint32_t sw[4];
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t* __restrict__ row = hist + (size_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i0 = base + 4 * t;
        uint32_t v[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = i0 + q < nb ? row[i0 + q] : 0u;
            s += v[q];
        }
        uint32_t inc = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += u;
        }
        if (lane == 63) sw[w] = inc;
        __syncthreads();
        const uint32_t before = (w > 0 ? sw[0] : 0u) + (w > 1 ? sw[1] : 0u) + (w > 2 ? sw[2] : 0u);
        const uint32_t total = sw[0] + sw[1] + sw[2] + sw[3];
        uint32_t run = carry + before + inc - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (i0 + q < nb) row[i0 + q] = run;
            run += v[q];
        }
        carry += total;
        __syncthreads();  // sw is rewritten by the next 1024 counts
    }
    if (t == 0) totals[blockIdx.x] = carry;
}

// ---- 2. finest cells over runs of the sorted points -------------------------------
//
// One wave per 64 N consecutive sorted points (N per lane, in order), so the
// work follows the points, not the cells: a cell's points are one run of
// equal keys.  A run inside one lane is finished by that lane: sums in
// sequence, and each point's slot is its rank among the run's distinct
// tickets.  A run across lanes is a segment of the wave: its sums come from a
// segmented scan, its top-8 from 8 rounds of a segmented minimum over the
// lanes' next candidates (each lane pops its candidate when the round's
// minimum has its ticket).  A run across waves leaves a partial record at
// each wave's end it crosses, and k_cells_join merges them.

// Position as an integer multiple of 2^-k (exact for |p| >= 2^(23-k)).
__device__ inline long long fixed(float v, float scale) {
    return __double2ll_rn((double)v * (double)scale);
}

constexpr uint32_t KEY_NONE = 0xffffffffu;  // no cell: past the end, or a graph without the pass
constexpr unsigned long long TOP_NONE = ~0ull;

struct __align__(16) CellPart {
    unsigned long long top[SLOTS];  // (ticket << 32 | index) ascending, TOP_NONE past the end
    float2 pos[SLOTS];              // their positions
    long long sx, sy;
    uint32_t key, cnt, pad0, pad1;
};
static_assert(sizeof(CellPart) == 160, "ten uint4 per partial");

// The fields of one graph the cell pass reads.
struct SegGraph {
    uint32_t passes, p0, f0, c0, gf, L;
    float seed, fxs;
};
__device__ inline SegGraph seg_graph(const GDesc* __restrict__ desc, uint32_t g) {
    const GDesc& d = desc[g];
    SegGraph s;
    s.passes = d.passes; s.p0 = d.p0; s.f0 = d.f0; s.c0 = d.c0; s.gf = d.gf; s.L = d.L;
    s.seed = d.seed; s.fxs = d.fx_scale;
    return s;
}

__device__ inline uint32_t seg_rm(const SegGraph& d, uint32_t f) {
    const uint32_t mc = f - d.f0;
    return compact1by1(mc >> 1) * d.gf + compact1by1(mc);
}

// Cross-lane moves by DPP (no LDS): row_shr:k = 0x110 + k, row_bcast:15 =
// 0x142 (rows 1 and 3 take lane 15 of the row before), row_bcast:31 = 0x143
// (rows 2 and 3 take lane 31), wave_shr:1 = 0x138.  A lane whose source is
// outside its row, or whose row is masked off, takes `old`.
template <int CTRL, int RM>
__device__ inline uint32_t dpp32(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ inline unsigned long long dpp64(unsigned long long old, unsigned long long v) {
    const uint32_t lo = dpp32<CTRL, RM>((uint32_t)old, (uint32_t)v);
    const uint32_t hi = dpp32<CTRL, RM>((uint32_t)(old >> 32), (uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline unsigned long long readlane64(unsigned long long v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}

// Segmented inclusive scans over the wave (Hillis-Steele inside 16-lane rows,
// then the row broadcasts); f = 1 opens a segment at the lane.
template <int CTRL, int RM>
__device__ inline void segmin_step(unsigned long long& v, uint32_t& f) {
    const unsigned long long vo = dpp64<CTRL, RM>(TOP_NONE, v);
    const uint32_t fo = dpp32<CTRL, RM>(0u, f);
    if (!f) v = vo < v ? vo : v;
    f |= fo;
}
__device__ inline unsigned long long segmin_scan(unsigned long long v, uint32_t f) {
    segmin_step<0x111, 0xf>(v, f);
    segmin_step<0x112, 0xf>(v, f);
    segmin_step<0x114, 0xf>(v, f);
    segmin_step<0x118, 0xf>(v, f);
    segmin_step<0x142, 0xa>(v, f);
    segmin_step<0x143, 0xc>(v, f);
    return v;
}
template <int CTRL, int RM>
__device__ inline void segsum_step(long long& x, long long& y, uint32_t& n, uint32_t& f) {
    const long long xo = (long long)dpp64<CTRL, RM>(0ull, (unsigned long long)x);
    const long long yo = (long long)dpp64<CTRL, RM>(0ull, (unsigned long long)y);
    const uint32_t no = dpp32<CTRL, RM>(0u, n);
    const uint32_t fo = dpp32<CTRL, RM>(0u, f);
    if (!f) { x += xo; y += yo; n += no; }
    f |= fo;
}
__device__ inline void segsum_scan(long long& x, long long& y, uint32_t& n, uint32_t f) {
    segsum_step<0x111, 0xf>(x, y, n, f);
    segsum_step<0x112, 0xf>(x, y, n, f);
    segsum_step<0x114, 0xf>(x, y, n, f);
    segsum_step<0x118, 0xf>(x, y, n, f);
    segsum_step<0x142, 0xa>(x, y, n, f);
    segsum_step<0x143, 0xc>(x, y, n, f);
}

template <int N>
__global__ void __launch_bounds__(256) k_cells_seg(const uint32_t* __restrict__ skey,
                                                   const uint32_t* __restrict__ sv,
                                                   const float2* __restrict__ spos, uint32_t np,
                                                   const uint32_t* __restrict__ blk_graph,
                                                   const GDesc* __restrict__ desc, uint32_t one_graph,
                                                   long long* __restrict__ sums,
                                                   int32_t* __restrict__ slot_idx,
                                                   float2* __restrict__ slot_pos,
                                                   CellPart* __restrict__ parts) {
    constexpr uint32_t WAVE_PTS = 64u * N;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t w0 = w * WAVE_PTS;
    if (w0 >= np) return;  // uniform over the wave
    const uint32_t base = w0 + lane * N;
    uint32_t key[N], idx[N];
    float2 p[N];
    if (base + N <= np) {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint4 a = reinterpret_cast<const uint4*>(skey + base)[q];
            const uint4 b = reinterpret_cast<const uint4*>(sv + base)[q];
            key[4 * q] = a.x; key[4 * q + 1] = a.y; key[4 * q + 2] = a.z; key[4 * q + 3] = a.w;
            idx[4 * q] = b.x; idx[4 * q + 1] = b.y; idx[4 * q + 2] = b.z; idx[4 * q + 3] = b.w;
        }
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
            const float4 a = reinterpret_cast<const float4*>(spos + base)[q];
            p[2 * q] = make_float2(a.x, a.y);
            p[2 * q + 1] = make_float2(a.z, a.w);
        }
    } else {
#pragma unroll
        for (int u = 0; u < N; ++u) {
            const uint32_t k = base + u;
            const bool in = k < np;
            key[u] = in ? skey[k] : KEY_NONE;
            idx[u] = in ? sv[k] : 0u;
            p[u] = in ? spos[k] : make_float2(0.f, 0.f);
        }
    }
    const uint32_t kprev = w0 > 0 ? skey[w0 - 1] : KEY_NONE;
    const uint32_t knext = w0 + WAVE_PTS < np ? skey[w0 + WAVE_PTS] : KEY_NONE;
    // graphs: keys ascend and every graph owns one key range, so a lane's
    // items belong to the graph of its first item or of its last one, except
    // in sets of graphs smaller than a lane
    uint32_t g[N];
#pragma unroll
    for (int u = 0; u < N; ++u)
        g[u] = (one_graph || key[u] == KEY_NONE) ? 0u : blk_graph[key[u] / CELL_BLOCK];
    uint32_t glast = g[0];
#pragma unroll
    for (int u = 1; u < N; ++u) glast = key[u] != KEY_NONE ? g[u] : glast;
    const SegGraph ga = seg_graph(desc, g[0]);
    const SegGraph gb = seg_graph(desc, glast);
    auto gof = [&](int u) -> SegGraph {
        return g[u] == g[0] ? ga : (g[u] == glast ? gb : seg_graph(desc, g[u]));
    };
    uint32_t tk[N];
    long long fx[N], fy[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
        const SegGraph gu = gof(u);
        if (key[u] != KEY_NONE && !((gu.passes >> P_MANYBODY) & 1u)) key[u] = KEY_NONE;
        tk[u] = ticket_bits(idx[u] - gu.p0, gu.seed);
        fx[u] = fixed(p[u].x, gu.fxs);
        fy[u] = fixed(p[u].y, gu.fxs);
    }
    uint32_t kl = __shfl_up(key[N - 1], 1, 64);
    uint32_t kn = __shfl_down(key[0], 1, 64);
    if (lane == 0) kl = kprev;
    if (lane == 63) kn = knext;
    // runs inside the lane
    uint32_t rid[N];
    rid[0] = 0;
#pragma unroll
    for (int u = 1; u < N; ++u) rid[u] = rid[u - 1] + (key[u] != key[u - 1] ? 1u : 0u);
    const uint32_t nrun = rid[N - 1] + 1;
    const bool cl = key[0] != KEY_NONE && key[0] == kl;      // the first run began before the lane
    const bool cr = key[N - 1] != KEY_NONE && key[N - 1] == kn;  // the last run goes on after it
    const bool whole = cl && cr && nrun == 1;
    const bool hasL = cr;            // L: the last run, continuing right
    const bool hasF = cl && !whole;  // F: the first run, from the left, ending in the lane
    // a ticket repeated inside a run keeps its smallest index, the others are
    // duplicates; rank = the run's distinct tickets below the point's
    bool dup[N];
    uint32_t rank[N];
    bool anyeq = false;
#pragma unroll
    for (int u = 0; u < N; ++u) {
        dup[u] = false;
        rank[u] = 0;
#pragma unroll
        for (int v = u + 1; v < N; ++v) anyeq |= tk[u] == tk[v];
    }
    if (__ballot(anyeq) != 0ull) {  // a repeated ticket somewhere in the wave: the full rule
#pragma unroll
        for (int u = 0; u < N; ++u) {
            bool dd = false;
#pragma unroll
            for (int v = 0; v < N; ++v)
                if (v != u) dd |= rid[v] == rid[u] && tk[v] == tk[u] && idx[v] < idx[u];
            dup[u] = dd;
        }
#pragma unroll
        for (int u = 0; u < N; ++u) {
            uint32_t r = 0;
#pragma unroll
            for (int v = 0; v < N; ++v)
                if (v != u) r += (rid[v] == rid[u] && !dup[v] && tk[v] < tk[u]) ? 1u : 0u;
            rank[u] = r;
        }
    } else {
#pragma unroll
        for (int u = 0; u < N; ++u)
#pragma unroll
            for (int v = u + 1; v < N; ++v)
                if (rid[u] == rid[v]) {
                    if (tk[v] < tk[u]) ++rank[u];
                    else ++rank[v];
                }
    }
    // runs finished inside the lane; sums of F and L
    long long Fx = 0, Fy = 0, Lx = 0, Ly = 0;
    uint32_t Fn = 0, Ln = 0;
    {
        long long ax = 0, ay = 0;
        uint32_t an = 0, ad = 0;
#pragma unroll
        for (int u = 0; u < N; ++u) {
            if (u == 0 || rid[u] != rid[u - 1]) { ax = 0; ay = 0; an = 0; ad = 0; }
            ax += fx[u];
            ay += fy[u];
            an += 1u;
            ad += dup[u] ? 0u : 1u;
            const bool isF = hasF && rid[u] == 0;
            const bool isL = hasL && rid[u] == nrun - 1;
            const bool last = (u == N - 1) || rid[u + 1] != rid[u];
            if (isF || isL || key[u] == KEY_NONE) {
                if (last && isF) { Fx = ax; Fy = ay; Fn = an; }
                if (last && isL) { Lx = ax; Ly = ay; Ln = an; }
                continue;
            }
            const SegGraph gu = gof(u);
            const uint32_t rm = seg_rm(gu, key[u]);
            const size_t sb = ((size_t)gu.f0 + rm) * SLOTS;
            if (!dup[u] && rank[u] < (uint32_t)SLOTS) {
                slot_idx[sb + rank[u]] = (int32_t)idx[u];
                slot_pos[sb + rank[u]] = p[u];
            }
            if (last) {
                const size_t c3 = 3 * ((size_t)gu.c0 + level_off(gu.L - 1) + rm);
                sums[c3] = ax;
                sums[c3 + 1] = ay;
                sums[c3 + 2] = (long long)an;
                for (uint32_t s = ad; s < (uint32_t)SLOTS; ++s) {
                    slot_idx[sb + s] = -1;
                    slot_pos[sb + s] = make_float2(0.f, 0.f);
                }
            }
        }
    }
    const unsigned long long mF = __ballot(hasF);
    const unsigned long long mL = __ballot(hasL);
    if ((mF | mL) == 0ull) {
        if (lane == 0) { parts[2 * w].key = KEY_NONE; parts[2 * w + 1].key = KEY_NONE; }
        return;
    }
    // segments: a lane whose L run starts in it (or that has none) opens one
    const uint32_t open = whole ? 0u : 1u;
    const unsigned long long above = lane < 63 ? (mF >> (lane + 1)) : 0ull;
    const uint32_t E = above ? lane + (uint32_t)__ffsll((long long)above) : 64u;  // the segment's F lane
    // the segment opened before the wave: lane 0 inside a run from the last wave
    const bool o0 = (__ballot(whole) & 1ull) != 0ull;
    const uint32_t E0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)E);
    const bool originL = o0 && lane < E0;
    const bool originF = lane == 0 || (o0 && lane == E0);
    const bool doneL = !originL && E < 64u;  // the L run began in this wave and ends in it
    const bool doneF = !originF;
    const unsigned long long mPF = __ballot(hasF && originF);  // a run from an earlier wave ends here
    // the partial an unfinished segment of this lane fills
    const uint32_t partL = E < 64u ? 2 * w : 2 * w + 1;
    // sums: segmented scan over the L parts, then the F lane adds its own
    {
        long long vx = hasL ? Lx : 0, vy = hasL ? Ly : 0;
        uint32_t vn = hasL ? Ln : 0u;
        segsum_scan(vx, vy, vn, open);
        const long long cx = (long long)dpp64<0x138, 0xf>(0ull, (unsigned long long)vx);
        const long long cy = (long long)dpp64<0x138, 0xf>(0ull, (unsigned long long)vy);
        const uint32_t cn = dpp32<0x138, 0xf>(0u, vn);
        if (hasF) {
            const long long tx = Fx + cx, ty = Fy + cy;
            const uint32_t tn = Fn + cn;
            if (doneF) {
                const uint32_t rm = seg_rm(ga, key[0]);
                const size_t c3 = 3 * ((size_t)ga.c0 + level_off(ga.L - 1) + rm);
                sums[c3] = tx;
                sums[c3 + 1] = ty;
                sums[c3 + 2] = (long long)tn;
            } else {
                parts[2 * w].key = key[0];
                parts[2 * w].sx = tx;
                parts[2 * w].sy = ty;
                parts[2 * w].cnt = tn;
            }
        }
        if (lane == 63) {
            parts[2 * w + 1].key = hasL ? key[N - 1] : KEY_NONE;
            if (hasL) {
                parts[2 * w + 1].sx = vx;
                parts[2 * w + 1].sy = vy;
                parts[2 * w + 1].cnt = vn;
            }
        }
        if (lane == 0 && mPF == 0ull) parts[2 * w].key = KEY_NONE;
    }
    // top-8: 8 rounds of a segmented minimum of the lanes' next candidates
    uint32_t cF8[N], cL8[N];  // each point's place in the F or L list (0xff: none)
#pragma unroll
    for (int u = 0; u < N; ++u) {
        cF8[u] = (hasF && rid[u] == 0 && !dup[u]) ? rank[u] : 0xffu;
        cL8[u] = (hasL && rid[u] == nrun - 1 && !dup[u]) ? rank[u] : 0xffu;
    }
    const uint32_t rmF = seg_rm(ga, key[0]), rmL = seg_rm(gb, key[N - 1]);
    const size_t sbF = ((size_t)ga.f0 + rmF) * SLOTS, sbL = ((size_t)gb.f0 + rmL) * SLOTS;
    uint32_t hF = 0, hL = 0;
#pragma unroll 1
    for (int r = 0; r < SLOTS; ++r) {
        uint32_t cFt = 0xffffffffu, cFi = 0xffffffffu, cLt = 0xffffffffu, cLi = 0xffffffffu;
        float2 pF = make_float2(0.f, 0.f), pL = make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < N; ++u) {
            const bool sf = cF8[u] == hF, sl = cL8[u] == hL;
            cFt = sf ? tk[u] : cFt; cFi = sf ? idx[u] : cFi; pF = sf ? p[u] : pF;
            cLt = sl ? tk[u] : cLt; cLi = sl ? idx[u] : cLi; pL = sl ? p[u] : pL;
        }
        const unsigned long long cF = ((unsigned long long)cFt << 32) | cFi;
        const unsigned long long cL = ((unsigned long long)cLt << 32) | cLi;
        const unsigned long long v = segmin_scan(cL, open);
        const unsigned long long carry = dpp64<0x138, 0xf>(TOP_NONE, v);
        const unsigned long long m = hasF ? (carry < cF ? carry : cF) : TOP_NONE;
        const unsigned long long mE = __shfl(m, (int)(E & 63u), 64);
        const unsigned long long m63 = readlane64(v, 63);
        const unsigned long long seg = E < 64u ? mE : m63;
        if (hasL && cL != TOP_NONE && seg != TOP_NONE && cLt == (uint32_t)(seg >> 32)) {
            if (cL == seg) {
                if (doneL) {
                    slot_idx[sbL + r] = (int32_t)cLi;
                    slot_pos[sbL + r] = pL;
                } else {
                    parts[partL].pos[r] = pL;
                }
            }
            ++hL;
        }
        if (hasF) {
            if (cF != TOP_NONE && m != TOP_NONE && cFt == (uint32_t)(m >> 32)) {
                if (cF == m) {
                    if (doneF) {
                        slot_idx[sbF + r] = (int32_t)cFi;
                        slot_pos[sbF + r] = pF;
                    } else {
                        parts[2 * w].pos[r] = pF;
                    }
                }
                ++hF;
            }
            if (doneF && m == TOP_NONE) {
                slot_idx[sbF + r] = -1;
                slot_pos[sbF + r] = make_float2(0.f, 0.f);
            }
            if (!doneF) parts[2 * w].top[r] = m;
        }
        if (lane == 63 && hasL) parts[2 * w + 1].top[r] = v;
    }
}

// One wave per wave of k_cells_seg whose first partial names a cell (a run
// that began in an earlier wave and ends in this one): merges the run's
// partials — the last one of each earlier wave it crosses, then this one —
// one per lane (lists merged in the lane above 64 of them), then 8 rounds of
// the wave's minimum.
template <int N>
__global__ void __launch_bounds__(256) k_cells_join(const CellPart* __restrict__ parts, uint32_t nw,
                                                    const uint32_t* __restrict__ cell_begin,
                                                    const uint32_t* __restrict__ blk_graph,
                                                    const GDesc* __restrict__ desc,
                                                    long long* __restrict__ sums,
                                                    int32_t* __restrict__ slot_idx,
                                                    float2* __restrict__ slot_pos) {
    constexpr uint32_t WAVE_PTS = 64u * N;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (w >= nw) return;
    const uint32_t f = parts[2 * w].key;
    if (f == KEY_NONE) return;
    const uint32_t wa = cell_begin[f] / WAVE_PTS;
    unsigned long long t8[SLOTS];
    float2 p8[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) { t8[s] = TOP_NONE; p8[s] = make_float2(0.f, 0.f); }
    long long sx = 0, sy = 0;
    uint32_t cnt = 0;
    for (uint32_t q = wa + lane; q <= w; q += 64) {
        const CellPart& P = q == w ? parts[2 * w] : parts[2 * q + 1];
        sx += P.sx;
        sy += P.sy;
        cnt += P.cnt;
        if (q < wa + 64) {
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) { t8[s] = P.top[s]; p8[s] = P.pos[s]; }
        } else {
            // a second list in this lane: merge (8 smallest distinct tickets)
#pragma unroll 1
            for (int s = 0; s < SLOTS; ++s) {
                unsigned long long x = P.top[s];
                float2 xp = P.pos[s];
                if (x == TOP_NONE) break;
                bool dupl = false;
#pragma unroll
                for (int k = 0; k < SLOTS; ++k) {
                    if ((uint32_t)(t8[k] >> 32) == (uint32_t)(x >> 32) && t8[k] != TOP_NONE) {
                        dupl = true;
                        if (x < t8[k]) { t8[k] = x; p8[k] = xp; }
                    }
                }
                if (dupl) continue;
#pragma unroll
                for (int k = 0; k < SLOTS; ++k) {
                    const bool sw = x < t8[k];
                    const unsigned long long a = t8[k];
                    const float2 ap = p8[k];
                    t8[k] = sw ? x : a;
                    p8[k] = sw ? xp : ap;
                    x = sw ? a : x;
                    xp = sw ? ap : xp;
                }
            }
        }
    }
    const uint32_t open = lane == 0 ? 1u : 0u;
    {
        segsum_scan(sx, sy, cnt, open);
        sx = (long long)readlane64((unsigned long long)sx, 63);
        sy = (long long)readlane64((unsigned long long)sy, 63);
        cnt = (uint32_t)__builtin_amdgcn_readlane((int)cnt, 63);
    }
    const GDesc& d = desc[blk_graph[f / CELL_BLOCK]];
    const uint32_t mc = f - d.f0;
    const uint32_t rm = compact1by1(mc >> 1) * d.gf + compact1by1(mc);
    const size_t sb = ((size_t)d.f0 + rm) * SLOTS;
    if (lane == 0) {
        const size_t c3 = 3 * ((size_t)d.c0 + level_off(d.L - 1) + rm);
        sums[c3] = sx;
        sums[c3 + 1] = sy;
        sums[c3 + 2] = (long long)cnt;
    }
    uint32_t h = 0;
#pragma unroll 1
    for (int r = 0; r < SLOTS; ++r) {
        unsigned long long c = TOP_NONE;
        float2 cp = make_float2(0.f, 0.f);
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            const bool sel = (uint32_t)s == h;
            c = sel ? t8[s] : c;
            cp = sel ? p8[s] : cp;
        }
        const unsigned long long mn = readlane64(segmin_scan(c, open), 63);
        if (mn == TOP_NONE) {
            if (lane == 0) {
                slot_idx[sb + r] = -1;
                slot_pos[sb + r] = make_float2(0.f, 0.f);
            }
            continue;
        }
        if (c != TOP_NONE && (uint32_t)(c >> 32) == (uint32_t)(mn >> 32)) {
            if (c == mn) {
                slot_idx[sb + r] = (int32_t)(uint32_t)c;
                slot_pos[sb + r] = cp;
            }
            ++h;
        }
    }
}

// ---- 3. pyramid ---------------------------------------------------------------

// An empty cell's centroid sits at EMPTY_AT in both coordinates with count 0:
// a far-field term then weighs exactly 0 with no test (0 * 1/l, l finite).
constexpr float EMPTY_AT = -1.0e18f;
__device__ inline float4 centroid(long long sx, long long sy, long long n, float inv) {
    if (n <= 0) return make_float4(EMPTY_AT, EMPTY_AT, 0.f, 0.f);
    const double dn = (double)n;
    return make_float4((float)((double)sx * (double)inv / dn), (float)((double)sy * (double)inv / dn),
                       (float)dn, 0.f);
}

// One block per tile of one graph's level `lin` (work item {graph slot, lin,
// T | nout << 16, tile}); produces the `nout` coarser levels of that tile down
// to one cell per tile.  At the finest level (write_in) an empty cell (end 0)
// has zero sums.  Only the finest stage's last level keeps its sums (the
// top-level stage starts from them); every level keeps its centroids.
// A finest-level tile of T >= 16 (whole blocks of CELL_BLOCK cells) whose
// blocks are all clean is skipped: it held no point at this pass nor at the
// last one, so every value it would write (empty centroids, zero sums) is
// already in place, written when it was last processed or when the packing
// marked every block dirty.  A processed tile clears its blocks' marks.
// Whether work item wk's tile must be written this pass: its graph runs the
// pass and, for a tracked finest-level tile, one of its blocks is dirty.
// Sets the tile's first block and block count (0 when untracked).
__device__ inline bool pyr_needed(uint4 wk, const GDesc& d, int write_in,
                                  const uint8_t* __restrict__ cell_dirty, uint32_t& blk0, uint32_t& nblk) {
    const int lin = (int)wk.y, T = (int)(wk.z & 0xffffu);
    const int tiles_x = (4 << lin) / T;
    const int tx = (int)wk.w % tiles_x, ty = (int)wk.w / tiles_x;
    // the tile's blocks of finest cells: T x T Morton-consecutive cells from an
    // aligned corner, so T * T / CELL_BLOCK whole blocks when T >= 16
    const bool tracked = write_in && T * T >= (int)CELL_BLOCK;
    blk0 = (d.f0 + morton2((uint32_t)(tx * T), (uint32_t)(ty * T))) / CELL_BLOCK;
    nblk = tracked ? (uint32_t)(T * T) / CELL_BLOCK : 0u;
    if (!(d.passes & (1u << P_MANYBODY))) return false;
    if (!tracked) return true;
    bool dirty = false;
    for (uint32_t q = 0; q < nblk; ++q) dirty |= cell_dirty[blk0 + q] != 0;
    return dirty;
}

// One tile of work item wk, by the whole workgroup (barriers inside).
__device__ __forceinline__ void pyr_tile(uint4 wk, const GDesc& d, const uint32_t* __restrict__ cell_end,
                                         long long* __restrict__ sums, float4* __restrict__ cent, int write_in,
                                         uint8_t* __restrict__ cell_occ, uint8_t* __restrict__ cell_dirty,
                                         uint32_t blk0, uint32_t nblk, long long* s0, long long* s1) {
    const int lin = (int)wk.y, T = (int)(wk.z & 0xffffu), nout = (int)(wk.z >> 16);
    const int gin = 4 << lin;
    const int tiles_x = gin / T;
    const int tx = (int)wk.w % tiles_x, ty = (int)wk.w / tiles_x;
    const size_t in0 = (size_t)d.c0 + level_off(lin);
    for (int c = threadIdx.x; c < T * T; c += blockDim.x) {
        const int x = c % T, y = c / T;
        const int X = tx * T + x, Y = ty * T + y;
        const size_t rm = in0 + (size_t)Y * gin + X;
        long long sx = 0, sy = 0, n = 0;
        if (!write_in || cell_end[d.f0 + morton2((uint32_t)X, (uint32_t)Y)] != 0) {
            sx = sums[3 * rm];
            sy = sums[3 * rm + 1];
            n = sums[3 * rm + 2];
        }
        s0[3 * c] = sx;
        s0[3 * c + 1] = sy;
        s0[3 * c + 2] = n;
        if (write_in) cent[rm] = centroid(sx, sy, n, d.fx_inv);
    }
    __syncthreads();
    long long* cur = s0;
    long long* nxt = s1;
    int size = T;
    for (int lev = 1; lev <= nout; ++lev) {
        int half = size / 2;
        int lout = lin - lev;
        int gout = gin >> lev;
        const size_t out0 = (size_t)d.c0 + level_off(lout);
        for (int c = threadIdx.x; c < half * half; c += blockDim.x) {
            int x = c % half, y = c / half;
            int c00 = (2 * y) * size + 2 * x, c10 = c00 + 1, c01 = c00 + size, c11 = c01 + 1;
            long long v[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                v[q] = (cur[3 * c00 + q] + cur[3 * c10 + q]) + (cur[3 * c01 + q] + cur[3 * c11 + q]);
            nxt[3 * c] = v[0];
            nxt[3 * c + 1] = v[1];
            nxt[3 * c + 2] = v[2];
            size_t rm = out0 + (size_t)(ty * half + y) * gout + (tx * half + x);
            if (write_in && lev == nout) {  // the top-level stage reads this level's sums
                sums[3 * rm] = v[0];
                sums[3 * rm + 1] = v[1];
                sums[3 * rm + 2] = v[2];
            }
            cent[rm] = centroid(v[0], v[1], v[2], d.fx_inv);
        }
        __syncthreads();
        long long* t = cur;
        cur = nxt;
        nxt = t;
        size = half;
    }
    // every lane read the marks before the barriers above
    if (threadIdx.x < nblk) {
        cell_occ[blk0 + threadIdx.x] = 0u;
        cell_dirty[blk0 + threadIdx.x] = 0u;
    }
}

// `span` work items per workgroup: with span 1 a workgroup takes its own tile
// (or leaves at once); a larger span (sets of many mostly empty tiles) lists
// the tiles of its range that need writing in LDS first and writes them one
// after another, so the launch does not dispatch a workgroup per clean tile.
__global__ void __launch_bounds__(256) k_pyramid(const uint4* __restrict__ work, uint32_t nwork,
                                                 uint32_t span, const GDesc* __restrict__ desc,
                                                 const uint32_t* __restrict__ cell_end,
                                                 long long* __restrict__ sums,
                                                 float4* __restrict__ cent, int write_in,
                                                 uint8_t* __restrict__ cell_occ,
                                                 uint8_t* __restrict__ cell_dirty) {
    __shared__ long long s0[32 * 32 * 3];
    __shared__ long long s1[16 * 16 * 3];
    if (span == 1) {
        const uint4 wk = work[blockIdx.x];
        const GDesc& d = desc[wk.x];
        uint32_t blk0, nblk;
        if (!pyr_needed(wk, d, write_in, cell_dirty, blk0, nblk)) return;  // uniform over the block
        pyr_tile(wk, d, cell_end, sums, cent, write_in, cell_occ, cell_dirty, blk0, nblk, s0, s1);
        return;
    }
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_n;
    const uint32_t w0 = blockIdx.x * span;
    for (uint32_t base = 0; base < span; base += 256) {
        if (threadIdx.x == 0) s_n = 0u;
        __syncthreads();
        const uint32_t w = w0 + base + threadIdx.x;
        if (base + threadIdx.x < span && w < nwork) {
            const uint4 wk = work[w];
            uint32_t blk0, nblk;
            if (pyr_needed(wk, desc[wk.x], write_in, cell_dirty, blk0, nblk)) s_list[atomicAdd(&s_n, 1u)] = w;
        }
        __syncthreads();
        const uint32_t n = s_n;
        for (uint32_t q = 0; q < n; ++q) {
            const uint4 wk = work[s_list[q]];
            const GDesc& d = desc[wk.x];
            uint32_t blk0, nblk;
            (void)pyr_needed(wk, d, write_in, cell_dirty, blk0, nblk);
            pyr_tile(wk, d, cell_end, sums, cent, write_in, cell_occ, cell_dirty, blk0, nblk, s0, s1);
        }
        __syncthreads();  // the list is rewritten next round
    }
}

// ---- 4. forces ------------------------------------------------------------------

// Clamped inverse-distance falloff of force-level.frag:43-56 /
// force-nearfield.frag:64-80 without the alpha*k*N factor, for l = |d|^2 > 0:
// contribution = c * d with c = 1/l, or l^(-3/4) below unit distance; with
// inv = 1/|d| both are inv * min(inv, sqrt(inv)).
__device__ inline float falloff_coef(float l) {
    const float inv = __builtin_amdgcn_rsqf(l);
    return inv * fminf(inv, __builtin_amdgcn_sqrtf(inv));
}

// One far-field term, zero where `use` is clear (outside the ring or the grid;
// an empty cell weighs 0 through its count and far-away centroid).  A ring
// cell lies at least one cell width from the query point (so l > 0), and with
// cells of width >= 2 (FAST) l >= 1 and the falloff is 1/l.
template <bool FAST>
__device__ inline void far_term(float4 c, float2 p, bool use, float& fx, float& fy) {
    const float dx = p.x - c.x, dy = p.y - c.y;
    const float l = dx * dx + dy * dy;
    const float coef = FAST ? __builtin_amdgcn_rcpf(l) : falloff_coef(l);
    const float w = use ? c.z * coef : 0.f;
    fx += w * dx;
    fy += w * dy;
}

// Level 0: every cell of the 4x4 grid outside the 3x3 neighbourhood
// (force-level.frag:70-79).  `cells` is the level (global or staged).  Rows
// are not unrolled, so only one row's loads are in flight per lane.
template <bool FAST>
__device__ inline void far_level0(const float4* cells, float2 p, int qx, int qy, float& fx, float& fy) {
#pragma unroll 1
    for (int cy = 0; cy < 4; ++cy) {
        const bool far_row = abs(cy - qy) > 1;
#pragma unroll
        for (int cx = 0; cx < 4; ++cx)
            far_term<FAST>(cells[cy * 4 + cx], p, far_row || abs(cx - qx) > 1, fx, fy);
    }
}

// Level >= 1: the aligned 6x6 child block minus the 3x3 (force-level.frag:80-95).
// The 3x3 sits at offset 1 + (q & 1) in each axis of the block, so the ring is
// exactly: rows {0, 4 - 3 py, 5} in full, and rows 1 + py .. 3 + py at columns
// {0, 4 - 3 px, 5} — 27 cells, visited without a mask.  Cell (cx, cy) is read
// at tile[(cy - y0) * w + (cx - x0)]: the level itself with x0 = y0 = 0, w = g,
// or a staged window of it.  Cells outside the grid are read at a clamped
// address and weighted 0.
template <bool FAST>
__device__ inline void far_level(const float4* tile, int x0, int y0, int w, int g, int qx, int qy,
                                 float2 p, float& fx, float& fy) {
    const int bx = (qx >> 1) * 2 - 2, by = (qy >> 1) * 2 - 2;
    const int px = qx & 1, py = qy & 1;
    int col[6];
    bool cok[6];
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        const int cx = bx + o;
        cok[o] = cx >= 0 && cx < g;
        col[o] = min(max(cx, 0), g - 1) - x0;
    }
    const int hcol = px ? col[1] : col[4];  // the hole rows' middle column, 4 - 3 px
    const bool hok = px ? cok[1] : cok[4];
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {  // full rows
        const int r = k == 0 ? 0 : (k == 1 ? 4 - 3 * py : 5);
        const int cy = by + r;
        const bool rok = cy >= 0 && cy < g;
        const int rb = (min(max(cy, 0), g - 1) - y0) * w;
#pragma unroll
        for (int ii = 0; ii < 6; ++ii) far_term<FAST>(tile[rb + col[ii]], p, rok && cok[ii], fx, fy);
    }
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {  // hole rows: the three columns outside the hole
        const int cy = by + 1 + py + k;
        const bool rok = cy >= 0 && cy < g;
        const int rb = (min(max(cy, 0), g - 1) - y0) * w;
        far_term<FAST>(tile[rb + col[0]], p, rok && cok[0], fx, fy);
        far_term<FAST>(tile[rb + hcol], p, rok && hok, fx, fy);
        far_term<FAST>(tile[rb + col[5]], p, rok && cok[5], fx, fy);
    }
}

// The same 27 ring cells read from a staged window padded with empty cells
// wherever the 6x6 block leaves the grid, so no term needs a mask: the window
// starts at (x0, y0) and is w cells wide.
template <bool FAST>
__device__ inline void far_level_padded(const float4* tile, int x0, int y0, int w, int qx, int qy, float2 p,
                                        float& fx, float& fy) {
    const int bx = (qx >> 1) * 2 - 2 - x0, by = (qy >> 1) * 2 - 2 - y0;
    const int px = qx & 1, py = qy & 1;
    const int hc = 4 - 3 * px;  // the hole rows' middle column
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {  // full rows
        const int r = k == 0 ? 0 : (k == 1 ? 4 - 3 * py : 5);
        const float4* row = tile + (by + r) * w + bx;
#pragma unroll
        for (int ii = 0; ii < 6; ++ii) far_term<FAST>(row[ii], p, true, fx, fy);
    }
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {  // hole rows: the three columns outside the hole
        const float4* row = tile + (by + 1 + py + k) * w + bx;
        far_term<FAST>(row[0], p, true, fx, fy);
        far_term<FAST>(row[hc], p, true, fx, fy);
        far_term<FAST>(row[5], p, true, fx, fy);
    }
}

// Near field at the finest level (force-nearfield.frag:101-163): the 3x3
// cells' sampled slots, Horvitz-Thompson weighted.  Counts come from `fin`
// (finest centroids, global or a staged window as in far_level).
__device__ __forceinline__ float2 near_field(const float4* fin, int x0, int y0, int w, const GDesc& d, int qx,
                                    int qy, uint32_t i, float2 p, float2 r,
                                    const int32_t* __restrict__ slot_idx,
                                    const float2* __restrict__ slot_pos) {
    const int gf = (int)d.gf;
    float wx = 0.f, wy = 0.f;
#pragma unroll 1
    for (int dj = -1; dj <= 1; ++dj) {
        const int cy = qy + dj;
        if (cy < 0 || cy >= gf) continue;
#pragma unroll 1
        for (int di = -1; di <= 1; ++di) {
            const int cx = qx + di;
            if (cx < 0 || cx >= gf) continue;
            float others = fin[(cy - y0) * w + (cx - x0)].z - ((di == 0 && dj == 0) ? 1.f : 0.f);
            if (others <= 0.f) continue;
            // the cell's 8 slots in 6 vector loads; filled slots lead, -1 after
            // them (slots past the first -1 and the point itself weigh 0)
            const size_t sb = ((size_t)d.f0 + cy * gf + cx) * SLOTS;
            const int4* __restrict__ si = reinterpret_cast<const int4*>(slot_idx + sb);
            const float4* __restrict__ sp = reinterpret_cast<const float4*>(slot_pos + sb);
            // coincident points share a cell: only the point's own cell can hold one
            const bool own = di == 0 && dj == 0;  // uniform over the wave
            const int4 ib = si[1];
            if (!own && __all(ib.w >= 0)) {
                // A neighbour cell with all 8 slots filled (8 distinct tickets:
                // a cell of 8 or more points may hold fewer) on every active
                // lane: every slot is sampled, none is the point itself (it lies
                // in its own cell) and none coincides with it (a different cell
                // holds a different position), so no term needs a mask.  1/l for
                // every term; if some lane of the wave has a pair below unit
                // distance, the cell is summed again with the clamped falloff,
                // which is the same value for every pair at or above it, so a
                // lane's bits do not depend on the other lanes of its wave.
                const float4 q[4] = {sp[0], sp[1], sp[2], sp[3]};
                float sxp = 0.f, syp = 0.f;
                bool small = false;
#pragma unroll
                for (int s = 0; s < SLOTS / 2; ++s) {
                    const float ax = p.x - q[s].x, ay = p.y - q[s].y, bx = p.x - q[s].z, by = p.y - q[s].w;
                    const float la = ax * ax + ay * ay, lb = bx * bx + by * by;
                    small = small || la < 1.f || lb < 1.f;
                    const float ia = __builtin_amdgcn_rsqf(la), ib2 = __builtin_amdgcn_rsqf(lb);
                    const float wa = ia * ia, wb = ib2 * ib2;
                    sxp += wa * ax;
                    syp += wa * ay;
                    sxp += wb * bx;
                    syp += wb * by;
                }
                if (__ballot(small) != 0ull) {
                    sxp = syp = 0.f;
#pragma unroll
                    for (int s = 0; s < SLOTS / 2; ++s) {
                        const float ax = p.x - q[s].x, ay = p.y - q[s].y, bx = p.x - q[s].z, by = p.y - q[s].w;
                        const float wa = falloff_coef(ax * ax + ay * ay), wb = falloff_coef(bx * bx + by * by);
                        sxp += wa * ax;
                        syp += wa * ay;
                        sxp += wb * bx;
                        syp += wb * by;
                    }
                }
                const float hw = others * __builtin_amdgcn_rcpf((float)SLOTS);  // as the masked path's
                wx += hw * sxp;
                wy += hw * syp;
                continue;
            }
            const int4 ia = si[0];
            const float4 q[4] = {sp[0], sp[1], sp[2], sp[3]};
            const int js[SLOTS] = {ia.x, ia.y, ia.z, ia.w, ib.x, ib.y, ib.z, ib.w};
            // One slot's pair: whether it is sampled and not the point itself,
            // and its offset (the point's jitter for a coincident point,
            // force-nearfield.frag:131-136: only the own cell can hold one).
            auto pair = [&](int s, bool& use, float& dx, float& dy, float& l) {
                const int j = js[s];
                use = (uint32_t)j != i;
                dx = p.x - q[s >> 1].x;
                dy = p.y - q[s >> 1].y;
                if (s & 1) {
                    dx = p.x - q[s >> 1].z;
                    dy = p.y - q[s >> 1].w;
                }
                l = dx * dx + dy * dy;
                if (own) {
                    const bool same = !(l > 0.f);
                    dx = same ? r.x : dx;
                    dy = same ? r.y : dy;
                    l = same ? r.x * r.x + r.y * r.y : l;
                }
            };
            // filled slots lead: those up to the first empty one are sampled
            float sxp = 0.f, syp = 0.f, sampled = 0.f;
            bool open = true, small = false;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) {
                bool use;
                float dx, dy, l;
                pair(s, use, dx, dy, l);
                open = open && js[s] >= 0;
                use = use && open;
                sampled += use ? 1.f : 0.f;
                use = use && l > 0.f;
                small = small || (use && l < 1.f);
                const float inv = __builtin_amdgcn_rsqf(l);
                const float wgt = use ? inv * inv : 0.f;
                sxp += wgt * dx;
                syp += wgt * dy;
            }
            // below unit distance the falloff is l^-3/4 (a square root): if a
            // lane of the wave has such a pair, the cell is summed again with
            // it (1/l = inv * inv above it, the same value)
            if (__ballot(small) != 0ull) {
                sxp = syp = 0.f;
                open = true;
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) {
                    bool use;
                    float dx, dy, l;
                    pair(s, use, dx, dy, l);
                    open = open && js[s] >= 0;
                    use = use && open && l > 0.f;
                    const float wgt = use ? falloff_coef(l) : 0.f;
                    sxp += wgt * dx;
                    syp += wgt * dy;
                }
            }
            if (sampled > 0.f) {
                // sampled is 1..8: a reciprocal (1 ulp) instead of an IEEE divide
                float hw = others * __builtin_amdgcn_rcpf(sampled);
                wx += hw * sxp;
                wy += hw * syp;
            }
        }
    }
    return make_float2(wx, wy);
}

// Scaling, the near-field step clamp (force-nearfield.frag:161) and the
// integrate pass.
__device__ inline void mb_finish(const GDesc& d, uint32_t i, float2 p, float2 r, float fx, float fy,
                                 float2 wn, const uint8_t* __restrict__ flags, float2* __restrict__ pos0,
                                 float2* __restrict__ pos1, float2* __restrict__ vout) {
    float wx = wn.x * d.ak, wy = wn.y * d.ak;
    wx += wx * r.x;
    wy += wy * r.y;
    const float maxstep = 2.0f * d.h[d.L - 1];
    float sp = sqrtf(wx * wx + wy * wy);
    if (sp > maxstep) {
        float s = maxstep / sp;
        wx *= s;
        wy *= s;
    }
    const float2 v = make_float2(d.ak * fx + wx, d.ak * fy + wy);
    if (vout) {
        vout[i] = v;
        return;
    }
    float2* __restrict__ pout = ((d.inbuf >> P_MANYBODY) & 1u) ? pos0 : pos1;
    pout[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, d.friction, d.S);
}

// TILE: staged cells per level (16 KB at 1024; Tuning::force_tile)

// One lane per point in cell order.  A block whose points all belong to one
// running graph stages, level by level, the union of its points' 6x6 windows
// into LDS (256 points in cell order cover a compact region, so the window is
// small and every lane's reads hit it); other blocks, and windows above
// MB_TILE cells, read the level from global memory.  Either way the same
// values are summed in the same order.
// Eight waves per SIMD: the compiler fits the kernel in 64 VGPRs without spills
// (88 unbounded, five waves).
template <int MB_TILE>
__global__ void __launch_bounds__(256, 8) k_mb_force(const uint32_t* __restrict__ sv,
                                                  const uint32_t* __restrict__ skey,
                                                  const float2* __restrict__ sorted_pos,
                                                  const float4* __restrict__ cent,
                                                  const int32_t* __restrict__ slot_idx,
                                                  const float2* __restrict__ slot_pos,
                                                  const float2* __restrict__ jitter,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ blk_graph,
                                                  const GDesc* __restrict__ desc, uint32_t np,
                                                  float2* __restrict__ pos0,
                                                  float2* __restrict__ pos1,
                                                  float2* __restrict__ vout, uint32_t one_graph) {
    __shared__ float4 s_tile[MB_TILE];
    __shared__ int s_box[4][4];
    __shared__ uint32_t s_g;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t k = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + t;
    const bool valid = k < np;
    const uint32_t kk = valid ? k : np - 1;
    const uint32_t i = sv[kk];
    const uint32_t g = one_graph ? 0u : blk_graph[skey[kk] / CELL_BLOCK];  // the graph of the point's cell
    const bool run = (desc[g].passes >> P_MANYBODY) & 1u;
    if (t == 0) s_g = g;
    __syncthreads();
    const bool uniform = __syncthreads_and(g == s_g && run);
    if (!uniform) {
        if (!valid || !run) return;
        const GDesc& d = desc[g];
        const float2 p = sorted_pos[k];
        const int L = (int)d.L;
        float fx = 0.f, fy = 0.f;
        const int gf = (int)d.gf;
        const int qx = cell_of(p.x, d.h[L - 1], gf), qy = cell_of(p.y, d.h[L - 1], gf);
        if (d.h[L - 1] >= 2.f) {
            far_level0<true>(cent + d.c0, p, qx >> (L - 1), qy >> (L - 1), fx, fy);
            for (int lvl = 1; lvl < L; ++lvl) {
                const int gl = 4 << lvl, k = L - 1 - lvl;
                far_level<true>(cent + d.c0 + level_off(lvl), 0, 0, gl, gl, qx >> k, qy >> k, p, fx, fy);
            }
        } else {
            far_level0<false>(cent + d.c0, p, qx >> (L - 1), qy >> (L - 1), fx, fy);
            for (int lvl = 1; lvl < L; ++lvl) {
                const int gl = 4 << lvl, k = L - 1 - lvl;
                far_level<false>(cent + d.c0 + level_off(lvl), 0, 0, gl, gl, qx >> k, qy >> k, p, fx, fy);
            }
        }
        const float2 r = jitter[i];
        const float2 wn = near_field(cent + d.c0 + level_off(L - 1), 0, 0, gf, d, qx, qy, i, p, r, slot_idx,
                                     slot_pos);
        mb_finish(d, i, p, r, fx, fy, wn, flags, pos0, pos1, vout);
        return;
    }
    const GDesc& d = desc[__builtin_amdgcn_readfirstlane(g)];
    const float2 p = sorted_pos[kk];
    const int L = (int)d.L;
    float fx = 0.f, fy = 0.f;
    const bool fast = d.h[L - 1] >= 2.f;  // uniform: one graph
    // The point's finest cell; its cell at a coarser level is this >> k
    // exactly (h[lvl] = h[L-1] 2^k, and floor(x / 2^k) = floor(x) >> k).
    const int gf = (int)d.gf;
    const int fqx = cell_of(p.x, d.h[L - 1], gf), fqy = cell_of(p.y, d.h[L - 1], gf);
    if (t < 16) s_tile[t] = cent[d.c0 + t];
    __syncthreads();
    if (fast) far_level0<true>(s_tile, p, fqx >> (L - 1), fqy >> (L - 1), fx, fy);
    else far_level0<false>(s_tile, p, fqx >> (L - 1), fqy >> (L - 1), fx, fy);
    int x0 = 0, y0 = 0, w = 0;
    bool staged = false;
    // The block's box of finest cells, once: every level's window is the
    // shifted box padded by one cell, widened to the aligned 6x6 blocks.
    {
        int bx0 = fqx, by0 = fqy;
        int bx1 = bx0, by1 = by0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            bx0 = min(bx0, __shfl_xor(bx0, o, 64));
            by0 = min(by0, __shfl_xor(by0, o, 64));
            bx1 = max(bx1, __shfl_xor(bx1, o, 64));
            by1 = max(by1, __shfl_xor(by1, o, 64));
        }
        if (lane == 0) {
            s_box[wave][0] = bx0;
            s_box[wave][1] = by0;
            s_box[wave][2] = bx1;
            s_box[wave][3] = by1;
        }
    }
    __syncthreads();
    const int fx0 = min(min(s_box[0][0], s_box[1][0]), min(s_box[2][0], s_box[3][0]));
    const int fy0 = min(min(s_box[0][1], s_box[1][1]), min(s_box[2][1], s_box[3][1]));
    const int fx1 = max(max(s_box[0][2], s_box[1][2]), max(s_box[2][2], s_box[3][2]));
    const int fy1 = max(max(s_box[0][3], s_box[1][3]), max(s_box[2][3], s_box[3][3]));
    for (int lvl = 1; lvl < L; ++lvl) {
        const int gl = 4 << lvl, k = L - 1 - lvl;
        const float4* __restrict__ lv = cent + d.c0 + level_off(lvl);
        const int qx = fqx >> k, qy = fqy >> k;
        // the window is not clipped to the grid: cells outside it stage as empty
        x0 = (((fx0 >> k) - 1) >> 1) * 2 - 2;
        y0 = (((fy0 >> k) - 1) >> 1) * 2 - 2;
        const int x1 = (((fx1 >> k) + 1) >> 1) * 2 + 3;
        const int y1 = (((fy1 >> k) + 1) >> 1) * 2 + 3;
        w = x1 - x0 + 1;
        const int area = w * (y1 - y0 + 1);
        staged = area <= MB_TILE;
        __syncthreads();  // the previous level's tile is no longer read
        if (staged) {
            // row = floor((c + 1/2) / w) through a float reciprocal: exact for c, w < 2^10
            const float iw = __builtin_amdgcn_rcpf((float)w);
            for (int c = t; c < area; c += 256) {
                const int row = (int)(((float)c + 0.5f) * iw);
                const int gx = x0 + (c - row * w), gy = y0 + row;
                s_tile[c] = (gx >= 0 && gx < gl && gy >= 0 && gy < gl) ? lv[gy * gl + gx]
                                                                        : make_float4(EMPTY_AT, EMPTY_AT, 0.f, 0.f);
            }
            __syncthreads();
            if (fast) far_level_padded<true>(s_tile, x0, y0, w, qx, qy, p, fx, fy);
            else far_level_padded<false>(s_tile, x0, y0, w, qx, qy, p, fx, fy);
        } else {
            if (fast) far_level<true>(lv, 0, 0, gl, gl, qx, qy, p, fx, fy);
            else far_level<false>(lv, 0, 0, gl, gl, qx, qy, p, fx, fy);
        }
    }
    if (!valid) return;
    const float2 r = jitter[i];
    const float4* fin = cent + d.c0 + level_off(L - 1);
    const float2 wn = staged ? near_field(s_tile, x0, y0, w, d, fqx, fqy, i, p, r, slot_idx, slot_pos)
                             : near_field(fin, 0, 0, gf, d, fqx, fqy, i, p, r, slot_idx, slot_pos);
    mb_finish(d, i, p, r, fx, fy, wn, flags, pos0, pos1, vout);
}

}  // namespace

// Digits of at most RB_MAXW bits, as few passes as cover key_range, for n pairs.
int radix_alloc(Ctx* c, RadixSort& rs, uint32_t n, uint64_t key_range, const std::vector<uint4>* seg_blk,
                const std::vector<uint32_t>* seg_key0) {
    int rc = CG_OK;
    int bits = 1;
    while (bits < 32 && (1ull << bits) < key_range) ++bits;
    rs.passes = (bits + RB_MAXW - 1) / RB_MAXW;
    rs.width = (bits + rs.passes - 1) / rs.passes;
    rs.n = n;
    if (seg_blk && !seg_blk->empty()) {
        rs.nb = (uint32_t)seg_blk->size();
        rs.blk = (uint4*)dalloc(c, seg_blk->size() * 16, &rc);
        rs.key0 = (uint32_t*)dalloc(c, seg_key0->size() * 4, &rc);
        if (rc) return rc;
        CG_HIP(c, hipMemcpyAsync(rs.blk, seg_blk->data(), seg_blk->size() * 16, hipMemcpyHostToDevice,
                                 c->stream));
        CG_HIP(c, hipMemcpyAsync(rs.key0, seg_key0->data(), seg_key0->size() * 4, hipMemcpyHostToDevice,
                                 c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));  // the host tables are the caller's temporaries
    } else {
        rs.nb = (uint32_t)(((size_t)(n ? n : 1) + RB_TILE - 1) / RB_TILE);
    }
    rs.n_hist = (1u << rs.width) * rs.nb;
    rs.hist = (uint32_t*)dalloc(c, (size_t)rs.n_hist * 4, &rc);
    if (rc) return rc;
    rs.tot = (uint32_t*)dalloc(c, (size_t)(1u << rs.width) * 4, &rc);
    return rc;
}

// single_gf != 0: the set is one graph at offset 0 with that finest grid (the
// plain sort; no host tables are built).
int mb_alloc(Ctx* c, uint32_t single_gf) {
    int rc = CG_OK;
    const size_t np = c->NP ? c->NP : 1;
    const size_t F = c->F ? c->F : CELL_BLOCK, C = c->C ? c->C : 1;
    for (int q = 0; q < 2; ++q) {
        c->rk[q] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
        c->rv[q] = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
    }
    (void)single_gf;
    // the counting sort writes its keys, indices and positions to index 1 of
    // each ping-pong pair (the visiting order and the payload it reads are
    // index 1 of the previous pass, its first step writes index 0)
    c->mb_rs = RadixSort();
    c->mb_rs.passes = 1;
    c->bin_rank = (uint32_t*)dalloc(c, np * 4, &rc); if (rc) return rc;
    c->cell_count = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    c->scan_sums = (uint32_t*)dalloc(c, (F / (256 * SCAN_CPT_MIN) + 2) * 4, &rc); if (rc) return rc;
    c->sorted = c->rv[c->mb_rs.passes & 1];
    for (int q = 0; q < 2; ++q) {
        c->rp[q] = (float2*)dalloc(c, np * 8, &rc);
        if (rc) return rc;
    }
    c->sorted_pos = c->rp[c->mb_rs.passes & 1];  // the sort's position payload
    c->cell_begin = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    c->cell_end = (uint32_t*)dalloc(c, F * 4, &rc); if (rc) return rc;
    const size_t nblk = F / CELL_BLOCK + 1;
    c->cell_occ = (uint8_t*)dalloc(c, nblk, &rc); if (rc) return rc;
    c->cell_dirty = (uint8_t*)dalloc(c, nblk, &rc); if (rc) return rc;
    c->slot_idx = (int32_t*)dalloc(c, F * SLOTS * 4, &rc); if (rc) return rc;
    c->slot_pos = (float2*)dalloc(c, F * SLOTS * 8, &rc); if (rc) return rc;
    // the partial cells at the two ends of each wave of the cell pass (waves of
    // 64 x 4 points at the least)
    c->cell_parts = (uint4*)dalloc(c, (np / 256 + 2) * 2 * sizeof(CellPart), &rc); if (rc) return rc;
    c->sums = (long long*)dalloc(c, C * 3 * sizeof(long long), &rc); if (rc) return rc;
    c->cent = (float4*)dalloc(c, C * sizeof(float4), &rc); if (rc) return rc;
    // the first tick visits the points in index order
    if (c->NP)
        CG_HIP(c, hipMemcpyAsync(c->sorted, c->present, (size_t)c->NP * 4, hipMemcpyDeviceToDevice,
                                 c->stream));
    // every cell starts empty and every block dirty, so the first pass writes
    // every pyramid tile; the previous keys read by its clear all name cell 0
    // (clearing a cell is always safe: the pass rewrites the occupied ones)
    CG_HIP(c, hipMemsetAsync(c->cell_end, 0, F * 4, c->stream));
    CG_HIP(c, hipMemsetAsync(c->cell_count, 0, F * 4, c->stream));
    CG_HIP(c, hipMemsetAsync(c->cell_occ, 0, nblk, c->stream));
    CG_HIP(c, hipMemsetAsync(c->cell_dirty, 1, nblk, c->stream));
    // the first pass's previous keys: a cell of each point's own graph (its
    // first), which names the point's graph to the count and is safe to clear
    if (c->G > 1 && c->NP)
        hipLaunchKernelGGL(k_first_keys, dim3(blocks_for(c->NP, 256)), dim3(256), 0, c->stream, c->present,
                           c->NP, c->gid, c->f0s, c->rk[c->mb_rs.passes & 1]);
    else
        CG_HIP(c, hipMemsetAsync(c->rk[c->mb_rs.passes & 1], 0, np * 4, c->stream));
    return rc;
}

// Sorts (k[0], v[0]) by key into k[passes & 1], v[passes & 1]; every pass is
// stable.  hist0_ready: the first digit's block histogram is already in place.
void radix_run(Ctx* c, const RadixSort& rs, uint32_t* const (&k)[2], uint32_t* const (&v)[2],
               bool hist0_ready, float2* const* pay) {
    const int W = rs.width;
    for (int p = 0; p < rs.passes; ++p) {
        if (p || !hist0_ready)
            hipLaunchKernelGGL(k_radix_up, dim3(rs.nb), dim3(256), 0, c->stream, k[p & 1], rs.n, p * W, W,
                               rs.nb, rs.hist, rs.blk, rs.key0);
        hipLaunchKernelGGL(k_scan_rows, dim3(1u << W), dim3(256), 0, c->stream, rs.hist, rs.nb, rs.tot);
        hipLaunchKernelGGL(k_radix_down, dim3(rs.nb), dim3(256), 0, c->stream, k[p & 1], v[p & 1], rs.n,
                           p * W, W, rs.nb, rs.hist, rs.tot, k[(p + 1) & 1], v[(p + 1) & 1],
                           pay ? pay[p & 1] : nullptr, pay ? pay[(p + 1) & 1] : nullptr, rs.blk, rs.key0);
    }
}

int mb_run(Ctx* c, float2* vout) {
    if (c->NP == 0) return CG_OK;
    const uint32_t np = c->NP, F = c->F;
    prof_begin(c, CG_STAGE_MB_SORT);
    // the previous pass's sorted keys (rk[1]) name the cells to clear before
    // this pass counts into them
    hipLaunchKernelGGL(k_mb_clear, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, c->rk[1], np, c->cell_end,
                       c->cell_dirty, c->cell_count);
    hipLaunchKernelGGL(k_bin_count, dim3(blocks_for(np, CB_TILE)), dim3(256), 0, c->stream, c->sorted, c->rk[1],
                       c->pos[0], c->pos[1], c->blk_graph, c->desc, np, c->cell_count, c->rk[0], c->rv[0], c->rp[0],
                       c->bin_rank,
                       c->cell_occ, c->cell_dirty, c->G == 1 ? 1u : 0u);
    // cells per scan thread: the fewest (4 .. 256) that keep the ranges to ~4096
    uint32_t cpt = SCAN_CPT_MIN;
    while (cpt < SCAN_CPT_MAX && (uint64_t)F > 4096ull * 256u * cpt) cpt *= 2;
    const uint32_t nr = (uint32_t)(((uint64_t)F + 256u * cpt - 1) / (256u * cpt));
    hipLaunchKernelGGL(k_count_sums, dim3(nr), dim3(256), 0, c->stream, c->cell_count, F, cpt, c->cell_occ,
                       c->scan_sums);
    // up to COUNT_OWN_PREFIX ranges each bounds block adds up the sums before it
    const uint32_t own = nr <= COUNT_OWN_PREFIX ? 1u : 0u;
    if (!own) hipLaunchKernelGGL(k_count_scan, dim3(1), dim3(256), 0, c->stream, c->scan_sums, nr);
    hipLaunchKernelGGL(k_count_bounds, dim3(nr), dim3(256), 0, c->stream, c->cell_count, F, cpt, c->cell_occ,
                       c->scan_sums, own, c->cell_begin, c->cell_end);
    hipLaunchKernelGGL(k_bin_place, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, c->rk[0], c->rv[0],
                       c->rp[0], c->bin_rank, np, c->cell_begin, c->rk[1], c->rv[1], c->rp[1]);
    prof_end(c, CG_STAGE_MB_SORT);
    prof_begin(c, CG_STAGE_MB_CELLS);
    // pyramid workgroups: one per tile up to ~4096 of them, then `span` tiles each
    auto span_for = [](uint32_t items) {
        uint32_t sp = 1;
        while (sp * 2 <= items / 4096u) sp *= 2;
        return sp;
    };
    // the cells from runs of the sorted points: one wave per 64 N of them,
    // then the merge of the runs that cross a wave's end
    {
        const uint32_t items = c->tune.cell_items;
        const uint32_t nw = (np + 64u * items - 1) / (64u * items);
        CellPart* parts = reinterpret_cast<CellPart*>(c->cell_parts);
        const uint32_t one = c->G == 1 ? 1u : 0u;
        if (items == 4) {
            hipLaunchKernelGGL(k_cells_seg<4>, dim3((nw + 3) / 4), dim3(256), 0, c->stream, c->rk[1], c->sorted,
                               c->sorted_pos, np, c->blk_graph, c->desc, one, c->sums, c->slot_idx, c->slot_pos,
                               parts);
            hipLaunchKernelGGL(k_cells_join<4>, dim3((nw + 3) / 4), dim3(256), 0, c->stream, parts, nw,
                               c->cell_begin, c->blk_graph, c->desc, c->sums, c->slot_idx, c->slot_pos);
        } else {
            hipLaunchKernelGGL(k_cells_seg<8>, dim3((nw + 3) / 4), dim3(256), 0, c->stream, c->rk[1], c->sorted,
                               c->sorted_pos, np, c->blk_graph, c->desc, one, c->sums, c->slot_idx, c->slot_pos,
                               parts);
            hipLaunchKernelGGL(k_cells_join<8>, dim3((nw + 3) / 4), dim3(256), 0, c->stream, parts, nw,
                               c->cell_begin, c->blk_graph, c->desc, c->sums, c->slot_idx, c->slot_pos);
        }
    }
    prof_end(c, CG_STAGE_MB_CELLS);
    prof_begin(c, CG_STAGE_MB_PYRAMID);
    // pyramid: finest -> tiles of up to 32x32 -> remaining top levels
    if (c->n_pyr[0]) {
        const uint32_t pspan = span_for(c->n_pyr[0]);
        hipLaunchKernelGGL(k_pyramid, dim3((c->n_pyr[0] + pspan - 1) / pspan), dim3(256), 0, c->stream, c->pyr[0],
                           c->n_pyr[0], pspan, c->desc, c->cell_end, c->sums, c->cent, 1, c->cell_occ,
                           c->cell_dirty);
    }
    if (c->n_pyr[1])
        hipLaunchKernelGGL(k_pyramid, dim3(c->n_pyr[1]), dim3(256), 0, c->stream, c->pyr[1], c->n_pyr[1], 1u,
                           c->desc, c->cell_end, c->sums, c->cent, 0, c->cell_occ, c->cell_dirty);
    prof_end(c, CG_STAGE_MB_PYRAMID);
    prof_begin(c, CG_STAGE_MB_FORCE);
    if (c->tune.force_tile == 512)
        hipLaunchKernelGGL(k_mb_force<512>, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, c->sorted,
                           c->rk[1], c->sorted_pos, c->cent, c->slot_idx, c->slot_pos, c->jitter, c->flags,
                           c->blk_graph, c->desc, np, c->pos[0], c->pos[1], vout, c->G == 1 ? 1u : 0u);
    else
        hipLaunchKernelGGL(k_mb_force<1024>, dim3(blocks_for(np, 256)), dim3(256), 0, c->stream, c->sorted,
                           c->rk[1], c->sorted_pos, c->cent, c->slot_idx, c->slot_pos, c->jitter, c->flags,
                           c->blk_graph, c->desc, np, c->pos[0], c->pos[1], vout, c->G == 1 ? 1u : 0u);
    prof_end(c, CG_STAGE_MB_FORCE);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Pyramid work lists for one graph slot (stage 0: finest level in tiles of up
// to 32x32; stage 1: the remaining top levels of graphs finer than 128^2).
void pyramid_work(uint32_t slot, uint32_t gf, uint32_t L, uint4* w0, uint32_t* n0, uint4* w1, uint32_t* n1) {
    const uint32_t lf = L - 1;
    const uint32_t T1 = std::min<uint32_t>(32u, gf / 4);
    uint32_t nout1 = 0;
    while ((1u << nout1) < T1) ++nout1;
    const uint32_t tiles = (gf / T1) * (gf / T1);  // at most (512 / 32)^2 = PYR_W0_MAX
    for (uint32_t t = 0; t < tiles; ++t) w0[t] = make_uint4(slot, lf, T1 | (nout1 << 16), t);
    *n0 = tiles;
    *n1 = 0;
    const uint32_t g2 = gf / T1;
    if (g2 > 4) {
        const uint32_t lin2 = lf - nout1, T2 = g2 / 4;
        uint32_t nout2 = 0;
        while ((1u << nout2) < T2) ++nout2;
        for (uint32_t t = 0; t < 16; ++t) w1[t] = make_uint4(slot, lin2, T2 | (nout2 << 16), t);
        *n1 = 16;
    }
}

void pyramid_work_host(uint32_t slot, uint32_t gf, uint32_t L, std::vector<uint4>& w0,
                       std::vector<uint4>& w1) {
    uint4 a[PYR_W0_MAX], b[PYR_W1_MAX];
    uint32_t na = 0, nb = 0;
    pyramid_work(slot, gf, L, a, &na, b, &nb);
    w0.insert(w0.end(), a, a + na);
    w1.insert(w1.end(), b, b + nb);
}

}  // namespace cg
