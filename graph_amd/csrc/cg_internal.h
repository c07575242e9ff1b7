// cg_internal.h — shared state and device helpers of the force-simulation library.
//
// One context holds a set of graphs packed into shared device arrays: points of
// graph g occupy [p0, p0 + n), its present points [k0, k0 + np) of the present
// list, its finest cells [f0, f0 + gf^2) of the finest-cell index space and its
// pyramid levels a block of the level arena starting at c0.  Every kernel looks
// up its graph's descriptor, so one launch per pass serves every graph, and a
// graph computes the same bits alone or packed with others.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <map>
#include <vector>

#include "../../include/cosmosgl.h"

namespace cg {

// Per-point flag bits (device byte array `flags`).
constexpr uint8_t F_ABSENT = 1;   // NaN position at set time (Points/index.ts:1313-1386)
constexpr uint8_t F_PINNED = 2;   // setPinnedPoints (Points/index.ts:1264-1304)

constexpr int MAX_LEVELS = 8;     // 4^2 .. 512^2 (ForceManyBody/index.ts:19, 646-656)
constexpr int SLOTS = 8;          // NEAR_FIELD_SLOTS (ForceManyBody/index.ts:35)
constexpr uint32_t CELL_BLOCK = 256;  // each graph's finest cells start on a multiple of this

// Link degree tiers: points with at most Tuning::lane_max entries run one
// lane each; up to Tuning::group_max entries one 64-lane workgroup each (at
// most 4 entries per lane); above that 256-lane workgroups of at most
// Tuning::hub_entries entries per lane: a point above 256 * hub_entries
// entries is cut into segments of at most that many.
constexpr uint32_t LANE_SPLIT = 8;  // lane-tier schedule classes: <= 2, <= LANE_SPLIT entries, then the rest
constexpr uint32_t T64_LIMIT = 256; // 64 lanes x 4 entries
constexpr int HUB_ROUNDS = 5;        // segment/fix-up rounds per pass for split points

// Free implementation parameters (cg_tuning in the header; chosen by the
// autotuner).  None changes the discretisation; the first three do not change
// any result bit, the link tiers move chains between the serial and the
// multiple-shooting solvers (results within the force tolerance).
struct Tuning {
    uint32_t graph_ticks = 8;      // ticks per captured HIP graph
    uint32_t cell_items = 8;       // sorted points per lane of the cell pass (4 or 8)
    uint32_t force_tile = 1024;    // staged cells per level in the force kernel (512 or 1024)
    uint32_t lane_max = 64;        // link entries per direction up to which a point runs one lane
    uint32_t group_max = 256;      // up to which it runs one 64-lane unit (<= T64_LIMIT)
    uint32_t hub_entries = 6;      // entries per lane of 256-lane units (4, 6 or 8)
};

// Passes of one tick, in order (src/index.ts:2012-2078).  Bit p of GDesc::passes
// says the graph runs pass p this tick; bit p of GDesc::inbuf names the position
// buffer (0/1) holding the graph's positions when pass p starts.
enum Pass : uint32_t {
    P_MOUSE = 0,
    P_GRAVITY = 1,
    P_CENTER = 2,
    P_MANYBODY = 3,
    P_LINK_IN = 4,
    P_LINK_OUT = 5,
    P_CLUSTER = 6,
    P_COLLISION = 7,
    P_COUNT = 8
};

// Level l of a graph's pyramid starts this many cells after c0 (levels 4^2, 8^2, ...).
__host__ __device__ inline uint32_t level_off(uint32_t l) { return 16u * ((1u << (2 * l)) - 1u) / 3u; }

// Per-graph descriptor, one per graph per tick (device table).
struct GDesc {
    uint32_t p0, n;         // packed point range
    uint32_t k0, np;        // packed present-list range
    uint32_t f0;            // first finest cell
    uint32_t c0;            // level arena base
    uint32_t gf, L;         // finest grid size, level count
    uint32_t passes;        // passes run this tick
    uint32_t inbuf;         // input position buffer per pass
    uint32_t col_c0;        // collision cell base (4 grids of col_G^2)
    uint32_t col_G;
    float h[MAX_LEVELS];    // S / (4 << l) as f32
    float S;
    float alpha;            // f32 copy of this tick's alpha
    float seed;             // near-field seed of this tick
    float ak;               // alpha * repulsion
    float friction;
    float gravity, center;
    float mouse_k, mx, my;
    float spring_alpha, dist, lo, span;
    float tol;              // multiple-shooting tolerance (2.5e-6 S)
    float col_strength, col_radius, col_padding, col_h;
    float fx_scale, fx_inv;  // 2^k and 2^-k: positions as int64 multiples of 2^-k in cell sums
    float cluster_k;         // simulationCluster
};

// Per-graph link CSR built once per cg_set_links, in the reference's entry
// order with graph-local indices (ForceLink/index.ts:40-83).
struct GraphLinks {
    uint32_t E[2] = {0, 0};
    uint4* ent_pt[2] = {nullptr, nullptr};   // {local neighbour, draw, strength, bias}
    uint32_t* off_pt[2] = {nullptr, nullptr};// [n+1]
    uint32_t* live[2] = {nullptr, nullptr};  // [n] entries toward present neighbours
    uint32_t* lrank[2] = {nullptr, nullptr}; // [E+1] live entries before each entry
    uint32_t* degree = nullptr;              // [n]
};

struct Graph {
    uint32_t id = 0;
    bool live = false;
    uint32_t n = 0, np = 0, m = 0;
    cg_sim_params params{};
    double alpha = 1.0;
    uint64_t rng = 0x5eed5eedULL;
    // host copies
    std::vector<float> xy;                   // positions as set (absent = NaN), until packed
    bool xy_pending = false;
    float extent = 0.f;                      // largest |coordinate| ever set or dragged (present points)
    std::vector<uint8_t> flags;
    std::vector<uint32_t> pinned;
    std::vector<float> jitter;
    std::vector<float> links, strength, rin, rout;
    bool links_set = false;
    std::vector<float> sizes;                // raw sizes (NaN = default); empty = none
    std::vector<int32_t> clusters;           // cluster per point (< 0: none); empty = none set
    std::vector<float> cluster_pos;          // [x, y] per cluster (< 0: centroid); empty = none set
    std::vector<float> cluster_strength;     // per point; empty = all 1
    bool has_clusters = false, has_cluster_pos = false;  // setPointClusters / setClusterPositions
    // derived
    uint32_t gf = 0, L = 0;
    GraphLinks lk;
    bool links_ready = false;
    uint32_t col_G = 0;
    float col_h = 0.f;
    bool col_dirty = true;
    std::vector<float> col_size;             // resolved sizes
    // device copies of the graph's per-point inputs, staged when they change;
    // the packer assembles the packed arrays from them on the device
    bool staged = false;
    uint8_t* d_flags = nullptr;              // [n]
    float2* d_jit = nullptr;                 // [n]
    float* d_size = nullptr;                 // [n] resolved collision sizes (collision graphs)
    uint32_t* d_pres = nullptr;              // [np] local indices of the present points
    // placement in the packed arrays (valid after pack)
    uint32_t slot = 0;                       // index in the descriptor table
    uint32_t p0 = 0, k0 = 0, f0 = 0, c0 = 0, col_c0 = 0;
    uint32_t cur = 0;                        // buffer holding current positions
    uint32_t w0[2] = {0, 0}, wn[2] = {0, 0}; // warm-start region per link direction
};

struct LinkDir {
    uint32_t E_live = 0;         // live entries over all graphs
    uint4* ent = nullptr;        // schedule-ordered live entries (global neighbour indices)
    uint32_t* sched = nullptr;   // present points: lane tier in three degree classes (index order), then by degree
    uint32_t* soff = nullptr;    // [n_sched+1]
    uint32_t n_sched = 0;
    uint32_t n_lane = 0;         // lane-tier points (first in the schedule)
    uint32_t* sgraph = nullptr;  // [n_lane] graph slot of each lane-tier point (sets of several graphs)
    // chunked-tier work units, in schedule order: one per 64-lane point, then
    // one per segment of each 256-lane point (one segment up to 256 x hub_entries)
    uint32_t n_units = 0, n_u64 = 0;
    uint32_t first_split_unit = 0;  // split points come last (degree order)
    uint4* unit_a = nullptr;     // {first entry, entries, -, graph}
    uint4* unit_b = nullptr;     // {schedule slot, warm offset, split point or NO_HUB, segment}
    uint2* hubs = nullptr;       // split points: {first unit, segments}
    uint32_t n_hubs = 0;
    float4* seg_out = nullptr;   // [2][2 n_units] by round parity: segment end, start used; end Jacobian
    uint32_t* hub_done = nullptr;  // [3 n_hubs] round that found it holding (0: open); largest corrections by parity (f32)
    float2* warm = nullptr;      // last converged chunk starts, per graph regions
    float4* wjac = nullptr;      // split segments: each lane's prefix Jacobian at that start
                                 // (lane 0: the segment start it converged from)
    float4* terms = nullptr;     // split segments: each lane's gathered terms (q.x, q.y, D, K),
                                 // written by round 0, read by the later rounds of a pass
    size_t warm_len = 0;
};

// A radix sort of (key, value) pairs over ping-pong buffers (k_manybody.hip).
struct RadixSort {
    int passes = 1, width = 8;      // digits of `width` bits cover every key
    uint32_t n = 0, nb = 0;         // pairs, blocks of RB_TILE pairs
    uint32_t n_hist = 0;            // (1 << width) * nb
    uint32_t* hist = nullptr;       // [n_hist] per-pass [digit][block] counts
    uint32_t* tot = nullptr;        // [1 << width] digit totals of the current pass
    // Graph segments (the many-body sort of a packed set, whose input is
    // graph-major): block b sorts positions [blk[b].x, blk[b].y) of its graph,
    // whose blocks are [blk[b].z, blk[b].w), on digits of key - key0[b], so a
    // pass covers the largest graph's key range rather than the whole set's.
    // Null: one segment, blocks of RB_TILE pairs from position 0, key0 = 0.
    uint4* blk = nullptr;
    uint32_t* key0 = nullptr;
};

struct Ctx {
    int device = 0;
    Tuning tune;
    uint8_t* scratch = nullptr;     // this device's static scratch array (api.hip), set up by cg_create
    // arena mode: dalloc carves from [arena, arena + arena_cap) and dfree is a
    // no-op (the allocation-free entry point's buffers)
    uint8_t* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;     // second stream (default priority) for the hub link tier
    hipStream_t ingest = nullptr;   // per-graph link builds (default priority): they run beside ticks
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint64_t budget = 0;
    uint64_t held = 0;              // live, pending and cached bytes (cg_device_bytes)
    std::string err;
    std::vector<std::pair<void*, size_t>> allocs;
    // Freed blocks, kept for reuse instead of hipFree (which synchronises the
    // device and costs ~0.15 ms): `pending` since the last settle (kernels
    // enqueued before the free may still read them), `cache` once the
    // context's streams have been seen idle (reusable).  Both count in held;
    // a budget check or a failed hipMalloc evicts them with hipFree.
    std::vector<std::pair<void*, size_t>> pending;
    std::multimap<size_t, void*> cache;
    uint64_t cached = 0;            // bytes in pending and cache

    std::vector<Graph> graphs;      // by id
    bool dirty = true;              // the graph set changed: repack before the next tick

    // packed sizes
    uint32_t G = 0;                 // live graphs
    uint32_t N = 0, NP = 0;         // points, present points
    uint32_t F = 0, C = 0;          // finest cells, level cells
    uint32_t CG = 0;                // collision cells (all grids)
    std::vector<uint32_t> order;    // live graph ids in packing order (slot -> id)

    // packed point state
    float2* pos[2] = {nullptr, nullptr};
    uint8_t* flags = nullptr;
    uint32_t* present = nullptr;
    uint32_t* gid = nullptr;        // graph slot of every packed point
    float2* jitter = nullptr;
    float* size = nullptr;          // resolved collision sizes
    GDesc* desc = nullptr;          // descriptor table of the tick being launched
    std::vector<GDesc> h_desc;      // placement part of every descriptor, by slot
    uint32_t* f0s = nullptr;        // [G] finest-cell bases
    uint32_t* blk_graph = nullptr;  // [F / CELL_BLOCK] graph slot of each block of finest cells
    uint32_t* colc0s = nullptr;     // [G] collision cell bases

    // many-body: radix sort of present points by finest cell
    uint32_t* rk[2] = {nullptr, nullptr};  // [NP] cell keys, ping-pong over the passes
    uint32_t* rv[2] = {nullptr, nullptr};  // [NP] point indices, ping-pong
    uint32_t* sorted = nullptr;     // rv[passes & 1]: points grouped by cell, next tick's visit order
    RadixSort mb_rs;
    uint32_t* cell_begin = nullptr; // [F] first sorted position of each finest cell
    uint32_t* cell_end = nullptr;   // [F] one past its last (0 = empty)
    // Per block of CELL_BLOCK finest cells (Morton-consecutive): a cell of it is
    // occupied this pass (cell_occ), or it held points this pass or the last one,
    // so its pyramid tile must be rewritten (cell_dirty).  A tile whose blocks
    // are all clean keeps last pass's (empty) centroids.
    uint8_t* cell_occ = nullptr;    // [F / CELL_BLOCK]
    uint8_t* cell_dirty = nullptr;  // [F / CELL_BLOCK]
    uint32_t* cell_count = nullptr; // [F] points per finest cell this pass (0 outside it)
    uint32_t* bin_rank = nullptr;   // [NP] each point's rank in its cell, in visiting order
    uint32_t* scan_sums = nullptr;  // [F / 1024 + 2] per-range counts of the count scan, then their prefixes
    long long* sums = nullptr;      // [3 C] fixed-point sum x, sum y, count per level cell
    float4* cent = nullptr;         // [C]
    float2* rp[2] = {nullptr, nullptr};  // [NP] position payload of the radix sort
    float2* sorted_pos = nullptr;   // [NP] positions in cell order (one of rp)
    int32_t* slot_idx = nullptr;    // [8 F] global point index or -1
    float2* slot_pos = nullptr;
    uint4* cell_parts = nullptr;    // [2 per wave of the cell pass] partial cells at a wave's two ends
    uint4* pyr[2] = {nullptr, nullptr};  // pyramid work lists {graph slot, level, tile, levels}
    uint32_t n_pyr[2] = {0, 0};
    double* center_part = nullptr;  // [2 chunks] per-chunk partial sums
    double* center_sum = nullptr;   // [3 G]
    uint4* center_chunks = nullptr; // {k begin, k end, graph slot, -} per chunk
    uint32_t* center_first = nullptr;  // [G + 1] first chunk of each graph
    uint32_t n_center_chunks = 0;

    // collision (graphs with simulationCollision on)
    uint32_t NPC = 0;               // their present points
    uint32_t* cpresent = nullptr;   // [NPC]
    uint32_t* col_keys[2] = {nullptr, nullptr};  // [NPC] base-grid cell keys, ping-pong
    uint32_t* col_vals[2] = {nullptr, nullptr};  // [NPC] point indices
    float2* col_pos[2] = {nullptr, nullptr};     // [NPC] positions riding the sort
    float* col_ssize = nullptr;     // [NPC] sizes in sorted order
    RadixSort col_rs;
    uint32_t* col_begin = nullptr;  // [CG / 4] per base cell
    uint32_t* col_end = nullptr;
    float4* col_cell = nullptr;     // [CG] the four grids' cell aggregates
    uint32_t* col_crowd = nullptr;  // [1 + cap]: count, then the cells deferred to one block each
    uint32_t col_crowd_cap = 0;

    // clusters: point -> packed cluster, per-cluster targets, present members by cluster
    uint32_t CC = 0;                // packed clusters
    int32_t* pcl = nullptr;         // [N] packed cluster of each point, -1 = none
    float* pcs = nullptr;           // [N] cluster strength of each point
    float2* cpos = nullptr;         // [CC] custom position (a negative coordinate: centroid)
    uint32_t* cl_members = nullptr; // present clustered points grouped by cluster, index order
    uint4* cl_chunks = nullptr;     // {member begin, member end, cluster, graph slot}
    uint32_t* cl_first = nullptr;   // [CC + 1] first chunk of each cluster
    uint32_t n_cl_chunks = 0;
    double* cl_part = nullptr;      // [2 chunks] per-chunk sums
    double* cl_sum = nullptr;       // [3 CC] sum x, sum y, count

    // links
    LinkDir dir[2];
    bool any_links = false;
    uint32_t* link_stats = nullptr;  // [16] sweep counters (cg_debug_link_stats)

    // stage timing (cg_set_profiling)
    uint32_t prof_mask = 0;
    std::vector<hipEvent_t> prof_pool;
    size_t prof_used = 0;
    hipEvent_t prof_open[CG_STAGES] = {};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_rec;

    // per-tick descriptor tables: device [desc_ticks * G], pinned host staging of
    // the same size, and the event after the last upload from that staging
    GDesc* desc_base = nullptr;
    GDesc* h_desc_pinned = nullptr;
    uint32_t desc_ticks = 0;
    hipEvent_t ev_desc = nullptr;
    bool desc_pending = false;
    // tick q of a launch of n ticks reads window slot q (desc_win + q G),
    // copied from the ring before the launch
    GDesc* desc_win = nullptr;

    // staging of cg_drag_points / cg_get_positions_at (grown on demand)
    uint32_t* io_idx = nullptr;
    float2* io_xy = nullptr;
    uint32_t io_cap = 0;

    // captured tick sequences: {passes run, ticks} -> executable graph
    bool capture = true;
    std::vector<std::pair<std::pair<uint32_t, uint32_t>, hipGraphExec_t>> graphs_exec;
};

void prof_begin(Ctx* c, int stage);
void prof_end(Ctx* c, int stage);

// error helpers (api.hip)
int fail(Ctx* c, int code, const std::string& msg);
int hip_fail(Ctx* c, hipError_t e, const char* what);
void* dalloc(Ctx* c, size_t bytes, int* rc);
void dfree(Ctx* c, void* p);
void dfree_idle(Ctx* c, void* p);  // a block no queued work uses: reusable at once
void cache_settle(Ctx* c);     // synchronises the streams; freed blocks become reusable

#define CG_HIP(c, expr)                                  \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return cg::hip_fail(c, e_, #expr); \
    } while (0)

// Launchers.  Input and output position buffers come from GDesc::inbuf; vout
// (non-null: evaluation mode) receives the pass's velocity instead of integrating.
int pw_run(Ctx* c, int pass, float2* vout);                       // k_pointwise.hip
int center_run(Ctx* c);
int cluster_run(Ctx* c, float2* vout);
void center_chunks_host(const std::vector<uint32_t>& k0, const std::vector<uint32_t>& np,
                        std::vector<uint4>& chunks, std::vector<uint32_t>& first);
int radix_alloc(Ctx* c, RadixSort& rs, uint32_t n, uint64_t key_range,
                const std::vector<uint4>* seg_blk = nullptr,
                const std::vector<uint32_t>* seg_key0 = nullptr);  // k_manybody.hip
void radix_run(Ctx* c, const RadixSort& rs, uint32_t* const (&k)[2], uint32_t* const (&v)[2],
               bool hist0_ready, float2* const* pay = nullptr);
int mb_alloc(Ctx* c, uint32_t single_gf = 0);
int mb_run(Ctx* c, float2* vout);
void pyramid_work_host(uint32_t slot, uint32_t gf, uint32_t L, std::vector<uint4>& w0,
                       std::vector<uint4>& w1);
constexpr uint32_t PYR_W0_MAX = 256, PYR_W1_MAX = 16;  // work items per graph of each stage
void pyramid_work(uint32_t slot, uint32_t gf, uint32_t L, uint4* w0, uint32_t* n0, uint4* w1, uint32_t* n1);
int link_build_graph(Ctx* c, Graph& g);                           // k_link.hip
void link_free_graph(Ctx* c, Graph& g);
int link_pack(Ctx* c, const LinkDir (&old)[2]);
// Copies every segment {source offset (bit 31: srcB), destination offset,
// count, -} of `segs` into dst0 (and dst1 unless null) in one launch on the
// library stream; `segs` must stay alive until the stream has run it.
template <typename T>
int seg_copy(Ctx* c, const T* srcA, const T* srcB, T* dst0, T* dst1, std::vector<uint4>& segs);
void link_free_dir(Ctx* c, LinkDir& L);
int link_run(Ctx* c, int dir, float2* vout);
int col_prepare_graph(Ctx* c, Graph& g);                          // k_collision.hip
int col_run(Ctx* c, float2* vout);
int col_alloc(Ctx* c);
int col_point_cells(Ctx* c, uint32_t* out);                       // [4 NPC] cells per point, per grid
constexpr uint32_t RADIX_TILE = 2048;  // pairs per block of the radix sort (k_manybody.hip)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

__host__ __device__ inline uint32_t part1by1(uint32_t x) {
    x &= 0x0000ffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__host__ __device__ inline uint32_t compact1by1(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0f0f0f0fu;
    x = (x | (x >> 4)) & 0x00ff00ffu;
    x = (x | (x >> 8)) & 0x0000ffffu;
    return x;
}
__host__ __device__ inline uint32_t morton2(uint32_t x, uint32_t y) {
    return part1by1(x) | (part1by1(y) << 1);
}

// clamp(int(floor(p / h)), 0, g-1) with an IEEE f32 division — the cell formula
// of calculate-level.vert:47-48 / force-level.frag:66 / build-grid.vert:50-55.
__device__ inline int cell_of(float p, float h, int g) {
    float f = floorf(__fdiv_rn(p, h));
    if (!(f >= 0.f)) return 0;
    if (f >= (float)(g - 1)) return g - 1;
    return (int)f;
}

// lowbias32 and the near-field ticket bits — build-nearfield-slots.vert:69-79.
// `index` is the point's index within its graph.
__device__ inline uint32_t ticket_bits(uint32_t index, float seed) {
    uint32_t h = index ^ __float_as_uint(seed);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    float t = (float)(h >> 8);
    float a = __fadd_rn(t, 0.5f);
    float b = __fdiv_rn(a, 16777216.0f);
    float c = __fmul_rn(b, 0.998f);
    float d = __fadd_rn(0.001f, c);
    return __float_as_uint(d);
}

// Integrate one point — update-position.frag:51-57: v *= friction; p += v;
// clamp to [0, S].  Separate roundings (no contraction), as the shader states them.
__device__ inline float2 integrate(float2 p, float2 v, float friction, float S) {
    float vx = __fmul_rn(v.x, friction), vy = __fmul_rn(v.y, friction);
    float x = __fadd_rn(p.x, vx), y = __fadd_rn(p.y, vy);
    x = fminf(fmaxf(x, 0.f), S);
    y = fminf(fmaxf(y, 0.f), S);
    return make_float2(x, y);
}

// Gravity velocity at p — force-gravity.frag:31-38: alpha*g*|d|*0.1 along d
// toward the centre of the space; the atan/cos/sin round trip is d/|d|.
__device__ inline float2 gravity_velocity(float2 p, const GDesc& d) {
    float2 v = make_float2(0.f, 0.f);
    float dx = d.S * 0.5f - p.x, dy = d.S * 0.5f - p.y;
    float dist = sqrtf(dx * dx + dy * dy);
    if (dist > 0.f) {
        float add = d.alpha * d.gravity * dist * 0.1f;
        v.x = add * (dx / dist);
        v.y = add * (dy / dist);
    }
    return v;
}

// Largest j in [0, G) with base[j] <= x (base ascending, base[0] == 0).
__device__ inline uint32_t find_graph(const uint32_t* __restrict__ base, uint32_t G, uint32_t x) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

inline unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
// workgroup i runs on XCD i % 8.  This maps i to a logical block so that XCD x
// runs one contiguous x-th share of [0, nb): blocks of nearby data then share
// an L2 instead of every L2 fetching the whole working set.
constexpr uint32_t XCDS = 8;
__device__ inline uint32_t xcd_block(uint32_t i, uint32_t nb) {
    const uint32_t q = nb / XCDS, r = nb % XCDS, x = i % XCDS, j = i / XCDS;
    return x * q + min(x, r) + j;
}

// The same in runs of C blocks: of each group of XCDS * C consecutive logical
// blocks, XCD x runs the x-th run of C, so every XCD keeps a share of each
// part of [0, nb) (work ordered by cost stays balanced) while nearby blocks
// still share an L2.  A last partial group keeps the round-robin order.
__device__ inline uint32_t xcd_run_block(uint32_t i, uint32_t nb, uint32_t C) {
    const uint32_t full = nb / (XCDS * C) * (XCDS * C);
    if (i >= full) return i;
    const uint32_t x = i % XCDS, j = i / XCDS;
    return ((j / C) * XCDS + x) * C + j % C;
}

}  // namespace cg
