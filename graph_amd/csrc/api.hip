// api.hip — the extern "C" boundary: context lifetime, per-graph ingest, the
// packer that lays every live graph out in the shared device arrays, the tick
// driver and the inspection entry points.
//
// The tick driver restates Graph.runSimulationStep (src/index.ts:2006-2099):
// [mouse -> integrate] -> gravity -> center -> many-body -> link IN ->
// link OUT -> collision, each force followed by its own integrate pass (fused
// into the force kernel here), then the alpha decay of Store.addAlpha
// (Store/index.ts:328-332).  Positions ping-pong between two device buffers;
// each graph records which buffer holds its current positions, so graphs that
// skip a pass (no links, collision off, not stepped) stay consistent.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <new>

#include "cg_internal.h"

struct cg_ctx {
    cg::Ctx c;
};

namespace cg {

// Static scratch of the allocation-free entry point (cg_eval_manybody_noalloc):
// every device the library runs on holds one copy until the process ends.  In
// host code `scratch` is only a handle for hipGetSymbolAddress; the device
// address of each device's copy is looked up once per device, with that device
// current, when a context is created on it, and used from then on.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[MAX_DEVICES] = {};

// The calling thread's current device must be `device`.
int scratch_setup(Ctx* c, int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(c, CG_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = (uint8_t*)ptr;
    }
    c->scratch = g_scratch_ptr[device];
    return CG_OK;
}

int fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
    return fail(c, CG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocation charged against the context's HBM budget.  Sets *rc only
// on failure, so a chain of calls can share one status.
// Bytes a context keeps in freed-but-cached blocks beyond which frees go
// straight to hipFree.
constexpr uint64_t CACHE_MAX = 8ull << 30;

namespace {
// hipFree of one cached block (largest first), or false when none is left;
// pending blocks only with `pending_too` (hipFree waits for the device).
bool evict_one(Ctx* c, bool pending_too) {
    if (!c->cache.empty()) {
        auto it = std::prev(c->cache.end());
        (void)hipFree(it->second);
        c->held -= it->first;
        c->cached -= it->first;
        c->cache.erase(it);
        return true;
    }
    if (pending_too && !c->pending.empty()) {
        (void)hipFree(c->pending.back().first);
        c->held -= c->pending.back().second;
        c->cached -= c->pending.back().second;
        c->pending.pop_back();
        return true;
    }
    return false;
}
}  // namespace

void* dalloc(Ctx* c, size_t bytes, int* rc) {
    if (c->arena) {  // carve from the static scratch: 256-byte aligned, never freed
        const size_t off = (c->arena_used + 255) & ~(size_t)255;
        if (off + (bytes ? bytes : 16) > c->arena_cap) {
            *rc = fail(c, CG_ERR_INVALID, "the input needs more than the " + std::to_string(c->arena_cap) +
                                              "-byte static scratch");
            return nullptr;
        }
        c->arena_used = off + (bytes ? bytes : 16);
        return c->arena + off;
    }
    const size_t want = ((bytes ? bytes : 16) + 255) & ~(size_t)255;
    // a cached block of at least the size, at most twice it (+64 KiB for small ones)
    auto it = c->cache.lower_bound(want);
    if (it != c->cache.end() && it->first <= std::max<size_t>(2 * want, want + (64u << 10))) {
        void* p = it->second;
        const size_t sz = it->first;
        c->cache.erase(it);
        c->cached -= sz;
        c->allocs.emplace_back(p, sz);
        return p;
    }
    while (c->budget && c->held + want > c->budget && evict_one(c, true)) {
    }
    if (c->budget && c->held + want > c->budget) {
        *rc = fail(c, CG_ERR_BUDGET, "device allocation of " + std::to_string(bytes) +
                                         " bytes exceeds the context budget");
        return nullptr;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        while (evict_one(c, true)) {
        }
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        *rc = hip_fail(c, e, ("hipMalloc of " + std::to_string(want) + " bytes (" + std::to_string(c->held) +
                              " held)").c_str());
        return nullptr;
    }
    c->held += want;
    c->allocs.emplace_back(p, want);
    return p;
}

namespace {
// Returns a block to the cache: pending until the streams have been
// synchronised, or reusable at once when no queued work uses it (`idle`).
void release_block(Ctx* c, void* p, bool idle) {
    if (!p || c->arena) return;
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        if (c->allocs[k].first == p) {
            const size_t sz = c->allocs[k].second;
            c->allocs[k] = c->allocs.back();
            c->allocs.pop_back();
            if (c->cached + sz <= CACHE_MAX) {
                if (idle) c->cache.emplace(sz, p);
                else c->pending.emplace_back(p, sz);
                c->cached += sz;
            } else {
                c->held -= sz;
                (void)hipFree(p);
            }
            return;
        }
    }
    (void)hipFree(p);
}
}  // namespace

void dfree(Ctx* c, void* p) { release_block(c, p, false); }

void dfree_idle(Ctx* c, void* p) { release_block(c, p, true); }

void cache_settle(Ctx* c) {
    if (c->pending.empty()) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ingest) (void)hipStreamSynchronize(c->ingest);
    for (auto& b : c->pending) c->cache.emplace(b.second, b.first);
    c->pending.clear();
}

void prof_begin(Ctx* c, int stage) {
    if (!(c->prof_mask & (1u << stage))) return;
    if (c->prof_used == c->prof_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        c->prof_pool.push_back(e);
    }
    hipEvent_t e = c->prof_pool[c->prof_used++];
    (void)hipEventRecord(e, c->stream);
    c->prof_open[stage] = e;
}

void prof_end(Ctx* c, int stage) {
    if (!(c->prof_mask & (1u << stage)) || !c->prof_open[stage]) return;
    if (c->prof_used == c->prof_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        c->prof_pool.push_back(e);
    }
    hipEvent_t e = c->prof_pool[c->prof_used++];
    (void)hipEventRecord(e, c->stream);
    c->prof_rec.push_back({stage, {c->prof_open[stage], e}});
    c->prof_open[stage] = nullptr;
}

namespace {

__global__ void k_set_word(uint32_t* __restrict__ a, uint32_t v) { *a = v; }

// drag-point.frag:28-31: listed points (graph-local indices of present points,
// filtered on the host) take new positions.
__global__ void k_drag(float2* __restrict__ pos, uint32_t p0, const uint32_t* __restrict__ idx,
                       const float2* __restrict__ xy, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    pos[p0 + idx[k]] = xy[k];
}

// Tracked positions: NaN past the point count (absent points hold NaN already).
__global__ void k_gather(const float2* __restrict__ pos, uint32_t p0, uint32_t n,
                         const uint32_t* __restrict__ idx, uint32_t count, float2* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    out[k] = idx[k] < n ? pos[p0 + idx[k]] : make_float2(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000));
}

// Present list, flags and graph slots of one graph at offset 0 (the
// allocation-free entry point): one block; thread t scans a contiguous chunk,
// a block scan of the chunk counts places each present point in index order.
// A point with a NaN coordinate is absent (position-utils.ts:17-24).
__global__ void __launch_bounds__(1024) k_present_list(const float2* __restrict__ pos, uint32_t n,
                                                       uint32_t* __restrict__ present,
                                                       uint8_t* __restrict__ flags,
                                                       uint32_t* __restrict__ gid) {
    __shared__ uint32_t s[1024];
    const uint32_t t = threadIdx.x, chunk = (n + 1023) / 1024;
    const uint32_t b = min(n, t * chunk), e = min(n, b + chunk);
    uint32_t cnt = 0;
    for (uint32_t i = b; i < e; ++i) {
        const float2 p = pos[i];
        const bool absent = isnan(p.x) || isnan(p.y);
        flags[i] = absent ? F_ABSENT : 0;
        gid[i] = 0;
        cnt += absent ? 0u : 1u;
    }
    s[t] = cnt;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t w = s[t] - cnt;
    for (uint32_t i = b; i < e; ++i) {
        const float2 p = pos[i];
        if (!(isnan(p.x) || isnan(p.y))) present[w++] = i;
    }
}

// A graph's staged per-point inputs and its place in the packed arrays.
struct GStage {
    const uint8_t* flags;
    const float2* jit;
    const float* size;        // null: no collision (size 0)
    const uint32_t* pres;
    uint32_t p0, k0, np, pad;
};

// The packed per-point arrays from every graph's staged inputs: point i
// belongs to the graph whose range [p0, p0 + n) holds it (p0s ascending).
__global__ void k_pack_points(const GStage* __restrict__ tab, const uint32_t* __restrict__ p0s, uint32_t G,
                              uint32_t N, uint8_t* __restrict__ flags, uint32_t* __restrict__ gid,
                              float2* __restrict__ jit, float* __restrict__ size) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t s = find_graph(p0s, G, i);
    const GStage& t = tab[s];
    const uint32_t j = i - t.p0;
    flags[i] = t.flags[j];
    gid[i] = s;
    jit[i] = t.jit[j];
    size[i] = t.size ? t.size[j] : 0.f;
}

// A present list: entry k belongs to the table entry whose [k0, k0 + np)
// holds it (k0s ascending); global index = p0 + the staged local index.
__global__ void k_pack_present(const GStage* __restrict__ tab, const uint32_t* __restrict__ k0s, uint32_t G,
                               uint32_t NP, uint32_t* __restrict__ present) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NP) return;
    const GStage& t = tab[find_graph(k0s, G, k)];
    present[k] = t.p0 + t.pres[k - t.k0];
}

constexpr uint64_t GRAPH_SEED = 0x5eed5eedULL;   // every graph's generator starts here
constexpr uint32_t DESC_ENTRIES = 1u << 16;      // descriptors per upload (ticks x graphs)
constexpr uint32_t DESC_MAX_TICKS = 256;

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
float rng_unit(Graph& g) { return (float)(splitmix(g.rng) >> 40) * 0x1.0p-24f; }

// Pyramid geometry — ForceManyBody/index.ts:646-656: finest = the smallest
// power of two >= 2*sqrt(n), clamped to [8, 512]; levels = log2(finest) - 1.
void geometry(uint32_t n, uint32_t* gf, uint32_t* L) {
    uint32_t p = 1;
    while ((uint64_t)(p / 2) * (p / 2) < n && p < (1u << 30)) p <<= 1;
    if (n == 0) p = 0;
    p = std::min<uint32_t>(512u, std::max<uint32_t>(8u, p));
    uint32_t lg = 0;
    while ((1u << lg) < p) ++lg;
    *gf = p;
    *L = lg - 1;
}

bool on(float v) { return v != 0.f && !std::isnan(v); }  // JS truthiness of a number

bool same_bits(float a, float b) {
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x == y;
}

// The inputs that shape a graph's collision grid or its membership in the
// collision arrays; a change repacks.
bool same_collision(const cg_sim_params& a, const cg_sim_params& b) {
    return on(a.collision) == on(b.collision) && same_bits(a.collision_radius, b.collision_radius) &&
           same_bits(a.collision_padding, b.collision_padding) &&
           same_bits(a.point_default_size, b.point_default_size) && same_bits(a.space_size, b.space_size);
}

bool collides(const Graph& g) { return on(g.params.collision) && g.np > 0; }

bool clustered(const Graph& g) { return g.has_clusters || g.has_cluster_pos; }

// Clusters/index.ts:60-64: the highest defined non-negative index plus one (at least 1).
uint32_t cluster_count(const Graph& g) {
    int32_t mx = 0;
    for (uint32_t i = 0; i < g.n && i < g.clusters.size(); ++i) mx = std::max(mx, g.clusters[i]);
    return (uint32_t)mx + 1;
}

// Finest cells a graph occupies in the packed cell arrays (rounded up to whole blocks).
uint64_t pad_cells(uint32_t gf) {
    const uint64_t n = (uint64_t)gf * gf;
    return (n + CELL_BLOCK - 1) / CELL_BLOCK * CELL_BLOCK;
}

Graph* lookup(Ctx* c, uint32_t id) {
    if (id >= c->graphs.size() || !c->graphs[id].live) return nullptr;
    return &c->graphs[id];
}

void reset_graph(Graph& g, uint32_t id) {
    g = Graph();
    g.id = id;
    g.live = true;
    g.rng = GRAPH_SEED;
    cg_default_params(&g.params);
}

// Graph 0 backs the single-graph entry points; it joins the set on first use.
Graph& default_graph(Ctx* c) {
    if (c->graphs.empty()) c->graphs.emplace_back();
    Graph& g = c->graphs[0];
    if (!g.live) {
        reset_graph(g, 0);
        c->dirty = true;
    }
    return g;
}

void set_params(Ctx* c, Graph& g, const cg_sim_params& p) {
    if (!same_collision(g.params, p)) {
        g.col_dirty = true;
        c->dirty = true;
    }
    g.params = p;
}

template <typename T>
void release(Ctx* c, T*& p) {
    dfree(c, (void*)p);
    p = nullptr;
}

void drop_graphs(Ctx* c) {
    for (auto& e : c->graphs_exec) (void)hipGraphExecDestroy(e.second);
    c->graphs_exec.clear();
}

// Whether the set's ticks replay captured HIP graphs.  Only below
// CAPTURE_MAX_POINTS: above it a tick's kernels are long enough that launching
// them directly costs nothing measurable (11.5M points: 62.9 against 62.4 ms per
// 16 ticks), and the executable graphs of a large set, destroyed at each
// repack, did not give back the freed packing: with captured ticks every repack
// of a 20 GB serving set (links with split hubs and collision) lost ~7 GB of
// device memory until the device ran out; with direct launches none.
constexpr uint32_t CAPTURE_MAX_POINTS = 8u << 20;
bool captures(const Ctx* c) { return c->capture && c->prof_mask == 0 && c->N < CAPTURE_MAX_POINTS; }

// Frees the packed arrays except the position buffers and the link schedules.
void free_packed(Ctx* c) {
    drop_graphs(c);
    release(c, c->desc_win);
    release(c, c->flags); release(c, c->present); release(c, c->gid); release(c, c->jitter);
    release(c, c->size); release(c, c->desc_base); c->desc = nullptr;
    release(c, c->f0s); release(c, c->colc0s); release(c, c->blk_graph);
    release(c, c->col_ssize);
    for (int b = 0; b < 2; ++b) {
        release(c, c->col_keys[b]); release(c, c->col_vals[b]); release(c, c->col_pos[b]);
        release(c, c->pyr[b]);
        c->n_pyr[b] = 0;
    }
    for (int q = 0; q < 2; ++q) { release(c, c->rk[q]); release(c, c->rv[q]); }
    c->sorted = nullptr;
    release(c, c->mb_rs.hist); release(c, c->mb_rs.tot); release(c, c->mb_rs.blk); release(c, c->mb_rs.key0);
    c->mb_rs = RadixSort();
    release(c, c->col_rs.hist); release(c, c->col_rs.tot); release(c, c->col_rs.blk); release(c, c->col_rs.key0);
    c->col_rs = RadixSort();
    release(c, c->cell_begin); release(c, c->cell_end); release(c, c->cell_occ); release(c, c->cell_dirty);
    release(c, c->cell_count); release(c, c->bin_rank); release(c, c->scan_sums);
    release(c, c->sums); release(c, c->cent); release(c, c->rp[0]); release(c, c->rp[1]); c->sorted_pos = nullptr;
    release(c, c->slot_idx); release(c, c->slot_pos); release(c, c->cell_parts);
    release(c, c->center_part); release(c, c->center_sum); release(c, c->center_chunks);
    release(c, c->center_first);
    c->n_center_chunks = 0;
    release(c, c->cpresent);
    release(c, c->pcl); release(c, c->pcs); release(c, c->cpos); release(c, c->cl_members);
    release(c, c->cl_chunks); release(c, c->cl_first); release(c, c->cl_part); release(c, c->cl_sum);
    c->CC = 0;
    c->n_cl_chunks = 0;
    release(c, c->col_begin); release(c, c->col_end); release(c, c->col_cell); release(c, c->col_crowd);
    c->col_crowd_cap = 0;
}

template <typename T>
int upload(Ctx* c, T* dst, const std::vector<T>& src) {
    if (!src.empty())
        CG_HIP(c, hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return CG_OK;
}

// Uploads one graph's per-point inputs (flags, jitter, resolved collision
// sizes, local present indices) to its own device arrays.  Runs when they
// changed since the last packing, so a repack uploads nothing for graphs that
// were already resident.
int stage_graph(Ctx* c, Graph& g) {
    dfree(c, g.d_flags); dfree(c, g.d_jit); dfree(c, g.d_size); dfree(c, g.d_pres);
    g.d_flags = nullptr; g.d_jit = nullptr; g.d_size = nullptr; g.d_pres = nullptr;
    int rc = CG_OK;
    const size_t n = g.n ? g.n : 1;
    g.d_flags = (uint8_t*)dalloc(c, n, &rc);
    g.d_jit = (float2*)dalloc(c, n * 8, &rc);
    g.d_pres = (uint32_t*)dalloc(c, (g.np ? g.np : 1) * 4, &rc);
    if (collides(g)) g.d_size = (float*)dalloc(c, n * 4, &rc);
    if (rc) return rc;
    std::vector<uint32_t> pres;
    pres.reserve(g.np);
    for (uint32_t i = 0; i < g.n; ++i)
        if (!(g.flags[i] & F_ABSENT)) pres.push_back(i);
    if (g.n) {
        CG_HIP(c, hipMemcpyAsync(g.d_flags, g.flags.data(), g.n, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipMemcpyAsync(g.d_jit, g.jitter.data(), (size_t)g.n * 8, hipMemcpyHostToDevice, c->stream));
        if (g.d_size)
            CG_HIP(c, hipMemcpyAsync(g.d_size, g.col_size.data(), (size_t)g.n * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (!pres.empty())
        CG_HIP(c, hipMemcpyAsync(g.d_pres, pres.data(), pres.size() * 4, hipMemcpyHostToDevice, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));  // the host vectors are temporaries
    g.staged = true;
    return CG_OK;
}

void unstage_graph(Ctx* c, Graph& g) {
    dfree(c, g.d_flags); dfree(c, g.d_jit); dfree(c, g.d_size); dfree(c, g.d_pres);
    g.d_flags = nullptr; g.d_jit = nullptr; g.d_size = nullptr; g.d_pres = nullptr;
    g.staged = false;
}

// What a repack replaces: the previous packing's device-resident state (the
// position buffers and the link schedules with their warm starts), the
// placement of every graph in it and the set's sizes.  A repack that fails
// restores it, so positions stay readable and a later repack starts from them.
struct PackSnap {
    float2* pos[2] = {nullptr, nullptr};
    LinkDir dir[2];
    std::vector<uint32_t> order;
    std::vector<GDesc> h_desc;
    uint32_t G = 0, N = 0, NP = 0, F = 0, C = 0, CG = 0, NPC = 0;
    struct Place {
        uint32_t slot, p0, k0, f0, c0, col_c0, cur, w0[2], wn[2];
    };
    std::vector<Place> place;  // by graph id
};

PackSnap pack_snapshot(const Ctx* c) {
    PackSnap s;
    s.pos[0] = c->pos[0];
    s.pos[1] = c->pos[1];
    s.dir[0] = c->dir[0];
    s.dir[1] = c->dir[1];
    s.order = c->order;
    s.h_desc = c->h_desc;
    s.G = c->G; s.N = c->N; s.NP = c->NP; s.F = c->F; s.C = c->C; s.CG = c->CG; s.NPC = c->NPC;
    for (const Graph& g : c->graphs)
        s.place.push_back({g.slot, g.p0, g.k0, g.f0, g.c0, g.col_c0, g.cur, {g.w0[0], g.w0[1]},
                           {g.wn[0], g.wn[1]}});
    return s;
}

// Undoes a failed repack: frees everything it allocated and puts the previous
// position buffers, link schedules and placement back.  The other packed
// arrays (scratch and host-derived tables) stay freed; `dirty` stays set, so
// the next step repacks from the restored state.
void pack_rollback(Ctx* c, const PackSnap& s) {
    (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    free_packed(c);
    for (int b = 0; b < 2; ++b)
        if (c->pos[b] != s.pos[b]) dfree(c, c->pos[b]);
    for (int d = 0; d < 2; ++d)
        if (c->dir[d].ent != s.dir[d].ent || c->dir[d].warm != s.dir[d].warm) link_free_dir(c, c->dir[d]);
    c->pos[0] = s.pos[0];
    c->pos[1] = s.pos[1];
    c->dir[0] = s.dir[0];
    c->dir[1] = s.dir[1];
    c->order = s.order;
    c->h_desc = s.h_desc;
    c->G = s.G; c->N = s.N; c->NP = s.NP; c->F = s.F; c->C = s.C; c->CG = s.CG; c->NPC = s.NPC;
    for (size_t id = 0; id < c->graphs.size() && id < s.place.size(); ++id) {
        Graph& g = c->graphs[id];
        const PackSnap::Place& q = s.place[id];
        g.slot = q.slot; g.p0 = q.p0; g.k0 = q.k0; g.f0 = q.f0; g.c0 = q.c0; g.col_c0 = q.col_c0;
        g.cur = q.cur;
        for (int d = 0; d < 2; ++d) {
            g.w0[d] = q.w0[d];
            g.wn[d] = q.wn[d];
        }
    }
    c->desc = nullptr;
    c->desc_pending = false;
    c->dirty = true;
}

// Lays every live graph out in fresh packed arrays (id order), carrying the
// positions of graphs that were already packed over on the device.  `old`
// holds the previous packing; the caller frees or restores it.
int pack_body(Ctx* c, const PackSnap& old) {
    CG_HIP(c, hipStreamSynchronize(c->stream));
    if (c->side) CG_HIP(c, hipStreamSynchronize(c->side));
    if (c->ingest) CG_HIP(c, hipStreamSynchronize(c->ingest));
    c->desc_pending = false;
    int rc = CG_OK;
    c->order.clear();
    for (auto& g : c->graphs)
        if (g.live) c->order.push_back(g.id);
    const uint32_t G = (uint32_t)c->order.size();
    for (uint32_t id : c->order) {
        Graph& g = c->graphs[id];
        if (collides(g) && (g.col_dirty || g.col_size.size() != g.n)) {
            rc = col_prepare_graph(c, g);
            if (rc) return rc;
            g.staged = false;
        }
        if (g.staged && collides(g) != (g.d_size != nullptr)) g.staged = false;
        if (!g.staged) {
            rc = stage_graph(c, g);
            if (rc) return rc;
        }
    }
    uint64_t N = 0, NP = 0, F = 0, C = 0, CG = 0, NPC = 0;
    for (uint32_t s = 0; s < G; ++s) {
        const Graph& g = c->graphs[c->order[s]];
        N += g.n;
        NP += g.np;
        F += pad_cells(g.gf);
        C += level_off(g.L);
        if (collides(g)) {
            CG += 4ull * g.col_G * g.col_G;
            NPC += g.np;
        }
    }
    if (N >= (1ull << 31) || F * SLOTS >= (1ull << 32) || C >= (1ull << 31) || CG >= (1ull << 28) ||
        4 * NPC >= (1ull << 31))
        return fail(c, CG_ERR_INVALID, "the graph set exceeds the 32-bit index space of one context");
    std::vector<uint32_t> old_p0(G);
    std::vector<GDesc> st(G);
    {
        uint32_t p0 = 0, k0 = 0, f0 = 0, c0 = 0, col0 = 0;
        for (uint32_t s = 0; s < G; ++s) {
            Graph& g = c->graphs[c->order[s]];
            old_p0[s] = g.p0;
            g.slot = s;
            g.p0 = p0; g.k0 = k0; g.f0 = f0; g.c0 = c0; g.col_c0 = col0;
            p0 += g.n; k0 += g.np; f0 += (uint32_t)pad_cells(g.gf); c0 += level_off(g.L);
            if (collides(g)) col0 += 4 * g.col_G * g.col_G;
            GDesc d{};
            d.p0 = g.p0; d.n = g.n; d.k0 = g.k0; d.np = g.np; d.f0 = g.f0; d.c0 = g.c0;
            d.gf = g.gf; d.L = g.L; d.col_c0 = g.col_c0;
            d.col_G = collides(g) ? g.col_G : 0;
            d.col_h = collides(g) ? g.col_h : 1.f;
            st[s] = d;
        }
    }
    float2* const* old_pos = old.pos;
    c->pos[0] = c->pos[1] = nullptr;
    free_packed(c);
    cache_settle(c);  // nothing in flight reads the freed arrays: reusable below
    c->dir[0] = c->dir[1] = LinkDir();
    c->G = G;
    c->N = (uint32_t)N;
    c->NP = (uint32_t)NP;
    c->F = (uint32_t)F;
    c->C = (uint32_t)C;
    c->CG = (uint32_t)CG;
    c->NPC = (uint32_t)NPC;
    const size_t nn = N ? N : 1, gg = G ? G : 1;
    c->pos[0] = (float2*)dalloc(c, nn * 8, &rc);
    c->pos[1] = (float2*)dalloc(c, nn * 8, &rc);
    c->flags = (uint8_t*)dalloc(c, nn, &rc);
    c->gid = (uint32_t*)dalloc(c, nn * 4, &rc);
    c->jitter = (float2*)dalloc(c, nn * 8, &rc);
    c->size = (float*)dalloc(c, nn * 4, &rc);
    c->present = (uint32_t*)dalloc(c, (NP ? NP : 1) * 4, &rc);
    c->cpresent = (uint32_t*)dalloc(c, (NPC ? NPC : 1) * 4, &rc);
    c->f0s = (uint32_t*)dalloc(c, gg * 4, &rc);
    c->colc0s = (uint32_t*)dalloc(c, gg * 4, &rc);
    c->blk_graph = (uint32_t*)dalloc(c, (F / CELL_BLOCK + 1) * 4, &rc);
    if (rc) return rc;

    // per-point arrays, assembled on the device from every graph's staged inputs
    std::vector<uint32_t> hf0(G), hcol(G), k0s(G), nps(G), hblk(F / CELL_BLOCK);
    std::vector<uint32_t> tp0(G), tk0(G), tck0;
    std::vector<GStage> tab(G), ctab;
    std::vector<uint4> w0, w1;
    uint32_t ck0 = 0;
    for (uint32_t s = 0; s < G; ++s) {
        const Graph& g = c->graphs[c->order[s]];
        tab[s] = GStage{g.d_flags, g.d_jit, g.d_size, g.d_pres, g.p0, g.k0, g.np, 0u};
        tp0[s] = g.p0;
        tk0[s] = g.k0;
        if (collides(g)) {
            ctab.push_back(GStage{g.d_flags, g.d_jit, g.d_size, g.d_pres, g.p0, ck0, g.np, 0u});
            tck0.push_back(ck0);
            ck0 += g.np;
        }
        hf0[s] = g.f0;
        std::fill(hblk.begin() + g.f0 / CELL_BLOCK, hblk.begin() + (g.f0 + pad_cells(g.gf)) / CELL_BLOCK, s);
        hcol[s] = g.col_c0;
        k0s[s] = g.k0;
        nps[s] = g.np;
        if (g.np) pyramid_work_host(s, g.gf, g.L, w0, w1);
    }
    {
        // the tables for the assembly kernels, freed once they have run
        const size_t tb = (gg + ctab.size() + 1) * sizeof(GStage), wb = (2 * gg + tck0.size() + 1) * 4;
        uint8_t* tmp = (uint8_t*)dalloc(c, tb + wb, &rc);
        if (rc) return rc;
        GStage* dtab = (GStage*)tmp;
        GStage* dctab = dtab + gg;
        uint32_t* dp0 = (uint32_t*)(tmp + tb);
        uint32_t* dk0 = dp0 + gg;
        uint32_t* dck0 = dk0 + gg;
        hipError_t e = hipSuccess;
        if (G) e = hipMemcpyAsync(dtab, tab.data(), G * sizeof(GStage), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && !ctab.empty())
            e = hipMemcpyAsync(dctab, ctab.data(), ctab.size() * sizeof(GStage), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && G) e = hipMemcpyAsync(dp0, tp0.data(), G * 4, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && G) e = hipMemcpyAsync(dk0, tk0.data(), G * 4, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && !tck0.empty())
            e = hipMemcpyAsync(dck0, tck0.data(), tck0.size() * 4, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && N)
            hipLaunchKernelGGL(k_pack_points, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, dtab, dp0, G,
                               (uint32_t)N, c->flags, c->gid, c->jitter, c->size);
        if (e == hipSuccess && NP)
            hipLaunchKernelGGL(k_pack_present, dim3(blocks_for(NP, 256)), dim3(256), 0, c->stream, dtab, dk0, G,
                               (uint32_t)NP, c->present);
        if (e == hipSuccess && NPC)
            hipLaunchKernelGGL(k_pack_present, dim3(blocks_for(NPC, 256)), dim3(256), 0, c->stream, dctab, dck0,
                               (uint32_t)ctab.size(), (uint32_t)NPC, c->cpresent);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // the host tables are temporaries
        dfree(c, tmp);
        if (e != hipSuccess) return hip_fail(c, e, "pack: assembling the per-point arrays");
    }
    rc = upload(c, c->f0s, hf0); if (rc) return rc;
    rc = upload(c, c->colc0s, hcol); if (rc) return rc;
    rc = upload(c, c->blk_graph, hblk); if (rc) return rc;

    // positions: host input for newly set graphs, device copy for the others
    std::vector<uint4> psegs;  // resident graphs: one launch writes both buffers
    for (uint32_t s = 0; s < G; ++s) {
        Graph& g = c->graphs[c->order[s]];
        if (g.n == 0) {
            g.xy_pending = false;
            g.cur = 0;
            continue;
        }
        const size_t bytes = (size_t)g.n * 8;
        if (g.xy_pending || !old_pos[0]) {
            for (int b = 0; b < 2; ++b)
                CG_HIP(c, hipMemcpyAsync(c->pos[b] + g.p0, g.xy.data(), bytes, hipMemcpyHostToDevice, c->stream));
        } else {
            psegs.push_back(make_uint4(old_p0[s] | (g.cur ? 0x80000000u : 0u), g.p0, g.n, 0u));
        }
        g.cur = 0;
    }
    rc = seg_copy<float2>(c, old_pos[0], old_pos[1], c->pos[0], c->pos[1], psegs);
    if (rc) return rc;

    // centroid chunks and pyramid work lists
    {
        std::vector<uint4> chunks;
        std::vector<uint32_t> first;
        center_chunks_host(k0s, nps, chunks, first);
        c->n_center_chunks = (uint32_t)chunks.size();
        c->center_chunks = (uint4*)dalloc(c, (chunks.size() + 1) * 16, &rc);
        c->center_first = (uint32_t*)dalloc(c, first.size() * 4, &rc);
        c->center_part = (double*)dalloc(c, (chunks.size() + 1) * 2 * sizeof(double), &rc);
        c->center_sum = (double*)dalloc(c, gg * 3 * sizeof(double), &rc);
        for (int b = 0; b < 2; ++b) {
            const auto& w = b ? w1 : w0;
            c->n_pyr[b] = (uint32_t)w.size();
            c->pyr[b] = (uint4*)dalloc(c, (w.size() + 1) * 16, &rc);
        }
        if (rc) return rc;
        rc = upload(c, c->center_chunks, chunks); if (rc) return rc;
        rc = upload(c, c->center_first, first); if (rc) return rc;
        rc = upload(c, c->pyr[0], w0); if (rc) return rc;
        rc = upload(c, c->pyr[1], w1); if (rc) return rc;
    }

    // clusters: packed indices, strengths, targets and member chunks (the
    // per-point arrays only when some graph has clusters: the pass reads them
    // for those graphs' points only)
    bool any_clustered = false;
    for (uint32_t s = 0; s < G; ++s) any_clustered |= clustered(c->graphs[c->order[s]]);
    {
        const size_t ncl_pts = any_clustered ? N : 0;
        std::vector<int32_t> hpcl(ncl_pts, -1);
        std::vector<float> hpcs(ncl_pts, 1.f);
        std::vector<float2> hcpos;
        std::vector<uint32_t> members, first;
        std::vector<uint4> chunks;
        std::vector<std::vector<uint32_t>> bucket;
        for (uint32_t s = 0; s < G; ++s) {
            const Graph& g = c->graphs[c->order[s]];
            if (!clustered(g)) continue;
            const uint32_t ncl = cluster_count(g), cl0 = (uint32_t)hcpos.size();
            for (uint32_t q = 0; q < ncl; ++q) {
                const bool have = 2 * (size_t)q + 1 < g.cluster_pos.size();
                hcpos.push_back(have ? make_float2(g.cluster_pos[2 * q], g.cluster_pos[2 * q + 1])
                                     : make_float2(-1.f, -1.f));
            }
            bucket.assign(ncl, {});
            for (uint32_t i = 0; i < g.n; ++i) {
                const int32_t q = i < g.clusters.size() ? g.clusters[i] : -1;
                if (i < g.cluster_strength.size()) hpcs[g.p0 + i] = g.cluster_strength[i];
                if (q < 0) continue;
                hpcl[g.p0 + i] = (int32_t)(cl0 + (uint32_t)q);
                if (!(g.flags[i] & F_ABSENT)) bucket[q].push_back(g.p0 + i);
            }
            for (uint32_t q = 0; q < ncl; ++q) {
                first.push_back((uint32_t)chunks.size());
                const uint32_t b = (uint32_t)members.size();
                members.insert(members.end(), bucket[q].begin(), bucket[q].end());
                const uint32_t e = (uint32_t)members.size();
                for (uint32_t k = b; k < e; k += 4096)
                    chunks.push_back(make_uint4(k, std::min(e, k + 4096), cl0 + q, s));
            }
        }
        first.push_back((uint32_t)chunks.size());
        c->CC = (uint32_t)hcpos.size();
        c->n_cl_chunks = (uint32_t)chunks.size();
        c->pcl = (int32_t*)dalloc(c, (ncl_pts ? ncl_pts : 1) * 4, &rc);
        c->pcs = (float*)dalloc(c, (ncl_pts ? ncl_pts : 1) * 4, &rc);
        c->cpos = (float2*)dalloc(c, (hcpos.size() + 1) * 8, &rc);
        c->cl_members = (uint32_t*)dalloc(c, (members.size() + 1) * 4, &rc);
        c->cl_chunks = (uint4*)dalloc(c, (chunks.size() + 1) * 16, &rc);
        c->cl_first = (uint32_t*)dalloc(c, first.size() * 4, &rc);
        c->cl_part = (double*)dalloc(c, (chunks.size() + 1) * 2 * sizeof(double), &rc);
        c->cl_sum = (double*)dalloc(c, (hcpos.size() + 1) * 3 * sizeof(double), &rc);
        if (rc) return rc;
        rc = upload(c, c->pcl, hpcl); if (rc) return rc;
        rc = upload(c, c->pcs, hpcs); if (rc) return rc;
        rc = upload(c, c->cpos, hcpos); if (rc) return rc;
        rc = upload(c, c->cl_members, members); if (rc) return rc;
        rc = upload(c, c->cl_chunks, chunks); if (rc) return rc;
        rc = upload(c, c->cl_first, first); if (rc) return rc;
    }

    // descriptor tables (device and pinned staging)
    c->desc_ticks = std::max<uint32_t>(1, std::min<uint32_t>(DESC_MAX_TICKS, DESC_ENTRIES / (uint32_t)gg));
    c->desc_base = (GDesc*)dalloc(c, (size_t)c->desc_ticks * gg * sizeof(GDesc), &rc);
    c->desc_win = (GDesc*)dalloc(c, (size_t)c->desc_ticks * gg * sizeof(GDesc), &rc);
    if (rc) return rc;
    if (c->h_desc_pinned) (void)hipHostFree(c->h_desc_pinned);
    c->h_desc_pinned = nullptr;
    CG_HIP(c, hipHostMalloc((void**)&c->h_desc_pinned, (size_t)c->desc_ticks * gg * sizeof(GDesc)));
    c->h_desc = st;
    c->desc = c->desc_base;
    rc = upload(c, c->desc_base, st); if (rc) return rc;

    rc = mb_alloc(c); if (rc) return rc;
    rc = col_alloc(c); if (rc) return rc;
    rc = link_pack(c, old.dir); if (rc) return rc;
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

// Repacks the graph set.  On success the previous packing is freed and the
// host copies of newly uploaded positions are released; on failure (budget,
// allocation, device error) the previous packing is restored and the error
// returned.
int pack(Ctx* c) {
    PackSnap old = pack_snapshot(c);
    const int rc = pack_body(c, old);
    if (rc) {
        const std::string msg = c->err;
        pack_rollback(c, old);
        c->err = msg;
        return rc;
    }
    dfree(c, old.pos[0]);
    dfree(c, old.pos[1]);
    for (int d = 0; d < 2; ++d) {
        LinkDir L = old.dir[d];
        link_free_dir(c, L);
    }
    cache_settle(c);  // the repack ended synchronised: the old packing is reusable
    for (uint32_t id : c->order) {
        Graph& g = c->graphs[id];
        if (g.xy_pending) {
            g.xy_pending = false;
            std::vector<float>().swap(g.xy);
        }
    }
    c->dirty = false;
    return CG_OK;
}

// The parameter-dependent part of a descriptor (alpha, seed and the force constants).
// `extent`: the largest |coordinate| the graph's present points may hold (the
// integrate pass clamps to [0, S], but positions as set or dragged are not).
void fill_dynamic(GDesc& d, const cg_sim_params& p, float extent, float a, float seed) {
    const float S = p.space_size;
    d.S = S;
    for (uint32_t l = 0; l < (uint32_t)MAX_LEVELS; ++l)
        d.h[l] = (l < d.L) ? (float)((double)S / (double)(4u << l)) : 1.f;
    d.alpha = a;
    d.seed = seed;
    d.ak = a * p.repulsion;
    d.friction = p.friction;
    d.gravity = p.gravity;
    d.center = p.center;
    d.mouse_k = p.repulsion_from_mouse;
    d.mx = p.mouse_position[0];
    d.my = p.mouse_position[1];
    d.spring_alpha = p.link_spring * a;
    d.dist = p.link_distance;
    d.lo = p.link_dist_random_variation_range[0];
    d.span = p.link_dist_random_variation_range[1] - p.link_dist_random_variation_range[0];
    d.tol = 2.5e-6f * S;  // a quarter of the 1e-5 * spaceSize force tolerance
    d.col_strength = p.collision;
    d.col_radius = p.collision_radius;
    d.col_padding = p.collision_padding;
    d.cluster_k = p.cluster;
    // fixed point of the cell sums: np * 16 X * 2^k < 2^63 with X = max(S, extent),
    // so no position a point can hold overflows a sum (k < 0 for huge extents)
    const double X = std::max<double>(std::fabs((double)S), (double)extent);
    int bits_s = 0, bits_n = 0;
    while (bits_s < 160 && std::ldexp(1.0, bits_s) < X) ++bits_s;
    while (bits_n < 32 && (1ull << bits_n) < (unsigned long long)d.np + 1) ++bits_n;
    const int k = std::max(-100, std::min(40, 59 - bits_s - bits_n));
    d.fx_scale = std::ldexp(1.f, k);
    d.fx_inv = std::ldexp(1.f, -k);
}

// Passes a graph runs in one tick — the gates of src/index.ts:2012-2078.
uint32_t graph_passes(const Graph& g) {
    const cg_sim_params& p = g.params;
    if (!p.enable_simulation || g.np == 0) return 0;
    uint32_t passes = 0;
    if (p.mouse_repulsion) passes |= 1u << P_MOUSE;
    if (p.run_forces) {
        if (on(p.gravity)) passes |= 1u << P_GRAVITY;
        if (on(p.center)) passes |= 1u << P_CENTER;
        passes |= 1u << P_MANYBODY;
        if (g.links_ready && g.m > 0) passes |= (1u << P_LINK_IN) | (1u << P_LINK_OUT);
        if (clustered(g)) passes |= 1u << P_CLUSTER;
        if (collides(g)) passes |= 1u << P_COLLISION;
    }
    return passes;
}

struct TickPlan {
    uint32_t any;   // union of the passes of every graph
};

int launch_tick(Ctx* c, const TickPlan& tp) {
    const uint32_t a = tp.any;
    int rc = CG_OK;
    if (a & (1u << P_MOUSE)) {
        prof_begin(c, CG_STAGE_MOUSE);
        rc = pw_run(c, P_MOUSE, nullptr);
        prof_end(c, CG_STAGE_MOUSE);
        if (rc) return rc;
    }
    if (a & (1u << P_GRAVITY)) {
        prof_begin(c, CG_STAGE_GRAVITY);
        rc = pw_run(c, P_GRAVITY, nullptr);
        prof_end(c, CG_STAGE_GRAVITY);
        if (rc) return rc;
    }
    if (a & (1u << P_CENTER)) {
        prof_begin(c, CG_STAGE_CENTER);
        rc = pw_run(c, P_CENTER, nullptr);
        prof_end(c, CG_STAGE_CENTER);
        if (rc) return rc;
    }
    if (a & (1u << P_MANYBODY)) {
        rc = mb_run(c, nullptr);
        if (rc) return rc;
    }
    for (int d = 0; d < 2; ++d) {
        if (!(a & (1u << (P_LINK_IN + d)))) continue;
        prof_begin(c, CG_STAGE_LINK_IN + d);
        rc = link_run(c, d, nullptr);
        prof_end(c, CG_STAGE_LINK_IN + d);
        if (rc) return rc;
    }
    if (a & (1u << P_CLUSTER)) {
        prof_begin(c, CG_STAGE_CLUSTER);
        rc = cluster_run(c, nullptr);
        prof_end(c, CG_STAGE_CLUSTER);
        if (rc) return rc;
    }
    if (a & (1u << P_COLLISION)) {
        prof_begin(c, CG_STAGE_COLLISION);
        rc = col_run(c, nullptr);
        prof_end(c, CG_STAGE_COLLISION);
        if (rc) return rc;
    }
    return CG_OK;
}

// n ticks as launched (or captured): tick q reads its descriptor table at
// window slot q, which the caller fills before the launch.
int launch_tick_seq(Ctx* c, const TickPlan& tp, uint32_t n) {
    for (uint32_t q = 0; q < n; ++q) {
        c->desc = c->desc_win + (size_t)q * c->G;
        int rc = launch_tick(c, tp);
        if (rc) return rc;
    }
    return CG_OK;
}

// The executable graph of n ticks of plan tp, captured on first use.  Tick q
// of it reads descriptor window slot q, which run_ticks fills before every
// launch, so one graph serves any n ticks of any chunk while the packing stands.
int tick_graph(Ctx* c, const TickPlan& tp, uint32_t n, hipGraphExec_t* out) {
    for (auto& e : c->graphs_exec)
        if (e.first.first == tp.any && e.first.second == n) {
            *out = e.second;
            return CG_OK;
        }
    CG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_tick_seq(c, tp, n);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
    c->graphs_exec.push_back({{tp.any, n}, x});
    *out = x;
    return CG_OK;
}


// Advances the graphs in `ids` by nticks.  The graph `sel` (or none) takes its
// near-field seeds from `seeds` and reports alpha after each tick in `trace`.
int run_ticks(Ctx* c, const std::vector<uint32_t>& ids, uint32_t nticks, int64_t sel,
              const float* seeds, double* trace) {
    int rc = CG_OK;
    if (c->dirty) {
        rc = pack(c);
        if (rc) return rc;
    }
    const uint32_t G = c->G;
    if (G == 0 || nticks == 0) return CG_OK;
    std::vector<uint8_t> run(G, 0);
    for (uint32_t id : ids) run[c->graphs[id].slot] = 1;
    const int64_t sel_slot = sel >= 0 ? (int64_t)c->graphs[sel].slot : -1;
    std::vector<TickPlan> plan;
    for (uint32_t t0 = 0; t0 < nticks; t0 += c->desc_ticks) {
        const uint32_t T = std::min(c->desc_ticks, nticks - t0);
        if (c->desc_pending) {
            CG_HIP(c, hipEventSynchronize(c->ev_desc));
            c->desc_pending = false;
        }
        plan.assign(T, TickPlan{0});
        for (uint32_t tt = 0; tt < T; ++tt) {
            GDesc* D = c->h_desc_pinned + (size_t)tt * G;
            uint32_t any = 0;
            for (uint32_t s = 0; s < G; ++s) {
                Graph& g = c->graphs[c->order[s]];
                const cg_sim_params& p = g.params;
                GDesc d = c->h_desc[s];
                const bool live = run[s] && p.enable_simulation;
                const uint32_t passes = live ? graph_passes(g) : 0u;
                uint32_t cur = g.cur, inbuf = 0;
                for (uint32_t q = 0; q < P_COUNT; ++q) {
                    if (cur) inbuf |= 1u << q;
                    if ((passes >> q) & 1u) cur ^= 1u;
                }
                float seed = 0.f;
                if (live) seed = ((int64_t)s == sel_slot && seeds) ? seeds[t0 + tt] : rng_unit(g);
                fill_dynamic(d, p, g.extent, (float)g.alpha, seed);
                d.passes = passes;
                d.inbuf = inbuf;
                D[s] = d;
                if (live) {
                    g.cur = cur;
                    if (p.run_forces) {
                        const double decay = 1.0 - std::pow(0.001, 1.0 / (double)p.decay);
                        g.alpha += (0.0 - g.alpha) * decay;
                        if (p.mouse_repulsion) g.alpha = std::max(g.alpha, 0.1);
                    }
                }
                any |= passes;
            }
            plan[tt] = TickPlan{any};
            if (trace && sel_slot >= 0) trace[t0 + tt] = c->graphs[sel].alpha;
        }
        CG_HIP(c, hipMemcpyAsync(c->desc_base, c->h_desc_pinned, (size_t)T * G * sizeof(GDesc),
                                 hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipEventRecord(c->ev_desc, c->stream));
        c->desc_pending = true;
        // stage timing brackets single launches, so it runs uncaptured
        const bool cap = captures(c);
        for (uint32_t tt = 0; tt < T;) {
            uint32_t r = 1;
            while (tt + r < T && plan[tt + r].any == plan[tt].any) ++r;
            for (uint32_t done = 0; done < r;) {
                // graph_ticks ticks per captured graph, one-tick graphs for the rest
                const uint32_t gt = c->tune.graph_ticks;
                const uint32_t n = (cap && r - done >= gt) ? gt : 1u;
                // the n ticks' descriptor tables into the window their kernels read
                CG_HIP(c, hipMemcpyAsync(c->desc_win, c->desc_base + (size_t)(tt + done) * G,
                                         (size_t)n * G * sizeof(GDesc), hipMemcpyDeviceToDevice, c->stream));
                if (cap) {
                    hipGraphExec_t x = nullptr;
                    rc = tick_graph(c, plan[tt], n, &x);
                    if (rc) break;
                    CG_HIP(c, hipGraphLaunch(x, c->stream));
                } else {
                    rc = launch_tick_seq(c, plan[tt], n);
                    if (rc) break;
                }
                done += n;
            }
            if (rc) {
                c->desc = c->desc_base;
                return rc;
            }
            tt += r;
        }
    }
    c->desc = c->desc_base;
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Writes one descriptor table (slot 0 of the device tables) synchronously.
int write_table(Ctx* c, const std::vector<GDesc>& table) {
    if (c->desc_pending) {
        CG_HIP(c, hipEventSynchronize(c->ev_desc));
        c->desc_pending = false;
    }
    CG_HIP(c, hipMemcpyAsync(c->desc_base, table.data(), table.size() * sizeof(GDesc),
                             hipMemcpyHostToDevice, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    c->desc = c->desc_base;
    return CG_OK;
}

int graph_set_positions(Ctx* c, Graph& g, const float* xy, uint32_t n) {
    if (n && !xy) return fail(c, CG_ERR_INVALID, "set_positions: null positions");
    if (n >= (1u << 31)) return fail(c, CG_ERR_INVALID, "set_positions: too many points");
    const bool same_n = n == g.n;
    g.n = n;
    // Points/position-utils.ts:17-24: a half-NaN position becomes fully NaN (absent)
    g.xy.assign(2 * (size_t)n, 0.f);
    g.flags.assign(n, 0);
    uint32_t np = 0;
    float ext = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        float px = xy[2 * i], py = xy[2 * i + 1];
        bool absent = std::isnan(px) || std::isnan(py);
        g.xy[2 * i] = absent ? NAN : px;
        g.xy[2 * i + 1] = absent ? NAN : py;
        if (absent) g.flags[i] = F_ABSENT;
        else {
            ++np;
            ext = std::max(ext, std::max(std::fabs(px), std::fabs(py)));
        }
    }
    g.extent = ext;
    for (uint32_t k : g.pinned)
        if (k < n) g.flags[k] |= F_PINNED;
    g.np = np;
    geometry(n, &g.gf, &g.L);
    // default jitter from the graph's generator: (U(-1,1)*1e-5, U(-1,1)*1e-5)
    g.jitter.resize(2 * (size_t)n);
    for (auto& v : g.jitter) v = (rng_unit(g) * 2.f - 1.f) * 0.00001f;
    if (!same_n) g.sizes.clear();
    g.xy_pending = true;
    g.col_dirty = true;
    g.staged = false;
    c->dirty = true;
    if (g.links_set) return link_build_graph(c, g);
    return CG_OK;
}

int graph_set_links(Ctx* c, Graph& g, const float* links, uint32_t m, const float* strength,
                    const float* rand_in, const float* rand_out) {
    if (m && !links) return fail(c, CG_ERR_INVALID, "set_links: null links");
    g.m = m;
    g.links.assign(links, links + 2 * (size_t)m);
    if (strength) g.strength.assign(strength, strength + m);
    else g.strength.clear();
    g.rin.resize(m);
    g.rout.resize(m);
    for (uint32_t l = 0; l < m; ++l) g.rin[l] = rand_in ? rand_in[l] : rng_unit(g);
    for (uint32_t l = 0; l < m; ++l) g.rout[l] = rand_out ? rand_out[l] : rng_unit(g);
    g.links_set = true;
    c->dirty = true;
    return link_build_graph(c, g);
}

int graph_set_jitter(Ctx* c, Graph& g, const float* xy, uint32_t n) {
    if (n != g.n || (n && !xy)) return fail(c, CG_ERR_INVALID, "set_jitter: expected one pair per point");
    g.jitter.assign(xy, xy + 2 * (size_t)n);
    g.staged = false;
    if (!c->dirty && n) {
        CG_HIP(c, hipMemcpyAsync(c->jitter + g.p0, g.jitter.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                                 c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return CG_OK;
}

int graph_set_point_sizes(Ctx* c, Graph& g, const float* sizes, uint32_t n) {
    if (sizes && n != g.n) return fail(c, CG_ERR_INVALID, "set_point_sizes: expected one size per point");
    if (sizes) g.sizes.assign(sizes, sizes + n);
    else g.sizes.clear();
    g.col_dirty = true;
    if (collides(g)) c->dirty = true;
    return CG_OK;
}

int graph_set_pinned(Ctx* c, Graph& g, const uint32_t* idx, uint32_t count) {
    if (count && !idx) return fail(c, CG_ERR_INVALID, "set_pinned: null indices");
    g.pinned.assign(idx, idx + count);
    for (auto& f : g.flags) f &= (uint8_t)~F_PINNED;
    for (uint32_t k = 0; k < count; ++k)
        if (idx[k] < g.n) g.flags[idx[k]] |= F_PINNED;
    g.staged = false;
    if (!c->dirty && g.n) {
        CG_HIP(c, hipMemcpyAsync(c->flags + g.p0, g.flags.data(), g.n, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return CG_OK;
}

int graph_set_point_clusters(Ctx* c, Graph& g, const int32_t* idx, uint32_t n) {
    if (idx) g.clusters.assign(idx, idx + n);
    else g.clusters.clear();
    g.has_clusters = idx != nullptr;
    c->dirty = true;
    return CG_OK;
}

int graph_set_cluster_positions(Ctx* c, Graph& g, const float* xy, uint32_t count) {
    if (xy) g.cluster_pos.assign(xy, xy + 2 * (size_t)count);
    else g.cluster_pos.clear();
    g.has_cluster_pos = xy != nullptr;
    c->dirty = true;
    return CG_OK;
}

int graph_set_point_cluster_strength(Ctx* c, Graph& g, const float* s, uint32_t n) {
    if (s) g.cluster_strength.assign(s, s + n);
    else g.cluster_strength.clear();
    c->dirty = true;
    return CG_OK;
}

int graph_get_positions(Ctx* c, Graph& g, float* xy, uint32_t n) {
    if (n != g.n || (n && !xy)) return fail(c, CG_ERR_INVALID, "get_positions: size mismatch");
    if (n == 0) return CG_OK;
    if (g.xy_pending) {
        std::memcpy(xy, g.xy.data(), (size_t)n * 8);
        return CG_OK;
    }
    CG_HIP(c, hipMemcpyAsync(xy, c->pos[g.cur] + g.p0, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int io_reserve(Ctx* c, uint32_t count) {
    if (count <= c->io_cap) return CG_OK;
    dfree(c, c->io_idx);
    dfree(c, c->io_xy);
    c->io_idx = nullptr;
    c->io_xy = nullptr;
    c->io_cap = 0;
    int rc = CG_OK;
    const uint32_t cap = std::max<uint32_t>(count, 1024);
    c->io_idx = (uint32_t*)dalloc(c, (size_t)cap * 4, &rc);
    c->io_xy = (float2*)dalloc(c, (size_t)cap * 8, &rc);
    if (rc) return rc;
    c->io_cap = cap;
    return CG_OK;
}

int graph_drag_points(Ctx* c, Graph& g, const uint32_t* idx, const float* xy, uint32_t count) {
    if (count && (!idx || !xy)) return fail(c, CG_ERR_INVALID, "drag_points: null input");
    // only present points below the point count move (absent points stay absent)
    std::vector<uint32_t> hi;
    std::vector<float> hxy;
    for (uint32_t k = 0; k < count; ++k) {
        if (idx[k] >= g.n || (g.flags[idx[k]] & F_ABSENT)) continue;
        hi.push_back(idx[k]);
        hxy.push_back(xy[2 * k]);
        hxy.push_back(xy[2 * k + 1]);
        g.extent = std::max(g.extent, std::max(std::fabs(xy[2 * k]), std::fabs(xy[2 * k + 1])));
    }
    if (hi.empty()) return CG_OK;
    if (g.xy_pending) {  // not on the device yet: edit the pending copy
        for (size_t k = 0; k < hi.size(); ++k) {
            g.xy[2 * (size_t)hi[k]] = hxy[2 * k];
            g.xy[2 * (size_t)hi[k] + 1] = hxy[2 * k + 1];
        }
        return CG_OK;
    }
    const uint32_t cnt = (uint32_t)hi.size();
    int rc = io_reserve(c, cnt);
    if (rc) return rc;
    CG_HIP(c, hipMemcpyAsync(c->io_idx, hi.data(), (size_t)cnt * 4, hipMemcpyHostToDevice, c->stream));
    CG_HIP(c, hipMemcpyAsync(c->io_xy, hxy.data(), (size_t)cnt * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_drag, dim3(blocks_for(cnt, 256)), dim3(256), 0, c->stream, c->pos[g.cur], g.p0,
                       c->io_idx, c->io_xy, cnt);
    CG_HIP(c, hipGetLastError());
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int graph_get_positions_at(Ctx* c, Graph& g, const uint32_t* idx, uint32_t count, float* xy) {
    if (count && (!idx || !xy)) return fail(c, CG_ERR_INVALID, "get_positions_at: null input");
    if (count == 0) return CG_OK;
    if (g.xy_pending) {
        for (uint32_t k = 0; k < count; ++k) {
            const bool ok = idx[k] < g.n;
            xy[2 * k] = ok ? g.xy[2 * (size_t)idx[k]] : NAN;
            xy[2 * k + 1] = ok ? g.xy[2 * (size_t)idx[k] + 1] : NAN;
        }
        return CG_OK;
    }
    int rc = io_reserve(c, count);
    if (rc) return rc;
    CG_HIP(c, hipMemcpyAsync(c->io_idx, idx, (size_t)count * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(count, 256)), dim3(256), 0, c->stream, c->pos[g.cur], g.p0,
                       g.n, c->io_idx, count, c->io_xy);
    CG_HIP(c, hipMemcpyAsync(xy, c->io_xy, (size_t)count * 8, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int pass_of_force(int32_t force) {
    switch (force) {
        case CG_FORCE_GRAVITY: return P_GRAVITY;
        case CG_FORCE_CENTER: return P_CENTER;
        case CG_FORCE_MOUSE: return P_MOUSE;
        case CG_FORCE_MANYBODY: return P_MANYBODY;
        case CG_FORCE_LINK_IN: return P_LINK_IN;
        case CG_FORCE_LINK_OUT: return P_LINK_OUT;
        case CG_FORCE_COLLISION: return P_COLLISION;
        case CG_FORCE_CLUSTER: return P_CLUSTER;
        default: return -1;
    }
}

}  // namespace
}  // namespace cg

namespace cg {
namespace {

// One force pass of graph `id` from its current state (cg_eval_force and
// cg_graph_eval_force); the other graphs of the set run nothing.
int eval_force(Ctx* c, uint32_t id, const cg_sim_params* p, int32_t force, float alpha, float seed, float* v) {
    Graph* gp = lookup(c, id);
    if (!gp) return fail(c, CG_ERR_INVALID, "eval_force: unknown graph id " + std::to_string(id));
    if (gp->n && !v) return fail(c, CG_ERR_INVALID, "eval_force: null output");
    const int pass = pass_of_force(force);
    if (pass < 0) return fail(c, CG_ERR_INVALID, "eval_force: unknown force");
    if (p) set_params(c, *gp, *p);
    int rc = CG_OK;
    if (c->dirty) {
        rc = pack(c);
        if (rc) return rc;
    }
    Graph& g = c->graphs[id];
    if (g.n == 0) return CG_OK;
    std::memset(v, 0, (size_t)g.n * 8);
    const bool runs = g.np && ((pass < P_LINK_IN) || (pass == P_COLLISION && collides(g)) ||
                               (pass == P_CLUSTER && clustered(g)) ||
                               ((pass == P_LINK_IN || pass == P_LINK_OUT) && g.links_ready && g.m > 0));
    if (!runs) return CG_OK;
    std::vector<GDesc> table(c->G);
    for (uint32_t s = 0; s < c->G; ++s) {
        const Graph& h = c->graphs[c->order[s]];
        GDesc d = c->h_desc[s];
        fill_dynamic(d, h.params, h.extent, alpha, seed);
        d.passes = (s == g.slot) ? (1u << pass) : 0u;
        d.inbuf = h.cur ? ((1u << P_COUNT) - 1) : 0u;
        table[s] = d;
    }
    rc = write_table(c, table);
    if (rc) return rc;
    float2* dv = (float2*)dalloc(c, (size_t)c->N * 8, &rc);
    if (rc) return rc;
    // A link pass leaves its converged chunk starts as the next tick's warm
    // starts; an evaluation must leave the simulation's state as it found it,
    // so they are saved and restored around it.
    LinkDir* L = (pass == P_LINK_IN || pass == P_LINK_OUT) ? &c->dir[pass - P_LINK_IN] : nullptr;
    const size_t wl = (L && L->warm) ? L->warm_len : 0;
    uint8_t* keep = wl ? (uint8_t*)dalloc(c, wl * (sizeof(float2) + sizeof(float4)), &rc) : nullptr;
    if (rc) {
        dfree(c, dv);
        return rc;
    }
    hipError_t e = hipMemsetAsync(dv, 0, (size_t)c->N * 8, c->stream);
    if (e == hipSuccess && keep) e = hipMemcpyAsync(keep, L->warm, wl * 8, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && keep)
        e = hipMemcpyAsync(keep + wl * 8, L->wjac, wl * 16, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) rc = hip_fail(c, e, "eval_force clear");
    if (!rc) {
        if (pass <= P_CENTER) rc = pw_run(c, pass, dv);
        else if (pass == P_MANYBODY) rc = mb_run(c, dv);
        else if (pass == P_COLLISION) rc = col_run(c, dv);
        else if (pass == P_CLUSTER) rc = cluster_run(c, dv);
        else rc = link_run(c, pass - P_LINK_IN, dv);
    }
    if (keep) {
        e = hipMemcpyAsync(L->warm, keep, wl * 8, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(L->wjac, keep + wl * 8, wl * 16, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess && rc == CG_OK) rc = hip_fail(c, e, "eval_force warm-start restore");
    }
    if (rc == CG_OK) {
        e = hipMemcpyAsync(v, dv + g.p0, (size_t)g.n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "eval_force readback");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(c, keep);
    dfree(c, dv);
    return rc;
}

}  // namespace
}  // namespace cg

using namespace cg;

#define CG_CTX(x)                       \
    if (!(x)) return CG_ERR_INVALID;    \
    Ctx* c = &(x)->c;                   \
    CG_HIP(c, hipSetDevice(c->device))

#define CG_GRAPH(g, id)                                                        \
    Graph* g = lookup(c, (id));                                                \
    if (!g) return fail(c, CG_ERR_INVALID, "unknown graph id " + std::to_string(id))

extern "C" {

void cg_default_tuning(cg_tuning* t) {
    if (!t) return;
    const Tuning d;
    t->graph_ticks = d.graph_ticks;
    t->cell_items = d.cell_items;
    t->force_tile = d.force_tile;
    t->link_lane_max = d.lane_max;
    t->link_group_max = d.group_max;
    t->link_hub_entries = d.hub_entries;
}

int cg_set_tuning(cg_ctx* x, const cg_tuning* t) {
    CG_CTX(x);
    if (!t) return fail(c, CG_ERR_INVALID, "cg_set_tuning: null tuning");
    if (t->graph_ticks < 1 || t->graph_ticks > 256)
        return fail(c, CG_ERR_INVALID, "cg_set_tuning: graph_ticks must be in [1, 256]");
    if (t->cell_items != 4 && t->cell_items != 8)
        return fail(c, CG_ERR_INVALID, "cg_set_tuning: cell_items must be 4 or 8");
    if (t->force_tile != 512 && t->force_tile != 1024)
        return fail(c, CG_ERR_INVALID, "cg_set_tuning: force_tile must be 512 or 1024");
    if (t->link_lane_max < 2 || t->link_lane_max >= t->link_group_max || t->link_group_max > T64_LIMIT)
        return fail(c, CG_ERR_INVALID, "cg_set_tuning: need 2 <= link_lane_max < link_group_max <= 256");
    if (t->link_hub_entries != 4 && t->link_hub_entries != 6 && t->link_hub_entries != 8)
        return fail(c, CG_ERR_INVALID, "cg_set_tuning: link_hub_entries must be 4, 6 or 8");
    Tuning n;
    n.graph_ticks = t->graph_ticks;
    n.cell_items = t->cell_items;
    n.force_tile = t->force_tile;
    n.lane_max = t->link_lane_max;
    n.group_max = t->link_group_max;
    n.hub_entries = t->link_hub_entries;
    const Tuning& o = c->tune;
    const bool relayout = n.lane_max != o.lane_max || n.group_max != o.group_max || n.hub_entries != o.hub_entries;
    c->tune = n;
    CG_HIP(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured graphs hold the old launch shapes
    if (relayout) c->dirty = true;  // the link schedules follow the tuning
    return CG_OK;
}

int cg_get_tuning(cg_ctx* x, cg_tuning* t) {
    if (!x || !t) return CG_ERR_INVALID;
    const Tuning& d = x->c.tune;
    t->graph_ticks = d.graph_ticks;
    t->cell_items = d.cell_items;
    t->force_tile = d.force_tile;
    t->link_lane_max = d.lane_max;
    t->link_group_max = d.group_max;
    t->link_hub_entries = d.hub_entries;
    return CG_OK;
}

void cg_default_params(cg_sim_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->space_size = 4096.f;
    p->decay = 5000.f;
    p->gravity = 0.25f;
    p->center = 0.f;
    p->repulsion = 1.0f;
    p->link_spring = 1.f;
    p->link_distance = 10.f;
    p->link_dist_random_variation_range[0] = 1.f;
    p->link_dist_random_variation_range[1] = 1.2f;
    p->repulsion_from_mouse = 2.f;
    p->friction = 0.85f;
    p->collision = 0.f;
    p->collision_radius = 0.f;
    p->collision_padding = 0.f;
    p->point_default_size = 4.f;
    p->enable_simulation = 1;
    p->run_forces = 1;
    p->mouse_repulsion = 0;
    p->cluster = 0.1f;
}

int cg_create(cg_ctx** out, const cg_create_params* params) {
    if (!out) return CG_ERR_INVALID;
    *out = nullptr;
    cg_ctx* x = new (std::nothrow) cg_ctx();
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    c->device = params ? params->device : 0;
    c->budget = params ? params->hbm_budget : 0;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->ingest, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_desc, hipEventDisableTiming);
    if (e != hipSuccess) {
        cg_destroy(x);
        return CG_ERR_DEVICE;
    }
    // the static scratch's address on this device (used only by the
    // allocation-free entry point, which reports its absence)
    if (scratch_setup(c, c->device) != CG_OK) c->scratch = nullptr;
    c->err.clear();
    *out = x;
    return CG_OK;
}

int cg_destroy(cg_ctx* x) {
    if (!x) return CG_OK;
    Ctx* c = &x->c;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ingest) (void)hipStreamSynchronize(c->ingest);
    drop_graphs(c);
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_desc) (void)hipEventDestroy(c->ev_desc);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->ingest) (void)hipStreamDestroy(c->ingest);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    c->allocs.clear();
    for (auto& a : c->pending) (void)hipFree(a.first);
    c->pending.clear();
    for (auto& a : c->cache) (void)hipFree(a.second);
    c->cache.clear();
    if (c->h_desc_pinned) (void)hipHostFree(c->h_desc_pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete x;
    return CG_OK;
}

const char* cg_last_error(const cg_ctx* x) { return x ? x->c.err.c_str() : "null context"; }

uint64_t cg_device_bytes(const cg_ctx* x) { return x ? x->c.held : 0; }

// ---- single-graph entry points (graph 0) ---------------------------------

int cg_set_positions(cg_ctx* x, const float* xy, uint32_t n) {
    CG_CTX(x);
    return graph_set_positions(c, default_graph(c), xy, n);
}

int cg_set_links(cg_ctx* x, const float* links, uint32_t m, const float* strength,
                 const float* rand_in, const float* rand_out) {
    CG_CTX(x);
    return graph_set_links(c, default_graph(c), links, m, strength, rand_in, rand_out);
}

int cg_set_jitter(cg_ctx* x, const float* xy, uint32_t n) {
    CG_CTX(x);
    return graph_set_jitter(c, default_graph(c), xy, n);
}

int cg_set_point_sizes(cg_ctx* x, const float* sizes, uint32_t n) {
    CG_CTX(x);
    return graph_set_point_sizes(c, default_graph(c), sizes, n);
}

int cg_set_pinned(cg_ctx* x, const uint32_t* idx, uint32_t count) {
    CG_CTX(x);
    return graph_set_pinned(c, default_graph(c), idx, count);
}

int cg_set_point_clusters(cg_ctx* x, const int32_t* clusters, uint32_t n) {
    CG_CTX(x);
    return graph_set_point_clusters(c, default_graph(c), clusters, n);
}

int cg_set_cluster_positions(cg_ctx* x, const float* xy, uint32_t count) {
    CG_CTX(x);
    return graph_set_cluster_positions(c, default_graph(c), xy, count);
}

int cg_set_point_cluster_strength(cg_ctx* x, const float* strength, uint32_t n) {
    CG_CTX(x);
    return graph_set_point_cluster_strength(c, default_graph(c), strength, n);
}

int cg_set_alpha(cg_ctx* x, double alpha) {
    if (!x) return CG_ERR_INVALID;
    default_graph(&x->c).alpha = alpha;
    return CG_OK;
}

int cg_get_alpha(cg_ctx* x, double* alpha) {
    if (!x || !alpha) return CG_ERR_INVALID;
    *alpha = default_graph(&x->c).alpha;
    return CG_OK;
}

int cg_step(cg_ctx* x, const cg_sim_params* p, uint32_t nticks, const float* seeds,
            double* alpha_trace) {
    if (!x || !p) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (!p->enable_simulation) return CG_OK;
    CG_HIP(c, hipSetDevice(c->device));
    Graph& g = default_graph(c);
    set_params(c, g, *p);
    return run_ticks(c, {0u}, nticks, 0, seeds, alpha_trace);
}

int cg_prepare(cg_ctx* x, const cg_sim_params* p) {
    if (!x || !p) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    CG_HIP(c, hipSetDevice(c->device));
    Graph& g = default_graph(c);
    set_params(c, g, *p);
    if (c->dirty) {
        int rc = pack(c);
        if (rc) return rc;
    }
    if (!captures(c) || !p->enable_simulation) return CG_OK;
    // the plan cg_step runs for graph 0 alone (the other graphs are not stepped)
    const TickPlan tp{graph_passes(c->graphs[0])};
    if (tp.any == 0) return CG_OK;
    hipGraphExec_t e = nullptr;
    int rc = tick_graph(c, tp, c->tune.graph_ticks, &e);
    if (!rc) rc = tick_graph(c, tp, 1, &e);
    c->desc = c->desc_base;
    return rc;
}

int cg_get_positions(cg_ctx* x, float* xy, uint32_t n) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) {
        if (n) return fail(c, CG_ERR_INVALID, "cg_get_positions: size mismatch");
        return CG_OK;
    }
    return graph_get_positions(c, c->graphs[0], xy, n);
}

int cg_drag_points(cg_ctx* x, const uint32_t* indices, const float* xy, uint32_t count) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) return fail(c, CG_ERR_STATE, "cg_drag_points: no graph");
    return graph_drag_points(c, c->graphs[0], indices, xy, count);
}

int cg_get_positions_at(cg_ctx* x, const uint32_t* indices, uint32_t count, float* xy) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) {
        for (uint32_t k = 0; xy && k < 2 * count; ++k) xy[k] = NAN;
        return CG_OK;
    }
    return graph_get_positions_at(c, c->graphs[0], indices, count, xy);
}

// ---- packed graph set ------------------------------------------------------

int cg_graph_add(cg_ctx* x, uint32_t* id) {
    CG_CTX(x);
    if (!id) return fail(c, CG_ERR_INVALID, "cg_graph_add: null id");
    if (c->graphs.empty()) c->graphs.emplace_back();  // id 0 stays with the single-graph calls
    const uint32_t nid = (uint32_t)c->graphs.size();
    c->graphs.emplace_back();
    reset_graph(c->graphs.back(), nid);
    c->dirty = true;
    *id = nid;
    return CG_OK;
}

int cg_graph_remove(cg_ctx* x, uint32_t id) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    CG_HIP(c, hipStreamSynchronize(c->stream));
    link_free_graph(c, *g);
    unstage_graph(c, *g);
    *g = Graph();
    g->id = id;
    g->live = false;
    c->dirty = true;
    return CG_OK;
}

int cg_graph_set_params(cg_ctx* x, uint32_t id, const cg_sim_params* p) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    if (!p) return fail(c, CG_ERR_INVALID, "cg_graph_set_params: null params");
    set_params(c, *g, *p);
    return CG_OK;
}

int cg_graph_set_seed(cg_ctx* x, uint32_t id, uint64_t seed) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    g->rng = seed;
    return CG_OK;
}

int cg_graph_set_positions(cg_ctx* x, uint32_t id, const float* xy, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_positions(c, *g, xy, n);
}

int cg_graph_set_links(cg_ctx* x, uint32_t id, const float* links, uint32_t m, const float* strength,
                       const float* rand_in, const float* rand_out) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_links(c, *g, links, m, strength, rand_in, rand_out);
}

int cg_graph_set_jitter(cg_ctx* x, uint32_t id, const float* xy, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_jitter(c, *g, xy, n);
}

int cg_graph_set_point_sizes(cg_ctx* x, uint32_t id, const float* sizes, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_point_sizes(c, *g, sizes, n);
}

int cg_graph_set_pinned(cg_ctx* x, uint32_t id, const uint32_t* idx, uint32_t count) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_pinned(c, *g, idx, count);
}

int cg_graph_set_point_clusters(cg_ctx* x, uint32_t id, const int32_t* clusters, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_point_clusters(c, *g, clusters, n);
}

int cg_graph_set_cluster_positions(cg_ctx* x, uint32_t id, const float* xy, uint32_t count) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_cluster_positions(c, *g, xy, count);
}

int cg_graph_set_point_cluster_strength(cg_ctx* x, uint32_t id, const float* strength, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_set_point_cluster_strength(c, *g, strength, n);
}

int cg_graph_set_alpha(cg_ctx* x, uint32_t id, double alpha) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    g->alpha = alpha;
    return CG_OK;
}

int cg_graph_get_alpha(cg_ctx* x, uint32_t id, double* alpha) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    if (!alpha) return fail(c, CG_ERR_INVALID, "cg_graph_get_alpha: null output");
    *alpha = g->alpha;
    return CG_OK;
}

int cg_graph_get_positions(cg_ctx* x, uint32_t id, float* xy, uint32_t n) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_get_positions(c, *g, xy, n);
}

int cg_graph_drag_points(cg_ctx* x, uint32_t id, const uint32_t* indices, const float* xy,
                         uint32_t count) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_drag_points(c, *g, indices, xy, count);
}

int cg_graph_get_positions_at(cg_ctx* x, uint32_t id, const uint32_t* indices, uint32_t count, float* xy) {
    CG_CTX(x);
    CG_GRAPH(g, id);
    return graph_get_positions_at(c, *g, indices, count, xy);
}

int cg_step_graphs(cg_ctx* x, const uint32_t* ids, uint32_t count, uint32_t nticks) {
    CG_CTX(x);
    if (count && !ids) return fail(c, CG_ERR_INVALID, "cg_step_graphs: null ids");
    std::vector<uint32_t> v(ids, ids + count);
    for (uint32_t id : v) {
        CG_GRAPH(g, id);
        (void)g;
    }
    return run_ticks(c, v, nticks, -1, nullptr, nullptr);
}

int cg_sync(cg_ctx* x) {
    CG_CTX(x);
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

// ---- inspection (graph 0) --------------------------------------------------


int cg_eval_force(cg_ctx* x, const cg_sim_params* p, int32_t force, float alpha, float seed, float* v) {
    if (!x || !p) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    CG_HIP(c, hipSetDevice(c->device));
    default_graph(c);
    return eval_force(c, 0, p, force, alpha, seed, v);
}

int cg_graph_eval_force(cg_ctx* x, uint32_t id, int32_t force, float alpha, float seed, float* v) {
    CG_CTX(x);
    return eval_force(c, id, nullptr, force, alpha, seed, v);
}

int cg_eval_manybody_noalloc(cg_ctx* x, const cg_sim_params* p, float alpha, float seed, const float* xy,
                             const float* jitter, uint32_t n, float* v) {
    CG_CTX(x);
    if (!p) return fail(c, CG_ERR_INVALID, "cg_eval_manybody_noalloc: null params");
    if (n == 0) return CG_OK;
    if (!xy || !v) return fail(c, CG_ERR_INVALID, "cg_eval_manybody_noalloc: null positions or output");
    if (n > CG_NOALLOC_MAX_POINTS)
        return fail(c, CG_ERR_INVALID, "cg_eval_manybody_noalloc: " + std::to_string(n) +
                                           " points exceed the static scratch's maximum of " +
                                           std::to_string(CG_NOALLOC_MAX_POINTS));
    if (!c->scratch)
        return fail(c, CG_ERR_STATE, "cg_eval_manybody_noalloc: no static scratch address on this device");
    // present count and extent on the host (the graph's descriptor needs both)
    uint32_t np = 0;
    float ext = 0.f;
    for (uint32_t i = 0; i < n; ++i) {
        const float px = xy[2 * (size_t)i], py = xy[2 * (size_t)i + 1];
        if (std::isnan(px) || std::isnan(py)) continue;
        ++np;
        ext = std::max(ext, std::max(std::fabs(px), std::fabs(py)));
    }
    std::memset(v, 0, (size_t)n * 8);
    if (np == 0) return CG_OK;
    uint32_t gf = 0, L = 0;
    geometry(n, &gf, &L);
    // A context of one graph at offset 0 whose buffers are carved from the
    // scratch: the same launchers and kernels as cg_eval_force, no allocation.
    Ctx s;
    s.device = c->device;
    s.stream = c->stream;
    s.arena = c->scratch;
    s.arena_cap = SCRATCH_BYTES;
    s.G = 1;
    s.N = n;
    s.NP = np;
    s.F = (uint32_t)pad_cells(gf);
    s.C = level_off(L);
    int rc = CG_OK;
    s.pos[0] = (float2*)dalloc(&s, (size_t)n * 8, &rc);
    s.pos[1] = s.pos[0];  // evaluation mode writes no positions
    s.jitter = (float2*)dalloc(&s, (size_t)n * 8, &rc);
    s.flags = (uint8_t*)dalloc(&s, n, &rc);
    s.gid = (uint32_t*)dalloc(&s, (size_t)n * 4, &rc);
    s.present = (uint32_t*)dalloc(&s, (size_t)np * 4, &rc);
    s.desc_win = (GDesc*)dalloc(&s, sizeof(GDesc), &rc);
    s.desc = s.desc_win;
    s.f0s = (uint32_t*)dalloc(&s, 4, &rc);
    s.blk_graph = (uint32_t*)dalloc(&s, (size_t)(s.F / CELL_BLOCK) * 4, &rc);
    uint4 w0[PYR_W0_MAX], w1[PYR_W1_MAX];
    pyramid_work(0, gf, L, w0, &s.n_pyr[0], w1, &s.n_pyr[1]);
    s.pyr[0] = (uint4*)dalloc(&s, (size_t)(s.n_pyr[0] + 1) * 16, &rc);
    s.pyr[1] = (uint4*)dalloc(&s, (size_t)(s.n_pyr[1] + 1) * 16, &rc);
    float2* dv = (float2*)dalloc(&s, (size_t)n * 8, &rc);
    if (rc) return fail(c, rc, s.err);
    GDesc d{};
    d.n = n;
    d.np = np;
    d.gf = gf;
    d.L = L;
    d.col_h = 1.f;
    fill_dynamic(d, *p, ext, alpha, seed);
    d.passes = 1u << P_MANYBODY;
    d.inbuf = 0;
    CG_HIP(c, hipMemcpyAsync(s.pos[0], xy, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    if (jitter)
        CG_HIP(c, hipMemcpyAsync(s.jitter, jitter, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    else
        CG_HIP(c, hipMemsetAsync(s.jitter, 0, (size_t)n * 8, c->stream));
    CG_HIP(c, hipMemsetAsync(s.f0s, 0, 4, c->stream));
    CG_HIP(c, hipMemsetAsync(s.blk_graph, 0, (size_t)(s.F / CELL_BLOCK) * 4, c->stream));
    CG_HIP(c, hipMemsetAsync(dv, 0, (size_t)n * 8, c->stream));
    CG_HIP(c, hipMemcpyAsync(s.desc_win, &d, sizeof(GDesc), hipMemcpyHostToDevice, c->stream));
    CG_HIP(c, hipMemcpyAsync(s.pyr[0], w0, (size_t)s.n_pyr[0] * 16, hipMemcpyHostToDevice, c->stream));
    if (s.n_pyr[1])
        CG_HIP(c, hipMemcpyAsync(s.pyr[1], w1, (size_t)s.n_pyr[1] * 16, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_present_list, dim3(1), dim3(1024), 0, c->stream, s.pos[0], n, s.present, s.flags,
                       s.gid);
    CG_HIP(c, hipGetLastError());
    rc = mb_alloc(&s, gf);  // carves the binning and pyramid buffers; first visit order = present list
    if (!rc) rc = mb_run(&s, dv);
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        return fail(c, rc, s.err);
    }
    CG_HIP(c, hipMemcpyAsync(v, dv, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int cg_debug_slots(cg_ctx* x, int32_t* slots, uint32_t capacity, uint32_t* finest,
                   uint32_t* levels) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) return fail(c, CG_ERR_STATE, "cg_debug_slots: no graph");
    const Graph& g = c->graphs[0];
    if (finest) *finest = g.gf;
    if (levels) *levels = g.L;
    const size_t need = (size_t)g.gf * g.gf * SLOTS;
    if (!slots) return CG_OK;
    if (capacity < need || !c->slot_idx || c->dirty)
        return fail(c, CG_ERR_INVALID, "cg_debug_slots: capacity or no many-body pass yet");
    // the kernels skip the slots of empty cells (end 0): report them as -1
    const size_t cells = (size_t)g.gf * g.gf;
    std::vector<uint32_t> cnt(cells);
    CG_HIP(c, hipMemcpyAsync(slots, c->slot_idx + (size_t)g.f0 * SLOTS, need * 4, hipMemcpyDeviceToHost,
                             c->stream));
    CG_HIP(c, hipMemcpyAsync(cnt.data(), c->cell_end + g.f0, cells * 4, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t rm = 0; rm < cells; ++rm) {
        const bool empty = cnt[morton2((uint32_t)(rm % g.gf), (uint32_t)(rm / g.gf))] == 0;
        for (size_t s = 0; s < SLOTS; ++s) {
            int32_t& v = slots[rm * SLOTS + s];
            v = (empty || v < 0) ? -1 : v - (int32_t)g.p0;
        }
    }
    return CG_OK;
}

int cg_debug_cells(cg_ctx* x, uint32_t level, uint32_t* counts, float* centroid, uint32_t* members,
                   uint32_t capacity) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) return fail(c, CG_ERR_STATE, "cg_debug_cells: no graph");
    const Graph& g = c->graphs[0];
    if (c->dirty || !c->cent || g.np == 0) return fail(c, CG_ERR_STATE, "cg_debug_cells: no many-body pass yet");
    if (level >= g.L) return fail(c, CG_ERR_INVALID, "cg_debug_cells: level out of range");
    const uint32_t gl = 4u << level;
    const size_t cells = (size_t)gl * gl;
    std::vector<float4> cent(cells);
    CG_HIP(c, hipMemcpyAsync(cent.data(), c->cent + g.c0 + level_off(level), cells * sizeof(float4),
                             hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> b, e, sv;
    const bool finest = level == g.L - 1;
    if (members) {
        if (!finest) return fail(c, CG_ERR_INVALID, "cg_debug_cells: members exist for the finest level only");
        if (capacity < g.np) return fail(c, CG_ERR_INVALID, "cg_debug_cells: capacity");
        const size_t F = (size_t)g.gf * g.gf;
        b.resize(F);
        e.resize(F);
        sv.resize(c->NP);
        CG_HIP(c, hipMemcpyAsync(b.data(), c->cell_begin + g.f0, F * 4, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipMemcpyAsync(e.data(), c->cell_end + g.f0, F * 4, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipMemcpyAsync(sv.data(), c->sorted, (size_t)c->NP * 4, hipMemcpyDeviceToHost, c->stream));
    }
    CG_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t rm = 0; rm < cells; ++rm) {
        const float4 q = cent[rm];
        if (counts) counts[rm] = (uint32_t)q.z;
        if (centroid) {
            centroid[2 * rm] = q.z > 0.f ? q.x : 0.f;
            centroid[2 * rm + 1] = q.z > 0.f ? q.y : 0.f;
        }
    }
    if (members) {
        // row-major cell order; the sorted array holds each finest cell's
        // points contiguously (Morton cell order, [begin, end) per cell)
        size_t w = 0;
        for (size_t rm = 0; rm < cells; ++rm) {
            const uint32_t m = morton2((uint32_t)(rm % gl), (uint32_t)(rm / gl));
            if (e[m] == 0) continue;
            for (uint32_t k = b[m]; k < e[m]; ++k) {
                if (w >= g.np || sv[k] < g.p0 || sv[k] >= g.p0 + g.n)
                    return fail(c, CG_ERR_STATE, "cg_debug_cells: a cell holds a point of another graph");
                members[w++] = sv[k] - g.p0;
            }
        }
        if (w != g.np) return fail(c, CG_ERR_STATE, "cg_debug_cells: cells do not cover every present point");
    }
    return CG_OK;
}

int cg_debug_collision_cells(cg_ctx* x, uint32_t* keys, uint32_t* points, uint32_t* cells, uint32_t* counts,
                             uint32_t capacity, uint32_t* count, uint32_t* grid) {
    CG_CTX(x);
    if (c->graphs.empty() || !c->graphs[0].live) return fail(c, CG_ERR_STATE, "cg_debug_collision_cells: no graph");
    const Graph& g = c->graphs[0];
    if (c->dirty || !collides(g) || !c->col_keys[0])
        return fail(c, CG_ERR_STATE, "cg_debug_collision_cells: no collision pass yet");
    if (grid) *grid = g.col_G;
    if (count) *count = g.np;
    if ((keys || points) && capacity < g.np) return fail(c, CG_ERR_INVALID, "cg_debug_collision_cells: capacity");
    const uint32_t npc = c->NPC;
    const int fin = c->col_rs.passes & 1;
    std::vector<uint32_t> hk(npc), hv(npc);
    CG_HIP(c, hipMemcpyAsync(hk.data(), c->col_keys[fin], (size_t)npc * 4, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipMemcpyAsync(hv.data(), c->col_vals[fin], (size_t)npc * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> pc;
    if (cells) {
        // each point's four cells from its current positions, by the kernels' cell function
        std::vector<GDesc> table(c->G);
        for (uint32_t s = 0; s < c->G; ++s) {
            const Graph& h = c->graphs[c->order[s]];
            GDesc d = c->h_desc[s];
            fill_dynamic(d, h.params, h.extent, (float)h.alpha, 0.f);
            d.passes = 0;
            d.inbuf = h.cur ? ((1u << P_COUNT) - 1) : 0u;
            table[s] = d;
        }
        int rc = write_table(c, table);
        if (rc) return rc;
        uint32_t* dc = (uint32_t*)dalloc(c, (size_t)npc * 16, &rc);
        if (rc) return rc;
        rc = col_point_cells(c, dc);
        pc.resize(4 * (size_t)npc);
        hipError_t e = rc ? hipSuccess : hipMemcpyAsync(pc.data(), dc, (size_t)npc * 16, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(c, dc);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(c, e, "cg_debug_collision_cells readback");
    }
    if (counts) {
        // the count channel of the four grids' cell aggregates
        const size_t nc = 4 * (size_t)g.col_G * g.col_G;
        std::vector<float4> agg(nc);
        CG_HIP(c, hipMemcpyAsync(agg.data(), c->col_cell + g.col_c0, nc * 16, hipMemcpyDeviceToHost, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t q = 0; q < nc; ++q) counts[q] = (uint32_t)agg[q].w;
    }
    CG_HIP(c, hipStreamSynchronize(c->stream));
    const uint32_t lo = g.col_c0 / 4, hi = lo + g.col_G * g.col_G;
    uint32_t w = 0;
    for (uint32_t k = 0; k < npc; ++k) {
        if (hk[k] < lo || hk[k] >= hi) continue;
        if (w < capacity) {
            if (keys) keys[w] = hk[k] - lo;
            if (points) points[w] = hv[k] - g.p0;
        }
        ++w;
    }
    if (w != g.np) return fail(c, CG_ERR_STATE, "cg_debug_collision_cells: sorted pairs do not cover the graph");
    if (cells) {
        for (size_t q = 0; q < 4 * (size_t)g.n; ++q) cells[q] = 0xffffffffu;
        // the collision present list is graph-major: graph 0's points come first
        uint32_t k0 = 0;
        for (uint32_t s = 0; s < g.slot; ++s)
            if (c->h_desc[s].col_G) k0 += c->h_desc[s].np;
        std::vector<uint32_t> cp(g.np);
        CG_HIP(c, hipMemcpy(cp.data(), c->cpresent + k0, (size_t)g.np * 4, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < g.np; ++k)
            for (int o = 0; o < 4; ++o) cells[4 * (size_t)(cp[k] - g.p0) + o] = pc[4 * (size_t)(k0 + k) + o];
    }
    return CG_OK;
}

int cg_debug_link_csr(cg_ctx* x, int32_t dir, uint32_t* offsets, uint32_t* nbr, float* bias,
                      float* strength, uint32_t* degree, uint32_t capacity, uint32_t* entries) {
    if (!x || dir < 0 || dir > 1) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    CG_HIP(c, hipSetDevice(c->device));
    if (c->graphs.empty() || !c->graphs[0].live) return fail(c, CG_ERR_STATE, "cg_debug_link_csr: no graph");
    const Graph& g = c->graphs[0];
    const GraphLinks& lk = g.lk;
    if (entries) *entries = lk.E[dir];
    if (!g.links_ready) return fail(c, CG_ERR_STATE, "cg_debug_link_csr: no links");
    const uint32_t n = g.n, E = lk.E[dir];
    if (offsets) CG_HIP(c, hipMemcpy(offsets, lk.off_pt[dir], ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
    if (degree) CG_HIP(c, hipMemcpy(degree, lk.degree, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (nbr || bias || strength) {
        if (capacity < E) return fail(c, CG_ERR_INVALID, "cg_debug_link_csr: capacity");
        std::vector<uint32_t> ent((size_t)E * 4);
        if (E) CG_HIP(c, hipMemcpy(ent.data(), lk.ent_pt[dir], (size_t)E * 16, hipMemcpyDeviceToHost));
        for (uint32_t e = 0; e < E; ++e) {
            if (nbr) nbr[e] = ent[4 * (size_t)e];
            if (strength) std::memcpy(&strength[e], &ent[4 * (size_t)e + 2], 4);
            if (bias) std::memcpy(&bias[e], &ent[4 * (size_t)e + 3], 4);
        }
    }
    return CG_OK;
}

int cg_debug_link_stats(cg_ctx* x, uint32_t* out16) {
    if (!x || !out16) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    std::memset(out16, 0, 16 * sizeof(uint32_t));
    if (!c->link_stats) return CG_OK;
    CG_HIP(c, hipSetDevice(c->device));
    CG_HIP(c, hipMemcpyAsync(out16, c->link_stats, 64, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipMemsetAsync(c->link_stats, 0, 64, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int cg_set_capture(cg_ctx* x, int32_t enable) {
    if (!x) return CG_ERR_INVALID;
    x->c.capture = enable != 0;
    return CG_OK;
}

int cg_set_profiling(cg_ctx* x, uint32_t mask) {
    if (!x) return CG_ERR_INVALID;
    x->c.prof_mask = mask & (((1u << CG_STAGES) - 1) | CG_PROFILE_LINK_STATS);
    return CG_OK;
}

int cg_get_profile(cg_ctx* x, double* ms, uint32_t* launches, uint32_t nstages) {
    CG_CTX(x);
    CG_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < nstages; ++s) {
        if (ms) ms[s] = 0.0;
        if (launches) launches[s] = 0;
    }
    for (auto& r : c->prof_rec) {
        float t = 0.f;
        CG_HIP(c, hipEventElapsedTime(&t, r.second.first, r.second.second));
        if ((uint32_t)r.first < nstages) {
            if (ms) ms[r.first] += t;
            if (launches) launches[r.first] += 1;
        }
    }
    c->prof_rec.clear();
    c->prof_used = 0;
    for (auto& e : c->prof_open) e = nullptr;
    return CG_OK;
}

}  // extern "C"
