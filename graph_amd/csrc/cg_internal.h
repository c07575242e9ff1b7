// cg_internal.h — shared state and device helpers of the force-simulation library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cosmosgl.h"

namespace cg {

// Per-point flag bits (device byte array `flags`).
constexpr uint8_t F_ABSENT = 1;   // NaN position at set time (Points/index.ts:1313-1386)
constexpr uint8_t F_PINNED = 2;   // setPinnedPoints (Points/index.ts:1264-1304)

constexpr int MAX_LEVELS = 8;     // 4^2 .. 512^2 (ForceManyBody/index.ts:19, 646-656)
constexpr int SLOTS = 8;          // NEAR_FIELD_SLOTS (ForceManyBody/index.ts:35)
constexpr uint32_t BIG_CELL = 64; // finest cells above this many points use a wave each

// Link degree tiers: points with at most LANE_MAX entries run one lane each;
// up to WAVE_MAX entries one 64-lane workgroup each; above, 1024 lanes each.
constexpr uint32_t LANE_MAX = 32;
constexpr uint32_t WAVE_MAX = 4096;

struct LinkDir {
    uint32_t E = 0;              // entries (valid links) in this direction
    uint32_t E_live = 0;         // entries whose neighbour is present
    uint4* ent_pt = nullptr;     // point-ordered entries {nbr, rand, strength, bias} (reference order)
    uint32_t* off_pt = nullptr;  // [n+1] point-ordered offsets
    uint4* ent = nullptr;        // schedule-ordered live entries used by the kernels
    uint32_t* sched = nullptr;   // schedule: present points by ascending live degree
    uint32_t* soff = nullptr;    // [n_sched+1] offsets of each schedule slot into ent
    uint32_t n_sched = 0;
    uint32_t n_lane = 0, n_wave = 0, n_block = 0;  // tier sizes, in schedule order
    uint32_t e_chunk0 = 0;       // first entry (in ent) of the chunked tiers
    uint32_t* poff = nullptr;    // [n_wave+n_block+1] step-major offsets into link_pre
    uint32_t* eslot = nullptr;   // schedule slot of each chunked-tier entry
    float2* warm = nullptr;      // per chunk-tier lane: last converged chunk start
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;     // second stream (default priority) for the hub link tier
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint64_t budget = 0;
    uint64_t held = 0;
    std::string err;
    std::vector<std::pair<void*, size_t>> allocs;

    uint32_t n = 0;       // points
    uint32_t np = 0;      // present points
    double alpha = 1.0;
    uint64_t rng = 0x5eed5eedULL;

    // host copies of the inputs needed to rebuild derived state
    std::vector<float> h_links, h_strength, h_rin, h_rout;
    uint32_t m = 0;
    bool links_set = false;
    std::vector<uint8_t> h_flags;
    std::vector<float> h_sizes;     // raw sizes (NaN = default), empty = none set

    // point state
    float2* pos[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t* flags = nullptr;
    uint32_t* present = nullptr;
    float2* jitter = nullptr;

    // many-body
    uint32_t gf = 0, L = 0;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint32_t* cell_begin = nullptr;
    uint32_t* cell_end = nullptr;
    double* sums[MAX_LEVELS] = {};
    float4* cent[MAX_LEVELS] = {};
    float2* sorted_pos = nullptr;
    int32_t* slot_idx = nullptr;
    float2* slot_pos = nullptr;
    uint32_t* big_list = nullptr;
    uint32_t* big_count = nullptr;
    double* center_part = nullptr;  // block partial sums for the centroid
    double* center_sum = nullptr;   // [3]

    // collision
    uint32_t col_G = 0;
    uint32_t* col_begin = nullptr;  // [G*G]
    uint32_t* col_end = nullptr;
    float4* col_cell = nullptr;     // [4][G*G] {sum x, sum y, sum size, count}
    float* col_size = nullptr;      // resolved sizes [n]
    float col_max_size = 0.f;
    float col_last_default = -1.f;

    // links
    uint32_t* degree = nullptr;
    LinkDir dir[2];
    bool links_ready = false;
    float4* link_pre = nullptr;     // per-entry (q.x, q.y, D, K) of the chunked tiers
    uint32_t* link_stats = nullptr; // multiple-shooting sweep counters (cg_debug_link_stats)

    // stage timing (cg_set_profiling)
    uint32_t prof_mask = 0;
    std::vector<hipEvent_t> prof_pool;
    size_t prof_used = 0;
    hipEvent_t prof_open[CG_STAGES] = {};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_rec;

    std::vector<uint32_t> h_pinned;  // pinned indices as given (re-applied when n changes)
};

void prof_begin(Ctx* c, int stage);
void prof_end(Ctx* c, int stage);

// error helpers (api.hip)
int fail(Ctx* c, int code, const std::string& msg);
int hip_fail(Ctx* c, hipError_t e, const char* what);
void* dalloc(Ctx* c, size_t bytes, int* rc);
void dfree(Ctx* c, void* p);

#define CG_HIP(c, expr)                                  \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return cg::hip_fail(c, e_, #expr); \
    } while (0)

// launchers (k_*.hip)
int mb_alloc(Ctx* c);                                  // k_manybody.hip
int mb_run(Ctx* c, const cg_sim_params* p, float alpha, float seed, const float2* pin,
           float2* pout, float2* vout, bool keys_ready);
int pw_run(Ctx* c, const cg_sim_params* p, int force, float alpha, const float2* pin,
           float2* pout, float2* vout, bool write_keys);   // k_pointwise.hip
int keys_only(Ctx* c, const float2* pin, float space_size);
int link_build(Ctx* c);                                // k_link.hip
int link_run(Ctx* c, const cg_sim_params* p, int dir, float alpha, const float2* pin,
             float2* pout, float2* vout);
int col_prepare(Ctx* c, const cg_sim_params* p);       // k_collision.hip
int col_run(Ctx* c, const cg_sim_params* p, float alpha, const float2* pin, float2* pout,
            float2* vout);

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

inline unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace cg
