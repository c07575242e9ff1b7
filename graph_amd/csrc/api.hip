// api.hip — the extern "C" boundary: context lifetime, ingest, the tick
// driver and the inspection entry points.
//
// The tick driver restates Graph.runSimulationStep (src/index.ts:2006-2099):
// [mouse -> integrate] -> gravity -> center -> many-body -> link IN ->
// link OUT -> collision, each force followed by its own integrate pass (fused
// into the force kernel here), then the alpha decay of Store.addAlpha
// (Store/index.ts:328-332).  Positions ping-pong between two device buffers.
#include <cmath>
#include <cstdio>
#include <new>

#include "cg_internal.h"

struct cg_ctx {
    cg::Ctx c;
};

namespace cg {

int fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
    return fail(c, CG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocation charged against the context's HBM budget.  Sets *rc only
// on failure, so a chain of calls can share one status.
void* dalloc(Ctx* c, size_t bytes, int* rc) {
    if (c->budget && c->held + bytes > c->budget) {
        *rc = fail(c, CG_ERR_BUDGET, "device allocation of " + std::to_string(bytes) +
                                         " bytes exceeds the context budget");
        return nullptr;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *rc = hip_fail(c, e, "hipMalloc");
        return nullptr;
    }
    c->held += bytes;
    c->allocs.emplace_back(p, bytes);
    return p;
}

void dfree(Ctx* c, void* p) {
    if (!p) return;
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        if (c->allocs[k].first == p) {
            c->held -= c->allocs[k].second;
            c->allocs[k] = c->allocs.back();
            c->allocs.pop_back();
            break;
        }
    }
    (void)hipFree(p);
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

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
float rng_unit(Ctx* c) { return (float)(splitmix(c->rng) >> 40) * 0x1.0p-24f; }

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

void free_points(Ctx* c) {
    dfree(c, c->pos[0]); dfree(c, c->pos[1]);
    dfree(c, c->flags); dfree(c, c->present); dfree(c, c->jitter);
    for (int b = 0; b < 2; ++b) { dfree(c, c->keys[b]); dfree(c, c->vals[b]); c->keys[b] = c->vals[b] = nullptr; }
    dfree(c, c->sort_tmp); dfree(c, c->cell_begin); dfree(c, c->cell_end);
    for (int l = 0; l < MAX_LEVELS; ++l) { dfree(c, c->sums[l]); dfree(c, c->cent[l]); c->sums[l] = nullptr; c->cent[l] = nullptr; }
    dfree(c, c->sorted_pos); dfree(c, c->slot_idx); dfree(c, c->slot_pos);
    dfree(c, c->big_list); dfree(c, c->big_count);
    dfree(c, c->col_begin); dfree(c, c->col_end); dfree(c, c->col_cell); dfree(c, c->col_size);
    c->pos[0] = c->pos[1] = nullptr;
    c->flags = nullptr; c->present = nullptr; c->jitter = nullptr;
    c->sort_tmp = nullptr; c->cell_begin = c->cell_end = nullptr;
    c->sorted_pos = nullptr; c->slot_idx = nullptr; c->slot_pos = nullptr;
    c->big_list = c->big_count = nullptr;
    c->col_begin = c->col_end = nullptr; c->col_cell = nullptr; c->col_size = nullptr;
    c->col_G = 0;
}

int upload_flags(Ctx* c) {
    CG_HIP(c, hipMemcpyAsync(c->flags, c->h_flags.data(), c->n, hipMemcpyHostToDevice, c->stream));
    return CG_OK;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

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
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete x;
        return CG_ERR_DEVICE;
    }
    *out = x;
    return CG_OK;
}

int cg_destroy(cg_ctx* x) {
    if (!x) return CG_OK;
    Ctx* c = &x->c;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    c->allocs.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete x;
    return CG_OK;
}

const char* cg_last_error(const cg_ctx* x) { return x ? x->c.err.c_str() : "null context"; }

uint64_t cg_device_bytes(const cg_ctx* x) { return x ? x->c.held : 0; }

int cg_set_positions(cg_ctx* x, const float* xy, uint32_t n) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (n && !xy) return fail(c, CG_ERR_INVALID, "cg_set_positions: null positions");
    CG_HIP(c, hipSetDevice(c->device));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    const bool same_n = (n == c->n) && c->pos[0];
    std::vector<float> keep_jitter;
    free_points(c);
    c->n = n;
    c->cur = 0;
    // Points/position-utils.ts:17-24: a half-NaN position becomes fully NaN (absent)
    std::vector<float> pos(2 * (size_t)n);
    std::vector<uint32_t> present;
    c->h_flags.assign(n, 0);
    present.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        float px = xy[2 * i], py = xy[2 * i + 1];
        bool absent = std::isnan(px) || std::isnan(py);
        pos[2 * i] = absent ? NAN : px;
        pos[2 * i + 1] = absent ? NAN : py;
        if (absent) c->h_flags[i] = F_ABSENT;
        else present.push_back(i);
    }
    for (uint32_t k : c->h_pinned)
        if (k < n) c->h_flags[k] |= F_PINNED;
    c->np = (uint32_t)present.size();
    geometry(n, &c->gf, &c->L);
    int rc = CG_OK;
    const size_t nn = n ? n : 1;
    c->pos[0] = (float2*)dalloc(c, nn * 8, &rc); if (rc) return rc;
    c->pos[1] = (float2*)dalloc(c, nn * 8, &rc); if (rc) return rc;
    c->flags = (uint8_t*)dalloc(c, nn, &rc); if (rc) return rc;
    c->present = (uint32_t*)dalloc(c, (c->np ? c->np : 1) * 4, &rc); if (rc) return rc;
    c->jitter = (float2*)dalloc(c, nn * 8, &rc); if (rc) return rc;
    if (n) {
        CG_HIP(c, hipMemcpyAsync(c->pos[0], pos.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipMemcpyAsync(c->pos[1], pos.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        rc = upload_flags(c); if (rc) return rc;
        if (c->np)
            CG_HIP(c, hipMemcpyAsync(c->present, present.data(), (size_t)c->np * 4, hipMemcpyHostToDevice, c->stream));
        // default jitter from the context generator: (U(-1,1)*1e-5, U(-1,1)*1e-5)
        std::vector<float> jit(2 * (size_t)n);
        for (auto& v : jit) v = (rng_unit(c) * 2.f - 1.f) * 0.00001f;
        CG_HIP(c, hipMemcpyAsync(c->jitter, jit.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (!same_n) c->h_sizes.clear();
    rc = mb_alloc(c); if (rc) return rc;
    if (c->links_set) { rc = link_build(c); if (rc) return rc; }
    return CG_OK;
}

int cg_set_links(cg_ctx* x, const float* links, uint32_t m, const float* strength,
                 const float* rand_in, const float* rand_out) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (m && !links) return fail(c, CG_ERR_INVALID, "cg_set_links: null links");
    CG_HIP(c, hipSetDevice(c->device));
    c->m = m;
    c->h_links.assign(links, links + 2 * (size_t)m);
    if (strength) c->h_strength.assign(strength, strength + m);
    else c->h_strength.clear();
    c->h_rin.resize(m);
    c->h_rout.resize(m);
    for (uint32_t l = 0; l < m; ++l) c->h_rin[l] = rand_in ? rand_in[l] : rng_unit(c);
    for (uint32_t l = 0; l < m; ++l) c->h_rout[l] = rand_out ? rand_out[l] : rng_unit(c);
    c->links_set = true;
    return link_build(c);
}

int cg_set_jitter(cg_ctx* x, const float* xy, uint32_t n) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (n != c->n || (n && !xy)) return fail(c, CG_ERR_INVALID, "cg_set_jitter: expected one pair per point");
    CG_HIP(c, hipSetDevice(c->device));
    if (n) {
        CG_HIP(c, hipMemcpyAsync(c->jitter, xy, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return CG_OK;
}

int cg_set_point_sizes(cg_ctx* x, const float* sizes, uint32_t n) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (sizes && n != c->n) return fail(c, CG_ERR_INVALID, "cg_set_point_sizes: expected one size per point");
    if (sizes) c->h_sizes.assign(sizes, sizes + n);
    else c->h_sizes.clear();
    c->col_last_default = -1.f;  // re-resolve sizes on the next collision pass
    return CG_OK;
}

int cg_set_pinned(cg_ctx* x, const uint32_t* idx, uint32_t count) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (count && !idx) return fail(c, CG_ERR_INVALID, "cg_set_pinned: null indices");
    CG_HIP(c, hipSetDevice(c->device));
    c->h_pinned.assign(idx, idx + count);
    for (auto& f : c->h_flags) f &= (uint8_t)~F_PINNED;
    for (uint32_t k = 0; k < count; ++k)
        if (idx[k] < c->n) c->h_flags[idx[k]] |= F_PINNED;
    if (c->n) {
        int rc = upload_flags(c); if (rc) return rc;
        CG_HIP(c, hipStreamSynchronize(c->stream));
    }
    return CG_OK;
}

int cg_set_alpha(cg_ctx* x, double alpha) {
    if (!x) return CG_ERR_INVALID;
    x->c.alpha = alpha;
    return CG_OK;
}

int cg_get_alpha(cg_ctx* x, double* alpha) {
    if (!x || !alpha) return CG_ERR_INVALID;
    *alpha = x->c.alpha;
    return CG_OK;
}

static bool on(float v) { return v != 0.f && !std::isnan(v); }  // JS truthiness of a number

int cg_step(cg_ctx* x, const cg_sim_params* p, uint32_t nticks, const float* seeds,
            double* alpha_trace) {
    if (!x || !p) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (!p->enable_simulation) return CG_OK;
    CG_HIP(c, hipSetDevice(c->device));
    const double decay_factor = 1.0 - std::pow(0.001, 1.0 / (double)p->decay);
    const bool have_links = c->links_ready && c->m > 0;  // linksTextureSize != 0
    int rc = CG_OK;
    for (uint32_t t = 0; t < nticks; ++t) {
        const float a = (float)c->alpha;
        const float seed = seeds ? seeds[t] : rng_unit(c);
        if (c->np && p->mouse_repulsion) {
            prof_begin(c, CG_STAGE_MOUSE);
            rc = pw_run(c, p, CG_FORCE_MOUSE, a, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr, false);
            prof_end(c, CG_STAGE_MOUSE);
            if (rc) return rc;
            c->cur ^= 1;
        }
        if (p->run_forces) {
            if (c->np) {
                const bool grav = on(p->gravity), cent = on(p->center);
                if (grav) {
                    prof_begin(c, CG_STAGE_GRAVITY);
                    rc = pw_run(c, p, CG_FORCE_GRAVITY, a, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr, !cent);
                    prof_end(c, CG_STAGE_GRAVITY);
                    if (rc) return rc;
                    c->cur ^= 1;
                }
                if (cent) {
                    prof_begin(c, CG_STAGE_CENTER);
                    rc = pw_run(c, p, CG_FORCE_CENTER, a, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr, true);
                    prof_end(c, CG_STAGE_CENTER);
                    if (rc) return rc;
                    c->cur ^= 1;
                }
                rc = mb_run(c, p, a, seed, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr, grav || cent);
                if (rc) return rc;
                c->cur ^= 1;
                if (have_links) {
                    for (int d = 0; d < 2; ++d) {
                        prof_begin(c, CG_STAGE_LINK_IN + d);
                        rc = link_run(c, p, d, a, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr);
                        prof_end(c, CG_STAGE_LINK_IN + d);
                        if (rc) return rc;
                        c->cur ^= 1;
                    }
                }
                if (on(p->collision)) {
                    prof_begin(c, CG_STAGE_COLLISION);
                    rc = col_run(c, p, a, c->pos[c->cur], c->pos[c->cur ^ 1], nullptr);
                    prof_end(c, CG_STAGE_COLLISION);
                    if (rc) return rc;
                    c->cur ^= 1;
                }
            }
            c->alpha += (0.0 - c->alpha) * decay_factor;
            if (p->mouse_repulsion) c->alpha = std::max(c->alpha, 0.1);
        }
        if (alpha_trace) alpha_trace[t] = c->alpha;
    }
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

int cg_get_positions(cg_ctx* x, float* xy, uint32_t n) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (n != c->n || (n && !xy)) return fail(c, CG_ERR_INVALID, "cg_get_positions: size mismatch");
    CG_HIP(c, hipSetDevice(c->device));
    if (n) {
        CG_HIP(c, hipMemcpyAsync(xy, c->pos[c->cur], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int cg_eval_force(cg_ctx* x, const cg_sim_params* p, int32_t force, float alpha, float seed,
                  float* v) {
    if (!x || !p) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (c->n && !v) return fail(c, CG_ERR_INVALID, "cg_eval_force: null output");
    CG_HIP(c, hipSetDevice(c->device));
    if (c->n == 0) return CG_OK;
    int rc = CG_OK;
    float2* dv = (float2*)dalloc(c, (size_t)c->n * 8, &rc);
    if (rc) return rc;
    CG_HIP(c, hipMemsetAsync(dv, 0, (size_t)c->n * 8, c->stream));
    const float2* pin = c->pos[c->cur];
    switch (force) {
        case CG_FORCE_GRAVITY:
        case CG_FORCE_CENTER:
        case CG_FORCE_MOUSE:
            rc = pw_run(c, p, force, alpha, pin, nullptr, dv, false);
            break;
        case CG_FORCE_MANYBODY:
            rc = mb_run(c, p, alpha, seed, pin, nullptr, dv, false);
            break;
        case CG_FORCE_LINK_IN:
        case CG_FORCE_LINK_OUT:
            rc = link_run(c, p, force - CG_FORCE_LINK_IN, alpha, pin, nullptr, dv);
            break;
        case CG_FORCE_COLLISION:
            rc = col_run(c, p, alpha, pin, nullptr, dv);
            break;
        default:
            rc = fail(c, CG_ERR_INVALID, "cg_eval_force: unknown force");
    }
    if (rc == CG_OK) {
        hipError_t e = hipMemcpyAsync(v, dv, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "cg_eval_force readback");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(c, dv);
    return rc;
}

int cg_debug_slots(cg_ctx* x, int32_t* slots, uint32_t capacity, uint32_t* finest,
                   uint32_t* levels) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    if (finest) *finest = c->gf;
    if (levels) *levels = c->L;
    const size_t need = (size_t)c->gf * c->gf * SLOTS;
    if (!slots) return CG_OK;
    if (capacity < need || !c->slot_idx) return fail(c, CG_ERR_INVALID, "cg_debug_slots: capacity");
    CG_HIP(c, hipSetDevice(c->device));
    CG_HIP(c, hipMemcpyAsync(slots, c->slot_idx, need * 4, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int cg_debug_link_csr(cg_ctx* x, int32_t dir, uint32_t* offsets, uint32_t* nbr, float* bias,
                      float* strength, uint32_t* degree, uint32_t capacity, uint32_t* entries) {
    if (!x || dir < 0 || dir > 1) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    const LinkDir& L = c->dir[dir];
    if (entries) *entries = L.E;
    if (!c->links_ready) return fail(c, CG_ERR_STATE, "cg_debug_link_csr: no links");
    CG_HIP(c, hipSetDevice(c->device));
    const uint32_t n = c->n;
    if (offsets) CG_HIP(c, hipMemcpy(offsets, L.off_pt, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
    if (degree) CG_HIP(c, hipMemcpy(degree, c->degree, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (nbr || bias || strength) {
        if (capacity < L.E) return fail(c, CG_ERR_INVALID, "cg_debug_link_csr: capacity");
        std::vector<uint32_t> ent((size_t)L.E * 4);
        if (L.E) CG_HIP(c, hipMemcpy(ent.data(), L.ent_pt, (size_t)L.E * 16, hipMemcpyDeviceToHost));
        for (uint32_t e = 0; e < L.E; ++e) {
            if (nbr) nbr[e] = ent[4 * (size_t)e];
            if (strength) std::memcpy(&strength[e], &ent[4 * (size_t)e + 2], 4);
            if (bias) std::memcpy(&bias[e], &ent[4 * (size_t)e + 3], 4);
        }
    }
    return CG_OK;
}

int cg_debug_link_stats(cg_ctx* x, uint32_t* out8) {
    if (!x || !out8) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    std::memset(out8, 0, 8 * sizeof(uint32_t));
    if (!c->link_stats) return CG_OK;
    CG_HIP(c, hipSetDevice(c->device));
    CG_HIP(c, hipMemcpyAsync(out8, c->link_stats, 32, hipMemcpyDeviceToHost, c->stream));
    CG_HIP(c, hipMemsetAsync(c->link_stats, 0, 64, c->stream));
    CG_HIP(c, hipStreamSynchronize(c->stream));
    return CG_OK;
}

int cg_set_profiling(cg_ctx* x, uint32_t mask) {
    if (!x) return CG_ERR_INVALID;
    x->c.prof_mask = mask & (((1u << CG_STAGES) - 1) | CG_PROFILE_LINK_STATS);
    return CG_OK;
}

int cg_get_profile(cg_ctx* x, double* ms, uint32_t* launches, uint32_t nstages) {
    if (!x) return CG_ERR_INVALID;
    Ctx* c = &x->c;
    CG_HIP(c, hipSetDevice(c->device));
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
