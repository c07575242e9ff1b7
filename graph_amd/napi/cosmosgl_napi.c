/*
 * cosmosgl_napi.c — Node-API binding of include/cosmosgl.h.
 *
 * The JavaScript side of cosmos.gl calls these exports in place of its
 * ForceX modules (see the C ABI header for the reference interface each entry
 * point replaces).  Typed arrays are passed by pointer, never copied here; the
 * library copies inputs itself.  Every failure throws a JS Error carrying
 * cg_last_error().
 *
 * Exports: create, destroy, setPositions, setLinks, setJitter, setPointSizes,
 * setPinned, setAlpha, getAlpha, step, getPositions, defaultParams.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cosmosgl.h"

#define CHECK(env, call)                                                     \
    do {                                                                     \
        if ((call) != napi_ok) {                                             \
            napi_throw_error((env), NULL, "cosmosgl: N-API call failed: " #call); \
            return NULL;                                                     \
        }                                                                    \
    } while (0)

static napi_value throw_rc(napi_env env, cg_ctx* ctx, int rc) {
    char msg[512];
    snprintf(msg, sizeof msg, "cosmosgl error %d: %s", rc, ctx ? cg_last_error(ctx) : "");
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static void finalize_ctx(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    /* a handle already destroyed explicitly holds NULL in its slot */
    cg_ctx** slot = (cg_ctx**)data;
    if (*slot) cg_destroy(*slot);
    *slot = NULL;
    free(slot);
}

static cg_ctx* get_ctx(napi_env env, napi_value v) {
    void* data = NULL;
    if (napi_get_value_external(env, v, &data) != napi_ok || !data || !*(cg_ctx**)data) {
        napi_throw_type_error(env, NULL, "cosmosgl: expected a live context handle");
        return NULL;
    }
    return *(cg_ctx**)data;
}

/* Typed array argument of the given element type (or null/undefined if allowed). */
static int get_typed(napi_env env, napi_value v, napi_typedarray_type want, void** ptr,
                     size_t* len, int nullable) {
    napi_valuetype vt;
    if (napi_typeof(env, v, &vt) != napi_ok) return 0;
    if (nullable && (vt == napi_null || vt == napi_undefined)) {
        *ptr = NULL;
        *len = 0;
        return 1;
    }
    bool is_ta = false;
    if (napi_is_typedarray(env, v, &is_ta) != napi_ok || !is_ta) return 0;
    napi_typedarray_type t;
    napi_value ab;
    size_t off;
    if (napi_get_typedarray_info(env, v, &t, len, ptr, &ab, &off) != napi_ok) return 0;
    return t == want;
}

static napi_value js_create(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_create_params cp = {0, 0};
    if (argc > 0) {
        int32_t dev = 0;
        if (napi_get_value_int32(env, argv[0], &dev) == napi_ok) cp.device = dev;
    }
    if (argc > 1) {
        double b = 0;
        if (napi_get_value_double(env, argv[1], &b) == napi_ok && b > 0) cp.hbm_budget = (uint64_t)b;
    }
    cg_ctx* ctx = NULL;
    int rc = cg_create(&ctx, &cp);
    if (rc != CG_OK) return throw_rc(env, NULL, rc);
    cg_ctx** slot = (cg_ctx**)malloc(sizeof(cg_ctx*));
    *slot = ctx;
    napi_value ext;
    CHECK(env, napi_create_external(env, slot, finalize_ctx, NULL, &ext));
    return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    void* data = NULL;
    if (argc < 1 || napi_get_value_external(env, argv[0], &data) != napi_ok || !data) {
        napi_throw_type_error(env, NULL, "cosmosgl: expected a context handle");
        return NULL;
    }
    cg_ctx** slot = (cg_ctx**)data;
    if (*slot) cg_destroy(*slot);
    *slot = NULL;
    return NULL;
}

static napi_value js_set_positions(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p;
    size_t len;
    if (argc < 2 || !get_typed(env, argv[1], napi_float32_array, &p, &len, 0)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setPositions: expected a Float32Array");
        return NULL;
    }
    int rc = cg_set_positions(ctx, (const float*)p, (uint32_t)(len / 2));
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_links(napi_env env, napi_callback_info info) {
    size_t argc = 5;
    napi_value argv[5];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void *l = NULL, *s = NULL, *ri = NULL, *ro = NULL;
    size_t nl = 0, ns = 0, ni = 0, no = 0;
    if (argc < 2 || !get_typed(env, argv[1], napi_float32_array, &l, &nl, 0) ||
        (argc > 2 && !get_typed(env, argv[2], napi_float32_array, &s, &ns, 1)) ||
        (argc > 3 && !get_typed(env, argv[3], napi_float32_array, &ri, &ni, 1)) ||
        (argc > 4 && !get_typed(env, argv[4], napi_float32_array, &ro, &no, 1))) {
        napi_throw_type_error(env, NULL,
                              "cosmosgl.setLinks: expected (handle, Float32Array links, strength?, randIn?, randOut?)");
        return NULL;
    }
    uint32_t m = (uint32_t)(nl / 2);
    /* GraphData.updateLinkStrength: a strength array of the wrong length is ignored */
    if (s && ns != m) s = NULL;
    if ((ri && ni != m) || (ro && no != m)) {
        napi_throw_range_error(env, NULL, "cosmosgl.setLinks: one draw per link expected");
        return NULL;
    }
    int rc = cg_set_links(ctx, (const float*)l, m, (const float*)s, (const float*)ri, (const float*)ro);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_jitter(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p;
    size_t len;
    if (argc < 2 || !get_typed(env, argv[1], napi_float32_array, &p, &len, 0)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setJitter: expected a Float32Array");
        return NULL;
    }
    int rc = cg_set_jitter(ctx, (const float*)p, (uint32_t)(len / 2));
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_point_sizes(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p = NULL;
    size_t len = 0;
    if (argc > 1 && !get_typed(env, argv[1], napi_float32_array, &p, &len, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setPointSizes: expected a Float32Array or null");
        return NULL;
    }
    int rc = cg_set_point_sizes(ctx, (const float*)p, (uint32_t)len);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_pinned(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p = NULL;
    size_t len = 0;
    if (argc > 1 && !get_typed(env, argv[1], napi_uint32_array, &p, &len, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setPinned: expected a Uint32Array or null");
        return NULL;
    }
    int rc = cg_set_pinned(ctx, (const uint32_t*)p, (uint32_t)len);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

// setPointClusters (Int32Array, negative = none), setClusterPositions
// (Float32Array [x, y] per cluster, negative = centroid), setPointClusterStrength
// (Float32Array); null clears each.
static napi_value js_set_point_clusters(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p = NULL;
    size_t len = 0;
    if (argc > 1 && !get_typed(env, argv[1], napi_int32_array, &p, &len, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setPointClusters: expected an Int32Array or null");
        return NULL;
    }
    int rc = cg_set_point_clusters(ctx, (const int32_t*)p, (uint32_t)len);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_cluster_positions(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p = NULL;
    size_t len = 0;
    if (argc > 1 && !get_typed(env, argv[1], napi_float32_array, &p, &len, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setClusterPositions: expected a Float32Array or null");
        return NULL;
    }
    int rc = cg_set_cluster_positions(ctx, (const float*)p, (uint32_t)(len / 2));
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_point_cluster_strength(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p = NULL;
    size_t len = 0;
    if (argc > 1 && !get_typed(env, argv[1], napi_float32_array, &p, &len, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.setPointClusterStrength: expected a Float32Array or null");
        return NULL;
    }
    int rc = cg_set_point_cluster_strength(ctx, (const float*)p, (uint32_t)len);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_set_alpha(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    double a = 1.0;
    if (argc < 2 || napi_get_value_double(env, argv[1], &a) != napi_ok) {
        napi_throw_type_error(env, NULL, "cosmosgl.setAlpha: expected a number");
        return NULL;
    }
    int rc = cg_set_alpha(ctx, a);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return NULL;
}

static napi_value js_get_alpha(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    double a = 0;
    int rc = cg_get_alpha(ctx, &a);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    napi_value out;
    CHECK(env, napi_create_double(env, a, &out));
    return out;
}

/* params: plain object with the reference's config key names. */
static void read_f32(napi_env env, napi_value obj, const char* key, float* dst) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
    napi_value v;
    double d;
    if (napi_get_named_property(env, obj, key, &v) != napi_ok) return;
    napi_valuetype vt;
    napi_typeof(env, v, &vt);
    if (vt == napi_undefined || vt == napi_null) return;  /* undefined = default */
    if (vt == napi_boolean) {
        bool b = false;
        napi_get_value_bool(env, v, &b);
        *dst = b ? 1.f : 0.f;
        return;
    }
    if (napi_get_value_double(env, v, &d) == napi_ok) *dst = (float)d;
}

static void read_bool(napi_env env, napi_value obj, const char* key, int32_t* dst) {
    float f = (float)*dst;
    read_f32(env, obj, key, &f);
    *dst = f != 0.f;
}

static void read_pair(napi_env env, napi_value obj, const char* key, float* dst) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return;
    napi_value v;
    bool arr = false;
    if (napi_get_named_property(env, obj, key, &v) != napi_ok) return;
    if (napi_is_array(env, v, &arr) != napi_ok || !arr) return;
    uint32_t n = 0;
    napi_get_array_length(env, v, &n);
    if (n != 2) return;
    for (uint32_t i = 0; i < 2; ++i) {
        napi_value e;
        double d;
        if (napi_get_element(env, v, i, &e) == napi_ok && napi_get_value_double(env, e, &d) == napi_ok)
            dst[i] = (float)d;
    }
}

static void read_params(napi_env env, napi_value obj, cg_sim_params* p) {
    cg_default_params(p);
    read_f32(env, obj, "spaceSize", &p->space_size);
    read_f32(env, obj, "simulationDecay", &p->decay);
    read_f32(env, obj, "simulationGravity", &p->gravity);
    read_f32(env, obj, "simulationCenter", &p->center);
    read_f32(env, obj, "simulationRepulsion", &p->repulsion);
    read_f32(env, obj, "simulationLinkSpring", &p->link_spring);
    read_f32(env, obj, "simulationLinkDistance", &p->link_distance);
    read_pair(env, obj, "simulationLinkDistRandomVariationRange", p->link_dist_random_variation_range);
    read_f32(env, obj, "simulationRepulsionFromMouse", &p->repulsion_from_mouse);
    read_f32(env, obj, "simulationFriction", &p->friction);
    read_f32(env, obj, "simulationCollision", &p->collision);
    read_f32(env, obj, "simulationCollisionRadius", &p->collision_radius);
    read_f32(env, obj, "simulationCollisionPadding", &p->collision_padding);
    read_f32(env, obj, "pointDefaultSize", &p->point_default_size);
    read_bool(env, obj, "enableSimulation", &p->enable_simulation);
    read_bool(env, obj, "runForces", &p->run_forces);
    read_bool(env, obj, "mouseRepulsion", &p->mouse_repulsion);
    read_pair(env, obj, "mousePosition", p->mouse_position);
    read_f32(env, obj, "simulationCluster", &p->cluster);
}

static napi_value js_step(napi_env env, napi_callback_info info) {
    size_t argc = 4;
    napi_value argv[4];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    cg_sim_params p;
    cg_default_params(&p);
    if (argc > 1) {
        napi_valuetype vt;
        napi_typeof(env, argv[1], &vt);
        if (vt == napi_object) read_params(env, argv[1], &p);
    }
    uint32_t nticks = 1;
    if (argc > 2) napi_get_value_uint32(env, argv[2], &nticks);
    void* seeds = NULL;
    size_t ns = 0;
    if (argc > 3 && !get_typed(env, argv[3], napi_float32_array, &seeds, &ns, 1)) {
        napi_throw_type_error(env, NULL, "cosmosgl.step: seeds must be a Float32Array or null");
        return NULL;
    }
    if (seeds && ns < nticks) {
        napi_throw_range_error(env, NULL, "cosmosgl.step: one seed per tick expected");
        return NULL;
    }
    napi_value ab, out;
    void* trace = NULL;
    CHECK(env, napi_create_arraybuffer(env, (size_t)nticks * sizeof(double), &trace, &ab));
    int rc = cg_step(ctx, &p, nticks, (const float*)seeds, (double*)trace);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    CHECK(env, napi_create_typedarray(env, napi_float64_array, nticks, ab, 0, &out));
    return out;
}

/* evalForce(h, params, force, alpha, seed, out: Float32Array(2n)) — one force
 * pass from the current positions, nothing integrated (cg_eval_force). */
static napi_value js_eval_force(napi_env env, napi_callback_info info) {
    size_t argc = 6;
    napi_value argv[6];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    if (argc < 6) {
        napi_throw_type_error(env, NULL, "cosmosgl.evalForce: (h, params, force, alpha, seed, out) expected");
        return NULL;
    }
    cg_sim_params p;
    cg_default_params(&p);
    napi_valuetype vt;
    napi_typeof(env, argv[1], &vt);
    if (vt == napi_object) read_params(env, argv[1], &p);
    int32_t force = 0;
    double alpha = 1.0, seed = 0.0;
    CHECK(env, napi_get_value_int32(env, argv[2], &force));
    CHECK(env, napi_get_value_double(env, argv[3], &alpha));
    CHECK(env, napi_get_value_double(env, argv[4], &seed));
    void* out;
    size_t len;
    if (!get_typed(env, argv[5], napi_float32_array, &out, &len, 0)) {
        napi_throw_type_error(env, NULL, "cosmosgl.evalForce: out must be a Float32Array");
        return NULL;
    }
    /* cg_eval_force writes one pair per point: the length is checked first by
     * cg_get_positions, which refuses a mismatched count before writing */
    int rc = cg_get_positions(ctx, (float*)out, (uint32_t)(len / 2));
    if (rc == CG_OK && len % 2) rc = CG_ERR_INVALID;
    if (rc == CG_OK) rc = cg_eval_force(ctx, &p, force, (float)alpha, (float)seed, (float*)out);
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return argv[5];
}

static napi_value js_get_positions(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
    cg_ctx* ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
    if (!ctx) return NULL;
    void* p;
    size_t len;
    if (argc < 2 || !get_typed(env, argv[1], napi_float32_array, &p, &len, 0)) {
        napi_throw_type_error(env, NULL, "cosmosgl.getPositions: expected a Float32Array");
        return NULL;
    }
    int rc = cg_get_positions(ctx, (float*)p, (uint32_t)(len / 2));
    if (rc != CG_OK) return throw_rc(env, ctx, rc);
    return argv[1];
}

static napi_value js_default_params(napi_env env, napi_callback_info info) {
    (void)info;
    cg_sim_params p;
    cg_default_params(&p);
    napi_value o, v;
    CHECK(env, napi_create_object(env, &o));
#define PUT(key, val)                                       \
    CHECK(env, napi_create_double(env, (double)(val), &v)); \
    CHECK(env, napi_set_named_property(env, o, key, v));
    PUT("spaceSize", p.space_size);
    PUT("simulationDecay", p.decay);
    PUT("simulationGravity", p.gravity);
    PUT("simulationCenter", p.center);
    PUT("simulationRepulsion", p.repulsion);
    PUT("simulationLinkSpring", p.link_spring);
    PUT("simulationLinkDistance", p.link_distance);
    PUT("simulationRepulsionFromMouse", p.repulsion_from_mouse);
    PUT("simulationFriction", p.friction);
    PUT("simulationCollision", p.collision);
    PUT("simulationCollisionPadding", p.collision_padding);
    PUT("pointDefaultSize", p.point_default_size);
    PUT("simulationCluster", p.cluster);
#undef PUT
    return o;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor d[] = {
        {"create", 0, js_create, 0, 0, 0, napi_enumerable, 0},
        {"destroy", 0, js_destroy, 0, 0, 0, napi_enumerable, 0},
        {"setPositions", 0, js_set_positions, 0, 0, 0, napi_enumerable, 0},
        {"setLinks", 0, js_set_links, 0, 0, 0, napi_enumerable, 0},
        {"setJitter", 0, js_set_jitter, 0, 0, 0, napi_enumerable, 0},
        {"setPointSizes", 0, js_set_point_sizes, 0, 0, 0, napi_enumerable, 0},
        {"setPinned", 0, js_set_pinned, 0, 0, 0, napi_enumerable, 0},
        {"setPointClusters", 0, js_set_point_clusters, 0, 0, 0, napi_enumerable, 0},
        {"setClusterPositions", 0, js_set_cluster_positions, 0, 0, 0, napi_enumerable, 0},
        {"setPointClusterStrength", 0, js_set_point_cluster_strength, 0, 0, 0, napi_enumerable, 0},
        {"setAlpha", 0, js_set_alpha, 0, 0, 0, napi_enumerable, 0},
        {"getAlpha", 0, js_get_alpha, 0, 0, 0, napi_enumerable, 0},
        {"step", 0, js_step, 0, 0, 0, napi_enumerable, 0},
        {"getPositions", 0, js_get_positions, 0, 0, 0, napi_enumerable, 0},
        {"evalForce", 0, js_eval_force, 0, 0, 0, napi_enumerable, 0},
        {"defaultParams", 0, js_default_params, 0, 0, 0, napi_enumerable, 0},
    };
    if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
    return exports;
}

NAPI_MODULE(cosmosgl, init)
