// k_pointwise.hip — per-point forces (gravity, center, mouse) fused with the
// integrate pass, the global centroid reduction, and the finest-cell key pass
// that feeds the many-body sort.
//
// Reference: ForceGravity/force-gravity.frag:24-40, ForceCenter/
// calculate-centermass.vert:11-30 + force-center.frag:22-38, ForceMouse/
// force-mouse.frag:21-36, Points/update-position.frag:26-59.
#include "cg_internal.h"

namespace cg {
namespace {

struct PwArgs {
    float alpha, gravity, center, mouse_k, friction, S;
    float mx, my;
    float hf;        // finest many-body cell size
    int gf;          // finest grid size
};

// Velocity of one force pass at position p.
template <int FORCE>
__device__ inline float2 pw_force(float2 p, const PwArgs& a, float cmx, float cmy) {
    float2 v = make_float2(0.f, 0.f);
    if (FORCE == CG_FORCE_GRAVITY) {
        // force-gravity.frag:31-38: alpha*g*|d|*0.1 along d; the atan/cos/sin
        // round trip is the unit vector d/|d|.
        float dx = a.S * 0.5f - p.x, dy = a.S * 0.5f - p.y;
        float dist = sqrtf(dx * dx + dy * dy);
        if (dist > 0.f) {
            float add = a.alpha * a.gravity * dist * 0.1f;
            v.x = add * (dx / dist);
            v.y = add * (dy / dist);
        }
    } else if (FORCE == CG_FORCE_CENTER) {
        // force-center.frag:27-35
        float dx = cmx - p.x, dy = cmy - p.y;
        float dist = sqrtf(dx * dx + dy * dy);
        if (dist > 0.f) {
            float add = a.alpha * a.center * dist * 0.01f;
            v.x = add * (dx / dist);
            v.y = add * (dy / dist);
        }
    } else if (FORCE == CG_FORCE_MOUSE) {
        // force-mouse.frag:26-33 (no alpha); atan2(0,0) taken as 0.
        float dx = a.mx - p.x, dy = a.my - p.y;
        float dist = sqrtf(dx * dx + dy * dy);
        float r = fmaxf(dist, 10.f);
        float add = 100.f * a.mouse_k / (r * r);
        float ux = 1.f, uy = 0.f;
        if (dist > 0.f) { ux = dx / dist; uy = dy / dist; }
        v.x = -add * ux;
        v.y = -add * uy;
    }
    return v;
}

// One thread per present point: force -> integrate -> (optional) finest cell key.
// pout == nullptr: evaluation only (velocity to vout).  FORCE < 0: key pass only.
template <int FORCE>
__global__ void __launch_bounds__(256) k_pointwise(const float2* __restrict__ pin,
                                                   float2* __restrict__ pout,
                                                   float2* __restrict__ vout,
                                                   const uint32_t* __restrict__ present,
                                                   const uint8_t* __restrict__ flags, uint32_t np,
                                                   PwArgs a, const double* __restrict__ csum,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    uint32_t i = present[k];
    float2 p = pin[i];
    float2 q = p;
    if (FORCE >= 0) {
        float cmx = 0.f, cmy = 0.f;
        if (FORCE == CG_FORCE_CENTER) {
            // centermass texel: (sum x, sum y, count) summed in f64, divided like the shader
            cmx = (float)csum[0] / (float)csum[2];
            cmy = (float)csum[1] / (float)csum[2];
        }
        float2 v = pw_force<FORCE>(p, a, cmx, cmy);
        if (vout) vout[i] = v;
        if (pout) {
            if (!(flags[i] & F_PINNED)) q = integrate(p, v, a.friction, a.S);
            pout[i] = q;
        }
    }
    if (keys) {
        uint32_t cx = (uint32_t)cell_of(q.x, a.hf, a.gf);
        uint32_t cy = (uint32_t)cell_of(q.y, a.hf, a.gf);
        keys[k] = morton2(cx, cy);
        vals[k] = i;
    }
}

// Centroid reduction: fixed-shape block partials in f64, then one block sums
// the partials in index order — the same sum every run (no float atomics).
__global__ void __launch_bounds__(256) k_center_partial(const float2* __restrict__ pin,
                                                        const uint32_t* __restrict__ present,
                                                        uint32_t np, double* __restrict__ part) {
    __shared__ double sx[256], sy[256];
    double ax = 0.0, ay = 0.0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < np; k += gridDim.x * 256) {
        float2 p = pin[present[k]];
        ax += (double)p.x;
        ay += (double)p.y;
    }
    sx[threadIdx.x] = ax;
    sy[threadIdx.x] = ay;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sx[threadIdx.x] += sx[threadIdx.x + s];
            sy[threadIdx.x] += sy[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sx[0];
        part[2 * blockIdx.x + 1] = sy[0];
    }
}

__global__ void k_center_final(const double* __restrict__ part, int nb, uint32_t np,
                               double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double x = 0.0, y = 0.0;
    for (int b = 0; b < nb; ++b) { x += part[2 * b]; y += part[2 * b + 1]; }
    out[0] = x;
    out[1] = y;
    out[2] = (double)np;
}

constexpr int CENTER_BLOCKS = 512;

}  // namespace

int pw_run(Ctx* c, const cg_sim_params* p, int force, float alpha, const float2* pin,
           float2* pout, float2* vout, bool write_keys) {
    if (c->np == 0) return CG_OK;
    PwArgs a;
    a.alpha = alpha;
    a.gravity = p->gravity;
    a.center = p->center;
    a.mouse_k = p->repulsion_from_mouse;
    a.friction = p->friction;
    a.S = p->space_size;
    a.mx = p->mouse_position[0];
    a.my = p->mouse_position[1];
    a.gf = (int)c->gf;
    a.hf = (float)((double)p->space_size / (double)c->gf);
    uint32_t* keys = write_keys ? c->keys[0] : nullptr;
    uint32_t* vals = write_keys ? c->vals[0] : nullptr;
    unsigned nb = blocks_for(c->np, 256);
    if (force == CG_FORCE_CENTER) {
        if (!c->center_part) {
            int rc = CG_OK;
            c->center_part = (double*)dalloc(c, sizeof(double) * 2 * CENTER_BLOCKS, &rc);
            if (rc) return rc;
            c->center_sum = (double*)dalloc(c, sizeof(double) * 4, &rc);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_center_partial, dim3(CENTER_BLOCKS), dim3(256), 0, c->stream, pin,
                           c->present, c->np, c->center_part);
        hipLaunchKernelGGL(k_center_final, dim3(1), dim3(64), 0, c->stream, c->center_part,
                           CENTER_BLOCKS, c->np, c->center_sum);
    }
    switch (force) {
        case CG_FORCE_GRAVITY:
            hipLaunchKernelGGL(k_pointwise<CG_FORCE_GRAVITY>, dim3(nb), dim3(256), 0, c->stream,
                               pin, pout, vout, c->present, c->flags, c->np, a, c->center_sum,
                               keys, vals);
            break;
        case CG_FORCE_CENTER:
            hipLaunchKernelGGL(k_pointwise<CG_FORCE_CENTER>, dim3(nb), dim3(256), 0, c->stream,
                               pin, pout, vout, c->present, c->flags, c->np, a, c->center_sum,
                               keys, vals);
            break;
        case CG_FORCE_MOUSE:
            hipLaunchKernelGGL(k_pointwise<CG_FORCE_MOUSE>, dim3(nb), dim3(256), 0, c->stream,
                               pin, pout, vout, c->present, c->flags, c->np, a, c->center_sum,
                               keys, vals);
            break;
        default:
            return fail(c, CG_ERR_INVALID, "pointwise: unknown force");
    }
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Key pass used when no per-point force runs right before the many-body pass.
int keys_only(Ctx* c, const float2* pin, float space_size) {
    if (c->np == 0) return CG_OK;
    PwArgs a{};
    a.gf = (int)c->gf;
    a.hf = (float)((double)space_size / (double)c->gf);
    hipLaunchKernelGGL(k_pointwise<-1>, dim3(blocks_for(c->np, 256)), dim3(256), 0, c->stream,
                       pin, (float2*)nullptr, (float2*)nullptr, c->present, c->flags, c->np, a,
                       (const double*)nullptr, c->keys[0], c->vals[0]);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

}  // namespace cg
