// k_pointwise.hip — per-point forces (gravity, center, mouse, clusters) fused
// with the integrate pass, and the centroid reductions of the center and
// cluster forces.
//
// Reference: ForceGravity/force-gravity.frag:24-40, ForceCenter/
// calculate-centermass.vert:11-30 + force-center.frag:22-38, ForceMouse/
// force-mouse.frag:21-36, Clusters/calculate-centermass.vert:21-53 +
// force-cluster.frag:28-53, Points/update-position.frag:26-59.
#include <algorithm>

#include "cg_internal.h"

namespace cg {
namespace {

constexpr uint32_t CENTER_CHUNK = 4096;  // present points per centroid partial

// Velocity of one force pass at position p.
template <int PASS>
__device__ inline float2 pw_force(float2 p, const GDesc& d, const double* __restrict__ csum,
                                  uint32_t g) {
    float2 v = make_float2(0.f, 0.f);
    if (PASS == P_GRAVITY) {
        v = gravity_velocity(p, d);
    } else if (PASS == P_CENTER) {
        // force-center.frag:27-35, centermass (sum x, sum y, count) summed in f64
        float cmx = (float)csum[3 * g] / (float)csum[3 * g + 2];
        float cmy = (float)csum[3 * g + 1] / (float)csum[3 * g + 2];
        float dx = cmx - p.x, dy = cmy - p.y;
        float dist = sqrtf(dx * dx + dy * dy);
        if (dist > 0.f) {
            float add = d.alpha * d.center * dist * 0.01f;
            v.x = add * (dx / dist);
            v.y = add * (dy / dist);
        }
    } else if (PASS == P_MOUSE) {
        // force-mouse.frag:26-33 (no alpha); atan2(0,0) taken as 0.
        float dx = d.mx - p.x, dy = d.my - p.y;
        float dist = sqrtf(dx * dx + dy * dy);
        float r = fmaxf(dist, 10.f);
        float add = 100.f * d.mouse_k / (r * r);
        float ux = 1.f, uy = 0.f;
        if (dist > 0.f) { ux = dx / dist; uy = dy / dist; }
        v.x = -add * ux;
        v.y = -add * uy;
    }
    return v;
}

// PW_ITEMS present points per lane (strided by the block, so each load is
// coalesced), each through force -> integrate; the lane's loads go out
// together, one memory latency per dependent step for all its points (4 for
// sets of at least PW_MANY points: the serving mix's gravity pass 101 -> 86
// us; 1 below, where fewer waves would leave the chip idler: 1M points 14 ->
// 15 us at 4).  vout: evaluation mode, velocity only.
constexpr uint32_t PW_MANY = 1u << 22;
template <int PASS, int PW_ITEMS>
__global__ void __launch_bounds__(256) k_pointwise(float2* __restrict__ pos0,
                                                   float2* __restrict__ pos1,
                                                   float2* __restrict__ vout,
                                                   const uint32_t* __restrict__ present,
                                                   const uint32_t* __restrict__ gid,
                                                   const uint8_t* __restrict__ flags,
                                                   const GDesc* __restrict__ desc, uint32_t np,
                                                   const double* __restrict__ csum, uint32_t one_graph) {
    const uint32_t k0 = blockIdx.x * blockDim.x * PW_ITEMS + threadIdx.x;
    uint32_t ii[PW_ITEMS], gg[PW_ITEMS];
#pragma unroll
    for (int u = 0; u < PW_ITEMS; ++u) ii[u] = present[min(k0 + u * blockDim.x, np - 1)];
#pragma unroll
    for (int u = 0; u < PW_ITEMS; ++u) gg[u] = one_graph ? 0u : gid[ii[u]];  // a set of one graph skips the slot load
    float2 pp[PW_ITEMS];
    uint8_t fl[PW_ITEMS];
#pragma unroll
    for (int u = 0; u < PW_ITEMS; ++u) {
        const GDesc& d = desc[gg[u]];
        pp[u] = ((d.inbuf >> PASS) & 1u) ? pos1[ii[u]] : pos0[ii[u]];
        fl[u] = vout ? 0 : flags[ii[u]];
    }
#pragma unroll
    for (int u = 0; u < PW_ITEMS; ++u) {
        if (k0 + u * blockDim.x >= np) break;
        const GDesc& d = desc[gg[u]];
        if (!(d.passes & (1u << PASS))) continue;
        const float2 v = pw_force<PASS>(pp[u], d, csum, gg[u]);
        if (vout) {
            vout[ii[u]] = v;
            continue;
        }
        float2* __restrict__ pout = ((d.inbuf >> PASS) & 1u) ? pos0 : pos1;
        pout[ii[u]] = (fl[u] & F_PINNED) ? pp[u] : integrate(pp[u], v, d.friction, d.S);
    }
}

// Centroid of each graph: one block per chunk of CENTER_CHUNK present points of
// one graph (fixed f64 tree), then one lane per graph adds its chunks in order.
// The shape depends only on the graph itself: the same sum alone or packed.
__global__ void __launch_bounds__(256) k_center_partial(const float2* __restrict__ pos0,
                                                        const float2* __restrict__ pos1,
                                                        const uint32_t* __restrict__ present,
                                                        const uint4* __restrict__ chunks,
                                                        const GDesc* __restrict__ desc,
                                                        double* __restrict__ part) {
    __shared__ double sx[256], sy[256];
    const uint4 ch = chunks[blockIdx.x];  // {k begin, k end, graph slot, -}
    const GDesc& d = desc[ch.z];
    const float2* __restrict__ pin = ((d.inbuf >> P_CENTER) & 1u) ? pos1 : pos0;
    double ax = 0.0, ay = 0.0;
    for (uint32_t k = ch.x + threadIdx.x; k < ch.y; k += 256) {
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

__global__ void k_center_final(const double* __restrict__ part, const uint32_t* __restrict__ first,
                               const GDesc* __restrict__ desc, uint32_t G,
                               double* __restrict__ out) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    double x = 0.0, y = 0.0;
    for (uint32_t b = first[g]; b < first[g + 1]; ++b) {
        x += part[2 * b];
        y += part[2 * b + 1];
    }
    out[3 * g] = x;
    out[3 * g + 1] = y;
    out[3 * g + 2] = (double)desc[g].np;
}

// Cluster centroids — Clusters/calculate-centermass.vert:21-53: sums of
// (x, y, 1) over the present points of each cluster.  One block per chunk of
// at most CENTER_CHUNK members of one cluster (fixed f64 tree), then one lane
// per cluster adds its chunks in order.
__global__ void __launch_bounds__(256) k_cluster_partial(const float2* __restrict__ pos0,
                                                         const float2* __restrict__ pos1,
                                                         const uint32_t* __restrict__ members,
                                                         const uint4* __restrict__ chunks,
                                                         const GDesc* __restrict__ desc,
                                                         double* __restrict__ part) {
    __shared__ double sx[256], sy[256];
    const uint4 ch = chunks[blockIdx.x];  // {member begin, member end, cluster, graph slot}
    const GDesc& d = desc[ch.w];
    const float2* __restrict__ pin = ((d.inbuf >> P_CLUSTER) & 1u) ? pos1 : pos0;
    double ax = 0.0, ay = 0.0;
    for (uint32_t k = ch.x + threadIdx.x; k < ch.y; k += 256) {
        float2 p = pin[members[k]];
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

__global__ void k_cluster_final(const double* __restrict__ part, const uint4* __restrict__ chunks,
                                const uint32_t* __restrict__ first, uint32_t ncl,
                                double* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncl) return;
    double x = 0.0, y = 0.0, n = 0.0;
    for (uint32_t b = first[c]; b < first[c + 1]; ++b) {
        x += part[2 * b];
        y += part[2 * b + 1];
        n += (double)(chunks[b].y - chunks[b].x);
    }
    out[3 * c] = x;
    out[3 * c + 1] = y;
    out[3 * c + 2] = n;
}

// force-cluster.frag:28-53: toward the cluster's custom position, or its
// centroid when a coordinate of the custom one is negative; no pull without a
// cluster or at zero distance.  v = alpha * dist * simulationCluster * strength
// along the unit vector (an empty cluster's centroid is 0/0, so no pull).
__global__ void __launch_bounds__(256) k_cluster_force(float2* __restrict__ pos0,
                                                       float2* __restrict__ pos1,
                                                       float2* __restrict__ vout,
                                                       const uint32_t* __restrict__ present,
                                                       const uint32_t* __restrict__ gid,
                                                       const uint8_t* __restrict__ flags,
                                                       const GDesc* __restrict__ desc, uint32_t np,
                                                       const int32_t* __restrict__ pcl,
                                                       const float* __restrict__ pcs,
                                                       const float2* __restrict__ cpos,
                                                       const double* __restrict__ csum) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const uint32_t i = present[k];
    const GDesc& d = desc[gid[i]];
    if (!(d.passes & (1u << P_CLUSTER))) return;
    const bool b = (d.inbuf >> P_CLUSTER) & 1u;
    const float2 p = (b ? pos1 : pos0)[i];
    float2 v = make_float2(0.f, 0.f);
    const int32_t cl = pcl[i];
    if (cl >= 0) {
        float2 t = cpos[cl];
        if (t.x < 0.f || t.y < 0.f) {
            const float n = (float)csum[3 * cl + 2];
            t = make_float2((float)csum[3 * cl] / n, (float)csum[3 * cl + 1] / n);
        }
        const float dx = t.x - p.x, dy = t.y - p.y;
        const float dist = sqrtf(dx * dx + dy * dy);
        if (dist > 0.f) {
            const float add = d.alpha * dist * d.cluster_k * pcs[i];
            v.x = add * (dx / dist);
            v.y = add * (dy / dist);
        }
    }
    if (vout) {
        vout[i] = v;
        return;
    }
    (b ? pos0 : pos1)[i] = (flags[i] & F_PINNED) ? p : integrate(p, v, d.friction, d.S);
}

}  // namespace

int cluster_run(Ctx* c, float2* vout) {
    if (c->NP == 0 || c->CC == 0) return CG_OK;
    if (c->n_cl_chunks)
        hipLaunchKernelGGL(k_cluster_partial, dim3(c->n_cl_chunks), dim3(256), 0, c->stream, c->pos[0],
                           c->pos[1], c->cl_members, c->cl_chunks, c->desc, c->cl_part);
    hipLaunchKernelGGL(k_cluster_final, dim3(blocks_for(c->CC, 64)), dim3(64), 0, c->stream, c->cl_part,
                       c->cl_chunks, c->cl_first, c->CC, c->cl_sum);
    hipLaunchKernelGGL(k_cluster_force, dim3(blocks_for(c->NP, 256)), dim3(256), 0, c->stream, c->pos[0],
                       c->pos[1], vout, c->present, c->gid, c->flags, c->desc, c->NP, c->pcl, c->pcs,
                       c->cpos, c->cl_sum);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

int center_run(Ctx* c) {
    if (c->NP == 0 || c->n_center_chunks == 0) return CG_OK;
    hipLaunchKernelGGL(k_center_partial, dim3(c->n_center_chunks), dim3(256), 0, c->stream, c->pos[0],
                       c->pos[1], c->present, c->center_chunks, c->desc, c->center_part);
    hipLaunchKernelGGL(k_center_final, dim3(blocks_for(c->G, 64)), dim3(64), 0, c->stream,
                       c->center_part, c->center_first, c->desc, c->G, c->center_sum);
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

int pw_run(Ctx* c, int pass, float2* vout) {
    if (c->NP == 0) return CG_OK;
    if (pass == P_CENTER) {
        int rc = center_run(c);
        if (rc) return rc;
    }
#define CG_PW(P)                                                                                         \
    if (c->NP >= PW_MANY)                                                                                \
        hipLaunchKernelGGL((k_pointwise<P, 4>), dim3(blocks_for(c->NP, 1024)), dim3(256), 0, c->stream,   \
                           c->pos[0], c->pos[1], vout, c->present, c->gid, c->flags, c->desc, c->NP,      \
                           c->center_sum, c->G == 1 ? 1u : 0u);                                           \
    else                                                                                                 \
        hipLaunchKernelGGL((k_pointwise<P, 1>), dim3(blocks_for(c->NP, 256)), dim3(256), 0, c->stream,    \
                           c->pos[0], c->pos[1], vout, c->present, c->gid, c->flags, c->desc, c->NP,      \
                           c->center_sum, c->G == 1 ? 1u : 0u)
    switch (pass) {
        case P_GRAVITY: CG_PW(P_GRAVITY); break;
        case P_CENTER: CG_PW(P_CENTER); break;
        case P_MOUSE: CG_PW(P_MOUSE); break;
        default: return fail(c, CG_ERR_INVALID, "pointwise: unknown pass");
    }
#undef CG_PW
    CG_HIP(c, hipGetLastError());
    return CG_OK;
}

// Chunk list of the centroid reduction: CENTER_CHUNK present points per chunk, per graph.
void center_chunks_host(const std::vector<uint32_t>& k0, const std::vector<uint32_t>& np,
                        std::vector<uint4>& chunks, std::vector<uint32_t>& first) {
    chunks.clear();
    first.assign(k0.size() + 1, 0);
    for (size_t g = 0; g < k0.size(); ++g) {
        first[g] = (uint32_t)chunks.size();
        for (uint32_t s = 0; s < np[g]; s += CENTER_CHUNK)
            chunks.push_back(make_uint4(k0[g] + s, k0[g] + std::min(np[g], s + CENTER_CHUNK), (uint32_t)g, 0));
    }
    first[k0.size()] = (uint32_t)chunks.size();
}

}  // namespace cg
