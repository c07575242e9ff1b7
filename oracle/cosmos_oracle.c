/*
 * cosmos_oracle.c — CPU restatement of cosmos.gl's force-simulation tick.
 *
 * TEST INFRASTRUCTURE ONLY.  The GPU results are compared with this file's:
 * only tests/, the smoke check and bench.py's cpu_baseline leg may load it.  The product path
 * (graph_amd/) never calls it and has no CPU fallback.
 *
 * Two kinds of functions live here:
 *   - integer references (bit-exact targets for the GPU): lowbias32 hash,
 *     near-field tickets, cell keys, pyramid geometry, top-8 slot peel,
 *     link CSR / degree / bias / strength;
 *   - FP64 evaluations of every force of the tick on the SAME discretisation
 *     (same cells, same slots), plus integration and alpha decay.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the cosmos.gl 3.4.1 source tree).  Build: gcc -O2 -ffp-contract=off
 * (no FMA contraction, no fast-math) so the f32 ticket arithmetic matches the
 * shader's IEEE single-precision steps.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the loops below.  Every parallel loop writes each output element
 * from exactly one iteration and keeps each sum inside one iteration, so the
 * results do not depend on the thread count. */
static int g_default_threads = 0;  /* the OpenMP default (OMP_NUM_THREADS), kept for n = 0 */
void or_set_threads(int n) {
    if (!g_default_threads) g_default_threads = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : g_default_threads);
}
int or_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* integer references                                                        */
/* ------------------------------------------------------------------------ */

/* lowbias32 — ForceManyBody/build-nearfield-slots.vert:69-74 */
uint32_t or_lowbias32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

static uint32_t f32_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ticket (hashValue) — build-nearfield-slots.vert:69-79.
 * h = lowbias32(uint(index) ^ floatBitsToUint(seed)); v = (float(h>>8)+0.5)/2^24;
 * v = 0.001 + v*0.998, every step an IEEE f32 operation. */
float or_ticket(uint32_t index, float seed) {
    uint32_t h = or_lowbias32(index ^ f32_bits(seed));
    volatile float t = (float)(h >> 8u);
    volatile float a = t + 0.5f;
    volatile float b = a / 16777216.0f;
    volatile float c = b * 0.998f;
    volatile float d = 0.001f + c;
    return d;
}

/* finest grid size and level count — ForceManyBody/index.ts:646-656.
 * g_f = min(512, max(8, 2^ceil(log2(2*sqrt(n))))), L = log2(g_f) - 1.
 * Integer form: the smallest power of two p with (p/2)^2 >= n. */
void or_geometry(uint32_t n, uint32_t* finest, uint32_t* levels) {
    uint32_t p = 1;
    while ((uint64_t)(p / 2) * (uint64_t)(p / 2) < (uint64_t)n && p < (1u << 30)) p <<= 1;
    if (n == 0) p = 0;          /* log2(0) = -inf -> 2^-inf = 0 */
    if (p < 8) p = 8;
    if (p > 512) p = 512;
    uint32_t lg = 0;
    while ((1u << lg) < p) ++lg;
    *finest = p;
    *levels = lg - 1;
}

/* cell index along one axis — calculate-level.vert:47-48, force-level.frag:66,
 * build-nearfield-slots.vert:82-83: clamp(int(floor(p / cellSize)), 0, g-1),
 * with p / cellSize an f32 division. */
int32_t or_cell(float p, float cell_size, int32_t g) {
    volatile float q = p / cell_size;
    float f = floorf(q);
    if (!(f >= 0.0f)) return 0;      /* also maps NaN to 0 (never binned anyway) */
    if (f >= (float)(g - 1)) return g - 1;
    return (int32_t)f;
}

/* Both axes' cells of n points at one grid — calculate-level.vert:47-48 (and
 * build-grid.vert:50-55 with an offset of ox, oy cells added to the position in
 * f32 first): cells[2i] = cx, cells[2i+1] = cy; absent points get -1. */
void or_cells(uint32_t n, const float* pos, const uint8_t* absent, float cell_size, int32_t g, float ox,
              float oy, int32_t* cells) {
    volatile float offx = ox * cell_size, offy = oy * cell_size;
    for (uint32_t i = 0; i < n; ++i) {
        if (absent && absent[i]) { cells[2 * i] = -1; cells[2 * i + 1] = -1; continue; }
        volatile float qx = pos[2 * i] + offx, qy = pos[2 * i + 1] + offy;
        cells[2 * i] = or_cell(qx, cell_size, g);
        cells[2 * i + 1] = or_cell(qy, cell_size, g);
    }
}

/* f32 cell size the reference uploads: adjustedSpaceSize / gridSize computed in
 * double and rounded to the f32 uniform (ForceManyBody/index.ts:483, 531, 583). */
float or_cell_size(float space_size, uint32_t g) {
    return (float)((double)space_size / (double)g);
}

/* Near-field slots for the finest grid — ForceManyBody/index.ts:511-559 and
 * build-nearfield-slots.vert:46-103.  Pass k keeps, per cell, the eligible
 * point with the smallest ticket, where eligible means ticket > ticket of
 * slot k-1 (and slot k-1 non-empty); ties in ticket go to the first drawn
 * point (lowest index).  So: order the cell's present points by (ticket, index),
 * keep only the first point of every run of equal tickets, take the first 8.
 * slots[cell*8 + k] = point index or -1; cells are row-major (cy*g + cx). */
typedef struct { uint32_t tb; uint32_t idx; } tk_t;
static int tk_cmp(const void* a, const void* b) {
    const tk_t* x = (const tk_t*)a; const tk_t* y = (const tk_t*)b;
    if (x->tb != y->tb) return x->tb < y->tb ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

void or_slots(uint32_t n, const float* pos, const uint8_t* absent, float space_size,
              float seed, int32_t* slots /* g_f*g_f*8 */) {
    uint32_t gf, L; or_geometry(n, &gf, &L); (void)L;
    float h = or_cell_size(space_size, gf);
    uint32_t cells = gf * gf;
    uint32_t* cnt = (uint32_t*)calloc(cells + 1, sizeof(uint32_t));
    uint32_t* cell_of = (uint32_t*)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        if (absent && absent[i]) { cell_of[i] = UINT32_MAX; continue; }
        int32_t cx = or_cell(pos[2 * i], h, (int32_t)gf);
        int32_t cy = or_cell(pos[2 * i + 1], h, (int32_t)gf);
        cell_of[i] = (uint32_t)cy * gf + (uint32_t)cx;
        cnt[cell_of[i] + 1]++;
    }
    for (uint32_t c = 0; c < cells; ++c) cnt[c + 1] += cnt[c];
    tk_t* buf = (tk_t*)malloc((size_t)(n ? n : 1) * sizeof(tk_t));
    uint32_t* fill = (uint32_t*)malloc((size_t)cells * sizeof(uint32_t));
    memcpy(fill, cnt, cells * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        if (cell_of[i] == UINT32_MAX) continue;
        tk_t t; t.tb = f32_bits(or_ticket(i, seed)); t.idx = i;
        buf[fill[cell_of[i]]++] = t;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t c = 0; c < cells; ++c) {
        uint32_t b = cnt[c], e = cnt[c + 1];
        qsort(buf + b, e - b, sizeof(tk_t), tk_cmp);
        int k = 0;
        uint32_t last_tb = 0; int have = 0;
        for (uint32_t j = b; j < e && k < 8; ++j) {
            if (have && buf[j].tb == last_tb) continue;  /* dropped by the peel */
            slots[(size_t)c * 8 + k++] = (int32_t)buf[j].idx;
            last_tb = buf[j].tb; have = 1;
        }
        for (; k < 8; ++k) slots[(size_t)c * 8 + k] = -1;
    }
    free(cnt); free(cell_of); free(buf); free(fill);
}

/* Link adjacency — GraphData/index.ts:286-289 (isPointIndex), :525-549
 * (_createAdjacencyLists), :551-568 (_calculateDegrees); ForceLink/index.ts:40-83.
 * links = [s0,t0,s1,t1,...] as f32 (the reference's Float32Array).
 * dir 0 = the INCOMING instance, which walks sourceIndexToTargetIndices
 *         (for point p as a source: its targets);
 * dir 1 = the OUTGOING instance, which walks targetIndexToSourceIndices.
 * Entries are point-ascending, then link-input order.
 * Outputs: offsets[n+1], nbr[E], link_of[E]; degree[n] (in+out over valid links).
 * Returns E. */
static int or_valid_index(float v, uint32_t n) {
    if (!(v >= 0.0f)) return 0;
    if (floorf(v) != v) return 0;
    return (double)v < (double)n;
}

uint32_t or_link_csr(uint32_t n, uint32_t m, const float* links, int dir,
                     uint32_t* offsets, uint32_t* nbr, uint32_t* link_of, uint32_t* degree) {
    memset(offsets, 0, (size_t)(n + 1) * sizeof(uint32_t));
    if (degree) memset(degree, 0, (size_t)n * sizeof(uint32_t));
    for (uint32_t l = 0; l < m; ++l) {
        float s = links[2 * l], t = links[2 * l + 1];
        if (!or_valid_index(s, n) || !or_valid_index(t, n)) continue;
        uint32_t si = (uint32_t)s, ti = (uint32_t)t;
        uint32_t owner = dir == 0 ? si : ti;
        offsets[owner + 1]++;
        if (degree) { degree[si]++; degree[ti]++; }
    }
    for (uint32_t p = 0; p < n; ++p) offsets[p + 1] += offsets[p];
    uint32_t E = offsets[n];
    uint32_t* fill = (uint32_t*)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    memcpy(fill, offsets, (size_t)n * sizeof(uint32_t));
    for (uint32_t l = 0; l < m; ++l) {
        float s = links[2 * l], t = links[2 * l + 1];
        if (!or_valid_index(s, n) || !or_valid_index(t, n)) continue;
        uint32_t si = (uint32_t)s, ti = (uint32_t)t;
        uint32_t owner = dir == 0 ? si : ti, other = dir == 0 ? ti : si;
        uint32_t e = fill[owner]++;
        nbr[e] = other;
        if (link_of) link_of[e] = l;
    }
    free(fill);
    return E;
}

/* Per-entry bias and strength as stored in the f32 textures —
 * ForceLink/index.ts:59-74.  degree = deg(connected q), connectedDegree = deg(p);
 * bias = deg q / (deg q + deg p) or 0.5; strength = sqrt(user strength if finite
 * and >= 0, else 1/max(min(deg q, deg p), 1)); both computed in double then
 * rounded to f32.  strength_in may be NULL (no user strengths). */
void or_link_props(uint32_t n, uint32_t E, const uint32_t* offsets, const uint32_t* nbr,
                   const uint32_t* link_of, const uint32_t* degree, const float* strength_in,
                   float* bias_out, float* strength_out) {
    for (uint32_t p = 0; p < n; ++p) {
        for (uint32_t e = offsets[p]; e < offsets[p + 1]; ++e) {
            uint32_t q = nbr[e];
            double dq = degree[q], dp = degree[p];
            double sum = dq + dp;
            double bias = sum != 0.0 ? dq / sum : 0.5;
            double mind = dq < dp ? dq : dp;
            double s;
            int have = 0;
            if (strength_in) {
                double u = strength_in[link_of[e]];
                if (isfinite(u) && u >= 0.0) { s = u; have = 1; }
            }
            if (!have) s = 1.0 / (mind > 1.0 ? mind : 1.0);
            bias_out[e] = (float)bias;
            strength_out[e] = (float)sqrt(s);
        }
    }
    (void)E;
}

/* ------------------------------------------------------------------------ */
/* FP64 force evaluations (velocity written by one force pass)               */
/* ------------------------------------------------------------------------ */

/* gravity — ForceGravity/force-gravity.frag:24-40: v = alpha*g*0.1*|d|*(cos,sin)(atan2 d),
 * d = (S/2,S/2) - p, zero when |d| = 0.  Closed form alpha*g*0.1*d. */
void or_gravity(uint32_t n, const double* pos, double S, double alpha, double g, double* v) {
    for (uint32_t i = 0; i < n; ++i) {
        double dx = S * 0.5 - pos[2 * i], dy = S * 0.5 - pos[2 * i + 1];
        double k = alpha * g * 0.1;
        v[2 * i] = k * dx; v[2 * i + 1] = k * dy;
    }
}

/* center — ForceCenter/calculate-centermass.vert:11-30, force-center.frag:22-38:
 * M = sum over present points of (x, y, 1); v = alpha*c*0.01*(M.xy/M.z - p). */
void or_center(uint32_t n, const double* pos, const uint8_t* absent, double alpha, double c, double* v) {
    double sx = 0, sy = 0, sn = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (absent && absent[i]) continue;
        sx += pos[2 * i]; sy += pos[2 * i + 1]; sn += 1.0;
    }
    double cx = sx / sn, cy = sy / sn;
    for (uint32_t i = 0; i < n; ++i) {
        double dx = cx - pos[2 * i], dy = cy - pos[2 * i + 1];
        double k = alpha * c * 0.01;
        v[2 * i] = k * dx; v[2 * i + 1] = k * dy;
    }
}

/* mouse — ForceMouse/force-mouse.frag:21-36 (no alpha):
 * d = m - p; r = max(|d|, 10); v = -(100*k/r^2)*(cos,sin)(atan2 d); atan2(0,0) = 0. */
void or_mouse(uint32_t n, const double* pos, double mx, double my, double k, double* v) {
    for (uint32_t i = 0; i < n; ++i) {
        double dx = mx - pos[2 * i], dy = my - pos[2 * i + 1];
        double d = sqrt(dx * dx + dy * dy);
        double r = d > 10.0 ? d : 10.0;
        double a = 100.0 * k / (r * r);
        double ux = 1.0, uy = 0.0;
        if (d > 0.0) { ux = dx / d; uy = dy / d; }
        v[2 * i] = -a * ux; v[2 * i + 1] = -a * uy;
    }
}

/* clusters — Clusters/calculate-centermass.vert:21-53 (sum of (x, y, 1) over
 * the present points of each cluster; unclustered points, cl < 0, contribute
 * nothing) and force-cluster.frag:28-53: the target is the custom position when
 * both of its coordinates are >= 0, else the centroid; when the distance d to it
 * is > 0, v = alpha * d * coef * strength along the unit vector.  An empty
 * cluster's centroid is 0/0: no pull. */
void or_cluster(uint32_t n, const double* pos, const uint8_t* absent, const int32_t* cl, uint32_t ncl,
                const double* cpos, const float* strength, double alpha, double coef, double* v) {
    double* sum = (double*)calloc(3 * (size_t)(ncl ? ncl : 1), sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
        if ((absent && absent[i]) || cl[i] < 0) continue;
        sum[3 * cl[i]] += pos[2 * i];
        sum[3 * cl[i] + 1] += pos[2 * i + 1];
        sum[3 * cl[i] + 2] += 1.0;
    }
    for (uint32_t i = 0; i < n; ++i) {
        v[2 * i] = 0.0;
        v[2 * i + 1] = 0.0;
        const int32_t c = cl[i];
        if (c < 0) continue;
        double tx, ty;
        if (cpos && cpos[2 * c] >= 0.0 && cpos[2 * c + 1] >= 0.0) {
            tx = cpos[2 * c];
            ty = cpos[2 * c + 1];
        } else {
            tx = sum[3 * c] / sum[3 * c + 2];
            ty = sum[3 * c + 1] / sum[3 * c + 2];
        }
        const double dx = tx - pos[2 * i], dy = ty - pos[2 * i + 1];
        const double d = sqrt(dx * dx + dy * dy);
        if (!(d > 0.0)) continue;
        const double k = alpha * d * coef * (strength ? (double)strength[i] : 1.0);
        v[2 * i] = k * dx / d;
        v[2 * i + 1] = k * dy / d;
    }
    free(sum);
}

/* clamped inverse-distance falloff shared by force-level.frag:43-56 (with the
 * cell count) and force-nearfield.frag:64-80 (count 1). */
static void falloff(double dx, double dy, double w, double* vx, double* vy) {
    double l = dx * dx + dy * dy;
    if (l <= 0.0) return;
    double d = sqrt(l);
    if (l < 1.0) l = sqrt(l);
    double a = w / sqrt(l);
    *vx += a * dx / d; *vy += a * dy / d;
}

/* many-body — ForceManyBody/index.ts:399-417, 464-638 with calculate-level.vert,
 * force-level.frag, build-nearfield-slots.vert, force-nearfield.frag.
 * pos: double xy (values representable in f32; cells use f32 arithmetic);
 * jitter: the per-point random vector r_i (already scaled by 1e-5);
 * seed: the tick's f32 seed.  v: double xy, 0 for absent points.
 * eval: optional list of point indices to evaluate (NULL = all), n_eval entries;
 * v is then written for those points only (indexed by point). */
void or_manybody(uint32_t n, const double* pos, const uint8_t* absent, const float* jitter,
                 float space_size, double alpha, double k, float seed,
                 const uint32_t* eval, uint32_t n_eval, double* v) {
    uint32_t gf, L; or_geometry(n, &gf, &L);
    /* aggregates per level (double sums) */
    double** agg = (double**)malloc(L * sizeof(double*));
    float* hs = (float*)malloc(L * sizeof(float));
    float* pf = (float*)malloc((size_t)(n ? n : 1) * 2 * sizeof(float));
    for (uint32_t i = 0; i < 2 * n; ++i) pf[i] = (float)pos[i];
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t lv = 0; lv < L; ++lv) {
        uint32_t g = 4u << lv;
        hs[lv] = or_cell_size(space_size, g);
        agg[lv] = (double*)calloc((size_t)g * g * 3, sizeof(double));
        for (uint32_t i = 0; i < n; ++i) {
            if (absent && absent[i]) continue;
            int32_t cx = or_cell(pf[2 * i], hs[lv], (int32_t)g);
            int32_t cy = or_cell(pf[2 * i + 1], hs[lv], (int32_t)g);
            double* a = agg[lv] + ((size_t)cy * g + cx) * 3;
            a[0] += pos[2 * i]; a[1] += pos[2 * i + 1]; a[2] += 1.0;
        }
    }
    int32_t* slots = (int32_t*)malloc((size_t)gf * gf * 8 * sizeof(int32_t));
    or_slots(n, pf, absent, space_size, seed, slots);
    double ak = alpha * k;
    uint32_t cnt = eval ? n_eval : n;
#pragma omp parallel for schedule(dynamic, 256)
    for (uint32_t t = 0; t < cnt; ++t) {
        uint32_t i = eval ? eval[t] : t;
        if (absent && absent[i]) { v[2 * i] = 0; v[2 * i + 1] = 0; continue; }
        double px = pos[2 * i], py = pos[2 * i + 1];
        double vx = 0, vy = 0;
        for (uint32_t lv = 0; lv < L; ++lv) {
            int32_t g = (int32_t)(4u << lv);
            int32_t qx = or_cell(pf[2 * i], hs[lv], g), qy = or_cell(pf[2 * i + 1], hs[lv], g);
            int32_t x0, y0, x1, y1;
            if (lv == 0) { x0 = 0; y0 = 0; x1 = g; y1 = g; }
            else { x0 = (qx / 2) * 2 - 2; y0 = (qy / 2) * 2 - 2; x1 = x0 + 6; y1 = y0 + 6; }
            for (int32_t cy = y0; cy < y1; ++cy) {
                for (int32_t cx = x0; cx < x1; ++cx) {
                    if (cx < 0 || cy < 0 || cx >= g || cy >= g) continue;
                    int32_t ddx = abs(cx - qx), ddy = abs(cy - qy);
                    if ((ddx > ddy ? ddx : ddy) <= 1) continue;
                    const double* a = agg[lv] + ((size_t)cy * g + cx) * 3;
                    if (a[2] <= 0.0) continue;
                    falloff(px - a[0] / a[2], py - a[1] / a[2], ak * a[2], &vx, &vy);
                }
            }
        }
        /* near field at the finest level */
        {
            int32_t g = (int32_t)gf;
            uint32_t lv = L - 1;
            int32_t qx = or_cell(pf[2 * i], hs[lv], g), qy = or_cell(pf[2 * i + 1], hs[lv], g);
            double rx = jitter ? jitter[2 * i] : 0.0, ry = jitter ? jitter[2 * i + 1] : 0.0;
            double wx = 0, wy = 0;
            for (int32_t dj = -1; dj <= 1; ++dj) {
                for (int32_t di = -1; di <= 1; ++di) {
                    int32_t cx = qx + di, cy = qy + dj;
                    if (cx < 0 || cy < 0 || cx >= g || cy >= g) continue;
                    const double* a = agg[lv] + ((size_t)cy * g + cx) * 3;
                    double others = a[2] - ((di == 0 && dj == 0) ? 1.0 : 0.0);
                    if (others <= 0.0) continue;
                    double sx = 0, sy = 0, sampled = 0;
                    for (int s = 0; s < 8; ++s) {
                        int32_t j = slots[((size_t)cy * g + cx) * 8 + s];
                        if (j < 0 || (uint32_t)j == i) continue;
                        sampled += 1.0;
                        double dx = px - pos[2 * j], dy = py - pos[2 * j + 1];
                        if (dx * dx + dy * dy <= 0.0) { dx = rx; dy = ry; }
                        falloff(dx, dy, ak, &sx, &sy);
                    }
                    if (sampled > 0.0) { wx += others / sampled * sx; wy += others / sampled * sy; }
                }
            }
            wx += wx * rx; wy += wy * ry;
            double maxstep = 2.0 * (double)hs[lv];
            double sp = sqrt(wx * wx + wy * wy);
            if (sp > maxstep) { wx *= maxstep / sp; wy *= maxstep / sp; }
            vx += wx; vy += wy;
        }
        v[2 * i] = vx; v[2 * i + 1] = vy;
    }
    for (uint32_t lv = 0; lv < L; ++lv) free(agg[lv]);
    free(agg); free(hs); free(pf); free(slots);
}

/* link spring, one direction — ForceLink/force-spring.ts:38-96 (Gauss-Seidel:
 * each term sees the velocity accumulated over the point's earlier links).
 * rand[e]: the entry's U(0,1) draw; bias/strength as stored (f32). */
void or_link(uint32_t n, const double* pos, const uint8_t* absent,
             const uint32_t* offsets, const uint32_t* nbr, const float* rand,
             const float* bias, const float* strength,
             double spring, double distance, double lo, double hi, double alpha,
             double* v) {
#pragma omp parallel for schedule(dynamic, 512)
    for (uint32_t p = 0; p < n; ++p) {
        double vx = 0, vy = 0;
        double px = pos[2 * p], py = pos[2 * p + 1];
        for (uint32_t e = offsets[p]; e < offsets[p + 1]; ++e) {
            uint32_t q = nbr[e];
            if (absent && absent[q]) continue;
            double D = ((double)rand[e] * (hi - lo) + lo) * distance;
            double x = pos[2 * q] - (px + vx), y = pos[2 * q + 1] - (py + vy);
            double l = sqrt(x * x + y * y);
            if (l < D * 0.99) l = D * 0.99;
            l = (l - D) / l;
            l *= spring * alpha;
            l *= strength[e];
            l *= bias[e];
            vx += x * l; vy += y * l;
        }
        v[2 * p] = vx; v[2 * p + 1] = vy;
    }
}

/* collision grid geometry — ForceCollision/index.ts:65-95.
 * max_size = max(pointDefaultSize, resolved sizes). */
void or_collision_grid(float space_size, double radius, double padding, double max_size,
                       uint32_t* grid, float* cell_size) {
    double eff = (radius > 0.0 ? radius : max_size * 0.5) + padding;
    double cs = eff * 2.0 > 8.0 ? eff * 2.0 : 8.0;
    double gg = floor((double)space_size / cs);
    if (gg < 1.0) gg = 1.0;
    if (gg > 512.0) gg = 512.0;
    *grid = (uint32_t)gg;
    *cell_size = (float)((double)space_size / gg);
}

/* collision — ForceCollision/index.ts:264-345, build-grid.vert:28-62,
 * force-collision-spatial.frag:41-168.  sizes: resolved per-point sizes. */
void or_collision(uint32_t n, const double* pos, const uint8_t* absent, const float* sizes,
                  float space_size, double alpha, double strength, double radius, double padding,
                  double max_size, double* v) {
    static const float OFF[4][2] = {{0.f, 0.f}, {0.5f, 0.f}, {0.f, 0.5f}, {0.5f, 0.5f}};
    uint32_t G; float cs;
    or_collision_grid(space_size, radius, padding, max_size, &G, &cs);
    double* grid = (double*)malloc((size_t)G * G * 4 * sizeof(double));
    for (uint32_t i = 0; i < 2 * n; ++i) v[i] = 0.0;
    for (int o = 0; o < 4; ++o) {
        memset(grid, 0, (size_t)G * G * 4 * sizeof(double));
        volatile float ox = OFF[o][0] * cs, oy = OFF[o][1] * cs;
        for (uint32_t i = 0; i < n; ++i) {
            if (absent && absent[i]) continue;
            float px = (float)pos[2 * i], py = (float)pos[2 * i + 1];
            volatile float qx = px + ox, qy = py + oy;
            int32_t cx = or_cell(qx, cs, (int32_t)G), cy = or_cell(qy, cs, (int32_t)G);
            double* c = grid + ((size_t)cy * G + cx) * 4;
            c[0] += pos[2 * i]; c[1] += pos[2 * i + 1]; c[2] += sizes[i]; c[3] += 1.0;
        }
#pragma omp parallel for schedule(static, 1024)
        for (uint32_t i = 0; i < n; ++i) {
            if (absent && absent[i]) continue;
            double px = pos[2 * i], py = pos[2 * i + 1];
            double cur_r = (radius > 0.0 ? radius : sizes[i] * 0.5) + padding;
            volatile float qx = (float)px + ox, qy = (float)py + oy;
            int32_t mx = or_cell(qx, cs, (int32_t)G), my = or_cell(qy, cs, (int32_t)G);
            double ux = 0, uy = 0, total = 0;
            for (int dx = -1; dx <= 1; ++dx) {
                for (int dy = -1; dy <= 1; ++dy) {
                    int32_t cx = mx + dx, cy = my + dy;
                    if (cx < 0 || cy < 0 || cx >= (int32_t)G || cy >= (int32_t)G) continue;
                    const double* c = grid + ((size_t)cy * G + cx) * 4;
                    if (c[3] < 0.5) continue;
                    double cnt = c[3], sx = c[0], sy = c[1], ss = c[2];
                    if (dx == 0 && dy == 0) { cnt -= 1.0; sx -= px; sy -= py; ss -= sizes[i]; }
                    if (cnt < 0.5) continue;
                    total += cnt;
                    double ax = sx / cnt, ay = sy / cnt, as = ss / cnt;
                    double oth_r = (radius > 0.0 ? radius : as * 0.5) + padding;
                    double rho = cur_r + oth_r;
                    double ddx = px - ax, ddy = py - ay;
                    double d = sqrt(ddx * ddx + ddy * ddy);
                    if (d < rho && d > 0.001) {
                        double ov = (rho - d) / rho;
                        double soft = sqrt(ov) * rho * 0.5;
                        double f = alpha * strength * soft * 0.25 * cnt;
                        if (f > rho * 0.5) f = rho * 0.5;
                        ux += f * ddx / d; uy += f * ddy / d;
                    } else if (d <= 0.001 && cnt > 0.0) {
                        double ang = (double)((float)i * 0.618033988749895f);
                        double f = alpha * strength * rho * 0.1;
                        if (f > rho * 0.3) f = rho * 0.3;
                        ux += f * cnt * cos(ang); uy += f * cnt * sin(ang);
                    }
                }
            }
            if (total > 2.0) { ux *= 2.0 / total; uy *= 2.0 / total; }
            double maxc = cur_r * 0.1;
            double corr = sqrt(ux * ux + uy * uy);
            if (corr > maxc) { ux *= maxc / corr; uy *= maxc / corr; }
            v[2 * i] += ux; v[2 * i + 1] += uy;
        }
    }
    free(grid);
}

/* integrate — Points/update-position.frag:26-59: pinned or absent points keep
 * their position; else v *= friction; p += v; clamp each coordinate to [0, S]. */
void or_integrate(uint32_t n, double* pos, const double* v, const uint8_t* pinned,
                  const uint8_t* absent, double friction, double S) {
    for (uint32_t i = 0; i < n; ++i) {
        if ((pinned && pinned[i]) || (absent && absent[i])) continue;
        for (int c = 0; c < 2; ++c) {
            double x = pos[2 * i + c] + v[2 * i + c] * friction;
            if (x < 0.0) x = 0.0;
            if (x > S) x = S;
            pos[2 * i + c] = x;
        }
    }
}

/* alpha decay — Store/index.ts:328-332, src/index.ts:2084-2087. */
double or_alpha_step(double alpha, double decay) {
    return alpha + (0.0 - alpha) * (1.0 - pow(0.001, 1.0 / decay));
}
