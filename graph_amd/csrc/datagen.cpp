// datagen.cpp — seeded synthetic inputs for the force-simulation workloads.
//
// Host-only C++.  Every generator draws from one documented PRNG
// (splitmix64-seeded xoshiro256**), so the same seed gives the same graph on
// every machine.  These produce the benchmark/test inputs (graphs, positions,
// radii, jitter, per-tick seeds, per-link random rest lengths); the reference
// draws the last three from npm `random`/`seedrandom`, which cannot run here,
// so they are inputs to the simulation library rather than part of it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // [0,1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // [0,1) as f32 with 24 random bits
    float uniformf() { return (float)(next() >> 40) * 0x1.0p-24f; }
    uint64_t below(uint64_t n) { return n ? (uint64_t)(uniform() * (double)n) % n : 0; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Walker alias table for O(1) weighted draws.
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double>& w) {
        size_t n = w.size();
        prob.assign(n, 0.0);
        alias.assign(n, 0);
        double sum = 0;
        for (double x : w) sum += x;
        std::vector<double> sc(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            sc[i] = w[i] * (double)n / sum;
            (sc[i] < 1.0 ? small : large).push_back((uint32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(); small.pop_back();
            uint32_t l = large.back(); large.pop_back();
            prob[s] = sc[s]; alias[s] = l;
            sc[l] = (sc[l] + sc[s]) - 1.0;
            (sc[l] < 1.0 ? small : large).push_back(l);
        }
        for (uint32_t i : large) prob[i] = 1.0;
        for (uint32_t i : small) prob[i] = 1.0;
    }
    uint32_t draw(Rng& r) const {
        uint32_t i = (uint32_t)r.below(prob.size());
        return r.uniform() < prob[i] ? i : alias[i];
    }
};

}  // namespace

extern "C" {

// Positions uniform in [0, S)^2 — interleaved f32 [x0, y0, x1, y1, ...].
void cgd_uniform_square(uint32_t n, float S, uint64_t seed, float* xy) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) {
        xy[2 * i] = r.uniformf() * S;
        xy[2 * i + 1] = r.uniformf() * S;
    }
}

// Positions uniform in the disk of radius R centred at (cx, cy).
void cgd_uniform_disk(uint32_t n, float cx, float cy, float R, uint64_t seed, float* xy) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) {
        double rad = R * std::sqrt(r.uniform());
        double th = 2.0 * M_PI * r.uniform();
        xy[2 * i] = (float)(cx + rad * std::cos(th));
        xy[2 * i + 1] = (float)(cy + rad * std::sin(th));
    }
}

// Barabasi-Albert preferential attachment.  Node t >= 1 links to min(m, t)
// distinct earlier nodes drawn proportional to degree (uniformly while no
// edge exists yet).  Links are [t, target] as f32 pairs.  Returns the link
// count, sum_{t=1}^{n-1} min(m, t); links must hold 2*that floats.
uint64_t cgd_barabasi_albert(uint32_t n, uint32_t m, uint64_t seed, float* links) {
    Rng r(seed);
    std::vector<uint32_t> ends;
    ends.reserve((size_t)n * m * 2);
    uint64_t cnt = 0;
    std::vector<uint32_t> picked;
    for (uint32_t t = 1; t < n; ++t) {
        uint32_t k = std::min(m, t);
        picked.clear();
        while (picked.size() < k) {
            uint32_t c = ends.empty() ? (uint32_t)r.below(t) : ends[r.below(ends.size())];
            if (std::find(picked.begin(), picked.end(), c) == picked.end()) picked.push_back(c);
        }
        for (uint32_t c : picked) {
            links[2 * cnt] = (float)t;
            links[2 * cnt + 1] = (float)c;
            ++cnt;
            ends.push_back(t);
            ends.push_back(c);
        }
    }
    return cnt;
}

uint64_t cgd_barabasi_albert_count(uint32_t n, uint32_t m) {
    uint64_t c = 0;
    for (uint32_t t = 1; t < n; ++t) c += std::min(m, t);
    return c;
}

// Chung-Lu power-law graph: expected weights w_i = kmin * (n / (i+1))^(1/(gamma-1)),
// and exactly `m` links whose endpoints are drawn independently proportional
// to w (self-loops redrawn; repeated pairs kept).  Point ids are shuffled so
// hubs are not clustered at low indices.
void cgd_chung_lu(uint32_t n, double gamma, double kmin, uint64_t m, uint64_t seed, float* links) {
    Rng r(seed);
    std::vector<double> w(n);
    for (uint32_t i = 0; i < n; ++i) w[i] = kmin * std::pow((double)n / (double)(i + 1), 1.0 / (gamma - 1.0));
    Alias al(w);
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (uint32_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[r.below(i)]);
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t a, b;
        do { a = al.draw(r); b = al.draw(r); } while (a == b && n > 1);
        links[2 * e] = (float)perm[a];
        links[2 * e + 1] = (float)perm[b];
    }
}

// Stochastic block model: `blocks` blocks with Zipf(s) sizes summing to n
// (every block at least one point), about deg_in*n_b/2 links inside each block
// and cross links so that `cross_frac` of all links join different blocks.
// Points of a block are contiguous.  Returns the link count (<= cap).
uint64_t cgd_sbm(uint32_t n, uint32_t blocks, double zipf_s, double deg_in, double cross_frac,
                 uint64_t seed, float* links, uint64_t cap) {
    Rng r(seed);
    std::vector<double> z(blocks);
    double zs = 0;
    for (uint32_t b = 0; b < blocks; ++b) { z[b] = 1.0 / std::pow((double)(b + 1), zipf_s); zs += z[b]; }
    std::vector<uint32_t> size(blocks, 1);
    uint64_t rem = n > blocks ? n - blocks : 0, given = 0;
    for (uint32_t b = 0; b < blocks; ++b) {
        uint32_t extra = (uint32_t)std::floor(z[b] / zs * (double)rem);
        size[b] += extra; given += extra;
    }
    for (uint32_t b = 0; given < rem; b = (b + 1) % blocks) { size[b]++; given++; }
    std::vector<uint32_t> start(blocks + 1, 0);
    for (uint32_t b = 0; b < blocks; ++b) start[b + 1] = start[b] + size[b];
    uint64_t cnt = 0, inside = 0;
    for (uint32_t b = 0; b < blocks && cnt < cap; ++b) {
        if (size[b] < 2) continue;
        uint64_t mb = (uint64_t)std::llround(deg_in * size[b] * 0.5);
        for (uint64_t e = 0; e < mb && cnt < cap; ++e) {
            uint32_t a, c;
            do { a = start[b] + (uint32_t)r.below(size[b]); c = start[b] + (uint32_t)r.below(size[b]); } while (a == c);
            links[2 * cnt] = (float)a; links[2 * cnt + 1] = (float)c; ++cnt; ++inside;
        }
    }
    uint64_t cross = cross_frac < 1.0 ? (uint64_t)std::llround((double)inside * cross_frac / (1.0 - cross_frac)) : 0;
    std::vector<uint32_t> block_of(n);
    for (uint32_t b = 0; b < blocks; ++b)
        for (uint32_t i = start[b]; i < start[b + 1]; ++i) block_of[i] = b;
    for (uint64_t e = 0; e < cross && cnt < cap && blocks > 1; ++e) {
        uint32_t a, c;
        do { a = (uint32_t)r.below(n); c = (uint32_t)r.below(n); } while (block_of[a] == block_of[c]);
        links[2 * cnt] = (float)a; links[2 * cnt + 1] = (float)c; ++cnt;
    }
    return cnt;
}

// f32 log-normal samples exp(mu + sigma * N(0,1)).
void cgd_lognormal(uint32_t n, double mu, double sigma, uint64_t seed, float* out) {
    Rng r(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = (float)std::exp(mu + sigma * r.normal());
}

// f32 uniform samples in [lo, hi).
void cgd_uniform(uint64_t n, double lo, double hi, uint64_t seed, float* out) {
    Rng r(seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = (float)(lo + (hi - lo) * r.uniform());
}

// Many-body jitter vectors the reference draws per texel:
// (U(-1,1) * 1e-5, U(-1,1) * 1e-5) — ForceManyBody/index.ts:133-138.
void cgd_jitter(uint32_t n, uint64_t seed, float* xy) {
    Rng r(seed);
    for (uint32_t i = 0; i < 2 * n; ++i) xy[i] = (float)((r.uniform() * 2.0 - 1.0) * 0.00001);
}

}  // extern "C"
