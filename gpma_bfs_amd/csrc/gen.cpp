// Seeded synthetic graph generators (host code). They stand in for the
// reference's dataset script (dataset/prepare_dataset.sh), which downloads
// SNAP pokec and shuffles it without a seed: every edge list here is a pure
// function of its parameters, so the GPU path, the oracle and the CPU
// baseline all see identical edges.
//
// Random source: SplitMix64 (Steele, Lea, Flood 2014), one stream per
// independent unit of work (an R-MAT draw, a source vertex, the shuffle), so
// results do not depend on thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gpma.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform double in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n) (multiply-shift; n < 2^64)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

uint64_t mix(uint64_t a, uint64_t b) {
    SplitMix64 m(a * 0xD1B54A32D192ED03ull ^ (b + 0x8CB92BA72F3D8DD7ull));
    m.next();
    return m.next();
}

unsigned nthreads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return std::min(t, 16u);
}

template <typename F>
void parallel_for(uint64_t n, F f) {
    unsigned T = nthreads();
    if (n < 4096 || T == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + T - 1) / T;
    for (unsigned t = 0; t < T; t++) {
        uint64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=, &f] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

// Fisher-Yates over parallel arrays with its own stream.
void shuffle_edges(uint32_t *src, uint32_t *dst, uint64_t n, uint64_t seed) {
    SplitMix64 r(mix(seed, 0x5348554646ull));
    for (uint64_t i = n; i > 1; i--) {
        uint64_t j = r.below(i);
        std::swap(src[i - 1], src[j]);
        std::swap(dst[i - 1], dst[j]);
    }
}

}  // namespace

extern "C" {

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004) without noise: each of `ndraws`
// draws descends `scale` levels, MSB first, picking quadrant (0,0) with
// probability a, (0,1) b, (1,0) c and (1,1) d = 1-a-b-c from one uniform
// double per level. Self-loops and duplicate edges are removed, the unique
// edges are taken in (src, dst) order and then shuffled. Writes at most `cap`
// edges and returns the number of unique edges in *n_out.
int gpma_gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c, uint64_t seed, uint32_t *src,
                  uint32_t *dst, uint64_t cap, uint64_t *n_out) {
    if (scale == 0 || scale > 32 || !n_out) return GPMA_ERR_ARG;
    std::vector<uint64_t> keys(ndraws);
    const double ab = a + b, abc = a + b + c;
    parallel_for(ndraws, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t k = lo; k < hi; k++) {
            SplitMix64 r(mix(seed, k));
            uint64_t s = 0, d = 0;
            for (uint32_t l = 0; l < scale; l++) {
                double u = r.uniform();
                uint64_t sb = 0, db = 0;
                if (u < a) {
                } else if (u < ab) {
                    db = 1;
                } else if (u < abc) {
                    sb = 1;
                } else {
                    sb = 1;
                    db = 1;
                }
                s = (s << 1) | sb;
                d = (d << 1) | db;
            }
            keys[k] = (s == d) ? ~0ull : ((s << 32) | d);
        }
    });
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
    uint64_t n = keys.size();
    *n_out = n;
    if (!src || !dst) return 0;
    if (cap < n) return GPMA_ERR_CAPACITY;
    for (uint64_t i = 0; i < n; i++) {
        src[i] = (uint32_t)(keys[i] >> 32);
        dst[i] = (uint32_t)keys[i];
    }
    shuffle_edges(src, dst, n, seed);
    return 0;
}

// Directed power-law graph with exactly E unique edges and no self-loops.
// Out-degrees: Pareto weights w = u^(-1/(alpha-1)), u uniform on (0,1],
// softly capped as w*cap/(w+cap), apportioned to integers by cumulative
// rounding so they sum to E (each degree is also clamped to N-1).
// Destinations: R-MAT column marginal over ceil(log2 N) bits, each bit 1 with
// probability p1 (0.24 = b+d for the usual a,b,c,d), rejected if >= N, equal
// to the source, or already drawn for that source (a vertex whose degree
// exceeds N/4 draws uniformly instead). The edge list is written source-major,
// then shuffled.
int gpma_gen_powerlaw(uint32_t N, uint64_t E, double alpha, double soft_cap, double p1, uint64_t seed, uint32_t *src,
                      uint32_t *dst) {
    if (N < 2 || !src || !dst || alpha <= 1.0) return GPMA_ERR_ARG;
    if (E > (uint64_t)N * (N - 1)) return GPMA_ERR_ARG;
    std::vector<double> w(N);
    const double ex = -1.0 / (alpha - 1.0);
    for (uint32_t v = 0; v < N; v++) {
        SplitMix64 r(mix(seed, 0x44454752ull + v));
        double u = 1.0 - r.uniform();  // (0, 1]
        double x = std::pow(u, ex);
        w[v] = x * soft_cap / (x + soft_cap);
    }
    // cumulative rounding in long double keeps sum(deg) == E exactly
    long double total = 0;
    for (uint32_t v = 0; v < N; v++) total += w[v];
    std::vector<uint64_t> deg(N);
    long double acc = 0;
    uint64_t prev = 0;
    for (uint32_t v = 0; v < N; v++) {
        acc += w[v];
        uint64_t cur = (v + 1 == N) ? E : (uint64_t)std::floor(acc * (long double)E / total);
        if (cur > E) cur = E;
        deg[v] = cur - prev;
        prev = cur;
    }
    // clamp to N-1, pushing the excess to following vertices (wrapping)
    uint64_t excess = 0;
    for (uint32_t v = 0; v < N; v++) {
        if (deg[v] > (uint64_t)N - 1) {
            excess += deg[v] - (N - 1);
            deg[v] = N - 1;
        }
    }
    for (uint32_t v = 0; excess; v = (v + 1) % N) {
        if (deg[v] < (uint64_t)N - 1) {
            deg[v]++;
            excess--;
        }
    }
    std::vector<uint64_t> off(N + 1, 0);
    for (uint32_t v = 0; v < N; v++) off[v + 1] = off[v] + deg[v];
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < N) bits++;
    // threshold for "bit = 1" on 16 random bits
    const uint32_t thr = (uint32_t)std::llround(p1 * 65536.0);
    parallel_for(N, [&](uint64_t lo, uint64_t hi) {
        std::vector<uint32_t> table;
        for (uint64_t v = lo; v < hi; v++) {
            uint64_t d = deg[v];
            if (!d) continue;
            uint64_t tsize = 16;
            while (tsize < 2 * d) tsize <<= 1;
            table.assign(tsize, 0xFFFFFFFFu);
            SplitMix64 r(mix(seed ^ 0x52455354ull, v));
            uint64_t base = off[v], got = 0;
            uint64_t rnd = 0;
            int left = 0;
            // a vertex wanting a large share of all vertices draws uniformly
            const bool uniform = d * 4 > N;
            while (got < d) {
                uint32_t x = 0;
                if (uniform) x = (uint32_t)r.below(N);
                else for (uint32_t l = 0; l < bits; l++) {
                    if (left == 0) {
                        rnd = r.next();
                        left = 4;
                    }
                    uint32_t chunk = (uint32_t)(rnd & 0xFFFF);
                    rnd >>= 16;
                    left--;
                    x = (x << 1) | (chunk < thr ? 1u : 0u);
                }
                if (x >= N || x == (uint32_t)v) continue;
                uint64_t hsh = ((uint64_t)x * 0x9E3779B97F4A7C15ull) >> 20;
                uint64_t p = hsh & (tsize - 1);
                bool dup = false;
                while (table[p] != 0xFFFFFFFFu) {
                    if (table[p] == x) {
                        dup = true;
                        break;
                    }
                    p = (p + 1) & (tsize - 1);
                }
                if (dup) continue;
                table[p] = x;
                src[base + got] = (uint32_t)v;
                dst[base + got] = x;
                got++;
            }
        }
    });
    shuffle_edges(src, dst, E, seed);
    return 0;
}

}  // extern "C"
