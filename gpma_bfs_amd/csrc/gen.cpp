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
#include <atomic>
#include <cstdio>
#include <string>
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

constexpr uint64_t RMAT_BINNED_SHUFFLE = 1ull << 26;

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

// Sorted unique keys, identical to std::sort + std::unique; large inputs
// are bucketed by their top bits and the buckets sorted in parallel.
void sort_unique(std::vector<uint64_t> &keys, uint32_t key_bits) {
    const uint64_t n = keys.size();
    if (n < (1ull << 22)) {
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        return;
    }
    const uint32_t bb = 16;
    const uint32_t shift = key_bits > bb ? key_bits - bb : 0;
    const uint64_t NBK = 1ull << bb;
    auto bucket = [&](uint64_t k) { return std::min<uint64_t>(k >> shift, NBK - 1); };
    const unsigned T = nthreads();
    const uint64_t chunk = (n + T - 1) / T;
    std::vector<uint64_t> cnt((uint64_t)T * NBK, 0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const uint64_t a = t * chunk, b = std::min(n, a + chunk);
            uint64_t *c = cnt.data() + (uint64_t)t * NBK;
            for (uint64_t i = a; i < b; i++) c[bucket(keys[i])]++;
        });
    for (auto &x : th) x.join();
    th.clear();
    std::vector<uint64_t> bstart(NBK + 1, 0);
    uint64_t run = 0;
    for (uint64_t b = 0; b < NBK; b++) {
        bstart[b] = run;
        for (unsigned t = 0; t < T; t++) {
            const uint64_t c = cnt[(uint64_t)t * NBK + b];
            cnt[(uint64_t)t * NBK + b] = run;
            run += c;
        }
    }
    bstart[NBK] = run;
    std::vector<uint64_t> tmp(n);
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const uint64_t a = t * chunk, b = std::min(n, a + chunk);
            uint64_t *c = cnt.data() + (uint64_t)t * NBK;
            for (uint64_t i = a; i < b; i++) tmp[c[bucket(keys[i])]++] = keys[i];
        });
    for (auto &x : th) x.join();
    th.clear();
    std::vector<uint64_t> ucount(NBK, 0);
    std::atomic<uint64_t> next{0};
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&] {
            for (uint64_t b; (b = next.fetch_add(1)) < NBK;) {
                auto lo = tmp.begin() + bstart[b], hi = tmp.begin() + bstart[b + 1];
                std::sort(lo, hi);
                ucount[b] = (uint64_t)(std::unique(lo, hi) - lo);
            }
        });
    for (auto &x : th) x.join();
    th.clear();
    std::vector<uint64_t> uoff(NBK + 1, 0);
    for (uint64_t b = 0; b < NBK; b++) uoff[b + 1] = uoff[b] + ucount[b];
    parallel_for(NBK, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t b = lo; b < hi; b++)
            std::copy(tmp.begin() + bstart[b], tmp.begin() + bstart[b] + ucount[b], keys.begin() + uoff[b]);
    });
    keys.resize(uoff[NBK]);
}

// Uniform shuffle of a large key list: every key goes to one of NBIN bins
// drawn from its own stream (index-keyed, so the result is independent of
// the thread count), keys keep their order inside a bin, each bin is then
// Fisher-Yates shuffled with its own stream, and the bins are concatenated.
// I.i.d. uniform bins followed by uniform in-bin shuffles give a uniform
// permutation (Sanders 1998, "Random permutations on distributed, external
// and hierarchical memory").
void shuffle_binned(std::vector<uint64_t> &keys, uint64_t seed) {
    const uint64_t n = keys.size();
    const uint64_t NBIN = 4096;
    const unsigned T = nthreads();
    const uint64_t chunk = (n + T - 1) / T;
    const uint64_t bseed = mix(seed, 0x42494E53ull);
    auto bin = [&](uint64_t i) { return (uint64_t)(((unsigned __int128)mix(bseed, i) * NBIN) >> 64); };
    std::vector<uint64_t> cnt((uint64_t)T * NBIN, 0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const uint64_t a = t * chunk, b = std::min(n, a + chunk);
            for (uint64_t i = a; i < b; i++) cnt[(uint64_t)t * NBIN + bin(i)]++;
        });
    for (auto &x : th) x.join();
    th.clear();
    std::vector<uint64_t> bstart(NBIN + 1, 0);
    uint64_t run = 0;
    for (uint64_t b = 0; b < NBIN; b++) {
        bstart[b] = run;
        for (unsigned t = 0; t < T; t++) {
            const uint64_t c = cnt[(uint64_t)t * NBIN + b];
            cnt[(uint64_t)t * NBIN + b] = run;
            run += c;
        }
    }
    bstart[NBIN] = run;
    std::vector<uint64_t> tmp(n);
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            const uint64_t a = t * chunk, b = std::min(n, a + chunk);
            uint64_t *c = cnt.data() + (uint64_t)t * NBIN;
            for (uint64_t i = a; i < b; i++) tmp[c[bin(i)]++] = keys[i];
        });
    for (auto &x : th) x.join();
    th.clear();
    const uint64_t sseed = mix(seed, 0x5348554646ull);
    parallel_for(NBIN, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t b = lo; b < hi; b++) {
            SplitMix64 r(mix(sseed, b));
            uint64_t *base = tmp.data() + bstart[b];
            for (uint64_t i = bstart[b + 1] - bstart[b]; i > 1; i--) std::swap(base[i - 1], base[r.below(i)]);
        }
    });
    keys.swap(tmp);
}

// Parser of whitespace-separated unsigned decimal integers over a buffer.
struct Scanner {
    const char *p, *e;
    // next integer (false at the end of the buffer or on a non-digit token)
    bool next(uint64_t *out, bool *bad) {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
        if (p == e) return false;
        if (*p < '0' || *p > '9') {
            *bad = true;
            return false;
        }
        uint64_t v = 0;
        while (p < e && *p >= '0' && *p <= '9') {
            v = v * 10 + (uint64_t)(*p - '0');
            if (v > 0xFFFFFFFFFFull) {
                *bad = true;
                return false;
            }
            p++;
        }
        if (p < e && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) {
            *bad = true;
            return false;
        }
        *out = v;
        return true;
    }
};

bool read_file(const char *path, std::string *buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf->append(tmp, n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

}  // namespace

namespace gpma {
void set_last_error(const std::string &msg);  // gpma.hip
}

extern "C" {

// The reference's graph text format (dataset/prepare_dataset.sh:34-37, read
// by gpma_bfs_demo.cu:6-30): a header "N E", then E lines "src dst" with
// 0-based node ids. Reads N and E, and with src and dst given (cap >= E
// entries) the edges in file order. Ids must be < 2^32.
int gpma_read_edge_text(const char *path, uint32_t *N, uint64_t *E, uint32_t *src, uint32_t *dst, uint64_t cap) {
    if (!path || !N || !E) {
        gpma::set_last_error("gpma_read_edge_text: null argument");
        return GPMA_ERR_ARG;
    }
    std::string buf;
    if (!read_file(path, &buf)) {
        gpma::set_last_error(std::string("gpma_read_edge_text: cannot read ") + path);
        return GPMA_ERR_ARG;
    }
    Scanner sc{buf.data(), buf.data() + buf.size()};
    bool bad = false;
    uint64_t n = 0, e = 0;
    if (!sc.next(&n, &bad) || !sc.next(&e, &bad) || n > 0xFFFFFFFFull) {
        gpma::set_last_error("gpma_read_edge_text: missing or bad \"N E\" header");
        return GPMA_ERR_ARG;
    }
    *N = (uint32_t)n;
    *E = e;
    if (!src || !dst) return GPMA_OK;
    if (cap < e) {
        gpma::set_last_error("gpma_read_edge_text: buffers hold " + std::to_string(cap) + " edges, file has " +
                             std::to_string(e));
        return GPMA_ERR_CAPACITY;
    }
    for (uint64_t i = 0; i < e; i++) {
        uint64_t a = 0, b = 0;
        if (!sc.next(&a, &bad) || !sc.next(&b, &bad) || a > 0xFFFFFFFFull || b > 0xFFFFFFFFull) {
            gpma::set_last_error("gpma_read_edge_text: edge " + std::to_string(i) + " missing or malformed");
            return GPMA_ERR_ARG;
        }
        src[i] = (uint32_t)a;
        dst[i] = (uint32_t)b;
    }
    return GPMA_OK;
}

// Writes the same format: "N E" then one "src dst" line per edge.
int gpma_write_edge_text(const char *path, uint32_t N, uint64_t E, const uint32_t *src, const uint32_t *dst) {
    if (!path || (E && (!src || !dst))) {
        gpma::set_last_error("gpma_write_edge_text: null argument");
        return GPMA_ERR_ARG;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        gpma::set_last_error(std::string("gpma_write_edge_text: cannot open ") + path);
        return GPMA_ERR_ARG;
    }
    std::string out;
    out.reserve(1 << 20);
    out += std::to_string(N) + " " + std::to_string(E) + "\n";
    bool ok = true;
    for (uint64_t i = 0; i < E && ok; i++) {
        char line[32];
        const int k = std::snprintf(line, sizeof(line), "%u %u\n", src[i], dst[i]);
        out.append(line, (size_t)k);
        if (out.size() >= (1u << 20)) {
            ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
            out.clear();
        }
    }
    if (ok && !out.empty()) ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        gpma::set_last_error(std::string("gpma_write_edge_text: write failed: ") + path);
        return GPMA_ERR_ARG;
    }
    return GPMA_OK;
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004) without noise: each of `ndraws`
// draws descends `scale` levels, MSB first, picking quadrant (0,0) with
// probability a, (0,1) b, (1,0) c and (1,1) d = 1-a-b-c from one uniform
// double per level. Self-loops and duplicate edges are removed, the unique
// edges are taken in (src, dst) order and then shuffled (Fisher-Yates below
// 2^26 edges, the binned parallel shuffle above). Writes at most `cap` edges
// and returns the number of unique edges in *n_out.
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
    sort_unique(keys, 32 + scale);
    if (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
    uint64_t n = keys.size();
    *n_out = n;
    if (!src || !dst) return 0;
    if (cap < n) return GPMA_ERR_CAPACITY;
    if (n < RMAT_BINNED_SHUFFLE) {
        for (uint64_t i = 0; i < n; i++) {
            src[i] = (uint32_t)(keys[i] >> 32);
            dst[i] = (uint32_t)keys[i];
        }
        shuffle_edges(src, dst, n, seed);
        return 0;
    }
    shuffle_binned(keys, seed);
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++) {
            src[i] = (uint32_t)(keys[i] >> 32);
            dst[i] = (uint32_t)keys[i];
        }
    });
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
