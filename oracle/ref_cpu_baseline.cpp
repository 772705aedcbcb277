// TEST INFRASTRUCTURE ONLY: a driver for the reference's own sequential
// PMA (cpu_baseline/containers/pma_dynamic_graph.hpp) and BFS
// (cpu_baseline/algorithms/bfs.hpp), compiled from the reference's headers
// where they lie (include path set by oracle/build.py). It runs the
// reference protocol of cpu_baseline/main.cpp:36-61 on an edge list read
// from a binary file instead of the pokec text file, and writes BFS levels,
// the live edge set and per-phase timings. Used to make golden fixtures and
// as the timed CPU baseline; never part of the product.
//
// Edge file: u32 N, u64 E, u32 src[E], u32 dst[E] (little endian).
// Usage: ref_cpu_baseline EDGES OUTDIR SLIDES START [timing-only: 0|1]
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

// same include order as cpu_baseline/main.cpp:1-3 (bfs.hpp relies on it)
#include "containers/pma_dynamic_graph.hpp"
#include "algorithms/bfs.hpp"

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void write_levels(const std::string &path, const std::vector<int> &lv) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return;
    std::vector<uint32_t> u(lv.begin(), lv.end());
    fwrite(u.data(), 4, u.size(), f);
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s EDGES OUTDIR SLIDES START [timing_only]\n", argv[0]);
        return 2;
    }
    const std::string outdir = argv[2];
    const int slides = atoi(argv[3]);
    const int start = atoi(argv[4]);
    const bool timing_only = argc > 5 && atoi(argv[5]) != 0;
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    uint32_t N = 0;
    uint64_t E = 0;
    if (fread(&N, 4, 1, f) != 1 || fread(&E, 8, 1, f) != 1) return 2;
    std::vector<uint32_t> src(E), dst(E);
    if (fread(src.data(), 4, E, f) != E || fread(dst.data(), 4, E, f) != E) return 2;
    fclose(f);
    std::vector<edge_type> edges(E);
    for (uint64_t i = 0; i < E; i++) edges[i] = edge_type((int)src[i], (int)dst[i]);

    // cpu_baseline/main.cpp:36-61
    double t0 = now_s();
    pma_dynamic_graph graph((int)N);
    double t1 = now_s();
    const int half = (int)(E / 2);
    graph.init_edge(edges.begin(), edges.begin() + half);
    double t2 = now_s();
    std::vector<int> lv;
    bfs<pma_dynamic_graph>(graph, start, lv);
    double t3 = now_s();
    long reach0 = 0;
    for (int x : lv) reach0 += x != 0;
    if (!timing_only) write_levels(outdir + "/levels_before.u32", lv);
    const int step = half / 100;
    double t4 = now_s();
    for (int i = 0; i < slides; i++) {
        graph.delete_edge(edges.begin() + step * i, edges.begin() + step * i + step);
        graph.insert_edge(edges.begin() + step * i + half, edges.begin() + step * i + step + half);
    }
    double t5 = now_s();
    bfs<pma_dynamic_graph>(graph, start, lv);
    double t6 = now_s();
    long reach1 = 0;
    for (int x : lv) reach1 += x != 0;
    if (!timing_only) {
        write_levels(outdir + "/levels_after.u32", lv);
        // live edges in (src, dst) order, walked through the CSR iterators
        FILE *o = fopen((outdir + "/live_after.u64").c_str(), "wb");
        if (o) {
            for (uint32_t v = 0; v < N; v++) {
                for (auto it = graph.get_edge_begin((int)v); it != graph.get_edge_end((int)v); ++it) {
                    uint64_t k = ((uint64_t)v << 32) | (uint32_t)(*it);
                    fwrite(&k, 8, 1, o);
                }
            }
            fclose(o);
        }
    }
    printf("{\"N\": %u, \"E\": %llu, \"half\": %d, \"step\": %d, \"slides\": %d, \"guard_init_s\": %.6f, "
           "\"init_edge_s\": %.6f, \"bfs0_s\": %.6f, \"slides_s\": %.6f, \"bfs1_s\": %.6f, \"reach_before\": %ld, "
           "\"reach_after\": %ld}\n",
           N, (unsigned long long)E, half, step, slides, t1 - t0, t2 - t1, t3 - t2, t5 - t4, t6 - t5, reach0, reach1);
    return 0;
}
