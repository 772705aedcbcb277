// The reference demo's protocol (gpma_bfs_demo.cu:6-116) written against the
// C++ facade (include/gpma.hpp): load an "N E" text graph with 0-based
// "src dst" lines, insert the first half as the window, BFS, slide the window
// 100 times (one update batch of `step` inserts then `step` deletes per
// slide), BFS again. Prints the same lines as the reference demo.
//
// Usage: gpma_bfs_demo GRAPH.txt START_NODE
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpma.hpp"

// gpma_bfs_demo.cu:6-30, through the library's reader of the same format
static void load_data(const char *path, std::vector<int> &xs, std::vector<int> &ys, int &node_size, int &edge_size) {
    uint32_t n = 0;
    uint64_t e = 0;
    if (gpma_read_edge_text(path, &n, &e, nullptr, nullptr, 0) != GPMA_OK) {
        printf("Open graph file failed.\n");
        exit(0);
    }
    node_size = (int)n;
    edge_size = (int)e;
    printf("node_num: %d, edge_num: %d\n", node_size, edge_size);
    std::vector<uint32_t> s(e), d(e);
    if (gpma_read_edge_text(path, &n, &e, s.data(), d.data(), e) != GPMA_OK) {
        fprintf(stderr, "%s\n", gpma_last_error());
        exit(1);
    }
    xs.assign(s.begin(), s.end());
    ys.assign(d.begin(), d.end());
    printf("Graph file is loaded.\n");
}

static long count_reach(const std::vector<SIZE_TYPE> &r) {
    long c = 0;
    for (SIZE_TYPE x : r) c += x != 0;
    return c;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        printf("Invalid arguments.\n");
        return -1;
    }
    int bfs_start_node = std::atoi(argv[2]);
    std::vector<int> host_x, host_y;
    int node_size = 0, edge_size = 0;
    load_data(argv[1], host_x, host_y, node_size, edge_size);
    int half = edge_size / 2;
    std::vector<KEY_TYPE> base_keys(half);
    for (int i = 0; i < half; i++) base_keys[i] = ((KEY_TYPE)host_x[i] << 32) + host_y[i];
    std::vector<VALUE_TYPE> base_values(half, 1);
    int num_slide = 100;
    int step = half / num_slide;

    GPMA_multidev<3, 1> gpma(node_size);
    gpma.update_batch(base_keys, base_values);
    std::vector<SIZE_TYPE> bfs_result(node_size);
    gpma_bfs(gpma, node_size, edge_size, bfs_start_node, bfs_result.data());
    printf("start from node %d, number of reachable nodes: %ld\n", bfs_start_node, count_reach(bfs_result));

    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < num_slide; i++) {
        std::vector<KEY_TYPE> hk(step * 2);
        for (int j = 0; j < step; j++) {
            int idx = half + i * step + j;
            hk[j] = ((KEY_TYPE)host_x[idx] << 32) + host_y[idx];
        }
        for (int j = 0; j < step; j++) {
            int idx = i * step + j;
            hk[j + step] = ((KEY_TYPE)host_x[idx] << 32) + host_y[idx];
        }
        std::vector<VALUE_TYPE> update_values(step * 2, VALUE_NONE);
        for (int j = 0; j < step; j++) update_values[j] = 1;
        gpma.update_batch(hk, update_values);
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("Graph is updated.\n");
    gpma_bfs(gpma, node_size, edge_size, bfs_start_node, bfs_result.data());
    printf("start from node %d, number of reachable nodes: %ld\n", bfs_start_node, count_reach(bfs_result));
    fprintf(stderr, "slides: %.3f s (%.1f M edge-updates/s, host batches incl. PCIe)\n", s,
            2.0 * step * num_slide / s / 1e6);
    return 0;
}
