// gpma.hpp - header-only C++ facade over gpma.h that keeps the reference's
// host-side names (gpma.cuh, gpma_bfs.cuh, multidev.cuh, utils.cuh), so a
// driver written against the reference compiles against this with its
// includes changed. Everything runs on one GPU through the C ABI; there is no
// DEV template parameter and no CPU instance.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gpma.h"

// utils.cuh:11,19-21
#ifndef MAKE_64b
#define MAKE_64b(high, low) (((uint64_t)(high) << 32) + (low))
#endif
using KEY_TYPE = uint64_t;
using VALUE_TYPE = char;
using SIZE_TYPE = uint32_t;

// gpma.cuh:15-19
constexpr KEY_TYPE KEY_NONE = GPMA_KEY_NONE;
constexpr KEY_TYPE KEY_MAX = GPMA_KEY_MAX;
constexpr SIZE_TYPE SIZE_NONE = 0xFFFFFFFFu;
constexpr VALUE_TYPE VALUE_NONE = 0;
constexpr KEY_TYPE COL_IDX_NONE = GPMA_COL_IDX_NONE;

namespace gpma_amd {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char *where) {
    if (rc != GPMA_OK) throw error(rc, std::string(where) + ": " + gpma_last_error());
}

}  // namespace gpma_amd

// The reference's vectors (utils.cuh:47-52); host vectors here, since the
// ABI stages host batches itself.
template <typename T>
using NATIVE_VEC = std::vector<T>;
using NATIVE_VEC_KEY_HOST = std::vector<KEY_TYPE>;
using NATIVE_VEC_VALUE_HOST = std::vector<VALUE_TYPE>;

// class GPMA (gpma.cuh:31-102), device-resident, move-only.
class GPMA {
  public:
    SIZE_TYPE row_num = 0;

    GPMA() = default;
    explicit GPMA(SIZE_TYPE rows, const gpma_config *cfg = nullptr) : row_num(rows) {
        gpma_amd::check(gpma_create(rows, cfg, &h_), "gpma_create");
    }
    ~GPMA() {
        if (h_) gpma_destroy(h_);
    }
    GPMA(const GPMA &) = delete;
    GPMA &operator=(const GPMA &) = delete;
    GPMA(GPMA &&o) noexcept : row_num(o.row_num), h_(o.h_) { o.h_ = nullptr; }
    GPMA &operator=(GPMA &&o) noexcept {
        std::swap(h_, o.h_);
        std::swap(row_num, o.row_num);
        return *this;
    }

    gpma_t handle() const { return h_; }
    gpma_csr csr() const {
        gpma_csr c{};
        gpma_amd::check(gpma_get_csr(h_, &c), "gpma_get_csr");
        return c;
    }
    // public members of the reference, as accessors
    const KEY_TYPE *keys() const { return csr().keys; }
    const VALUE_TYPE *values() const { return (const VALUE_TYPE *)csr().values; }
    const SIZE_TYPE *row_offset() const { return csr().row_offset; }
    SIZE_TYPE segment_length() const { return csr().segment_length; }
    SIZE_TYPE tree_height() const { return csr().tree_height; }
    int get_size() const { return (int)csr().capacity; }  // gpma.cuh:52

    // GPMA::mirror() (gpma.cuh:90-101): host copies of the whole state
    void mirror(std::vector<KEY_TYPE> &k, std::vector<VALUE_TYPE> &v, std::vector<SIZE_TYPE> &r) const {
        gpma_csr c = csr();
        k.resize(c.capacity);
        v.resize(c.capacity);
        r.resize((size_t)row_num + 1);
        gpma_amd::check(gpma_copy_layout(h_, k.data(), (uint8_t *)v.data(), r.data()), "gpma_copy_layout");
    }
    void print_status(const std::string &prefix = "DBG") const {
        gpma_csr c = csr();
        std::printf("%s: GPMA: capacity=%llu seg_length=%u tree_height=%u row_num=%u\n", prefix.c_str(),
                    (unsigned long long)c.capacity, c.segment_length, c.tree_height, c.row_num);
    }

  private:
    gpma_t h_ = nullptr;
};

// The name the reference's drivers log (gpma_bfs_demo.cu:63).
inline GPMA init_csr_gpma(SIZE_TYPE row_num) { return GPMA(row_num); }

// update_gpma(gpma, keys, values) (gpma.cuh:968-969). The reference sorts and
// compacts the caller's vectors in place; they are left untouched here.
template <typename KVec, typename VVec>
inline void update_gpma(GPMA &g, KVec &update_keys, VVec &update_values) {
    if (update_keys.size() != update_values.size())
        throw std::invalid_argument("update_gpma: keys and values differ in size");
    gpma_amd::check(gpma_update(g.handle(), (const uint64_t *)update_keys.data(),
                                (const uint8_t *)update_values.data(), update_keys.size(), 0),
                    "gpma_update");
}

// gpma_bfs(keys, values, row_offsets, node_size, edge_size, start, results)
// (gpma_bfs.cuh:511-515): device CSR view in, host results out.
inline void gpma_bfs(const KEY_TYPE *keys, const VALUE_TYPE *values, const SIZE_TYPE *row_offsets,
                     SIZE_TYPE node_size, SIZE_TYPE edge_size, SIZE_TYPE start_node, SIZE_TYPE *results) {
    gpma_amd::check(gpma_csr_bfs(keys, (const uint8_t *)values, row_offsets, node_size, edge_size, start_node,
                                 results),
                    "gpma_csr_bfs");
}

// BFS on a GPMA into host results[node_size] (shape of gpma_bfs.cuh:517-533).
inline void gpma_bfs(const GPMA &g, SIZE_TYPE node_size, SIZE_TYPE edge_size, SIZE_TYPE start_node,
                     SIZE_TYPE *results) {
    (void)edge_size;
    if (node_size != g.row_num) throw std::invalid_argument("gpma_bfs: node_size != row_num");
    gpma_amd::check(gpma_bfs(g.handle(), start_node, results, 0, nullptr), "gpma_bfs");
}

// GPMA_multidev<cpus, gpus> (multidev.cuh:77-142), restricted to one GPU:
// every key goes to the single device-resident GPMA, whatever `cpus` says,
// so the demo driver's call shape (construct, update_batch, gpma_bfs)
// compiles unchanged.
template <size_t cpus, size_t gpus>
class GPMA_multidev {
    static_assert(gpus == 1, "this build runs on one GPU");

  public:
    explicit GPMA_multidev(SIZE_TYPE row_num) : g_(row_num) {}
    template <typename KVec, typename VVec>
    void update_batch(KVec &keys, VVec &values) {
        if (keys.size() != values.size()) throw std::invalid_argument("keys.size() != values.size()");
        update_gpma(g_, keys, values);
    }
    const GPMA &device() const { return g_; }
    GPMA &device() { return g_; }

  private:
    GPMA g_;
};

template <size_t cpus, size_t gpus>
inline void gpma_bfs(const GPMA_multidev<cpus, gpus> &m, SIZE_TYPE node_size, SIZE_TYPE edge_size,
                     SIZE_TYPE start_node, SIZE_TYPE *results) {
    gpma_bfs(m.device(), node_size, edge_size, start_node, results);
}
