// TEST INFRASTRUCTURE ONLY. This file is the CPU oracle for the GPMA update
// and BFS path. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as a checker or a timed baseline.
// The product (gpma_bfs_amd/) never links, loads or calls it.
//
// What it restates: the host (DEV == CPU) branch of the reference's
// update_gpma pipeline in gpma.cuh, step by step, so that the physical layout
// (keys[], values[], row_offset[], segment_length, tree_height) it produces is
// the layout the reference produces on the same update sequence, plus the
// single-device BFS semantics of gpma_bfs.cuh.
//
// Pinning (see tests/test_oracle.py):
//   * recalc_density, fls, resize sizing: closed-form values of the reference
//     formulas (gpma.cuh:104-126,147-168,906-934) at the survey's shapes.
//   * layout: the hand-traced 4-node states in the reference's notes
//     (gpma.bfs.notes.pdf pp.2-4, reproduced as fixtures).
//   * live-key set: closed form of the sliding window after every batch.
//   * BFS levels: scipy's unweighted shortest paths and the reference's own
//     cpu_baseline (compiled from its sources into oracle/_ref/).
//   * redistribution: the per-item slot formula against the per-segment
//     closed form, by brute force over small (W, m).
//
// Deliberate, documented differences from the reference:
//   * with orc_set_fix_root(1) (off by default), a root rebalance that
//     would fail after a resize grows the array instead (the product's
//     behaviour); by default the pending inserts are dropped exactly as
//     gpma.cuh:1034-1044 does.
//   * duplicate keys inside one batch: the locate loop runs in parallel like
//     the reference's, but only the last record of a key writes its value
//     byte, the result of a sequential pass (the reference races on that
//     byte). Workloads contain no intra-batch duplicates.
//
// Built with OpenMP (oracle/build.py), the locate loop, the per-window
// rebalance loop and the batch sort run on every host core, as the
// reference's CPU path does; results do not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace {

// Stable sort of an index permutation (the batch sort, thrust::sort_by_key
// in the reference): the libstdc++ parallel stable sort when built with
// OpenMP (same result: it is stable), else std::stable_sort.
template <typename Less>
void stable_sort_order(std::vector<size_t> &order, Less less) {
#ifdef _OPENMP
    __gnu_parallel::stable_sort(order.begin(), order.end(), less);
#else
    std::stable_sort(order.begin(), order.end(), less);
#endif
}


bool g_fix_root = false;

using key_t_ = uint64_t;
using val_t = uint8_t;  // the reference's VALUE_TYPE is char; only 0 / non-0 matters
using size_type = uint32_t;

// gpma.cuh:15-19
constexpr key_t_ KEY_NONE = ~0ull;
constexpr key_t_ KEY_MAX = KEY_NONE - 1;
constexpr size_type SIZE_NONE = 0xFFFFFFFFu;
constexpr val_t VALUE_NONE = 0;
constexpr key_t_ COL_IDX_NONE = 0xFFFFFFFFull;

// gpma.cuh:104-126 (1-based index of the highest set bit, 0 for 0)
size_type fls(size_type x) {
    if (!x) return 0;
    return 32u - (size_type)__builtin_clz(x);
}

struct Gpma {
    std::vector<key_t_> keys;
    std::vector<val_t> values;
    size_type segment_length = 0;
    size_type tree_height = 0;
    std::vector<size_type> lower_element, upper_element;
    size_type row_num = 0;
    std::vector<size_type> row_offset;
    // diagnostics (not part of the reference state)
    uint64_t dropped = 0;       // updates lost to a failed post-resize root rebalance
    uint64_t resizes = 0;
    uint64_t bulk_loads = 0;
    std::vector<uint64_t> level_windows;  // windows tried per level, cumulative
    std::vector<uint64_t> level_success;  // windows rebalanced per level, cumulative
    std::vector<uint64_t> level_slots;    // slots touched per level, cumulative
};

// gpma.cuh:147-168: thresholds linear in the level, computed in double in
// exactly this expression order, then trimmed against the level below.
void recalculate_density(Gpma &g) {
    const double lower_leaf = 0.08, lower_root = 0.42, upper_leaf = 0.92, upper_root = 0.84;
    g.lower_element.assign(g.tree_height + 1, 0);
    g.upper_element.assign(g.tree_height + 1, 0);
    size_type level_length = g.segment_length;
    for (size_type i = 0; i <= g.tree_height; i++) {
        double dl = lower_root + (lower_leaf - lower_root) * (g.tree_height - i) / g.tree_height;
        double du = upper_root + (upper_leaf - upper_root) * (g.tree_height - i) / g.tree_height;
        g.lower_element[i] = (size_type)std::ceil(dl * level_length);
        g.upper_element[i] = (size_type)std::floor(du * level_length);
        if (i > 0) {
            g.lower_element[i] = std::max(g.lower_element[i], 2 * g.lower_element[i - 1]);
            g.upper_element[i] = std::min(g.upper_element[i], 2 * g.upper_element[i - 1]);
        }
        level_length <<= 1;
    }
}

// gpma.cuh:636-654: item i of m sorted items of a window of width W goes to
// slot (f*i/m)*S + (f*i mod m)/f with f = W/S, 64-bit arithmetic; a guard
// key records its slot in row_offset[row+1].
inline void redispatch_one(size_t i, const key_t_ *tk, const val_t *tv, key_t_ *keys, val_t *values,
                           size_type width, size_type S, size_type m, size_type *row_offset, size_type base) {
    uint64_t frac = width / S;
    uint64_t deno = m;
    size_type seg_idx = (size_type)(frac * i / deno);
    size_type seg_lane = (size_type)(frac * i % deno / frac);
    size_type loc = seg_idx * S + seg_lane;
    key_t_ k = tk[i];
    keys[loc] = k;
    values[loc] = tv[i];
    if ((k & COL_IDX_NONE) == COL_IDX_NONE) {
        size_type row = (size_type)(k >> 32);
        row_offset[row + 1] = loc + base;
    }
}

// gpma.cuh:188-240 (locate_leaf_loop_body + handle_del_mod). Parallel over
// the updates like the reference's OpenMP loop (gpma.cuh:234); the input is
// sorted, and of several records of one key only the last writes the value,
// which is the write a sequential pass leaves (locate reads no values).
void locate_leaf_batch(Gpma &g, const key_t_ *uk, const val_t *uv, size_type n, size_type *leaf) {
    const size_type S = g.segment_length;
#pragma omp parallel for schedule(static)
    for (size_type i = 0; i < n; i++) {
        key_t_ key = uk[i];
        val_t value = uv[i];
        size_type prefix = 0;
        size_type current_bit = S << g.tree_height >> 1;
        while (S <= current_bit) {
            if (g.keys[prefix | current_bit] <= key) prefix |= current_bit;
            current_bit >>= 1;
        }
        size_type out = prefix;
        if (value == VALUE_NONE) out = SIZE_NONE;
        const bool last_of_key = i + 1 == n || uk[i + 1] != key;
        for (size_type j = 0; j < S; j++) {
            if (g.keys[prefix + j] == key) {
                if (last_of_key) g.values[prefix + j] = value;
                out = SIZE_NONE;
                break;
            }
        }
        leaf[i] = out;
    }
}

// gpma.cuh:836-844: stable removal of finished (SIZE_NONE) updates.
void compact_insertions(std::vector<size_type> &nodes, std::vector<key_t_> &uk, std::vector<val_t> &uv, size_type &n) {
    size_type w = 0;
    for (size_type i = 0; i < n; i++) {
        if (nodes[i] == SIZE_NONE) continue;
        nodes[w] = nodes[i];
        uk[w] = uk[i];
        uv[w] = uv[i];
        w++;
    }
    n = w;
}

// gpma.cuh:846-879 with cpu_alg.hpp:7-30,78-88: run-length encode the sorted
// window ids, exclusive scan of run lengths, offset[u] = n.
size_type group_insertion_by_node(const std::vector<size_type> &nodes, size_type n, std::vector<size_type> &uniq,
                                  std::vector<size_type> &offset) {
    uniq.clear();
    offset.clear();
    if (n == 0) {
        // cpu_rle_simple returns 0 runs; the exclusive sum of 0 items writes
        // nothing and update_offset[0] = update_size = 0.
        offset.push_back(0);
        return 0;
    }
    size_type run_begin = 0;
    for (size_type i = 0; i < n; i++) {
        if (nodes[run_begin] != nodes[i]) {
            uniq.push_back(nodes[run_begin]);
            offset.push_back(run_begin);
            run_begin = i;
        }
    }
    uniq.push_back(nodes[run_begin]);
    offset.push_back(run_begin);
    offset.push_back(n);
    return (size_type)uniq.size();
}

// gpma.cuh:665-726 (rebalancing_impl_cpu): for every window, compact the
// live items (gaps and tombstones dropped) into a temporary, and if
// lower <= live + pending <= upper, append the pending inserts, sort, clear
// the window's keys and redistribute. A failing window is left untouched.
void rebalance_batch(Gpma &g, size_type level, std::vector<size_type> &nodes, const std::vector<key_t_> &uk,
                     const std::vector<val_t> &uv, const std::vector<size_type> &uniq,
                     const std::vector<size_type> &offset, size_type nuniq, size_type lower, size_type upper) {
    const size_type S = g.segment_length;
    const size_type width = S << level;
    if (g.level_windows.size() <= level) {
        g.level_windows.resize(level + 1, 0);
        g.level_success.resize(level + 1, 0);
        g.level_slots.resize(level + 1, 0);
    }
    g.level_windows[level] += nuniq;
    g.level_slots[level] += (uint64_t)nuniq * width;
    uint64_t success = 0;
    // windows are disjoint: parallel like the reference's OpenMP loop
    // (gpma.cuh:669-678)
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : success)
    for (size_type u = 0; u < nuniq; u++) {
        std::vector<key_t_> tk(width);
        std::vector<val_t> tv(width);
        std::vector<size_t> order;
        size_type base = uniq[u];
        key_t_ *key = g.keys.data() + base;
        val_t *value = g.values.data() + base;
        // compact_kernel_cpu, gpma.cuh:602-614
        size_type compacted = 0;
        for (size_type i = 0; i < width; i++) {
            if (key[i] == KEY_NONE || value[i] == VALUE_NONE) continue;
            tk[compacted] = key[i];
            tv[compacted] = value[i];
            compacted++;
        }
        size_type a = offset[u], b = offset[u + 1];
        size_type interval = b - a;
        size_type merge = compacted + interval;
        if (lower <= merge && merge <= upper) {
            success++;
            for (size_type i = 0; i < interval; i++) {
                tk[compacted + i] = uk[a + i];
                tv[compacted + i] = uv[a + i];
            }
            for (size_type i = a; i < b; i++) nodes[i] = SIZE_NONE;
            // thrust::sort_by_key(thrust::host, ...) on unique keys
            order.resize(merge);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return tk[x] < tk[y]; });
            std::vector<key_t_> sk(merge);
            std::vector<val_t> sv(merge);
            for (size_type i = 0; i < merge; i++) {
                sk[i] = tk[order[i]];
                sv[i] = tv[order[i]];
            }
            std::fill(key, key + width, KEY_NONE);
            for (size_type i = 0; i < merge; i++)
                redispatch_one(i, sk.data(), sv.data(), key, value, width, S, merge, g.row_offset.data(), base);
        }
    }
    g.level_success[level] += success;
}

// gpma.cuh:881-904
void up_level_batch(std::vector<size_type> &nodes, size_type n, size_type width) {
    for (size_type i = 0; i < n; i++) nodes[i] &= ~width;
}

// gpma.cuh:906-934: compact the whole array (stable), then size it to the
// smallest power of two >= 4 whose floor(0.84 * size) holds the live items
// plus `increased`, and derive S and h from the size.
size_type resize_gpma(Gpma &g, size_type increased) {
    g.resizes++;
    size_type w = 0;
    const size_t C = g.keys.size();
    for (size_t i = 0; i < C; i++) {
        if (g.keys[i] == KEY_NONE || g.values[i] == VALUE_NONE) continue;
        g.keys[w] = g.keys[i];
        g.values[w] = g.values[i];
        w++;
    }
    size_type compacted = w;
    std::fill(g.keys.begin() + compacted, g.keys.end(), KEY_NONE);
    size_type merge = compacted + increased;
    size_type tree_size = 4;
    while (std::floor(0.84 * tree_size) < merge) tree_size *= 2;
    g.segment_length = 1u << (fls(fls(tree_size)) - 1);
    g.tree_height = fls(tree_size / g.segment_length) - 1;
    // thrust resize: new key slots are KEY_NONE, new value slots are 0.
    g.keys.resize(tree_size, KEY_NONE);
    g.values.resize(tree_size, 0);
    recalculate_density(g);
    return compacted;
}

// gpma.cuh:936-966: bulk path for >= 5,000,000 pending inserts. Note the
// reference redistributes into value-initialised temporaries, so every gap
// of the new array holds key 0 with value 0 (not KEY_NONE).
void significant_insert(Gpma &g, const std::vector<key_t_> &uk, const std::vector<val_t> &uv, size_type n) {
    g.bulk_loads++;
    size_type valid = resize_gpma(g, n);
    for (size_type i = 0; i < n; i++) {
        g.keys[valid + i] = uk[i];
        g.values[valid + i] = uv[i];
    }
    size_type merge = valid + n;
    std::vector<size_t> order(merge);
    std::iota(order.begin(), order.end(), 0);
    stable_sort_order(order, [&](size_t x, size_t y) { return g.keys[x] < g.keys[y]; });
    std::vector<key_t_> sk(merge);
    std::vector<val_t> sv(merge);
    for (size_type i = 0; i < merge; i++) {
        sk[i] = g.keys[order[i]];
        sv[i] = g.values[order[i]];
    }
    size_type C = (size_type)g.keys.size();
    std::vector<key_t_> nk(C, 0);
    std::vector<val_t> nv(C, 0);
    for (size_type i = 0; i < merge; i++)
        redispatch_one(i, sk.data(), sv.data(), nk.data(), nv.data(), C, g.segment_length, merge, g.row_offset.data(), 0);
    g.keys.swap(nk);
    g.values.swap(nv);
}

// gpma.cuh:968-1050
void update_gpma(Gpma &g, std::vector<key_t_> uk, std::vector<val_t> uv) {
    size_type n = (size_type)uk.size();
    // step 1: sort_by_key (stable for the radix sort thrust uses on u64 keys)
    {
        std::vector<size_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        stable_sort_order(order, [&](size_t x, size_t y) { return uk[x] < uk[y]; });
        std::vector<key_t_> k2(n);
        std::vector<val_t> v2(n);
        for (size_type i = 0; i < n; i++) {
            k2[i] = uk[order[i]];
            v2[i] = uv[order[i]];
        }
        uk.swap(k2);
        uv.swap(v2);
    }
    // step 2: locate + in-place delete / modify
    std::vector<size_type> nodes(n);
    locate_leaf_batch(g, uk.data(), uv.data(), n, nodes.data());
    // step 3: extract insertions, group by leaf
    size_type update_size = n;
    compact_insertions(nodes, uk, uv, update_size);
    std::vector<size_type> uniq, offset;
    size_type nuniq = group_insertion_by_node(nodes, update_size, uniq, offset);
    // step 4: bulk path
    if (update_size >= 5u * 1000u * 1000u) {
        significant_insert(g, uk, uv, update_size);
        return;
    }
    // step 5: level loop
    for (size_type level = 0; level <= g.tree_height && update_size; level++) {
        size_type lo = g.lower_element[level], hi = g.upper_element[level];
        rebalance_batch(g, level, nodes, uk, uv, uniq, offset, nuniq, lo, hi);
        compact_insertions(nodes, uk, uv, update_size);
        up_level_batch(nodes, update_size, g.segment_length << level);
        nuniq = group_insertion_by_node(nodes, update_size, uniq, offset);
    }
    // step 6: grow / shrink and rebalance the root
    if (update_size > 0) {
        size_type live = resize_gpma(g, update_size);
        size_type level = g.tree_height;
        size_type lo = g.lower_element[level], hi = g.upper_element[level];
        if (g_fix_root) {
            const size_type merge = live + update_size;
            while (merge > hi) {
                size_type T = (size_type)g.keys.size() * 2;
                g.segment_length = 1u << (fls(fls(T)) - 1);
                g.tree_height = fls(T / g.segment_length) - 1;
                g.keys.resize(T, KEY_NONE);
                g.values.resize(T, 0);
                recalculate_density(g);
                level = g.tree_height;
                lo = g.lower_element[level];
                hi = g.upper_element[level];
            }
            if (merge < lo) lo = 0;
        }
        rebalance_batch(g, level, nodes, uk, uv, uniq, offset, nuniq, lo, hi);
        // Pending updates that still fail here are silently lost by the
        // reference (gpma.cuh:1043 is not checked); count them.
        uint64_t left = 0;
        for (size_type i = 0; i < update_size; i++)
            if (nodes[i] != SIZE_NONE) left++;
        g.dropped += left;
    }
}

// gpma.cuh:54-82: the 4-slot seed tree, then one update of all row guards.
Gpma *create(size_type row_num) {
    Gpma *g = new Gpma();
    g->row_num = row_num;
    g->row_offset.assign((size_t)row_num + 1, 0);
    g->keys.assign(4, KEY_NONE);
    g->values.assign(4, 0);
    g->segment_length = 2;
    g->tree_height = 1;
    g->keys[0] = g->keys[2] = KEY_MAX;
    g->values[0] = g->values[2] = 1;
    recalculate_density(*g);
    std::vector<key_t_> guards(row_num);
    for (size_type v = 0; v < row_num; v++) guards[v] = ((key_t_)v << 32) + COL_IDX_NONE;
    std::vector<val_t> ones(row_num, 1);
    update_gpma(*g, guards, ones);
    return g;
}

}  // namespace

extern "C" {

void *orc_create(uint32_t row_num) { return create(row_num); }

// Threads the parallel loops use (1 without OpenMP).
int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_fix_root(int on) { g_fix_root = on != 0; }
void orc_destroy(void *h) { delete (Gpma *)h; }

int orc_update(void *h, const uint64_t *keys, const uint8_t *values, uint64_t n) {
    Gpma &g = *(Gpma *)h;
    std::vector<key_t_> k(keys, keys + n);
    std::vector<val_t> v(values, values + n);
    update_gpma(g, std::move(k), std::move(v));
    return 0;
}

// info[0]=capacity, [1]=segment_length, [2]=tree_height, [3]=row_num,
// [4]=dropped, [5]=resizes, [6]=bulk loads
void orc_info(void *h, uint64_t *info) {
    Gpma &g = *(Gpma *)h;
    info[0] = g.keys.size();
    info[1] = g.segment_length;
    info[2] = g.tree_height;
    info[3] = g.row_num;
    info[4] = g.dropped;
    info[5] = g.resizes;
    info[6] = g.bulk_loads;
}

// per-level cumulative window statistics: out[3*l+0..2] = tried, succeeded, slots
uint32_t orc_level_stats(void *h, uint64_t *out, uint32_t cap_levels) {
    Gpma &g = *(Gpma *)h;
    uint32_t L = (uint32_t)g.level_windows.size();
    for (uint32_t l = 0; l < L && l < cap_levels; l++) {
        out[3 * l] = g.level_windows[l];
        out[3 * l + 1] = g.level_success[l];
        out[3 * l + 2] = g.level_slots[l];
    }
    return L;
}

void orc_thresholds(void *h, uint32_t *lower, uint32_t *upper) {
    Gpma &g = *(Gpma *)h;
    for (size_t i = 0; i < g.lower_element.size(); i++) {
        lower[i] = g.lower_element[i];
        upper[i] = g.upper_element[i];
    }
}

void orc_layout(void *h, uint64_t *keys, uint8_t *values, uint32_t *row_offset) {
    Gpma &g = *(Gpma *)h;
    std::memcpy(keys, g.keys.data(), g.keys.size() * 8);
    std::memcpy(values, g.values.data(), g.values.size());
    std::memcpy(row_offset, g.row_offset.data(), g.row_offset.size() * 4);
}

// sorted live keys: every slot whose key is not KEY_NONE and whose value is
// not 0 (guards and the two KEY_MAX sentinels included)
uint64_t orc_live(void *h, uint64_t *out) {
    Gpma &g = *(Gpma *)h;
    uint64_t n = 0;
    for (size_t i = 0; i < g.keys.size(); i++) {
        if (g.keys[i] == KEY_NONE || g.values[i] == VALUE_NONE) continue;
        if (out) out[n] = g.keys[i];
        n++;
    }
    return n;
}

// gpma_bfs.cuh:446-509 semantics on one device: level[start] = 1, level l+1
// for nodes first reached from level l over live slots of
// [row_offset[v], row_offset[v+1]); guards, gaps and tombstones skipped.
// Returns the reachable count (gpma_bfs_demo.cu:80).
uint32_t orc_bfs(void *h, uint32_t start, uint32_t *levels) {
    Gpma &g = *(Gpma *)h;
    const uint32_t N = g.row_num;
    std::fill(levels, levels + N, 0u);
    if (start >= N) return 0;
    std::vector<uint32_t> q;
    q.reserve(N);
    q.push_back(start);
    levels[start] = 1;
    uint32_t reach = 1;
    for (size_t head = 0; head < q.size(); head++) {
        uint32_t v = q[head];
        uint32_t lv = levels[v];
        for (uint32_t s = g.row_offset[v]; s < g.row_offset[v + 1]; s++) {
            uint32_t dst = (uint32_t)(g.keys[s] & COL_IDX_NONE);
            if (dst == 0xFFFFFFFFu || g.values[s] == VALUE_NONE) continue;
            if (levels[dst] == 0) {
                levels[dst] = lv + 1;
                q.push_back(dst);
                reach++;
            }
        }
    }
    return reach;
}

// Textbook BFS over an explicit edge list (cpu_baseline/algorithms/bfs.hpp:7-29
// level convention), independent of any GPMA layout.
uint32_t orc_bfs_edges(uint32_t N, const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t start,
                       uint32_t *levels) {
    std::vector<uint64_t> off((size_t)N + 1, 0);
    for (uint64_t e = 0; e < E; e++) off[src[e] + 1]++;
    for (uint32_t v = 0; v < N; v++) off[v + 1] += off[v];
    std::vector<uint32_t> adj(E);
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (uint64_t e = 0; e < E; e++) adj[pos[src[e]]++] = dst[e];
    std::fill(levels, levels + N, 0u);
    if (start >= N) return 0;
    std::vector<uint32_t> q;
    q.reserve(N);
    q.push_back(start);
    levels[start] = 1;
    uint32_t reach = 1;
    for (size_t head = 0; head < q.size(); head++) {
        uint32_t v = q[head];
        for (uint64_t p = off[v]; p < off[v + 1]; p++) {
            uint32_t d = adj[p];
            if (levels[d] == 0) {
                levels[d] = levels[v] + 1;
                q.push_back(d);
                reach++;
            }
        }
    }
    return reach;
}

// Textbook BFS (cpu_baseline/algorithms/bfs.hpp:7-29 semantics) from each of
// nstart sources over one adjacency built once: levels[i * N + v]. Returns
// the sum of the reachable counts.
uint64_t orc_bfs_edges_multi(uint32_t N, const uint32_t *src, const uint32_t *dst, uint64_t E,
                             const uint32_t *starts, uint32_t nstart, uint32_t *levels) {
    std::vector<uint64_t> off((size_t)N + 1, 0);
    for (uint64_t e = 0; e < E; e++) off[src[e] + 1]++;
    for (uint32_t v = 0; v < N; v++) off[v + 1] += off[v];
    std::vector<uint32_t> adj(E);
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (uint64_t e = 0; e < E; e++) adj[pos[src[e]]++] = dst[e];
    uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)nstart; i++) {
        uint32_t *lv = levels + (uint64_t)i * N;
        std::fill(lv, lv + N, 0u);
        const uint32_t start = starts[i];
        if (start >= N) continue;
        std::vector<uint32_t> q;
        q.reserve(1024);
        q.push_back(start);
        lv[start] = 1;
        for (size_t head = 0; head < q.size(); head++) {
            const uint32_t v = q[head];
            for (uint64_t p = off[v]; p < off[v + 1]; p++) {
                const uint32_t d = adj[p];
                if (lv[d] == 0) {
                    lv[d] = lv[v] + 1;
                    q.push_back(d);
                }
            }
        }
        total += q.size();
    }
    return total;
}

// Redistribution slot of item i (gpma.cuh:636-654), exposed for brute-force checks.
uint32_t orc_redispatch_slot(uint32_t width, uint32_t S, uint32_t m, uint32_t i) {
    uint64_t frac = width / S;
    uint64_t deno = m;
    uint32_t seg_idx = (uint32_t)(frac * i / deno);
    uint32_t seg_lane = (uint32_t)(frac * i % deno / frac);
    return seg_idx * S + seg_lane;
}

uint32_t orc_fls(uint32_t x) { return fls(x); }

// Threshold table for an arbitrary (S, h), gpma.cuh:147-168.
void orc_density_table(uint32_t S, uint32_t h, uint32_t *lower, uint32_t *upper) {
    Gpma g;
    g.segment_length = S;
    g.tree_height = h;
    recalculate_density(g);
    for (uint32_t i = 0; i <= h; i++) {
        lower[i] = g.lower_element[i];
        upper[i] = g.upper_element[i];
    }
}

}  // extern "C"
