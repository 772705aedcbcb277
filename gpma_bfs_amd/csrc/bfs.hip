// Frontier BFS over the gapped packed array.
//
// Reference semantics (gpma_bfs.cuh:149-395,446-509): level[start] = 1;
// each round expands every frontier vertex v over the slots
// [row_offset[v], row_offset[v+1]), skipping guards and gaps (low word
// 0xFFFFFFFF) and tombstones (value 0), and gives level l+1 to every
// destination not yet labelled. The reference materialises an edge queue
// (gather) and culls it with hashes, a non-atomic bitmap and a CAS on the
// level (contract); here one kernel per round expands and labels directly:
// a destination is claimed by the CAS on its level word, so each vertex
// enters exactly one frontier.
//
// Load balance: a workgroup takes 256 frontier vertices, scans their slot
// counts in LDS and walks the concatenated slot ranges 256 at a time; each
// lane finds its owner vertex by binary search over the scanned offsets, so
// consecutive lanes read consecutive slots of the same adjacency range and
// one hub's range is spread over the whole workgroup.
#include "common.h"
#include "kernels.h"

namespace gpma {

namespace {

constexpr uint32_t BFS_THREADS = 256;

__global__ void k_bfs_init(uint32_t *levels, uint32_t N, uint32_t start, uint32_t *frontier) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x)
        levels[i] = (i == start) ? 1u : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) frontier[0] = start;
}

__global__ __launch_bounds__(BFS_THREADS) void k_bfs_expand(const gkey *__restrict__ keys,
                                                            const gval *__restrict__ values,
                                                            const uint32_t *__restrict__ row_offset,
                                                            const uint32_t *__restrict__ front, uint32_t nf,
                                                            uint32_t *__restrict__ next, uint32_t *__restrict__ next_cnt,
                                                            uint32_t *__restrict__ levels, uint32_t level) {
    __shared__ uint32_t s_beg[BFS_THREADS];
    __shared__ uint32_t s_ofs[BFS_THREADS + 1];
    __shared__ uint32_t scratch[BFS_THREADS / WAVE + 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t newlev = level + 1;
    for (uint64_t tile = (uint64_t)blockIdx.x * BFS_THREADS; tile < nf; tile += (uint64_t)gridDim.x * BFS_THREADS) {
        const uint64_t fi = tile + tid;
        uint32_t b = 0, len = 0;
        if (fi < nf) {
            const uint32_t v = front[fi];
            b = row_offset[v];
            const uint32_t e = row_offset[v + 1];
            len = e > b ? e - b : 0;
        }
        uint32_t total;
        const uint32_t ofs = block_exclusive_scan(len, scratch, &total);
        s_beg[tid] = b;
        s_ofs[tid] = ofs;
        if (tid == 0) s_ofs[BFS_THREADS] = total;
        __syncthreads();
        for (uint32_t j0 = 0; j0 < total; j0 += BFS_THREADS) {
            const uint32_t j = j0 + tid;
            bool claimed = false;
            uint32_t dst = 0;
            if (j < total) {
                // owner: last t with s_ofs[t] <= j
                uint32_t lo = 0, hi = BFS_THREADS;
                while (hi - lo > 1) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (s_ofs[mid] <= j) lo = mid;
                    else hi = mid;
                }
                const uint64_t slot = (uint64_t)s_beg[lo] + (j - s_ofs[lo]);
                const gkey k = keys[slot];
                const gval v = values[slot];
                dst = (uint32_t)k;
                if (dst != COL_NONE && v != 0 && levels[dst] == 0) {
                    claimed = atomicCAS(&levels[dst], 0u, newlev) == 0u;
                }
            }
            const uint64_t m = ballot(claimed);
            if (m) {
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if (lane_id() == leader) base = atomicAdd(next_cnt, (uint32_t)__popcll(m));
                base = __shfl(base, leader, WAVE);
                if (claimed) next[base + (uint32_t)__popcll(m & lanemask_lt())] = dst;
            }
        }
        __syncthreads();
    }
}

}  // namespace

void launch_bfs_init(uint32_t *levels, uint32_t N, uint32_t start, uint32_t *frontier, hipStream_t st) {
    uint32_t g = (uint32_t)std::min<uint64_t>(2048, ((uint64_t)N + 255) / 256);
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_bfs_init, dim3(g), dim3(256), 0, st, levels, N, start, frontier);
}

void launch_bfs_expand(const gkey *keys, const gval *values, const uint32_t *row_offset, const uint32_t *front,
                       uint32_t nf, uint32_t *next, uint32_t *next_cnt, uint32_t *levels, uint32_t level,
                       hipStream_t st) {
    if (!nf) return;
    uint32_t g = (uint32_t)std::min<uint64_t>(4096, ((uint64_t)nf + BFS_THREADS - 1) / BFS_THREADS);
    hipLaunchKernelGGL(k_bfs_expand, dim3(g), dim3(BFS_THREADS), 0, st, keys, values, row_offset, front, nf, next,
                       next_cnt, levels, level);
}

}  // namespace gpma
