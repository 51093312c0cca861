// ge_tree_common.hpp -- device helpers shared by the quadtree build
// (ge_tree.hip) and the octree build (ge_octree.hip): the leaf rule, the
// wave-parallel std::partition_point, the node-start table and its
// exclusive scan.  Included inside an anonymous namespace of each
// translation unit.
#pragma once

#include "ge_internal.hpp"

namespace ge {
namespace {

enum : uint32_t { kKindInternal = 0, kKindLeaf = 1, kKindWaveSub = 2, kKindPending = 3, kKindDead = 4 };

constexpr int kTreeBlock = 256;
constexpr int kTreeWaves = kTreeBlock / 64;
constexpr uint32_t kTreeGridMax = 8192;
constexpr uint32_t kWaveSubMax = 64;
constexpr uint32_t kFrontMax = 1024;  // LDS frontier of a workgroup subtree
constexpr uint32_t kBlockListMax = 2048;  // internal nodes a workgroup subtree lists for the mass pass
constexpr int kWaveLevels = 64;       // levels of a wave subtree

// leaf_cap 0: the reference's rule (quadtree.cpp:16-17).  leaf_cap 1..64: a
// node of at most leaf_cap particles is a leaf (autotuned trees; not in the
// reference), with the same 1e-6 size floor so coincident particles end.
__device__ __forceinline__ bool leaf_rule(uint32_t cnt, float size, uint32_t leaf_cap) {
    if (leaf_cap) return cnt <= leaf_cap || size <= 0.000001f;
    return ((cnt <= 4 && size <= 2.0f) || cnt == 1) || size <= 0.000001f;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ float readlane_f(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}


// The same halving search over predicate bits held in a 64-bit mask
// (bit k = predicate of particle k of a wave subtree).
__device__ __forceinline__ uint32_t mask_part_point(uint64_t m, uint32_t first, uint32_t last) {
    uint32_t len = last - first;
    if (len == 0) return first;
    // bits [first, last) already partitioned (ones, then zeros): the halving
    // search returns the first zero, i.e. first + the number of ones
    const uint64_t r = (m >> first) & (len >= 64 ? ~0ull : ((1ull << len) - 1ull));
    if ((r & (r + 1ull)) == 0ull) return first + (uint32_t)__popcll(r);
    while (len > 0) {
        uint32_t half = len >> 1;
        uint32_t middle = first + half;
        if ((m >> middle) & 1ull) {
            first = middle + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

// partition_point evaluated by a whole wave: per round, lane j (1..63) takes
// node j of the next six levels of the halving search's decision tree (heap
// order), loads its probe, and the ballot lets every lane follow the path the
// sequential search would take.  Uniform result, identical to part_point.
// `off`: v holds the elements from index off on (v[i - off] is element i), so
// a range staged in LDS is read without rebasing the pointer.
__device__ uint32_t wave_part_point(const float *__restrict__ v, uint32_t first, uint32_t last, float mid,
                                    uint32_t off = 0) {
    const uint32_t j = lane_id();
    uint32_t f = first, len = last - first;
    while (len > 0) {
        bool pred = false;
        if (j >= 1) {
            const int d = 31 - __clz(j);
            uint32_t fj = f, lj = len;
            bool ok = true;
            for (int k = d - 1; k >= 0; --k) {
                if (lj == 0) { ok = false; break; }
                const uint32_t half = lj >> 1;
                if ((j >> k) & 1u) { fj = fj + half + 1; lj = lj - half - 1; }
                else { lj = half; }
            }
            if (ok && lj > 0) pred = v[fj + (lj >> 1) - off] < mid;
        }
        const uint64_t bal = __ballot(pred);
        uint32_t node = 1;
        for (int k = 0; k < 6 && len > 0; ++k) {
            const uint32_t half = len >> 1;
            const uint32_t p = (uint32_t)((bal >> node) & 1ull);
            if (p) { f = f + half + 1; len = len - half - 1; }
            else { len = half; }
            node = 2 * node + p;
        }
    }
    return f;
}

// The same search by groups of G lanes (G = 32 or 16), each group over its
// own range [first, last) (uniform within the group): log2(G) halving levels
// per round.  Per group, identical to part_point.  Called with all lanes.
template <int G>
__device__ uint32_t group_part_point(const float *__restrict__ v, uint32_t first, uint32_t last, float mid,
                                     uint32_t off = 0) {
    constexpr int D = G == 32 ? 5 : 4;
    static_assert(G == 32 || G == 16, "group of 32 or 16 lanes");
    const uint32_t lane = lane_id();
    const uint32_t j = lane & (uint32_t)(G - 1);
    const uint32_t shift = lane & ~(uint32_t)(G - 1);
    uint32_t f = first, len = last - first;
    while (__ballot(len > 0)) {
        bool pred = false;
        if (j >= 1 && len > 0) {
            const int d = 31 - __clz(j);
            uint32_t fj = f, lj = len;
            bool ok = true;
            for (int k = d - 1; k >= 0; --k) {
                if (lj == 0) { ok = false; break; }
                const uint32_t half = lj >> 1;
                if ((j >> k) & 1u) { fj = fj + half + 1; lj = lj - half - 1; }
                else { lj = half; }
            }
            if (ok && lj > 0) pred = v[fj + (lj >> 1) - off] < mid;
        }
        const uint64_t gb = __ballot(pred) >> shift;
        uint32_t node = 1;
        for (int k = 0; k < D && len > 0; ++k) {
            const uint32_t half = len >> 1;
            const uint32_t p = (uint32_t)((gb >> node) & 1ull);
            if (p) { f = f + half + 1; len = len - half - 1; }
            else { len = half; }
            node = 2 * node + p;
        }
    }
    return f;
}

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ uint64_t rfl64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// bits strictly above `lane`
__device__ __forceinline__ uint64_t above(uint32_t lane) { return lane >= 63 ? 0ull : (~0ull << (lane + 1)); }

__device__ __forceinline__ float shfl_f(float v, uint32_t src) { return __shfl(v, (int)src, 64); }


struct StartTable {
    uint32_t *count;    // C, then its exclusive prefix P (n + 1 entries)
    uint32_t *level;    // F
    uint32_t leaf_cap;  // leaf rule selector (see leaf_rule)
};

__device__ __forceinline__ void register_start(const StartTable &st, uint32_t a, uint32_t level) {
    atomicAdd(&st.count[a], 1u);
    atomicMin(&st.level[a], level);
}

// ---- exclusive scan of the start counts over n + 1 entries
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t &block_total) {
    // wave inclusive scan, then across the 16 waves
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= (uint32_t)d) incl += u;
    }
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < kScanThreads / 64; ++k) {
            const uint32_t t = sh[k];
            sh[k] = run;
            run += t;
        }
        sh[kScanThreads / 64] = run;
    }
    __syncthreads();
    block_total = sh[kScanThreads / 64];
    const uint32_t r = sh[w] + incl - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t *__restrict__ c, uint32_t len,
                                                            uint32_t *__restrict__ part) {
    __shared__ uint32_t sh[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; ++k) v += (base + k < len) ? c[base + k] : 0u;
    uint32_t tot;
    block_excl_scan(v, sh, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_parts(uint32_t *__restrict__ part, uint32_t nparts,
                                                             uint32_t *__restrict__ counters, uint32_t node_cap,
                                                             uint32_t top_cap) {
    __shared__ uint32_t sh[kScanThreads / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nparts; b0 += kScanThreads) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, sh, tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        // carry = total number of nodes
        counters[kCntNodeTotal] = carry <= node_cap ? carry : node_cap;
        if (carry > node_cap) counters[kCntError] |= kErrNodeCap;
        counters[kCntTopTotal] = min(counters[kCntTopAlloc], top_cap);
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t *__restrict__ c, uint32_t len,
                                                             const uint32_t *__restrict__ part) {
    __shared__ uint32_t sh[kScanThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < len) ? c[base + k] : 0u;
        sum += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(sum, sh, tot) + part[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < len) c[base + k] = run;
        run += v[k];
    }
}

uint32_t wave_grid(uint64_t waves) {
    uint64_t blocks = (waves + kTreeWaves - 1) / kTreeWaves;
    if (blocks < 1) blocks = 1;
    if (blocks > kTreeGridMax) blocks = kTreeGridMax;
    return (uint32_t)blocks;
}

}  // namespace
}  // namespace ge
