// ge_sort.hip -- stable LSD radix sort of packed 64-bit sort words
// (key << 28 | slot), used for the Morton order (morton.cpp:40-63) and the
// SPH cell order (neighborSearch.h:636-720).  Three kernels per 8-bit digit:
// per-tile digit counts, one-workgroup scan of the (digit, tile) table, and a
// stable scatter ranked with wave ballots.  Because the slot sits in the low
// bits and LSD passes are stable, ties keep ascending slot order.
#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kSortThreads * kItems;
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kSortThreads) void k_digit_count(const uint64_t *__restrict__ in,
                                                              uint32_t n, int shift, uint32_t dmask,
                                                              uint32_t *__restrict__ hist,
                                                              uint32_t nb) {
    __shared__ uint32_t cnt[256];
    const uint32_t t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        uint64_t i = base + (uint64_t)k * kSortThreads + t;
        if (i < n) {
            uint32_t d = (uint32_t)(in[i] >> shift) & dmask;
            atomicAdd(&cnt[d], 1u);
        }
    }
    __syncthreads();
    // digit-major layout so one linear exclusive scan yields scatter offsets
    hist[(uint64_t)t * nb + blockIdx.x] = cnt[t];
}

// Exclusive scan of len entries by one workgroup: each thread owns a
// contiguous chunk; chunk sums are scanned across the workgroup.
__global__ __launch_bounds__(kScanThreads) void k_scan_table(uint32_t *__restrict__ data, uint32_t len) {
    __shared__ uint32_t sums[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (len + kScanThreads - 1) / kScanThreads;
    const uint32_t lo = t * chunk;
    const uint32_t hi = min(len, lo + chunk);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += data[i];
    sums[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over the chunk sums
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        uint32_t v = t >= off ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    uint32_t run = t == 0 ? 0u : sums[t - 1];
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(const uint64_t *__restrict__ in,
                                                          uint64_t *__restrict__ out, uint32_t n,
                                                          int shift, int bits,
                                                          const uint32_t *__restrict__ offs,
                                                          uint32_t nb) {
    __shared__ uint32_t wave_cnt[kSortWaves][256];
    __shared__ uint32_t blk_off[256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t w = t >> 6;
    const uint32_t dmask = (1u << bits) - 1u;
#pragma unroll
    for (int k = 0; k < kSortWaves; ++k) wave_cnt[k][t] = 0;
    blk_off[t] = (t <= dmask) ? offs[(uint64_t)t * nb + blockIdx.x] : 0u;
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wbase = (uint64_t)blockIdx.x * kTile + (uint64_t)w * 64 * kItems;
    uint64_t key[kItems];
    uint32_t rank[kItems];
    uint32_t dig[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        uint64_t i = wbase + (uint64_t)k * 64 + lane;
        bool valid = i < n;
        key[k] = valid ? in[i] : 0ull;
        uint32_t d = (uint32_t)(key[k] >> shift) & dmask;
        dig[k] = d;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        uint32_t base = valid ? wave_cnt[w][d] : 0u;
        bool leader = valid && ((peers & lt_mask) == 0ull);
        if (leader) wave_cnt[w][d] = base + (uint32_t)__popcll(peers);
        rank[k] = base + before;
        if (!valid) dig[k] = 0xFFFFFFFFu;
    }
    __syncthreads();
    {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kSortWaves; ++k) {
            uint32_t c = wave_cnt[k][t];
            wave_cnt[k][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        uint32_t d = dig[k];
        if (d != 0xFFFFFFFFu) {
            uint32_t pos = blk_off[d] + wave_cnt[w][d] + rank[k];
            out[pos] = key[k];
        }
    }
}

}  // namespace

// Sorts n packed words by bits [lo_bit, hi_bit).  On return `data` points to
// the sorted array (the two buffers may have been swapped).
void radix_sort_packed(ge_ctx *c, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                       int hi_bit) {
    if (n <= 1) return;
    const uint32_t nb = blocks_for(n, kTile);
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        int bits = hi_bit - shift < 8 ? hi_bit - shift : 8;
        uint32_t dmask = (1u << bits) - 1u;
        hipLaunchKernelGGL(k_digit_count, dim3(nb), dim3(kSortThreads), 0, c->stream, data, n, shift,
                           dmask, c->b.hist, nb);
        hipLaunchKernelGGL(k_scan_table, dim3(1), dim3(kScanThreads), 0, c->stream, c->b.hist,
                           (uint32_t)(256u * nb));
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kSortThreads), 0, c->stream, data, alt, n, shift,
                           bits, c->b.hist, nb);
        uint64_t *t = data;
        data = alt;
        alt = t;
    }
}

uint32_t radix_hist_entries(uint32_t cap) { return 256u * blocks_for(cap, kTile); }

}  // namespace ge
