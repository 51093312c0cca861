// ge_sort.hip -- stable LSD radix sort of packed 64-bit sort words
// (key << 28 | slot), used for the Morton order (morton.cpp:40-63) and the
// SPH cell order (neighborSearch.h:636-720), and of (63-bit key, slot) pairs
// for the 3D Morton order (morton.cpp:115-137).  Three kernels per 8-bit digit:
// per-tile digit counts, one-workgroup scan of the (tile, digit) table, and a
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
                                                              uint32_t *__restrict__ hist) {
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
    hist[(uint64_t)blockIdx.x * 256 + t] = cnt[t];  // tile-major table
}

// Turns the (tile, digit) count table into scatter offsets: offset of digit d
// in tile b = sum of all counts of smaller digits + counts of d in tiles < b.
// Thread (g, d) owns digit d over a quarter g of the tiles; every pass reads
// the table with coalesced, independent loads.
__global__ __launch_bounds__(kScanThreads) void k_scan_table(uint32_t *__restrict__ hist, uint32_t nb) {
    __shared__ uint32_t part[4][256];
    __shared__ uint32_t dsum[256];
    const uint32_t d = threadIdx.x & 255u;
    const uint32_t g = threadIdx.x >> 8;
    const uint32_t b0 = (uint32_t)((uint64_t)nb * g / 4), b1 = (uint32_t)((uint64_t)nb * (g + 1) / 4);
    uint32_t sum = 0;
    uint32_t b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = hist[(uint64_t)(b + k) * 256 + d];
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += v[k];
    }
    for (; b < b1; ++b) sum += hist[(uint64_t)b * 256 + d];
    part[g][d] = sum;
    __syncthreads();
    if (g == 0) {
        const uint32_t t0 = part[0][d], t1 = part[1][d], t2 = part[2][d], t3 = part[3][d];
        part[0][d] = 0;
        part[1][d] = t0;
        part[2][d] = t0 + t1;
        part[3][d] = t0 + t1 + t2;
        dsum[d] = t0 + t1 + t2 + t3;
    }
    __syncthreads();
    // exclusive scan of the 256 digit totals
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t v = 0;
        if (g == 0 && d >= off) v = dsum[d - off];
        __syncthreads();
        if (g == 0) dsum[d] += v;
        __syncthreads();
    }
    uint32_t run = (d == 0 ? 0u : dsum[d - 1]) + part[g][d];
    b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = hist[(uint64_t)(b + k) * 256 + d];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            hist[(uint64_t)(b + k) * 256 + d] = run;
            run += v[k];
        }
    }
    for (; b < b1; ++b) {
        const uint32_t v = hist[(uint64_t)b * 256 + d];
        hist[(uint64_t)b * 256 + d] = run;
        run += v;
    }
}

// KV: a 32-bit value travels with each key (the 3D sort, whose 63-bit keys
// leave no room for the slot in the word).
template <bool KV>
__global__ __launch_bounds__(kSortThreads) void k_scatter(const uint64_t *__restrict__ in,
                                                          uint64_t *__restrict__ out, uint32_t n,
                                                          int shift, int bits,
                                                          const uint32_t *__restrict__ offs,
                                                          const uint32_t *__restrict__ vin,
                                                          uint32_t *__restrict__ vout) {
    __shared__ uint32_t wave_cnt[kSortWaves][256];
    __shared__ uint32_t blk_off[256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t w = t >> 6;
    const uint32_t dmask = (1u << bits) - 1u;
#pragma unroll
    for (int k = 0; k < kSortWaves; ++k) wave_cnt[k][t] = 0;
    blk_off[t] = (t <= dmask) ? offs[(uint64_t)blockIdx.x * 256 + t] : 0u;
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wbase = (uint64_t)blockIdx.x * kTile + (uint64_t)w * 64 * kItems;
    uint64_t key[kItems];
    uint32_t rank[kItems];
    uint32_t dig[kItems];
    uint32_t val[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        uint64_t i = wbase + (uint64_t)k * 64 + lane;
        bool valid = i < n;
        key[k] = valid ? in[i] : 0ull;
        if (KV) val[k] = valid ? vin[i] : 0u;
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
            if (KV) vout[pos] = val[k];
        }
    }
}

}  // namespace

// Sorts n packed words by bits [lo_bit, hi_bit).  On return `data` points to
// the sorted array (the two buffers may have been swapped).
void radix_sort_words(hipStream_t stream, uint32_t *hist, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                      int hi_bit) {
    if (n <= 1) return;
    const uint32_t nb = blocks_for(n, kTile);
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        int bits = hi_bit - shift < 8 ? hi_bit - shift : 8;
        uint32_t dmask = (1u << bits) - 1u;
        hipLaunchKernelGGL(k_digit_count, dim3(nb), dim3(kSortThreads), 0, stream, data, n, shift, dmask, hist);
        hipLaunchKernelGGL(k_scan_table, dim3(1), dim3(kScanThreads), 0, stream, hist, nb);
        hipLaunchKernelGGL(k_scatter<false>, dim3(nb), dim3(kSortThreads), 0, stream, data, alt, n, shift, bits, hist,
                           nullptr, nullptr);
        uint64_t *t = data;
        data = alt;
        alt = t;
    }
}

void radix_sort_packed(ge_ctx *c, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                       int hi_bit) {
    radix_sort_words(c->stream, c->b.hist, data, alt, n, lo_bit, hi_bit);
}

// Sorts n (key, value) pairs by key bits [lo_bit, hi_bit), stably; on return
// `keys` / `vals` point to the sorted arrays.  `hist` holds
// radix_hist_entries(n) entries.
void radix_sort_kv(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
                   uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit) {
    if (n <= 1) return;
    const uint32_t nb = blocks_for(n, kTile);
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        const int bits = hi_bit - shift < 8 ? hi_bit - shift : 8;
        const uint32_t dmask = (1u << bits) - 1u;
        hipLaunchKernelGGL(k_digit_count, dim3(nb), dim3(kSortThreads), 0, stream, keys, n, shift, dmask, hist);
        hipLaunchKernelGGL(k_scan_table, dim3(1), dim3(kScanThreads), 0, stream, hist, nb);
        hipLaunchKernelGGL(k_scatter<true>, dim3(nb), dim3(kSortThreads), 0, stream, keys, keys_alt, n, shift, bits,
                           hist, vals, vals_alt);
        uint64_t *t = keys;
        keys = keys_alt;
        keys_alt = t;
        uint32_t *tv = vals;
        vals = vals_alt;
        vals_alt = tv;
    }
}

uint32_t radix_hist_entries(uint32_t cap) { return 256u * blocks_for(cap, kTile); }

}  // namespace ge
