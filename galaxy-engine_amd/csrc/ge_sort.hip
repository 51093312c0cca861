// ge_sort.hip -- stable LSD radix sort of packed 64-bit sort words
// (key << 28 | slot), used for the Morton order (morton.cpp:40-63) and the
// SPH cell order (neighborSearch.h:636-720), and of (63-bit key, slot) pairs
// for the 3D Morton order (morton.cpp:115-137).  Because the slot sits in the
// low bits and LSD passes are stable, ties keep ascending slot order.
//
// One read of the keys counts the digits of every 8-bit pass per tile; one
// small kernel turns the counts into each pass's digit bases (a permutation
// does not change a digit's total).  Each pass is then one launch: a
// workgroup takes the next tile from a ticket counter, ranks its keys with
// wave ballots, publishes its digit counts and finds the counts of the tiles
// before it by decoupled look-back (a tile only waits on tiles whose tickets
// were taken earlier, and those publish before they wait, so every wait ends;
// each wait is also bounded by a spin limit that raises an error flag), then
// scatters.  Above 2^30 keys (the published counts are 30 bits) the sort uses
// three kernels per pass instead: per-tile digit counts, a one-workgroup scan
// of the (tile, digit) table and the same ranked scatter.
#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kSortThreads * kItems;
constexpr int kScanThreads = 1024;
constexpr int kMaxPasses = 8;                      // 64 key bits in 8-bit digits
constexpr uint32_t kFlagAgg = 1u << 30;            // status: this tile's own counts
constexpr uint32_t kFlagInc = 2u << 30;            // status: counts of tiles 0..this
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kSpinMax = 1u << 20;            // polls of one look-back wait
constexpr int kLookWindow = 8;                     // predecessor tiles read per look-back round trip
// tiles of the single-launch form (larger tiles shorten the look-back but
// leave most CUs idle at a million keys and crowd the LDS digit counters)
constexpr int kOsThreads = 256;
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsTile = kOsThreads * kItems;
constexpr uint32_t kOneSweepMax = 1u << 30;

__device__ __forceinline__ uint32_t ld_status(const uint32_t *a) {
    return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint32_t *a, uint32_t v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit counts of one tile for every pass at once; the table is pass-major:
// hist[(p * nb + tile) * 256 + d].
__global__ __launch_bounds__(kOsThreads) void k_digit_count_all(const uint64_t *__restrict__ in, uint32_t n,
                                                                int lo_bit, int npass, uint32_t last_mask,
                                                                uint32_t nb, uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[kMaxPasses][256];
    const uint32_t t = threadIdx.x;
    for (uint32_t e = t; e < kMaxPasses * 256; e += kOsThreads) (&cnt[0][0])[e] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kOsTile;
#pragma unroll 4
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kOsThreads + t;
        if (i < n) {
            const uint64_t w = in[i] >> lo_bit;
            for (int p = 0; p < npass; ++p) {
                const uint32_t d = (uint32_t)(w >> (8 * p)) & (p == npass - 1 ? last_mask : 0xFFu);
                atomicAdd(&cnt[p][d], 1u);
            }
        }
    }
    __syncthreads();
    if (t < 256)
        for (int p = 0; p < npass; ++p) hist[((uint64_t)p * nb + blockIdx.x) * 256 + t] = cnt[p][t];
}

// One workgroup per pass: the digit totals of the pass (summed over tiles by
// four thread groups), their exclusive scan into bases[p * 256 + d]; then the
// pass's table is cleared to serve as its look-back status, and its ticket
// counter is reset.
__global__ __launch_bounds__(kScanThreads) void k_digit_bases(uint32_t *__restrict__ hist, uint32_t nb,
                                                              uint32_t *__restrict__ bases,
                                                              uint32_t *__restrict__ tickets) {
    __shared__ uint32_t part[4][256];
    __shared__ uint32_t dsum[256];
    const uint32_t p = blockIdx.x;
    uint32_t *table = hist + (uint64_t)p * nb * 256;
    const uint32_t d = threadIdx.x & 255u;
    const uint32_t g = threadIdx.x >> 8;
    const uint32_t b0 = (uint32_t)((uint64_t)nb * g / 4), b1 = (uint32_t)((uint64_t)nb * (g + 1) / 4);
    uint32_t sum = 0;
    uint32_t b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = table[(uint64_t)(b + k) * 256 + d];
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += v[k];
    }
    for (; b < b1; ++b) sum += table[(uint64_t)b * 256 + d];
    part[g][d] = sum;
    __syncthreads();
    if (g == 0) dsum[d] = part[0][d] + part[1][d] + part[2][d] + part[3][d];
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        uint32_t v = 0;
        if (g == 0 && d >= off) v = dsum[d - off];
        __syncthreads();
        if (g == 0) dsum[d] += v;
        __syncthreads();
    }
    if (g == 0) bases[p * 256 + d] = d == 0 ? 0u : dsum[d - 1];
    // every thread clears the entries it read
    for (b = b0; b < b1; ++b) table[(uint64_t)b * 256 + d] = 0u;
    if (threadIdx.x == 0) tickets[p] = 0u;
}

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

// One pass of the single-launch form: ticketed tile, ballot ranks, look-back
// over the earlier tiles' published counts, scatter.
template <bool KV>
__global__ __launch_bounds__(kOsThreads) void k_onesweep(const uint64_t *__restrict__ in,
                                                           uint64_t *__restrict__ out, uint32_t n, int shift,
                                                           int bits, uint32_t *__restrict__ status,
                                                           const uint32_t *__restrict__ bases,
                                                           uint32_t *__restrict__ ticket, uint32_t *__restrict__ err,
                                                           const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ vout) {
    __shared__ uint32_t wave_cnt[kOsWaves][256];
    __shared__ uint32_t blk_off[256];
    __shared__ uint32_t tile_s;
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t w = t >> 6;
    const uint32_t dmask = (1u << bits) - 1u;
    if (t == 0) tile_s = atomicAdd(ticket, 1u);
    for (uint32_t e = t; e < kOsWaves * 256; e += kOsThreads) (&wave_cnt[0][0])[e] = 0;
    __syncthreads();
    const uint32_t tile = tile_s;

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wbase = (uint64_t)tile * kOsTile + (uint64_t)w * 64 * kItems;
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
    uint32_t count = 0;  // this tile's count of digit t (threads t < 256)
    if (t < 256) {
#pragma unroll
        for (int k = 0; k < kOsWaves; ++k) {
            const uint32_t c = wave_cnt[k][t];
            wave_cnt[k][t] = count;
            count += c;
        }
    }
    if (t <= dmask) {
        uint32_t excl = 0;
        uint32_t *mine = status + (uint64_t)tile * 256 + t;
        if (tile > 0) {
            st_status(mine, kFlagAgg | count);
            // look back kLookWindow tiles per round trip; a tile not yet
            // published ends the round and is read again
            int64_t j = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (j >= 0) {
                uint32_t v[kLookWindow];
#pragma unroll
                for (int k = 0; k < kLookWindow; ++k)
                    v[k] = (j - k >= 0) ? ld_status(status + (uint64_t)(j - k) * 256 + t) : kFlagInc;
                int k = 0;
                bool inc = false;
                for (; k < kLookWindow && j - k >= 0; ++k) {
                    const uint32_t f = v[k] & ~kValMask;
                    if (f == 0u) break;
                    excl += v[k] & kValMask;
                    if (f == kFlagInc) {
                        inc = true;
                        break;
                    }
                }
                if (inc) break;
                j -= k;
                if (k == 0) {  // the nearest unread tile has published nothing yet
                    if (++spins > kSpinMax) {  // never expected: flag it and stop waiting
                        atomicOr(err, kErrSort);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
        }
        st_status(mine, kFlagInc | (excl + count));
        blk_off[t] = bases[t] + excl;
    } else if (t < 256) {
        blk_off[t] = 0u;
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

// The single-launch passes over bits [lo_bit, hi_bit): `keys` (and `vals`
// with KV) end up sorted.  hist holds radix_hist_entries(n) entries.
template <bool KV>
void onesweep(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
              uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit, uint32_t *err) {
    const uint32_t nb = blocks_for(n, kOsTile);
    const int npass = (hi_bit - lo_bit + 7) / 8;
    const int last_bits = hi_bit - (lo_bit + 8 * (npass - 1));
    uint32_t *bases = hist + (uint64_t)kMaxPasses * 256 * nb;
    uint32_t *tickets = bases + kMaxPasses * 256;
    if (!err) err = tickets + kMaxPasses;
    hipLaunchKernelGGL(k_digit_count_all, dim3(nb), dim3(kOsThreads), 0, stream, keys, n, lo_bit, npass,
                       (1u << last_bits) - 1u, nb, hist);
    hipLaunchKernelGGL(k_digit_bases, dim3(npass), dim3(kScanThreads), 0, stream, hist, nb, bases, tickets);
    for (int p = 0; p < npass; ++p) {
        const int shift = lo_bit + 8 * p;
        const int bits = p == npass - 1 ? last_bits : 8;
        hipLaunchKernelGGL(k_onesweep<KV>, dim3(nb), dim3(kOsThreads), 0, stream, keys, keys_alt, n, shift, bits,
                           hist + (uint64_t)p * nb * 256, bases + p * 256, tickets + p, err, vals, vals_alt);
        uint64_t *t = keys;
        keys = keys_alt;
        keys_alt = t;
        if (KV) {
            uint32_t *tv = vals;
            vals = vals_alt;
            vals_alt = tv;
        }
    }
}

}  // namespace

// Sorts n packed words by bits [lo_bit, hi_bit).  On return `data` points to
// the sorted array (the two buffers may have been swapped).
void radix_sort_words(hipStream_t stream, uint32_t *hist, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                      int hi_bit, uint32_t *err) {
    if (n <= 1 || hi_bit <= lo_bit) return;
    if (n < kOneSweepMax) {
        uint32_t *none = nullptr, *none2 = nullptr;
        onesweep<false>(stream, hist, data, alt, none, none2, n, lo_bit, hi_bit, err);
        return;
    }
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
    radix_sort_words(c->stream, c->b.hist, data, alt, n, lo_bit, hi_bit, c->b.counters + kCntError);
}

// Sorts n (key, value) pairs by key bits [lo_bit, hi_bit), stably; on return
// `keys` / `vals` point to the sorted arrays.  `hist` holds
// radix_hist_entries(n) entries.
void radix_sort_kv(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
                   uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit, uint32_t *err) {
    if (n <= 1 || hi_bit <= lo_bit) return;
    if (n < kOneSweepMax) {
        onesweep<true>(stream, hist, keys, keys_alt, vals, vals_alt, n, lo_bit, hi_bit, err);
        return;
    }
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

// per-pass tile tables of the single-launch form, then its digit bases,
// ticket counters and error word (the three-kernel form uses the first table)
uint32_t radix_hist_entries(uint32_t cap) {
    const uint32_t one = kMaxPasses * 256u * blocks_for(cap, kOsTile) + kMaxPasses * 256u + 16u;
    const uint32_t three = 256u * blocks_for(cap, kTile);
    return one > three ? one : three;
}

}  // namespace ge
