// ge_sort.hip -- stable LSD radix sort of packed 64-bit sort words
// (key << 28 | slot), used for the Morton order (morton.cpp:40-63) and the
// SPH cell order (neighborSearch.h:636-720), and of (63-bit key, slot) pairs
// for the 3D Morton order (morton.cpp:115-137).  Because the slot sits in the
// low bits and LSD passes are stable, ties keep ascending slot order.
//
// 9-bit digits (four passes over a 36-bit Morton key).  One read of the keys
// counts the digits of every pass (a thread counts runs of equal digits over
// its consecutive keys, so nearly sorted input, the usual case frame to
// frame, costs few LDS atomics) and adds the workgroup's counts to per-pass
// totals; one small kernel turns the totals into each pass's digit bases (a
// permutation does not change a digit's total).  Each pass is then one
// launch: a workgroup takes the next tile from a ticket counter, ranks its
// keys with wave ballots, sorts the tile by digit in LDS, publishes its digit
// counts and finds the counts of the tiles before it by decoupled look-back
// (a tile only waits on tiles whose tickets were taken earlier, and those
// publish before they wait, so every wait ends; each wait is also bounded by
// a spin limit that raises an error flag), then writes each digit's run of
// the LDS-sorted tile to its place, so consecutive lanes write consecutive
// words.  Look-back statuses are 32-bit words (two flag bits, a 30-bit
// count), cleared by the counting kernel, which runs first: the look-back
// is bound by the bytes of its coherent round trips, which a generation tag
// in each status would double.
#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kDigitBits = 9;
constexpr int kRadix = 1 << kDigitBits;
constexpr int kOsThreads = kRadix;                 // one thread per digit in the look-back
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsItems = 8;
constexpr int kOsTile = kOsThreads * kOsItems;
constexpr int kOsPasses = (64 + kDigitBits - 1) / kDigitBits;
constexpr uint32_t kFlagAgg = 1u << 30;            // status: this tile's own count
constexpr uint32_t kFlagInc = 2u << 30;            // status: count of tiles 0..this
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kSpinMax = 1u << 20;            // polls of one look-back wait
constexpr int kLookWindow = 8;                     // predecessor tiles read per look-back round trip
constexpr uint32_t kOneSweepMax = 1u << 30;        // published counts are 30 bits

// hist layout (uint32 units): digit totals [pass][digit], digit bases
// [pass][digit] and tickets [pass] at fixed offsets (a buffer serves sorts
// of any length), then the look-back statuses [pass][tile][digit]
constexpr uint32_t kSortHeader = (2u * kOsPasses * kRadix + kOsPasses + 63u) & ~63u;

struct SortMem {
    uint32_t *totals, *bases, *tickets;
    uint32_t *status;
};

__host__ __device__ inline SortMem sort_mem(uint32_t *hist) {
    SortMem m;
    m.totals = hist;
    m.bases = m.totals + kOsPasses * kRadix;
    m.tickets = m.bases + kOsPasses * kRadix;
    m.status = hist + kSortHeader;
    return m;
}

__device__ __forceinline__ uint32_t ld_status(const uint32_t *a) {
    return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint32_t *a, uint32_t v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit counts of every pass: each thread takes kOsItems consecutive keys and
// adds one LDS count per run of equal digits; the workgroup then adds its
// non-zero counts to the per-pass totals.  Workgroup b also clears tile b's
// look-back statuses of every pass.
__global__ __launch_bounds__(kOsThreads) void k_digit_totals(const uint64_t *__restrict__ in, uint32_t n,
                                                             int lo_bit, int npass, uint32_t last_mask,
                                                             uint32_t *__restrict__ totals,
                                                             uint32_t *__restrict__ status, uint32_t nb) {
    __shared__ uint32_t cnt[kOsPasses][kRadix];
    const uint32_t t = threadIdx.x;
    for (int p = 0; p < npass; ++p) st_status(status + ((uint64_t)p * nb + blockIdx.x) * kRadix + t, 0u);
    for (uint32_t e = t; e < kOsPasses * kRadix; e += kOsThreads) (&cnt[0][0])[e] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kOsTile + (uint64_t)t * kOsItems;
    uint64_t w[kOsItems];
#pragma unroll
    for (int k = 0; k < kOsItems; ++k) w[k] = base + k < n ? in[base + k] >> lo_bit : 0ull;
    const uint32_t valid = base >= n ? 0u : (uint32_t)min((uint64_t)kOsItems, n - base);
    for (int p = 0; p < npass; ++p) {
        const uint32_t dm = p == npass - 1 ? last_mask : (uint32_t)(kRadix - 1);
        uint32_t cur = 0, run = 0;
#pragma unroll
        for (int k = 0; k < kOsItems; ++k) {
            if ((uint32_t)k < valid) {
                const uint32_t d = (uint32_t)(w[k] >> (kDigitBits * p)) & dm;
                if (run && d != cur) {
                    atomicAdd(&cnt[p][cur], run);
                    run = 0;
                }
                cur = d;
                ++run;
            }
        }
        if (run) atomicAdd(&cnt[p][cur], run);
    }
    __syncthreads();
    for (int p = 0; p < npass; ++p) {
        const uint32_t v = cnt[p][t];
        if (v) atomicAdd(&totals[p * kRadix + t], v);
    }
}

// exclusive scan of one value per thread over the workgroup (kOsThreads)
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t *sh) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= (uint32_t)d) incl += u;
    }
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int k = 0; k < kOsWaves; ++k) before += (uint32_t)k < w ? sh[k] : 0u;
    __syncthreads();
    return before + incl - v;
}

// One workgroup: each pass's digit bases from its totals; the totals are
// cleared for the next sort and the tickets reset.
__global__ __launch_bounds__(kOsThreads) void k_digit_bases(uint32_t *__restrict__ hist, int npass) {
    __shared__ uint32_t sh[kOsWaves];
    const SortMem m = sort_mem(hist);
    const uint32_t t = threadIdx.x;
    for (int p = 0; p < npass; ++p) {
        const uint32_t v = m.totals[p * kRadix + t];
        m.bases[p * kRadix + t] = wg_excl_scan(v, sh);
        m.totals[p * kRadix + t] = 0u;
    }
    if (t < (uint32_t)kOsPasses) m.tickets[t] = 0u;
}

// One pass: ticketed tile, ballot ranks, LDS sort of the tile by digit,
// look-back over the earlier tiles' published counts, run-wise write-out.
template <bool KV>
__global__ __launch_bounds__(kOsThreads) void k_onesweep(const uint64_t *__restrict__ in,
                                                           uint64_t *__restrict__ out, uint32_t n, int shift,
                                                           int bits, int pass, uint32_t *__restrict__ hist,
                                                           uint32_t nb, uint32_t *__restrict__ err,
                                                           const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ vout) {
    __shared__ uint32_t wave_cnt[kOsWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];  // first LDS slot of each digit in the sorted tile
    __shared__ uint32_t blk_off[kRadix];
    __shared__ uint64_t skey[kOsTile];
    __shared__ uint32_t sval[KV ? kOsTile : 1];
    __shared__ uint32_t sh[kOsWaves];
    __shared__ uint32_t tile_s;
    const SortMem m = sort_mem(hist);
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t w = t >> 6;
    const uint32_t dmask = (1u << bits) - 1u;
    if (t == 0) tile_s = atomicAdd(m.tickets + pass, 1u);
    for (uint32_t e = t; e < kOsWaves * kRadix; e += kOsThreads) (&wave_cnt[0][0])[e] = 0;
    __syncthreads();
    const uint32_t tile = tile_s;

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wbase = (uint64_t)tile * kOsTile + (uint64_t)w * 64 * kOsItems;
    uint64_t key[kOsItems];
    uint32_t rank[kOsItems];
    uint32_t dig[kOsItems];
    uint32_t val[kOsItems];
#pragma unroll
    for (int k = 0; k < kOsItems; ++k) {
        const uint64_t i = wbase + (uint64_t)k * 64 + lane;
        key[k] = i < n ? in[i] : 0ull;
        if (KV) val[k] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kOsItems; ++k) {
        const bool valid = wbase + (uint64_t)k * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[k] >> shift) & dmask;
        dig[k] = d;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t base = valid ? wave_cnt[w][d] : 0u;
        const bool leader = valid && ((peers & lt_mask) == 0ull);
        if (leader) wave_cnt[w][d] = base + (uint32_t)__popcll(peers);
        rank[k] = base + before;
        if (!valid) dig[k] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // thread t owns digit t: the tile's count, the waves' exclusive offsets
    uint32_t count = 0;
#pragma unroll
    for (int k = 0; k < kOsWaves; ++k) {
        const uint32_t c = wave_cnt[k][t];
        wave_cnt[k][t] = count;
        count += c;
    }
    uint32_t *mine = m.status + ((uint64_t)pass * nb + tile) * kRadix + t;
    const bool live = t <= dmask;
    if (live && tile > 0) st_status(mine, kFlagAgg | count);
    // the tile sorted by digit in LDS (while the predecessors publish)
    dstart[t] = wg_excl_scan(count, sh);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kOsItems; ++k) {
        const uint32_t d = dig[k];
        if (d != 0xFFFFFFFFu) {
            const uint32_t slot = dstart[d] + wave_cnt[w][d] + rank[k];
            skey[slot] = key[k];
            if (KV) sval[slot] = val[k];
        }
    }
    // look back kLookWindow tiles per round trip; a tile not yet published
    // (status still cleared) ends the round and is read again
    uint32_t excl = 0;
    if (live && tile > 0) {
        int64_t j = (int64_t)tile - 1;
        uint32_t spins = 0;
        const uint32_t *col = m.status + (uint64_t)pass * nb * kRadix + t;
        while (j >= 0) {
            uint32_t v[kLookWindow];
#pragma unroll
            for (int k = 0; k < kLookWindow; ++k)
                v[k] = (j - k >= 0) ? ld_status(col + (uint64_t)(j - k) * kRadix) : kFlagInc;
            int k = 0;
            bool inc = false;
            for (; k < kLookWindow && j - k >= 0; ++k) {
                const uint32_t lo = v[k];
                const uint32_t f = lo & ~kValMask;
                if (f == 0u) break;
                excl += lo & kValMask;
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
    if (live) st_status(mine, kFlagInc | (excl + count));
    blk_off[t] = live ? m.bases[pass * kRadix + t] + excl : 0u;
    __syncthreads();
    const uint64_t t0 = (uint64_t)tile * kOsTile;
    const uint32_t nvalid = t0 >= n ? 0u : (uint32_t)min((uint64_t)kOsTile, n - t0);
#pragma unroll
    for (int k = 0; k < kOsItems; ++k) {
        const uint32_t i = t + (uint32_t)k * kOsThreads;
        if (i < nvalid) {
            const uint64_t kk = skey[i];
            const uint32_t d = (uint32_t)(kk >> shift) & dmask;
            const uint32_t pos = blk_off[d] + (i - dstart[d]);
            out[pos] = kk;
            if (KV) vout[pos] = sval[i];
        }
    }
}

__global__ void k_sort_refuse(uint32_t *err) { atomicOr(err, kErrSort); }

// The passes over bits [lo_bit, hi_bit): `keys` (and `vals` with KV) end up
// sorted.  hist holds radix_hist_entries(n) entries, zeroed once when
// allocated.
template <bool KV>
void onesweep(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
              uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit, uint32_t *err) {
    if (n >= kOneSweepMax) {  // beyond every context's particle cap
        hipLaunchKernelGGL(k_sort_refuse, dim3(1), dim3(1), 0, stream, err);
        return;
    }
    const uint32_t nb = blocks_for(n, kOsTile);
    const int npass = (hi_bit - lo_bit + kDigitBits - 1) / kDigitBits;
    const int last_bits = hi_bit - (lo_bit + kDigitBits * (npass - 1));
    const SortMem m = sort_mem(hist);
    hipLaunchKernelGGL(k_digit_totals, dim3(nb), dim3(kOsThreads), 0, stream, keys, n, lo_bit, npass,
                       (1u << last_bits) - 1u, m.totals, m.status, nb);
    hipLaunchKernelGGL(k_digit_bases, dim3(1), dim3(kOsThreads), 0, stream, hist, npass);
    for (int p = 0; p < npass; ++p) {
        const int shift = lo_bit + kDigitBits * p;
        const int bits = p == npass - 1 ? last_bits : kDigitBits;
        hipLaunchKernelGGL(k_onesweep<KV>, dim3(nb), dim3(kOsThreads), 0, stream, keys, keys_alt, n, shift, bits, p,
                           hist, nb, err, vals, vals_alt);
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
    uint32_t *none = nullptr, *none2 = nullptr;
    onesweep<false>(stream, hist, data, alt, none, none2, n, lo_bit, hi_bit, err);
}

void radix_sort_packed(ge_ctx *c, uint64_t *&data, uint64_t *&alt, uint32_t n, int lo_bit,
                       int hi_bit) {
    radix_sort_words(c->stream, c->b.hist, data, alt, n, lo_bit, hi_bit, c->b.counters + kCntError);
}

// Sorts n (key, value) pairs by key bits [lo_bit, hi_bit), stably; on return
// `keys` / `vals` point to the sorted arrays.
void radix_sort_kv(hipStream_t stream, uint32_t *hist, uint64_t *&keys, uint64_t *&keys_alt, uint32_t *&vals,
                   uint32_t *&vals_alt, uint32_t n, int lo_bit, int hi_bit, uint32_t *err) {
    if (n <= 1 || hi_bit <= lo_bit) return;
    onesweep<true>(stream, hist, keys, keys_alt, vals, vals_alt, n, lo_bit, hi_bit, err);
}

// header and statuses; hipMemset to zero once after allocation (the totals
// must start at zero)
uint32_t radix_hist_entries(uint32_t cap) {
    const uint64_t nb = blocks_for(cap, kOsTile);
    return (uint32_t)(kSortHeader + (uint64_t)kOsPasses * nb * kRadix);
}

}  // namespace ge
