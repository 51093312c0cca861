// ge_constraints.hip -- Physics::constraints (GE/src/Physics/physics.cpp:768-950)
// on the device (include/ge_constraints.h).
//
// Per solve:
//   1. ids -> particle indices: the (id, index) words of the particles are
//      radix-sorted and each constraint binary-searches its two ids
//      (idToIndexTable, physics.cpp:777-787);
//   2. constraints with a missing particle or already broken are removed,
//      order kept (std::remove_if, physics.cpp:805-817), by a block-count /
//      one-workgroup-scan / scatter compaction;
//   3. the frame counter advances and every fifth frame the triangle rule
//      breaks constraints whose ends share no neighbour (physics.cpp:819-862):
//      the adjacency is the sorted (end, other end) words, each constraint
//      merges its two sorted neighbour lists;
//   4. 15 substeps; each substep runs the colours in turn, one launch per
//      colour, so no two constraints of a launch touch one particle and the
//      updates need no atomics.
// Colouring (Jones-Plassmann): constraints are adjacent when they share a
// particle; in round k an uncoloured constraint takes colour k when its
// (hash, index) priority beats every neighbour still uncoloured at the start
// of the round.  It runs after constraints are uploaded; removals keep a
// colouring valid.  Every step but the colouring's rounds runs without a
// host round trip: counts live on the device and the launches are sized by
// the host's upper bound.
#include <cstring>
#include <vector>

#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kCBlock = 256;
constexpr uint32_t kMaxColors = 64;

__global__ void k_set_u32(uint32_t *__restrict__ p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

struct CView {  // the particle fields the solver reads and writes
    float *x, *y, *vx, *vy, *ax, *ay, *m;
    const uint32_t *flags, *label;
};

// ---- 1. id map ------------------------------------------------------------

__global__ __launch_bounds__(kCBlock) void k_id_words(const uint32_t *__restrict__ id, uint32_t n,
                                                      uint64_t *__restrict__ w) {
    const uint32_t i = blockIdx.x * kCBlock + threadIdx.x;
    if (i < n) w[i] = ((uint64_t)id[i] << 32) | i;
}

__device__ __forceinline__ uint32_t find_id(const uint64_t *__restrict__ w, uint32_t n, uint32_t id) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)(w[mid] >> 32) < id) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && (uint32_t)(w[lo] >> 32) == id) ? (uint32_t)w[lo] : kNone;
}

// the index of each end, and whether the constraint stays (both ends exist,
// not broken)
__global__ __launch_bounds__(kCBlock) void k_map_ids(const ge_constraint *__restrict__ cons,
                                                     const uint32_t *__restrict__ count, const uint64_t *__restrict__ w,
                                                     uint32_t n, uint2 *__restrict__ cidx, uint32_t *__restrict__ keep) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= *count) return;
    const ge_constraint k = cons[c];
    const uint32_t a = find_id(w, n, k.id1), b = find_id(w, n, k.id2);
    cidx[c] = make_uint2(a, b);
    keep[c] = (a != kNone && b != kNone && !k.is_broken) ? 1u : 0u;
}

// ---- 2. order-keeping compaction -------------------------------------------

__global__ __launch_bounds__(kCBlock) void k_block_count(const uint32_t *__restrict__ keep,
                                                         const uint32_t *__restrict__ count,
                                                         uint32_t *__restrict__ block_cnt) {
    __shared__ uint32_t ws[kCBlock / 64];
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    const bool k = c < *count && keep[c];
    const uint64_t m = __ballot(k);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < kCBlock / 64; ++q) t += ws[q];
        block_cnt[blockIdx.x] = t;
    }
}

// exclusive scan of nb entries in place, total in a[nb]; one workgroup
__global__ __launch_bounds__(1024) void k_scan1(uint32_t nb, uint32_t *__restrict__ a) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < nb ? a[j] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const uint32_t add = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (j < nb) a[j] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[nb] = carry;
}

__global__ __launch_bounds__(kCBlock) void k_compact(const ge_constraint *__restrict__ src,
                                                     const uint2 *__restrict__ isrc, const uint32_t *__restrict__ csrc,
                                                     const uint32_t *__restrict__ keep, const uint32_t *__restrict__ count,
                                                     const uint32_t *__restrict__ block_off, ge_constraint *__restrict__ dst,
                                                     uint2 *__restrict__ idst, uint32_t *__restrict__ cdst) {
    __shared__ uint32_t ws[kCBlock / 64];
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    const bool k = c < *count && keep[c];
    const uint64_t m = __ballot(k);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (!k) return;
    uint32_t o = block_off[blockIdx.x];
    for (uint32_t q = 0; q < w; ++q) o += ws[q];
    o += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    dst[o] = src[c];
    idst[o] = isrc[c];
    cdst[o] = csrc[c];
}

__global__ void k_set_count(uint32_t *__restrict__ count, const uint32_t *__restrict__ total,
                            uint32_t *__restrict__ frame) {
    if (threadIdx.x != 0) return;
    const uint32_t t = *total;
    *count = t;
    if (t > 0) *frame = *frame + 1u;  // myVar.frameCount++ (physics.cpp:821)
}

// ---- 3. triangle rule -------------------------------------------------------

// (end, other end) words of every constraint, both directions
__global__ __launch_bounds__(kCBlock) void k_adj_words(const uint2 *__restrict__ cidx,
                                                       const uint32_t *__restrict__ count, uint32_t bound,
                                                       uint64_t *__restrict__ w) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= bound) return;
    const uint32_t m = *count;
    if (c < m) {
        const uint2 e = cidx[c];
        w[2 * c] = ((uint64_t)e.x << 32) | e.y;
        w[2 * c + 1] = ((uint64_t)e.y << 32) | e.x;
    } else {  // past the live constraints: sort to the end
        w[2 * c] = ~0ull;
        w[2 * c + 1] = ~0ull;
    }
}

__device__ __forceinline__ uint32_t lower_hi(const uint64_t *__restrict__ w, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)(w[mid] >> 32) < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// a constraint is part of a triangle when its ends share a neighbour
// (physics.cpp:838-858); only on frames where the counter is a multiple of 5
__global__ __launch_bounds__(kCBlock) void k_triangles(ge_constraint *__restrict__ cons, const uint2 *__restrict__ cidx,
                                                       const uint32_t *__restrict__ count,
                                                       const uint32_t *__restrict__ frame,
                                                       const uint64_t *__restrict__ w, uint32_t nw) {
    if (*frame % 5u != 0u) return;
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= *count) return;
    if (cons[c].is_broken) return;
    const uint2 e = cidx[c];
    uint32_t i = lower_hi(w, nw, e.x), j = lower_hi(w, nw, e.y);
    bool tri = false;
    while (i < nw && j < nw && (uint32_t)(w[i] >> 32) == e.x && (uint32_t)(w[j] >> 32) == e.y) {
        const uint32_t a = (uint32_t)w[i], b = (uint32_t)w[j];
        if (a == b) {
            tri = true;
            break;
        }
        if (a < b) ++i;
        else ++j;
    }
    if (!tri) cons[c].is_broken = 1;
}

// ---- colouring ----------------------------------------------------------------

__device__ __forceinline__ uint32_t prio(uint32_t c) {  // a fixed hash of the index
    uint32_t h = c * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h;
}

// incidence words (particle, constraint) of both ends
__global__ __launch_bounds__(kCBlock) void k_inc_words(const uint2 *__restrict__ cidx, uint32_t m,
                                                       uint64_t *__restrict__ w, uint32_t *__restrict__ color,
                                                       uint32_t *__restrict__ rnd) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= m) return;
    const uint2 e = cidx[c];
    w[2 * c] = ((uint64_t)e.x << 32) | c;
    w[2 * c + 1] = ((uint64_t)e.y << 32) | c;
    color[c] = kNone;
    rnd[c] = kNone;
}

__device__ __forceinline__ bool beats(uint32_t c, uint32_t d) {  // c's priority above d's
    const uint32_t pc = prio(c), pd = prio(d);
    return pc > pd || (pc == pd && c > d);
}

// Round `round` of the colouring: an uncoloured constraint whose priority
// beats every neighbour still uncoloured at the start of the round (rnd[d]
// unset or set in this round) joins the round's independent set and takes
// the smallest colour none of its earlier-coloured neighbours holds (so at
// most max degree + 1 colours).
__global__ __launch_bounds__(kCBlock) void k_color_round(const uint2 *__restrict__ cidx, uint32_t m,
                                                         const uint64_t *__restrict__ w, uint32_t nw, uint32_t round,
                                                         uint32_t *__restrict__ rnd, uint32_t *__restrict__ color,
                                                         uint32_t *__restrict__ left) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= m || rnd[c] != kNone) return;
    const uint2 e = cidx[c];
    bool top = true;
    uint64_t used = 0;  // colours of neighbours coloured in earlier rounds
    for (int s = 0; s < 2 && top; ++s) {
        const uint32_t p = s ? e.y : e.x;
        for (uint32_t k = lower_hi(w, nw, p); k < nw && (uint32_t)(w[k] >> 32) == p; ++k) {
            const uint32_t d = (uint32_t)w[k];
            if (d == c) continue;
            const uint32_t rd = rnd[d];
            if (rd == kNone || rd == round) {
                if (beats(d, c)) {
                    top = false;
                    break;
                }
            } else if (color[d] < 64u) {
                used |= 1ull << color[d];
            }
        }
    }
    if (top) {
        rnd[c] = round;
        // the smallest free colour; 64 (an error) when all 64 are taken
        color[c] = ~used ? (uint32_t)__builtin_ctzll(~used) : 64u;
    } else {
        atomicAdd(left, 1u);
    }
}

__global__ __launch_bounds__(kCBlock) void k_max_color(const uint32_t *__restrict__ color, uint32_t m,
                                                       uint32_t *__restrict__ top) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c < m) atomicMax(top, color[c]);
}

// (colour, index) words
__global__ __launch_bounds__(kCBlock) void k_color_words(const uint32_t *__restrict__ color,
                                                         const uint32_t *__restrict__ count, uint32_t bound,
                                                         uint64_t *__restrict__ w) {
    const uint32_t c = blockIdx.x * kCBlock + threadIdx.x;
    if (c >= bound) return;
    w[c] = c < *count ? (((uint64_t)color[c] << 32) | c) : ~0ull;
}

// order[k] = constraint of sorted word k; start[q] = first word of colour q
__global__ __launch_bounds__(kCBlock) void k_color_table(const uint64_t *__restrict__ w,
                                                         const uint32_t *__restrict__ count, uint32_t bound,
                                                         uint32_t ncol, uint32_t *__restrict__ order,
                                                         uint32_t *__restrict__ start) {
    const uint32_t k = blockIdx.x * kCBlock + threadIdx.x;
    const uint32_t m = *count;
    if (k < m) order[k] = (uint32_t)w[k];
    if (k <= ncol) start[k] = k == ncol ? m : lower_hi(w, m, k);
    (void)bound;
}

// ---- 4. one colour of one substep (physics.cpp:867-946) ----------------------

struct CParams {
    float stiff_mult, res_mult, damping, ratio;
    int unbreakable, forces;
    int periodic;
    float dw, dh;
};

__device__ __forceinline__ const ge_material *find_mat(const ge_material *mats, uint32_t nmat, uint32_t id) {
    for (uint32_t k = 0; k < nmat; ++k)
        if (mats[k].id == id) return &mats[k];
    return nullptr;
}

__global__ __launch_bounds__(kCBlock) void k_cons_color(ge_constraint *__restrict__ cons,
                                                        const uint2 *__restrict__ cidx,
                                                        const uint32_t *__restrict__ order,
                                                        const uint32_t *__restrict__ start, uint32_t q, CView s,
                                                        const ge_material *__restrict__ mats, uint32_t nmat,
                                                        CParams P) {
#pragma clang fp contract(off)
    const uint32_t lo = start[q], hi = start[q + 1];
    const uint32_t k = lo + blockIdx.x * kCBlock + threadIdx.x;
    if (k >= hi) return;
    const uint32_t c = order[k];
    ge_constraint K = cons[c];
    if (K.is_broken) return;
    const uint2 e = cidx[c];
    if (e.x == kNone || e.y == kNone) {
        cons[c].is_broken = 1;
        return;
    }
    const uint32_t i = e.x, j = e.y;
    const ge_material *mi = s.label ? find_mat(mats, nmat, s.label[i]) : nullptr;
    const ge_material *mj = s.label ? find_mat(mats, nmat, s.label[j]) : nullptr;
    float dx = s.x[j] - s.x[i], dy = s.y[j] - s.y[i];
    if (P.periodic) {  // physics.cpp:893-896
        dx = fmodf(dx + P.dw * 1.5f, P.dw) - P.dw * 0.5f;
        dy = fmodf(dy + P.dh * 1.5f, P.dh) - P.dh * 0.5f;
    }
    const float len = sqrtf(dx * dx + dy * dy);
    if (len < 0.0001f) return;
    const float dirx = dx / len, diry = dy / len;
    K.displacement = len - K.rest_length;
    if (!P.unbreakable && mi && mj) {
        if (!mi->is_plastic || !mj->is_plastic) {
            K.is_plastic = 0;
            const float limit = (K.resistance * P.res_mult) * K.rest_length;
            if (fabsf(K.displacement) >= limit) K.is_broken = 1;
        } else {
            K.is_plastic = 1;
            if (fabsf(K.displacement) >= K.plasticity_point * K.original_length) K.rest_length += K.displacement;
            const float brk = (K.original_length + K.original_length * (K.resistance * P.res_mult)) *
                              (mi->constraint_plastic_point_mult + mj->constraint_plastic_point_mult) * 0.5f;
            if (fabsf(K.rest_length) >= brk) K.is_broken = 1;
        }
        if ((s.flags[i] | s.flags[j]) & GE_FLAG_HOT_POINT) K.is_broken = 1;
    }
    if (P.forces && !K.is_broken) {
        const float mi_ = s.m[i], mj_ = s.m[j];
        const float sd = K.stiffness * K.displacement;
        const float sfx = ((sd * dirx) * mi_) * P.stiff_mult, sfy = ((sd * diry) * mi_) * P.stiff_mult;
        const float rvx = s.vx[j] - s.vx[i], rvy = s.vy[j] - s.vy[i];
        const float dd = -P.damping * (rvx * dirx + rvy * diry);
        const float dfx = (dd * dirx) * mi_, dfy = (dd * diry) * mi_;
        const float tx = sfx + dfx, ty = sfy + dfy;
        s.ax[i] = s.ax[i] + tx / mi_;
        s.ay[i] = s.ay[i] + ty / mi_;
        s.ax[j] = s.ax[j] - tx / mj_;
        s.ay[j] = s.ay[j] - ty / mj_;
        const float cf = K.stiffness * P.ratio * P.stiff_mult;
        const float cx = (dirx * K.displacement) * cf, cy = (diry * K.displacement) * cf;
        const float ms = mi_ + mj_;
        const float wi = mj_ / ms, wj = mi_ / ms;
        s.x[i] = s.x[i] + cx * wi;
        s.y[i] = s.y[i] + cy * wi;
        s.x[j] = s.x[j] - cx * wj;
        s.y[j] = s.y[j] - cy * wj;
    }
    cons[c] = K;
}

int failc(ge_ctx *c, int code, const char *msg) {
    c->err = msg;
    return code;
}

CView view_of(const State &s) {
    return CView{s.x, s.y, s.vx, s.vy, s.ax, s.ay, s.m, s.flags, s.label};
}

// incidence words, sorted (by particle)
void incidence(ge_ctx *c, uint32_t m) {
    Buffers &b = c->b;
    hipLaunchKernelGGL(k_inc_words, dim3(blocks_for(m, kCBlock)), dim3(kCBlock), 0, c->stream, b.cidx, m, b.cw_a,
                       b.ccolor, b.ckeep);  // the keep flags are free until the next solve maps ids again
    radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, 2 * m, 32, 64, b.counters + kCntError);
}

}  // namespace

// Jones-Plassmann colouring of the current constraints (needs their index
// map); host loop over rounds, one counter read per round
int cons_color(ge_ctx *c) {
    Buffers &b = c->b;
    uint32_t m = 0;
    if (hipMemcpyAsync(&m, b.ccount, sizeof m, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return failc(c, GE_E_HIP, "constraint count copy failed");
    c->cons_bound = m;
    c->ncolors = 0;
    if (m == 0) {
        c->cons_colored = true;
        return GE_OK;
    }
    incidence(c, m);
    for (uint32_t round = 0;; ++round) {
        if (round >= 4096) return failc(c, GE_E_CAPACITY, "constraint colouring did not finish");
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, b.cleft, 0u);
        hipLaunchKernelGGL(k_color_round, dim3(blocks_for(m, kCBlock)), dim3(kCBlock), 0, c->stream, b.cidx, m,
                           b.cw_a, 2 * m, round, b.ckeep, b.ccolor, b.cleft);
        uint32_t left = 0;
        if (hipMemcpyAsync(&left, b.cleft, sizeof left, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return failc(c, GE_E_HIP, "colouring counter copy failed");
        if (left == 0) break;
    }
    // the number of colours: one past the largest
    hipLaunchKernelGGL(k_max_color, dim3(blocks_for(m, kCBlock)), dim3(kCBlock), 0, c->stream, b.ccolor, m, b.cleft);
    uint32_t top = 0;
    if (hipMemcpyAsync(&top, b.cleft, sizeof top, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return failc(c, GE_E_HIP, "colour count copy failed");
    if (top >= kMaxColors) return failc(c, GE_E_CAPACITY, "constraint graph needs more than 64 colours");
    c->ncolors = top + 1;
    c->cons_colored = true;
    return GE_OK;
}

// Physics::constraints once
int cons_solve(ge_ctx *c, const ge_step_params *p, bool map_done) {
    Buffers &b = c->b;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n, bound = c->cons_bound;
    if (bound == 0 || n == 0) return GE_OK;
    const uint32_t nbc = blocks_for(bound, kCBlock);
    // 1. ids -> indices, keep flags
    if (!map_done) {
        hipLaunchKernelGGL(k_id_words, dim3(blocks_for(n, kCBlock)), dim3(kCBlock), 0, c->stream, s.id, n, b.cw_a);
        radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, n, 32, 64, b.counters + kCntError);
        hipLaunchKernelGGL(k_map_ids, dim3(nbc), dim3(kCBlock), 0, c->stream, b.cons, b.ccount, b.cw_a, n, b.cidx,
                           b.ckeep);
    }
    // 2. remove, order kept
    hipLaunchKernelGGL(k_block_count, dim3(nbc), dim3(kCBlock), 0, c->stream, b.ckeep, b.ccount, b.cblock);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, c->stream, nbc, b.cblock);
    hipLaunchKernelGGL(k_compact, dim3(nbc), dim3(kCBlock), 0, c->stream, b.cons, b.cidx, b.ccolor, b.ckeep, b.ccount,
                       b.cblock, b.cons_tmp, b.cidx_tmp, b.ccolor_tmp);
    std::swap(b.cons, b.cons_tmp);
    std::swap(b.cidx, b.cidx_tmp);
    std::swap(b.ccolor, b.ccolor_tmp);
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, c->stream, b.ccount, b.cblock + nbc, b.cframe);
    // 3. triangle rule on every fifth counted frame
    hipLaunchKernelGGL(k_adj_words, dim3(nbc), dim3(kCBlock), 0, c->stream, b.cidx, b.ccount, bound, b.cw_a);
    radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, 2 * bound, 0, 64, b.counters + kCntError);
    hipLaunchKernelGGL(k_triangles, dim3(nbc), dim3(kCBlock), 0, c->stream, b.cons, b.cidx, b.ccount, b.cframe, b.cw_a,
                       2 * bound);
    // colour order of the remaining constraints
    hipLaunchKernelGGL(k_color_words, dim3(nbc), dim3(kCBlock), 0, c->stream, b.ccolor, b.ccount, bound, b.cw_a);
    radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, bound, 0, 40, b.counters + kCntError);
    hipLaunchKernelGGL(k_color_table, dim3(blocks_for((uint64_t)bound + kMaxColors + 1, kCBlock)), dim3(kCBlock), 0,
                       c->stream, b.cw_a, b.ccount, bound, c->ncolors, b.corder, b.cstart);
    // 4. substeps
    CParams P;
    P.stiff_mult = c->cons_p.stiffness_mult;
    P.res_mult = c->cons_p.resistance_mult;
    P.damping = c->cons_p.damping;
    P.ratio = c->cons_p.correction_ratio;
    P.unbreakable = c->cons_p.unbreakable;
    P.forces = c->cons_p.forces && p->dt > 0.0f;
    P.periodic = p->periodic;
    P.dw = p->domain_w;
    P.dh = p->domain_h;
    const CView v = view_of(s);
    const ge_material *mats = c->thermal ? b.materials : nullptr;
    const uint32_t nmat = c->thermal ? c->nmat : 0;
    for (int step = 0; step < 15; ++step)
        for (uint32_t q = 0; q < c->ncolors; ++q)
            hipLaunchKernelGGL(k_cons_color, dim3(nbc), dim3(kCBlock), 0, c->stream, b.cons, b.cidx, b.corder,
                               b.cstart, q, v, mats, nmat, P);
    return GE_OK;
}

// index map + colouring when constraints were added
int cons_prepare(ge_ctx *c) {
    if (c->cons_colored || c->cons_bound == 0 || c->n == 0) return GE_OK;
    Buffers &b = c->b;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n, bound = c->cons_bound;
    hipLaunchKernelGGL(k_id_words, dim3(blocks_for(n, kCBlock)), dim3(kCBlock), 0, c->stream, s.id, n, b.cw_a);
    radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, n, 32, 64, b.counters + kCntError);
    hipLaunchKernelGGL(k_map_ids, dim3(blocks_for(bound, kCBlock)), dim3(kCBlock), 0, c->stream, b.cons, b.ccount,
                       b.cw_a, n, b.cidx, b.ckeep);
    return cons_color(c);
}

}  // namespace ge

using namespace ge;

extern "C" {

void ge_default_constraint_params(ge_constraint_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->stiffness_mult = 1.0f;       // parameters.h:221-222
    p->resistance_mult = 1.0f;
    p->unbreakable = 0;             // parameters.h:230
    p->forces = 1;
    p->damping = 0.001f;            // physics.h:25
    p->correction_ratio = 0.013333f;  // physics.h:27
}

int ge_constraints_capacity(const ge_ctx *c, uint32_t *cap) {
    if (!c || !cap) return GE_E_INVALID;
    *cap = c->cons_cap;
    return GE_OK;
}

int ge_set_constraint_params(ge_ctx *c, const ge_constraint_params *p) {
    if (!c || !p) return GE_E_INVALID;
    c->cons_p = *p;
    return GE_OK;
}

int ge_constraints_upload(ge_ctx *c, const ge_constraint *k, uint32_t count) {
    if (!c || (count && !k)) return GE_E_INVALID;
    if (count > c->cons_cap) return failc(c, GE_E_CAPACITY, "more constraints than max_constraints");
    (void)hipSetDevice(c->device);
    if ((count && hipMemcpyAsync(c->b.cons, k, sizeof(ge_constraint) * count, hipMemcpyHostToDevice, c->stream) !=
                      hipSuccess) ||
        hipMemcpyAsync(c->b.ccount, &count, sizeof count, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return failc(c, GE_E_HIP, "constraint copy failed");
    c->cons_bound = count;
    c->cons_colored = false;
    c->layout_version += 1;  // a captured frame holds the colour launches
    return GE_OK;
}

int ge_constraints_download(ge_ctx *c, ge_constraint *out, uint32_t cap, uint32_t *count) {
    if (!c || !count || (cap && !out)) return GE_E_INVALID;
    if (c->cons_cap == 0) {
        *count = 0;
        return GE_OK;
    }
    (void)hipSetDevice(c->device);
    uint32_t m = 0;
    if (hipMemcpyAsync(&m, c->b.ccount, sizeof m, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return failc(c, GE_E_HIP, "constraint count copy failed");
    *count = m;
    const uint32_t w = m < cap ? m : cap;
    if (w && (hipMemcpyAsync(out, c->b.cons, sizeof(ge_constraint) * w, hipMemcpyDeviceToHost, c->stream) !=
                  hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
        return failc(c, GE_E_HIP, "constraint copy failed");
    return GE_OK;
}

int ge_constraints_clear(ge_ctx *c) { return ge_constraints_upload(c, nullptr, 0); }

int ge_constraints_solve(ge_ctx *c, const ge_step_params *p) {
    if (!c || !p) return GE_E_INVALID;
    if (c->cons_bound == 0 || c->n == 0) return GE_OK;
    if (c->nsys > 1) return failc(c, GE_E_UNSUPPORTED, "constraints run on single-system contexts");
    (void)hipSetDevice(c->device);
    if (int r = cons_prepare(c)) return r;
    if (int r = cons_solve(c, p, false)) return r;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return failc(c, GE_E_HIP, "constraint solve failed");
    return GE_OK;
}

int ge_create_constraints(ge_ctx *c, const ge_step_params *p, uint32_t *created) {
    if (!c || !p) return GE_E_INVALID;
    if (created) *created = 0;
    if (!c->b.ccreate) return failc(c, GE_E_INVALID, "needs a context with enable_sph and max_constraints > 0");
    if (c->nsys > 1) return failc(c, GE_E_UNSUPPORTED, "constraints run on single-system contexts");
    if (!(p->sph_cell > 0.0f && p->sph_search_radius > 0.0f && p->sph_search_radius <= 8.0f * p->sph_cell))
        return failc(c, GE_E_INVALID, "sph_cell and sph_search_radius must be > 0 (radius at most 8 cells)");
    if (c->n == 0) return GE_OK;
    (void)hipSetDevice(c->device);
    launch_integrate_start_bbox(c, 0.0f, 0, 0.0f, 0);  // the grid frame: the current extremes
    if (int r = sph_new_grid(c, p)) return r;
    uint32_t h = 0;
    if (int r = sph_create_constraints(c, p, &h)) return r;
    if (created) *created = h;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return failc(c, GE_E_HIP, "constraint creation failed");
    uint32_t e = 0;
    if (hipMemcpy(&e, c->b.counters + kCntError, sizeof e, hipMemcpyDeviceToHost) == hipSuccess && (e & kErrConsCap)) {
        uint32_t z = e & ~kErrConsCap;
        (void)hipMemcpy(c->b.counters + kCntError, &z, sizeof z, hipMemcpyHostToDevice);
        return failc(c, GE_E_CAPACITY, "created constraints beyond max_constraints (the excess was dropped)");
    }
    return GE_OK;
}

int ge_constraints_order(ge_ctx *c, uint32_t *order, uint32_t cap, uint32_t *color_start, uint32_t color_cap,
                         uint32_t *ncolors) {
    if (!c || !ncolors || (cap && !order) || (color_cap && !color_start)) return GE_E_INVALID;
    (void)hipSetDevice(c->device);
    if (int r = cons_prepare(c)) return r;
    *ncolors = c->ncolors;
    const uint32_t bound = c->cons_bound;
    if (bound == 0) return GE_OK;
    Buffers &b = c->b;
    const uint32_t nbc = blocks_for(bound, kCBlock);
    hipLaunchKernelGGL(k_color_words, dim3(nbc), dim3(kCBlock), 0, c->stream, b.ccolor, b.ccount, bound, b.cw_a);
    radix_sort_words(c->stream, b.chist, b.cw_a, b.cw_b, bound, 0, 40, b.counters + kCntError);
    hipLaunchKernelGGL(k_color_table, dim3(blocks_for((uint64_t)bound + kMaxColors + 1, kCBlock)), dim3(kCBlock), 0,
                       c->stream, b.cw_a, b.ccount, bound, c->ncolors, b.corder, b.cstart);
    const uint32_t w = bound < cap ? bound : cap;
    const uint32_t wc = c->ncolors + 1 < color_cap ? c->ncolors + 1 : color_cap;
    if ((w && hipMemcpyAsync(order, b.corder, sizeof(uint32_t) * w, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (wc && hipMemcpyAsync(color_start, b.cstart, sizeof(uint32_t) * wc, hipMemcpyDeviceToHost, c->stream) !=
                   hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return failc(c, GE_E_HIP, "order copy failed");
    return GE_OK;
}

}  // extern "C"
