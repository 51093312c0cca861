// ge_merger.hip -- NeighborSearch::UpdateNeighbors (GE/include/Particles/
// neighborSearch.h:48-200) and Physics::mergerSolver (GE/src/Physics/
// physics.cpp:954-1067) on the device, for ge_merge (ge_abi.hip).
//
// The reference merges greedily and serially: particle i (from the last to
// the first) takes the first neighbour, scanning its neighbour list
// backwards, that is not dark, not already merged away and overlaps it
// (distance <= totalRadius_i + totalRadius_n); the heavier absorbs the
// lighter.  Positions and radii do not change during the loop, so the
// "overlap graph" (neighbour-list pairs that overlap) is fixed; a merge only
// touches the two particles of one edge, so the greedy result inside one
// connected component of that graph depends on nothing outside it.  Here:
//   1. the grid of UpdateNeighbors over the non-dark particles (cells by the
//      reference's bounds, (int)((p - min) / cell) clamped), as sorted
//      (cell, index) words: within a cell ascending index, the order a
//      serial fill gives;
//   2. per particle, its neighbour count within densityRadius (the render
//      field `neighbors`) and its overlap edges in neighbour-list order;
//   3. connected components by a lock-free union-find (roots link to the
//      smaller index, so parent pointers only decrease);
//   4. one thread per component replays the serial loop over the
//      component's particles, descending index, edges backwards;
//   5. merged-away indices come back to the host, which replays the
//      swap-with-back deletions on indices and returns the permutation.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ge_internal.hpp"

namespace ge {
namespace {

constexpr int kMBlock = 256;
constexpr int kCellKeyBits = 36;  // cell keys below 2^36 - 1 leave room for 28 index bits

struct MGrid {
    float min_x, min_y, cell;
    int w, h;
};

__global__ __launch_bounds__(kMBlock) void k_m_radius(const float *__restrict__ size, uint32_t n, float half,
                                                      float mult, float *__restrict__ rad) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i < n) rad[i] = size[i] * half * mult;  // globalLogic.cpp:1225-1227
}

// min / max of the non-dark positions: per block, then one block
__global__ __launch_bounds__(kMBlock) void k_m_bounds(const float *__restrict__ x, const float *__restrict__ y,
                                                      const uint32_t *__restrict__ flags, uint32_t n,
                                                      float4 *__restrict__ part) {
    __shared__ float red[4][kMBlock / 64];
    float a = 3.4e38f, b = 3.4e38f, c = -3.4e38f, d = -3.4e38f;
    for (uint32_t i = blockIdx.x * kMBlock + threadIdx.x; i < n; i += gridDim.x * kMBlock) {
        if (flags[i] & GE_FLAG_DARK) continue;
        const float px = x[i], py = y[i];
        a = px < a ? px : a;
        b = py < b ? py : b;
        c = px > c ? px : c;
        d = py > d ? py : d;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float ta = __shfl_xor(a, o, 64), tb = __shfl_xor(b, o, 64), tc = __shfl_xor(c, o, 64),
                    td = __shfl_xor(d, o, 64);
        a = ta < a ? ta : a;
        b = tb < b ? tb : b;
        c = tc > c ? tc : c;
        d = td > d ? td : d;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; red[3][w] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kMBlock / 64; ++k) {
            a = red[0][k] < a ? red[0][k] : a;
            b = red[1][k] < b ? red[1][k] : b;
            c = red[2][k] > c ? red[2][k] : c;
            d = red[3][k] > d ? red[3][k] : d;
        }
        part[blockIdx.x] = make_float4(a, b, c, d);
    }
}

// calculateGridBounds (neighborSearch.h:21-46) from the partials
__global__ void k_m_grid(const float4 *__restrict__ part, uint32_t np, float cell, MGrid *__restrict__ g,
                         uint32_t *__restrict__ err) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    float a = 3.4e38f, b = 3.4e38f, c = -3.4e38f, d = -3.4e38f;
    for (uint32_t k = 0; k < np; ++k) {
        const float4 q = part[k];
        a = q.x < a ? q.x : a;
        b = q.y < b ? q.y : b;
        c = q.z > c ? q.z : c;
        d = q.w > d ? q.w : d;
    }
    a -= cell;
    b -= cell;
    c += cell;
    d += cell;
    const float fw = (c - a) / cell, fh = (d - b) / cell;
    MGrid r;
    r.min_x = a;
    r.min_y = b;
    r.cell = cell;
    r.w = 1;
    r.h = 1;
    if (!(fw < 2.0e9f && fh < 2.0e9f) || (double)((int)fw + 1) * (double)((int)fh + 1) >= 68719476735.0) {
        *err = 1u;  // the int cell counts or the 36-bit cell keys would overflow
    } else {
        r.w = max(1, (int)fw + 1);
        r.h = max(1, (int)fh + 1);
    }
    *g = r;
}

__device__ __forceinline__ void cell_of(const MGrid &g, float px, float py, int &cx, int &cy) {
#pragma clang fp contract(off)
    cx = (int)((px - g.min_x) / g.cell);
    cy = (int)((py - g.min_y) / g.cell);
    cx = max(0, min(cx, g.w - 1));
    cy = max(0, min(cy, g.h - 1));
}

__global__ __launch_bounds__(kMBlock) void k_m_words(const float *__restrict__ x, const float *__restrict__ y,
                                                     const uint32_t *__restrict__ flags, uint32_t n,
                                                     const MGrid *__restrict__ gp, uint64_t *__restrict__ w) {
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    if (flags[i] & GE_FLAG_DARK) {
        w[i] = ~0ull;  // past every cell
        return;
    }
    const MGrid g = *gp;
    int cx, cy;
    cell_of(g, x[i], y[i], cx, cy);
    const uint64_t key = (uint64_t)cy * (uint64_t)g.w + (uint64_t)cx;
    w[i] = (key << kIdxBits) | i;
}

__device__ __forceinline__ uint32_t lower_cell(const uint64_t *__restrict__ w, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((w[mid] >> kIdxBits) < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Visits the neighbour list of UpdateNeighbors in its order: cells
// (dy, dx) = (-1..1, -1..1), ascending index inside a cell, self skipped,
// distance^2 < densityRadius^2; f(j) for each entry.
template <class F>
__device__ __forceinline__ void for_neighbors(const float *__restrict__ x, const float *__restrict__ y,
                                              const MGrid &g, const uint64_t *__restrict__ w, uint32_t n, uint32_t i,
                                              float dr2, F &&f) {
#pragma clang fp contract(off)
    const float px = x[i], py = y[i];
    int cx, cy;
    cell_of(g, px, py, cx, cy);
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int a = cx + dx, b = cy + dy;
            if (a < 0 || a >= g.w || b < 0 || b >= g.h) continue;
            const uint64_t key = (uint64_t)b * (uint64_t)g.w + (uint64_t)a;
            for (uint32_t k = lower_cell(w, n, key); k < n && (w[k] >> kIdxBits) == key; ++k) {
                const uint32_t j = (uint32_t)(w[k] & kIdxMask);
                if (j == i) continue;
                const float ex = px - x[j], ey = py - y[j];
                if (ex * ex + ey * ey < dr2) f(j);
            }
        }
}

__device__ __forceinline__ bool overlaps(const float *__restrict__ x, const float *__restrict__ y,
                                         const float *__restrict__ rad, uint32_t i, uint32_t j) {
#pragma clang fp contract(off)
    const float dx = x[j] - x[i], dy = y[j] - y[i];
    const float cr = rad[i] + rad[j];
    return dx * dx + dy * dy <= cr * cr;
}

// pass 0: neighbour counts (render field) and overlap-edge counts; pass 1:
// the overlap edges in neighbour-list order
__global__ __launch_bounds__(kMBlock) void k_m_edges(const float *__restrict__ x, const float *__restrict__ y,
                                                     const uint32_t *__restrict__ flags, const float *__restrict__ rad,
                                                     uint32_t n, const MGrid *__restrict__ gp,
                                                     const uint64_t *__restrict__ w, float dr2, int pass,
                                                     int32_t *__restrict__ nbrs, uint32_t *__restrict__ ecnt,
                                                     const uint32_t *__restrict__ eoff, uint32_t *__restrict__ edges) {
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    if (flags[i] & GE_FLAG_DARK) {  // no list; the count stays (neighborSearch.h:115-120)
        if (pass == 0) ecnt[i] = 0;
        return;
    }
    const MGrid g = *gp;
    uint32_t cnt = 0, e = 0;
    const uint32_t base = pass ? eoff[i] : 0u;
    for_neighbors(x, y, g, w, n, i, dr2, [&](uint32_t j) {
        ++cnt;
        if ((flags[j] & GE_FLAG_DARK) || !overlaps(x, y, rad, i, j)) return;
        if (pass) edges[base + e] = j;
        ++e;
    });
    if (pass == 0) {
        if (nbrs) nbrs[i] = (int32_t)cnt;
        ecnt[i] = e;
    }
}

// exclusive scan of n counts into off (n + 1 entries), one workgroup
__global__ __launch_bounds__(1024) void k_m_scan(const uint32_t *__restrict__ cnt, uint32_t n,
                                                 uint32_t *__restrict__ off) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < n ? cnt[j] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const uint32_t add = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (j < n) off[j] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) off[n] = carry;
}

__device__ __forceinline__ uint32_t uf_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t uf_find(const uint32_t *parent, uint32_t x) {
    uint32_t p = uf_load(parent + x);
    while (p != x) {  // parent pointers only decrease, so this ends
        x = p;
        p = uf_load(parent + x);
    }
    return x;
}

__global__ __launch_bounds__(kMBlock) void k_m_uf_init(uint32_t *__restrict__ parent, uint32_t *__restrict__ del,
                                                       uint32_t n) {
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    parent[i] = i;
    del[i] = 0u;
}

// union of every edge's ends: the larger root links under the smaller
__global__ __launch_bounds__(kMBlock) void k_m_union(const uint32_t *__restrict__ eoff,
                                                     const uint32_t *__restrict__ edges, uint32_t n,
                                                     uint32_t *__restrict__ parent) {
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    for (uint32_t k = eoff[i]; k < eoff[i + 1]; ++k) {
        const uint32_t j = edges[k];
        if (j < i) continue;  // each undirected edge once
        for (;;) {
            uint32_t a = uf_find(parent, i), b = uf_find(parent, j);
            if (a == b) break;
            if (a < b) {
                const uint32_t t = a;
                a = b;
                b = t;
            }
            if (atomicCAS(parent + a, a, b) == a) break;  // a was still a root: link it
        }
    }
}

// (root, index) words of the particles that have overlap edges
__global__ __launch_bounds__(kMBlock) void k_m_comp_words(const uint32_t *__restrict__ eoff,
                                                          const uint32_t *__restrict__ parent, uint32_t n,
                                                          uint64_t *__restrict__ w) {
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n) return;
    w[i] = (eoff[i + 1] > eoff[i]) ? (((uint64_t)uf_find(parent, i) << 32) | i) : ~0ull;
}

// one thread per component: the serial loop (physics.cpp:963-1024) over its
// particles from the highest index down
__global__ __launch_bounds__(kMBlock) void k_m_merge(const uint64_t *__restrict__ w, uint32_t n,
                                                     const uint32_t *__restrict__ eoff,
                                                     const uint32_t *__restrict__ edges, float *__restrict__ m,
                                                     float *__restrict__ vx, float *__restrict__ vy,
                                                     float *__restrict__ prev_size, uint32_t *__restrict__ del) {
#pragma clang fp contract(off)
    const uint32_t k = blockIdx.x * kMBlock + threadIdx.x;
    if (k >= n || w[k] == ~0ull) return;
    const uint32_t root = (uint32_t)(w[k] >> 32);
    if (k > 0 && (uint32_t)(w[k - 1] >> 32) == root && w[k - 1] != ~0ull) return;  // not the first member
    uint32_t last = k;
    while (last + 1 < n && w[last + 1] != ~0ull && (uint32_t)(w[last + 1] >> 32) == root) ++last;
    for (uint32_t q = last + 1; q-- > k;) {
        const uint32_t i = (uint32_t)w[q];
        if (del[i]) continue;
        for (uint32_t e = eoff[i + 1]; e-- > eoff[i];) {
            const uint32_t j = edges[e];
            if (del[j]) continue;
            const float mP = m[i], mN = m[j];
            const float pi = prev_size[i], pj = prev_size[j];
            const float full = sqrtf(pi * pi + pj * pj);
            const float mx = pi < pj ? pj : pi;
            if (mP >= mN) {
                const float ms = mP + mN;
                m[i] = ms;
                vx[i] = (vx[i] * mP + vx[j] * mN) / ms;
                vy[i] = (vy[i] * mP + vy[j] * mN) / ms;
                prev_size[i] = mx + (full - mx) * 0.25f;
                del[j] = 1u;
            } else {
                const float ms = mP + mN;
                m[j] = ms;
                vx[j] = (vx[j] * mN + vx[i] * mP) / ms;
                vy[j] = (vy[j] * mN + vy[i] * mP) / ms;
                prev_size[j] = mx + (full - mx) * 0.25f;
                del[i] = 1u;
            }
            break;
        }
    }
}

// merged-away indices, ascending (block counts, scan, scatter)
__global__ __launch_bounds__(kMBlock) void k_m_del_count(const uint32_t *__restrict__ del, uint32_t n,
                                                         uint32_t *__restrict__ bc) {
    __shared__ uint32_t ws[kMBlock / 64];
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    const uint64_t b = __ballot(i < n && del[i]);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < kMBlock / 64; ++q) t += ws[q];
        bc[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kMBlock) void k_m_del_list(const uint32_t *__restrict__ del, uint32_t n,
                                                        const uint32_t *__restrict__ boff, uint32_t *__restrict__ out) {
    __shared__ uint32_t ws[kMBlock / 64];
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    const bool d = i < n && del[i];
    const uint64_t b = __ballot(d);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) ws[wv] = (uint32_t)__popcll(b);
    __syncthreads();
    if (!d) return;
    uint32_t o = boff[blockIdx.x];
    for (uint32_t q = 0; q < wv; ++q) o += ws[q];
    o += (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    out[o] = i;
}

// new slot k takes old slot perm[k], every field
__global__ __launch_bounds__(kMBlock) void k_m_take(State src, State dst, const uint32_t *__restrict__ perm,
                                                    uint32_t m) {
    const uint32_t k = blockIdx.x * kMBlock + threadIdx.x;
    if (k >= m) return;
    const uint32_t j = perm[k];
    dst.x[k] = src.x[j]; dst.y[k] = src.y[j]; dst.vx[k] = src.vx[j]; dst.vy[k] = src.vy[j];
    dst.ax[k] = src.ax[j]; dst.ay[k] = src.ay[j]; dst.pvx[k] = src.pvx[j]; dst.pvy[k] = src.pvy[j];
    dst.m[k] = src.m[j]; dst.flags[k] = src.flags[j]; dst.id[k] = src.id[j];
    if (src.sph_mass) {
        dst.sph_mass[k] = src.sph_mass[j]; dst.rest_dens[k] = src.rest_dens[j]; dst.stiff[k] = src.stiff[j];
        dst.visc[k] = src.visc[j]; dst.cohesion[k] = src.cohesion[j]; dst.dens[k] = src.dens[j];
        dst.press[k] = src.press[j]; dst.press_tmp[k] = src.press_tmp[j]; dst.pred_dens[k] = src.pred_dens[j];
        dst.ppx[k] = src.ppx[j]; dst.ppy[k] = src.ppy[j]; dst.pfx[k] = src.pfx[j]; dst.pfy[k] = src.pfy[j];
    }
    if (src.color) {
        dst.color[k] = src.color[j]; dst.pcolor[k] = src.pcolor[j]; dst.scolor[k] = src.scolor[j];
        dst.size[k] = src.size[j]; dst.prev_size[k] = src.prev_size[j]; dst.nbrs[k] = src.nbrs[j];
    }
    if (src.temp) {
        dst.temp[k] = src.temp[j]; dst.label[k] = src.label[j]; dst.ke[k] = src.ke[j]; dst.prev_ke[k] = src.prev_ke[j];
    }
    if (src.hx) {
        dst.hx[k] = src.hx[j]; dst.hy[k] = src.hy[j]; dst.ox[k] = src.ox[j]; dst.oy[k] = src.oy[j];
    }
}

// the largest 2 * totalRadius of the non-dark particles, as float bits
// (positive floats order like their bits; candidates <= 0 leave the max 0)
__global__ __launch_bounds__(kMBlock) void k_m_max_diameter(const float *__restrict__ size,
                                                            const uint32_t *__restrict__ flags, uint32_t n, float half,
                                                            float mult, uint32_t *__restrict__ out) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kMBlock + threadIdx.x;
    if (i >= n || (flags[i] & GE_FLAG_DARK)) return;
    const float d = size[i] * half * mult * 2.0f;
    if (d > 0.0f) atomicMax(out, __float_as_uint(d));
}

template <typename T>
hipError_t grow(T *&p, size_t &cap, size_t need) {
    if (need <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = need + need / 4 + 1024;
    return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * cap);
}

}  // namespace

void merger_free(ge_ctx *c) {
    MergerBuf &M = c->merger;
    for (void *p : {(void *)M.wa, (void *)M.wb, (void *)M.hist, (void *)M.rad, (void *)M.del, (void *)M.ecnt,
                    (void *)M.eoff, (void *)M.edges, (void *)M.parent, (void *)M.part, (void *)M.grid,
                    (void *)M.flag, (void *)M.dlist, (void *)M.bc, (void *)M.perm})
        if (p) (void)hipFree(p);
    M = MergerBuf{};
}

// globalLogic.cpp:1272-1292: clamp(max 2 * totalRadius, 2, 80)
int merger_cell_size(ge_ctx *c, const ge_merger_params *mp, float *cell) {
    const State &s = c->b.st[c->b.cur];
    uint32_t *d = nullptr;
    uint32_t bits = 0;
    if (hipMalloc(&d, sizeof(uint32_t)) != hipSuccess) {
        c->err = "cell size scratch allocation failed";
        return GE_E_NOMEM;
    }
    bool ok = hipMemcpyAsync(d, &bits, sizeof bits, hipMemcpyHostToDevice, c->stream) == hipSuccess;
    if (ok && c->n)
        hipLaunchKernelGGL(k_m_max_diameter, dim3(blocks_for(c->n, kMBlock)), dim3(kMBlock), 0, c->stream, s.size,
                           s.flags, c->n, mp->texture_half_size, mp->size_multiplier, d);
    ok = ok && hipMemcpyAsync(&bits, d, sizeof bits, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
    (void)hipFree(d);
    if (!ok) {
        c->err = "cell size reduction failed";
        return GE_E_HIP;
    }
    float v;
    std::memcpy(&v, &bits, sizeof v);
    *cell = std::clamp(v, 2.0f, 80.0f);
    return GE_OK;
}

// steps 1-4 and the merged-away list; *dels gets the sorted indices
int merger_run(ge_ctx *c, const ge_merger_params *mp, std::vector<uint32_t> &dels) {
    MergerBuf &M = c->merger;
    const State &s = c->b.st[c->b.cur];
    const uint32_t n = c->n;
    dels.clear();
    if (n == 0) return GE_OK;
    auto need = [&](hipError_t e) { return e == hipSuccess; };
    const uint32_t nbm = blocks_for(n, kMBlock);
    const uint32_t nparts = std::min<uint32_t>(nbm, 1024u);
    if (!M.cap || M.cap < n) {
        merger_free(c);
        M.cap = c->cap;
        const size_t cap = M.cap;
        if (!(need(hipMalloc(&M.wa, sizeof(uint64_t) * cap)) && need(hipMalloc(&M.wb, sizeof(uint64_t) * cap)) &&
              need(hipMalloc(&M.hist, sizeof(uint32_t) * radix_hist_entries((uint32_t)cap))) &&
              need(hipMemset(M.hist, 0, sizeof(uint32_t) * radix_hist_entries((uint32_t)cap))) &&
              need(hipMalloc(&M.rad, sizeof(float) * cap)) && need(hipMalloc(&M.del, sizeof(uint32_t) * cap)) &&
              need(hipMalloc(&M.ecnt, sizeof(uint32_t) * cap)) &&
              need(hipMalloc(&M.eoff, sizeof(uint32_t) * (cap + 1))) &&
              need(hipMalloc(&M.parent, sizeof(uint32_t) * cap)) && need(hipMalloc(&M.part, sizeof(float4) * 1024)) &&
              need(hipMalloc(&M.grid, sizeof(MGrid))) && need(hipMalloc(&M.flag, sizeof(uint32_t) * 4)) &&
              need(hipMalloc(&M.dlist, sizeof(uint32_t) * cap)) &&
              need(hipMalloc(&M.bc, sizeof(uint32_t) * (cap / kMBlock + 2))) &&
              need(hipMalloc(&M.perm, sizeof(uint32_t) * cap)))) {
            merger_free(c);
            c->err = "merger buffer allocation failed";
            return GE_E_NOMEM;
        }
    }
    const uint32_t zero[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(M.flag, zero, sizeof zero, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        c->err = "merger flag copy failed";
        return GE_E_HIP;
    }
    float *size = s.size, *prev = s.prev_size;
    hipLaunchKernelGGL(k_m_radius, dim3(nbm), dim3(kMBlock), 0, c->stream, size, n, mp->texture_half_size,
                       mp->size_multiplier, M.rad);
    hipLaunchKernelGGL(k_m_bounds, dim3(nparts), dim3(kMBlock), 0, c->stream, s.x, s.y, s.flags, n, M.part);
    hipLaunchKernelGGL(k_m_grid, dim3(1), dim3(64), 0, c->stream, M.part, nparts, mp->cell_size,
                       reinterpret_cast<MGrid *>(M.grid), M.flag);
    hipLaunchKernelGGL(k_m_words, dim3(nbm), dim3(kMBlock), 0, c->stream, s.x, s.y, s.flags, n,
                       reinterpret_cast<const MGrid *>(M.grid), M.wa);
    radix_sort_words(c->stream, M.hist, M.wa, M.wb, n, kIdxBits, kIdxBits + kCellKeyBits, c->b.counters + kCntError);
    const float dr2 = mp->density_radius * mp->density_radius;
    const MGrid *gp = reinterpret_cast<const MGrid *>(M.grid);
    hipLaunchKernelGGL(k_m_edges, dim3(nbm), dim3(kMBlock), 0, c->stream, s.x, s.y, s.flags, M.rad, n, gp, M.wa, dr2,
                       0, s.nbrs, M.ecnt, nullptr, nullptr);
    hipLaunchKernelGGL(k_m_scan, dim3(1), dim3(1024), 0, c->stream, M.ecnt, n, M.eoff);
    uint32_t ne = 0, gerr = 0;
    if (hipMemcpyAsync(&ne, M.eoff + n, sizeof ne, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&gerr, M.flag, sizeof gerr, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "merger count copy failed";
        return GE_E_HIP;
    }
    if (gerr) {
        c->err = "merger grid: coordinates span more cells than 36-bit cell keys hold";
        return GE_E_RANGE;
    }
    if (ne == 0) return GE_OK;
    if (grow(M.edges, M.edge_cap, ne) != hipSuccess) {
        c->err = "merger edge allocation failed";
        return GE_E_NOMEM;
    }
    hipLaunchKernelGGL(k_m_edges, dim3(nbm), dim3(kMBlock), 0, c->stream, s.x, s.y, s.flags, M.rad, n, gp, M.wa, dr2,
                       1, nullptr, M.ecnt, M.eoff, M.edges);
    hipLaunchKernelGGL(k_m_uf_init, dim3(nbm), dim3(kMBlock), 0, c->stream, M.parent, M.del, n);
    hipLaunchKernelGGL(k_m_union, dim3(nbm), dim3(kMBlock), 0, c->stream, M.eoff, M.edges, n, M.parent);
    hipLaunchKernelGGL(k_m_comp_words, dim3(nbm), dim3(kMBlock), 0, c->stream, M.eoff, M.parent, n, M.wb);
    uint64_t *cw = M.wb, *ca = M.wa;  // the cell words are no longer needed
    radix_sort_words(c->stream, M.hist, cw, ca, n, 0, 64, c->b.counters + kCntError);
    hipLaunchKernelGGL(k_m_merge, dim3(nbm), dim3(kMBlock), 0, c->stream, cw, n, M.eoff, M.edges, s.m, s.vx, s.vy,
                       prev, M.del);
    hipLaunchKernelGGL(k_m_del_count, dim3(nbm), dim3(kMBlock), 0, c->stream, M.del, n, M.bc);
    hipLaunchKernelGGL(k_m_scan, dim3(1), dim3(1024), 0, c->stream, M.bc, nbm, M.bc + 0);
    hipLaunchKernelGGL(k_m_del_list, dim3(nbm), dim3(kMBlock), 0, c->stream, M.del, n, M.bc, M.dlist);
    uint32_t nd = 0;
    if (hipMemcpyAsync(&nd, M.bc + nbm, sizeof nd, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "merger deletion count copy failed";
        return GE_E_HIP;
    }
    dels.resize(nd);
    if (nd && (hipMemcpyAsync(dels.data(), M.dlist, sizeof(uint32_t) * nd, hipMemcpyDeviceToHost, c->stream) !=
                   hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess)) {
        c->err = "merger deletion list copy failed";
        return GE_E_HIP;
    }
    return GE_OK;
}

// swap-with-back deletions replayed on indices (physics.cpp:1042-1066): the
// new slot k holds old slot perm[k]; moves every field into the other
// state buffer
int merger_apply(ge_ctx *c, const std::vector<uint32_t> &dels, uint32_t &new_n) {
    const uint32_t n = c->n;
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    uint32_t size_now = n;
    for (size_t k = dels.size(); k-- > 0;) {
        const uint32_t idx = dels[k];
        if (idx >= size_now) continue;
        std::swap(perm[idx], perm[size_now - 1]);
        --size_now;
    }
    new_n = size_now;
    MergerBuf &M = c->merger;
    if (size_now && hipMemcpyAsync(M.perm, perm.data(), sizeof(uint32_t) * size_now, hipMemcpyHostToDevice,
                                   c->stream) != hipSuccess) {
        c->err = "merger permutation copy failed";
        return GE_E_HIP;
    }
    const int cur = c->b.cur;
    if (size_now)
        hipLaunchKernelGGL(k_m_take, dim3(blocks_for(size_now, kMBlock)), dim3(kMBlock), 0, c->stream, c->b.st[cur],
                           c->b.st[1 - cur], M.perm, size_now);
    c->b.cur = 1 - cur;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "merger permutation failed";
        return GE_E_HIP;
    }
    return GE_OK;
}

}  // namespace ge
