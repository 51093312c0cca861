// ge_tree.hip -- quadtree build reproducing Quadtree / Node::Node /
// Node::subGridMaker (GE/src/Physics/quadtree.cpp:5-107) node for node:
// the same float partition points (std::partition_point's halving search),
// the same leaf rule (quadtree.cpp:16-17), the same DFS pre-order layout with
// `next` skip counts, and the same mass / centre-of-mass rounding order
// (quadtree.h:53-111).
//
// Device schedule (no host round trips, fixed launch count):
//   1. k_tree_init: the root.
//   2. kTopLevels launches of k_tree_level: every pending node of level L is
//      classified; big internal nodes get their children appended to level
//      L+1; leaves and small nodes (<= kSubtreeMax particles, or anything on
//      the last level) become "in-thread subtrees" whose node count one thread
//      computes by replaying the recursion.  Counts climb to the parents with
//      the last-child-finishes pattern (integer atomics, order independent).
//   3. k_tree_offsets: each top node walks to the root summing earlier
//      siblings' subtree sizes: its pre-order index.
//   4. k_tree_emit: leaves and in-thread subtrees write their nodes at their
//      pre-order indices; masses climb to internal top nodes, again
//      last-child-finishes, summed in the reference's child order.
#include "ge_internal.hpp"

namespace ge {
namespace {

enum : uint32_t { kKindInternal = 0, kKindLeaf = 1, kKindSubtree = 2, kKindPending = 3, kKindDead = 4 };

constexpr int kTreeBlock = 256;
constexpr uint32_t kTreeGrid = 1024;

__device__ __forceinline__ bool leaf_rule(uint32_t cnt, float size) {
    // quadtree.cpp:16-17
    return ((cnt <= 4 && size <= 2.0f) || cnt == 1) || size <= 0.000001f;
}

// std::partition_point with predicate v[i] < mid (lower_bound halving)
__device__ __forceinline__ uint32_t part_point(const float *__restrict__ v, uint32_t first, uint32_t last,
                                               float mid) {
    uint32_t len = last - first;
    while (len > 0) {
        uint32_t half = len >> 1;
        uint32_t middle = first + half;
        if (v[middle] < mid) {
            first = middle + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

// subGridMaker's three searches (quadtree.cpp:30-65)
__device__ __forceinline__ void split(const float *__restrict__ x, const float *__restrict__ y, float px,
                                      float py, float size, uint32_t s, uint32_t e, uint32_t b[5]) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half;
    b[0] = s;
    b[4] = e;
    b[2] = part_point(y, s, e, midy);
    b[1] = part_point(x, s, b[2], midx);
    b[3] = part_point(x, b[2], e, midx);
}

// computeLeafMass (quadtree.h:53-67)
__device__ __forceinline__ void leaf_mass(const float *__restrict__ x, const float *__restrict__ y,
                                          const float *__restrict__ m, uint32_t s, uint32_t e, float &gm,
                                          float &cx, float &cy) {
#pragma clang fp contract(off)
    gm = 0.0f; cx = 0.0f; cy = 0.0f;
    for (uint32_t i = s; i < e; ++i) {
        const float mi = m[i];
        gm = gm + mi;
        cx = cx + x[i] * mi;
        cy = cy + y[i] * mi;
    }
    if (gm > 0) { cx = cx / gm; cy = cy / gm; }
}

struct Frame {
    float px, py, size;
    uint32_t b[5];
    uint32_t q;
    uint32_t out;
};

struct TreeOut {
    NodeRec *nodes;
    float4 *geom;
    uint32_t cap;
    int debug;
};

__device__ __forceinline__ void write_node(const TreeOut &o, uint32_t idx, float cx, float cy, float gm,
                                           float px, float py, float size, uint32_t next, uint32_t s,
                                           uint32_t e, uint32_t mask) {
    if (idx >= o.cap) return;  // caller raises kErrNodeCap
    NodeRec r;
    r.comx = cx; r.comy = cy; r.mass = gm; r.size_sq = size * size;
    r.next = next; r.start = s; r.end = e; r.child_mask = mask;
    o.nodes[idx] = r;
    if (o.debug) o.geom[idx] = make_float4(px, py, size, 0.0f);
}

// Replays Node::Node's recursion for one subtree inside one thread.  With
// EMIT=false it only counts nodes; with EMIT=true it writes them at
// pre-order indices starting at `out0` and returns the root's mass / COM.
template <bool EMIT>
__device__ uint32_t subtree(const float *__restrict__ x, const float *__restrict__ y,
                            const float *__restrict__ m, float px, float py, float size, uint32_t s,
                            uint32_t e, uint32_t out0, const TreeOut &o, uint32_t *err, float &rgm,
                            float &rcx, float &rcy) {
#pragma clang fp contract(off)
    uint32_t counter = out0;  // next pre-order index
    // visit the root
    if (leaf_rule(e - s, size)) {
        if (EMIT) {
            leaf_mass(x, y, m, s, e, rgm, rcx, rcy);
            write_node(o, counter, rcx, rcy, rgm, px, py, size, 0, s, e, 0);
        }
        return 1;
    }
    Frame st[kMaxDepth];
    int sp = 0;
    st[0].px = px; st[0].py = py; st[0].size = size; st[0].q = 0; st[0].out = counter;
    split(x, y, px, py, size, s, e, st[0].b);
    ++counter;
    while (sp >= 0) {
        Frame &f = st[sp];
        uint32_t q = f.q;
        while (q < 4 && !(f.b[q + 1] > f.b[q])) ++q;
        if (q < 4) {
            f.q = q + 1;
            const float half = f.size * 0.5f;
            const float cpx = f.px + ((q & 1) ? half : 0.0f);
            const float cpy = f.py + ((q & 2) ? half : 0.0f);
            const uint32_t cs = f.b[q], ce = f.b[q + 1];
            const uint32_t cout = counter++;
            if (leaf_rule(ce - cs, half)) {
                if (EMIT) {
                    float gm, cx, cy;
                    leaf_mass(x, y, m, cs, ce, gm, cx, cy);
                    write_node(o, cout, cx, cy, gm, cpx, cpy, half, 0, cs, ce, 0);
                }
            } else {
                if (sp + 1 >= kMaxDepth) {
                    atomicOr(err, kErrDepth);
                    return counter - out0;
                }
                Frame &c = st[++sp];
                c.px = cpx; c.py = cpy; c.size = half; c.q = 0; c.out = cout;
                split(x, y, cpx, cpy, half, cs, ce, c.b);
            }
        } else {
            // all children done: computeInternalMass in subGrids[lr][ud] order
            // (q = 0, 2, 1, 3) and calculateNextNeighbor
            if (EMIT) {
                uint32_t cidx[4] = {kNone, kNone, kNone, kNone};
                uint32_t mask = 0;
                uint32_t c = f.out + 1;
                for (int k = 0; k < 4; ++k) {
                    if (f.b[k + 1] > f.b[k]) {
                        cidx[k] = c;
                        mask |= 1u << k;
                        c += (c < o.cap ? o.nodes[c].next : 0u) + 1;
                    }
                }
                float gm = 0.0f, cx = 0.0f, cy = 0.0f;
                const int order[4] = {0, 2, 1, 3};
                for (int k = 0; k < 4; ++k) {
                    const uint32_t ci = cidx[order[k]];
                    if (ci == kNone || ci >= o.cap) continue;
                    const NodeRec r = o.nodes[ci];
                    gm = gm + r.mass;
                    cx = cx + r.comx * r.mass;
                    cy = cy + r.comy * r.mass;
                }
                if (gm > 0) { cx = cx / gm; cy = cy / gm; }
                write_node(o, f.out, cx, cy, gm, f.px, f.py, f.size, c - f.out - 1, f.b[0], f.b[4], mask);
                if (sp == 0) { rgm = gm; rcx = cx; rcy = cy; }
            }
            --sp;
        }
    }
    return counter - out0;
}

__global__ void k_tree_init(TopNodes T, const float *__restrict__ bb, uint32_t n,
                            uint32_t *__restrict__ counters) {
    const uint32_t t = threadIdx.x;
    if (t < kCntSlots - kCntLevel0) counters[kCntLevel0 + t] = 0;
    if (t == 0) {
        counters[kCntError] &= ~(kErrNodeCap | kErrDepth);
        T.px[0] = bb[0]; T.py[0] = bb[1]; T.size[0] = bb[2];
        T.start[0] = 0; T.end[0] = n; T.parent[0] = kNone; T.kind[0] = kKindPending;
    }
    __syncthreads();
    if (t == 0) {
        counters[kCntLevel0 + 0] = 1;
        counters[kCntLevelBase0 + 0] = 0;
    }
}

// Report a finished subtree's node count to its ancestors.
__device__ void propagate_count(TopNodes &T, uint32_t node, uint32_t count) {
    uint32_t p = T.parent[node];
    while (p != kNone) {
        __hip_atomic_fetch_add(&T.accum[p], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t old = __hip_atomic_fetch_sub(&T.pending[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old != 1u) return;
        count = 1u + __hip_atomic_load(&T.accum[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&T.count[p], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p = T.parent[p];
    }
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_level(TopNodes T, const float *__restrict__ x,
                                                           const float *__restrict__ y,
                                                           uint32_t *__restrict__ counters, int L,
                                                           int last_level) {
    const uint32_t base = counters[kCntLevelBase0 + L];
    uint32_t cnt = counters[kCntLevel0 + L];
    if (base >= T.cap) cnt = 0;
    else if (cnt > T.cap - base) cnt = T.cap - base;
    const uint32_t next_base = base + cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[kCntLevelBase0 + L + 1] = next_base;
    TreeOut none{nullptr, nullptr, 0, 0};
    for (uint32_t t = blockIdx.x * kTreeBlock + threadIdx.x; t < cnt; t += gridDim.x * kTreeBlock) {
        const uint32_t node = base + t;
        if (T.kind[node] == kKindDead) continue;
        const float px = T.px[node], py = T.py[node], size = T.size[node];
        const uint32_t s = T.start[node], e = T.end[node];
        float dummy0, dummy1, dummy2;
        if (leaf_rule(e - s, size)) {
            T.kind[node] = kKindLeaf;
            T.count[node] = 1;
            propagate_count(T, node, 1);
            continue;
        }
        bool in_thread = (e - s) <= kSubtreeMax || L == last_level;
        uint32_t b[5];
        uint32_t k = 0;
        if (!in_thread) {
            split(x, y, px, py, size, s, e, b);
            for (int q = 0; q < 4; ++q) k += (b[q + 1] > b[q]) ? 1u : 0u;
            const uint32_t off = atomicAdd(&counters[kCntLevel0 + L + 1], k);
            const uint32_t first = next_base + off;
            if (first + k > T.cap) {
                for (uint32_t c = first; c < T.cap && c < first + k; ++c) T.kind[c] = kKindDead;
                in_thread = true;
            } else {
                T.kind[node] = kKindInternal;
                T.pending[node] = k;
                T.pending2[node] = k;
                T.accum[node] = 0;
                const float half = size * 0.5f;
                uint32_t c = first;
                for (int q = 0; q < 4; ++q) {
                    if (b[q + 1] > b[q]) {
                        T.px[c] = px + ((q & 1) ? half : 0.0f);
                        T.py[c] = py + ((q & 2) ? half : 0.0f);
                        T.size[c] = half;
                        T.start[c] = b[q];
                        T.end[c] = b[q + 1];
                        T.parent[c] = node;
                        T.kind[c] = kKindPending;
                        T.child[4 * node + q] = c;
                        ++c;
                    } else {
                        T.child[4 * node + q] = kNone;
                    }
                }
            }
        }
        if (in_thread) {
            const uint32_t c = subtree<false>(x, y, nullptr, px, py, size, s, e, 0, none,
                                              counters + kCntError, dummy0, dummy1, dummy2);
            T.kind[node] = kKindSubtree;
            T.count[node] = c;
            propagate_count(T, node, c);
        }
    }
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_offsets(TopNodes T, uint32_t *__restrict__ counters,
                                                             uint32_t node_cap) {
    const uint32_t total = min(counters[kCntLevelBase0 + kTopLevels], T.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t nn = T.count[0];
        counters[kCntTopTotal] = total;
        counters[kCntNodeTotal] = nn <= node_cap ? nn : node_cap;
        if (nn > node_cap) atomicOr(&counters[kCntError], kErrNodeCap);
    }
    for (uint32_t t = blockIdx.x * kTreeBlock + threadIdx.x; t < total; t += gridDim.x * kTreeBlock) {
        if (T.kind[t] == kKindDead) continue;
        uint32_t o = 0, nd = t;
        while (T.parent[nd] != kNone) {
            const uint32_t p = T.parent[nd];
            o += 1;
            for (int q = 0; q < 4; ++q) {
                const uint32_t ch = T.child[4 * p + q];
                if (ch == nd) break;
                if (ch != kNone) o += T.count[ch];
            }
            nd = p;
        }
        T.offset[t] = o;
    }
}

// Report a finished node's mass to its ancestors; the last child to report
// computes the parent as computeInternalMass does (children q = 0, 2, 1, 3).
__device__ void propagate_mass(TopNodes &T, const TreeOut &o, uint32_t node) {
#pragma clang fp contract(off)
    uint32_t p = T.parent[node];
    while (p != kNone) {
        __threadfence();
        uint32_t old = __hip_atomic_fetch_sub(&T.pending2[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old != 1u) return;
        __threadfence();
        float gm = 0.0f, cx = 0.0f, cy = 0.0f;
        uint32_t mask = 0;
        const int order[4] = {0, 2, 1, 3};
        for (int k = 0; k < 4; ++k) {
            const uint32_t ch = T.child[4 * p + order[k]];
            if (ch == kNone) continue;
            mask |= 1u << order[k];
            const uint32_t oc = T.offset[ch];
            if (oc >= o.cap) continue;
            const float *rec = reinterpret_cast<const float *>(&o.nodes[oc]);
            const float cm = atomic_load_f32(rec + 2);
            const float ccx = atomic_load_f32(rec + 0);
            const float ccy = atomic_load_f32(rec + 1);
            gm = gm + cm;
            cx = cx + ccx * cm;
            cy = cy + ccy * cm;
        }
        if (gm > 0) { cx = cx / gm; cy = cy / gm; }
        write_node(o, T.offset[p], cx, cy, gm, T.px[p], T.py[p], T.size[p], T.count[p] - 1u, T.start[p],
                   T.end[p], mask);
        p = T.parent[p];
    }
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_emit(TopNodes T, const float *__restrict__ x,
                                                          const float *__restrict__ y,
                                                          const float *__restrict__ m,
                                                          uint32_t *__restrict__ counters, TreeOut o) {
    const uint32_t total = counters[kCntTopTotal];
    for (uint32_t t = blockIdx.x * kTreeBlock + threadIdx.x; t < total; t += gridDim.x * kTreeBlock) {
        const uint32_t kind = T.kind[t];
        if (kind != kKindLeaf && kind != kKindSubtree) continue;
        const uint32_t out = T.offset[t];
        if (out + T.count[t] > o.cap) {
            atomicOr(&counters[kCntError], kErrNodeCap);
        }
        float gm, cx, cy;
        subtree<true>(x, y, m, T.px[t], T.py[t], T.size[t], T.start[t], T.end[t], out, o,
                      counters + kCntError, gm, cx, cy);
        propagate_mass(T, o, t);
    }
}

}  // namespace

void launch_tree_build(ge_ctx *c, const float *bb_dev) {
    const State &s = c->b.st[c->b.cur];
    TopNodes &T = c->b.top;
    hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(kTreeBlock), 0, c->stream, T, bb_dev, c->n, c->b.counters);
    for (int L = 0; L < kTopLevels; ++L) {
        // level L holds at most min(4^L, n) nodes
        uint64_t maxn = 1;
        for (int k = 0; k < L && maxn < c->n; ++k) maxn *= 4;
        if (maxn > c->n) maxn = c->n;
        uint32_t grid = blocks_for(maxn, kTreeBlock);
        if (grid > kTreeGrid) grid = kTreeGrid;
        hipLaunchKernelGGL(k_tree_level, dim3(grid), dim3(kTreeBlock), 0, c->stream, T, s.x, s.y, c->b.counters,
                           L, kTopLevels - 1);
    }
    uint32_t tg = blocks_for(T.cap, kTreeBlock);
    if (tg > kTreeGrid) tg = kTreeGrid;
    hipLaunchKernelGGL(k_tree_offsets, dim3(tg), dim3(kTreeBlock), 0, c->stream, T, c->b.counters, c->b.node_cap);
    TreeOut o{c->b.nodes, c->b.node_geom, c->b.node_cap, c->debug ? 1 : 0};
    hipLaunchKernelGGL(k_tree_emit, dim3(tg), dim3(kTreeBlock), 0, c->stream, T, s.x, s.y, s.m, c->b.counters, o);
}

}  // namespace ge
