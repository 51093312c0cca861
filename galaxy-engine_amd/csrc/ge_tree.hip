// ge_tree.hip -- quadtree build reproducing Quadtree / Node::Node /
// Node::subGridMaker (GE/src/Physics/quadtree.cpp:5-107) node for node:
// the same float partition points (std::partition_point's halving search),
// the same leaf rule (quadtree.cpp:16-17), the same DFS pre-order layout with
// `next` skip counts, and the same mass / centre-of-mass rounding order
// (quadtree.h:53-111).
//
// Device schedule (no host round trips, fixed launch count):
//   1. k_tree_init: the root.
//   2. kTopLevels launches of k_tree_level, one wave per pending node of
//      level L.  A node with more than kBlockSubMax particles is split with a
//      speculative binary search: each round the wave evaluates the predicate
//      at the 63 probes of the next six levels of partition_point's decision
//      tree, then follows the decision path; the result is exactly the
//      sequential search's.  Its children go to level L+1.  Smaller nodes are
//      handed on (step 3) or, at <= 64 particles, finished as wave subtrees.
//   3. k_tree_blocks: one workgroup per subtree of at most kBlockSubMax
//      particles runs that subtree's level loop itself, frontier in LDS and a
//      workgroup barrier per level instead of a launch per level.
//   A "wave subtree" (<= 64 particles) is replayed by one wave from ballot
//   masks: one particle per lane, every partition point a bit search on a
//   64-bit mask, the recursion stack in LDS.  Subtree node counts climb to the
//   parents with the last-child-finishes pattern (integer atomics).
//   4. k_tree_offsets: each top node walks to the root summing earlier
//      siblings' subtree sizes: its pre-order index.
//   5. k_tree_emit: leaves and wave subtrees write their nodes at their
//      pre-order indices; masses climb to the internal nodes, again
//      last-child-finishes, summed in the reference's child order.
#include "ge_internal.hpp"

namespace ge {
namespace {

enum : uint32_t { kKindInternal = 0, kKindLeaf = 1, kKindWaveSub = 2, kKindPending = 3, kKindDead = 4 };

constexpr int kTreeBlock = 256;
constexpr int kTreeWaves = kTreeBlock / 64;
constexpr uint32_t kTreeGridMax = 8192;
constexpr uint32_t kWaveSubMax = 64;
constexpr uint32_t kFrontMax = 1024;  // LDS frontier of a workgroup subtree

__device__ __forceinline__ bool leaf_rule(uint32_t cnt, float size) {
    // quadtree.cpp:16-17
    return ((cnt <= 4 && size <= 2.0f) || cnt == 1) || size <= 0.000001f;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ float readlane_f(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
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

// The same halving search over predicate bits held in a 64-bit mask
// (bit k = predicate of particle k of a wave subtree).
__device__ __forceinline__ uint32_t mask_part_point(uint64_t m, uint32_t first, uint32_t last) {
    uint32_t len = last - first;
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
__device__ uint32_t wave_part_point(const float *__restrict__ v, uint32_t first, uint32_t last, float mid) {
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
            if (ok && lj > 0) pred = v[fj + (lj >> 1)] < mid;
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

struct TreeOut {
    NodeRec *nodes;
    float4 *geom;
    uint32_t cap;
    int debug;
};

__device__ __forceinline__ void write_node(const TreeOut &o, uint32_t idx, float cx, float cy, float gm,
                                           float px, float py, float size, uint32_t next, uint32_t s,
                                           uint32_t e, uint32_t mask) {
    if (idx >= o.cap) return;  // kErrNodeCap is raised by the caller
    NodeRec r;
    r.comx = cx; r.comy = cy; r.mass = gm; r.size_sq = size * size;
    r.next = next; r.start = s; r.end = e; r.child_mask = mask;
    o.nodes[idx] = r;
    if (o.debug) o.geom[idx] = make_float4(px, py, size, 0.0f);
}

// ---------------------------------------------------------------------------
// Wave subtree: particles [s0, s0 + cnt) with cnt <= 64, one per lane.  The
// recursion of Node::Node is replayed uniformly by the wave; its stack lives
// in LDS.  EMIT=false counts nodes; EMIT=true writes them from index out0 and
// returns the root's mass and centre of mass.
struct WFrame {
    float px, py, size;
    uint32_t b[5];
    uint32_t q, out, nacc;
    float cm[4], cx[4], cy[4];
};

template <bool EMIT>
__device__ uint32_t wave_subtree(const float *__restrict__ x, const float *__restrict__ y,
                                 const float *__restrict__ m, float px, float py, float size, uint32_t s0,
                                 uint32_t cnt, uint32_t out0, const TreeOut &o, WFrame *stk,
                                 uint32_t *err, float &rgm, float &rcx, float &rcy) {
#pragma clang fp contract(off)
    const uint32_t lane = lane_id();
    const bool have = lane < cnt;
    const float lx = have ? x[s0 + lane] : 0.0f;
    const float ly = have ? y[s0 + lane] : 0.0f;
    const float lm = (EMIT && have) ? m[s0 + lane] : 0.0f;
    const uint64_t valid = __ballot(have);
    uint32_t counter = out0;

    // leaf mass over relative range [a, b): computeLeafMass order
    auto leaf_mass = [&](uint32_t a, uint32_t b, float &gm, float &cx, float &cy) {
        gm = 0.0f; cx = 0.0f; cy = 0.0f;
        for (uint32_t i = a; i < b; ++i) {
            const float mi = readlane_f(lm, i);
            gm = gm + mi;
            cx = cx + readlane_f(lx, i) * mi;
            cy = cy + readlane_f(ly, i) * mi;
        }
        if (gm > 0) { cx = cx / gm; cy = cy / gm; }
    };
    auto split = [&](float fpx, float fpy, float fsize, uint32_t a, uint32_t b, uint32_t *bb) {
        const float half = fsize * 0.5f;
        const float midx = fpx + half, midy = fpy + half;
        const uint64_t mx = __ballot(have && lx < midx) & valid;
        const uint64_t my = __ballot(have && ly < midy) & valid;
        bb[0] = a;
        bb[4] = b;
        bb[2] = mask_part_point(my, a, b);
        bb[1] = mask_part_point(mx, a, bb[2]);
        bb[3] = mask_part_point(mx, bb[2], b);
    };

    if (leaf_rule(cnt, size)) {
        if (EMIT) {
            leaf_mass(0, cnt, rgm, rcx, rcy);
            if (lane == 0) write_node(o, counter, rcx, rcy, rgm, px, py, size, 0, s0, s0 + cnt, 0);
        }
        return 1;
    }
    // Every lane writes the same (uniform) values into the LDS frames, so each
    // lane only ever reads back its own writes: no cross-lane ordering needed.
    int sp = 0;
    {
        WFrame &f = stk[0];
        uint32_t bb[5];
        split(px, py, size, 0, cnt, bb);
        f.px = px; f.py = py; f.size = size; f.q = 0; f.out = counter; f.nacc = 0;
        for (int k = 0; k < 5; ++k) f.b[k] = bb[k];
        ++counter;
    }
    while (sp >= 0) {
        WFrame &f = stk[sp];
        uint32_t q = f.q;
        while (q < 4 && !(f.b[q + 1] > f.b[q])) ++q;
        if (q < 4) {
            const float half = f.size * 0.5f;
            const float cpx = f.px + ((q & 1) ? half : 0.0f);
            const float cpy = f.py + ((q & 2) ? half : 0.0f);
            const uint32_t a = f.b[q], b = f.b[q + 1];
            const uint32_t cout = counter++;
            f.q = q + 1;
            if (leaf_rule(b - a, half)) {
                f.nacc += 1;
                if (EMIT) {
                    float gm, cx, cy;
                    leaf_mass(a, b, gm, cx, cy);
                    if (lane == 0) write_node(o, cout, cx, cy, gm, cpx, cpy, half, 0, s0 + a, s0 + b, 0);
                    f.cm[q] = gm; f.cx[q] = cx; f.cy[q] = cy;
                }
            } else {
                if (sp + 1 >= kMaxDepth) {
                    if (lane == 0) atomicOr(err, kErrDepth);
                    return counter - out0;
                }
                uint32_t bb[5];
                split(cpx, cpy, half, a, b, bb);
                WFrame &c = stk[sp + 1];
                c.px = cpx; c.py = cpy; c.size = half; c.q = 0; c.out = cout; c.nacc = 0;
                for (int k = 0; k < 5; ++k) c.b[k] = bb[k];
                ++sp;
            }
        } else {
            // computeInternalMass (children q = 0, 2, 1, 3) and next
            const uint32_t next = f.nacc;
            if (EMIT) {
                float gm = 0.0f, cx = 0.0f, cy = 0.0f;
                uint32_t mask = 0;
                const int order[4] = {0, 2, 1, 3};
                for (int k = 0; k < 4; ++k) {
                    const int qq = order[k];
                    if (!(f.b[qq + 1] > f.b[qq])) continue;
                    mask |= 1u << qq;
                    const float cm = f.cm[qq];
                    gm = gm + cm;
                    cx = cx + f.cx[qq] * cm;
                    cy = cy + f.cy[qq] * cm;
                }
                if (gm > 0) { cx = cx / gm; cy = cy / gm; }
                if (lane == 0) write_node(o, f.out, cx, cy, gm, f.px, f.py, f.size, next, s0 + f.b[0], s0 + f.b[4], mask);
                if (sp == 0) {
                    rgm = gm; rcx = cx; rcy = cy;
                } else {
                    WFrame &p = stk[sp - 1];
                    const uint32_t pq = p.q - 1;  // the quadrant this frame fills
                    p.cm[pq] = gm; p.cx[pq] = cx; p.cy[pq] = cy;
                }
            }
            if (sp > 0) stk[sp - 1].nacc += 1 + next;
            --sp;
        }
    }
    return counter - out0;
}

// ---------------------------------------------------------------------------
__global__ void k_tree_init(TopNodes T, const float *__restrict__ bb, uint32_t n,
                            uint32_t *__restrict__ counters) {
    const uint32_t t = threadIdx.x;
    if (t < kCntSlots - kCntLevel0) counters[kCntLevel0 + t] = 0;
    if (t == 0) {
        counters[kCntError] &= ~(kErrNodeCap | kErrDepth | kErrFrontier);
        counters[kCntBlockRoots] = 0;
        counters[kCntTopAlloc] = 0;
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
    for (int hops = 0; p != kNone && p < T.cap && hops < 4096; ++hops) {
        __hip_atomic_fetch_add(&T.accum[p], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t old = __hip_atomic_fetch_sub(&T.pending[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old != 1u) return;
        count = 1u + __hip_atomic_load(&T.accum[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&T.count[p], count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p = T.parent[p];
    }
}

// Split a node with the whole wave: subGridMaker's three searches
// (quadtree.cpp:30-65).
__device__ __forceinline__ void wave_split(const float *__restrict__ x, const float *__restrict__ y, float px,
                                           float py, float size, uint32_t s, uint32_t e, uint32_t b[5]) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half;
    b[0] = s;
    b[4] = e;
    b[2] = wave_part_point(y, s, e, midy);
    b[1] = wave_part_point(x, s, b[2], midx);
    b[3] = wave_part_point(x, b[2], e, midx);
}

// Lane 0: record children of an internal node at indices first.. (q order).
__device__ void write_children(TopNodes &T, uint32_t node, float px, float py, float size, const uint32_t b[5],
                               uint32_t first, uint32_t k, uint32_t *out_list) {
    T.kind[node] = kKindInternal;
    T.pending[node] = k;
    T.pending2[node] = k;
    T.accum[node] = 0;
    const float half = size * 0.5f;
    uint32_t c = first, j = 0;
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
            if (out_list) out_list[j++] = c;
            ++c;
        } else {
            T.child[4 * node + q] = kNone;
        }
    }
}

__device__ __forceinline__ uint32_t child_count(const uint32_t b[5]) {
    uint32_t k = 0;
    for (int q = 0; q < 4; ++q) k += (b[q + 1] > b[q]) ? 1u : 0u;
    return k;
}

// Leaf or wave subtree: finish the node now.  Returns false if the node needs
// splitting.  Uniform across the wave.
__device__ bool finish_small(TopNodes &T, const float *__restrict__ x, const float *__restrict__ y,
                             uint32_t node, float px, float py, float size, uint32_t s, uint32_t e, WFrame *stk,
                             uint32_t *err) {
    const uint32_t c = e - s;
    const uint32_t lane = lane_id();
    if (leaf_rule(c, size)) {
        if (lane == 0) {
            T.kind[node] = kKindLeaf;
            T.count[node] = 1;
            propagate_count(T, node, 1);
        }
        return true;
    }
    if (c <= kWaveSubMax) {
        const TreeOut none{nullptr, nullptr, 0, 0};
        float d0, d1, d2;
        const uint32_t k = wave_subtree<false>(x, y, nullptr, px, py, size, s, c, 0, none, stk, err, d0, d1, d2);
        if (lane == 0) {
            T.kind[node] = kKindWaveSub;
            T.count[node] = k;
            propagate_count(T, node, k);
        }
        return true;
    }
    return false;
}

__device__ __forceinline__ void push_block_root(TopNodes &T, uint32_t *counters, uint32_t node) {
    const uint32_t i = atomicAdd(&counters[kCntBlockRoots], 1u);
    if (i < T.cap) T.broot[i] = node;
    else atomicOr(&counters[kCntError], kErrFrontier);
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_level(TopNodes T, const float *__restrict__ x,
                                                           const float *__restrict__ y,
                                                           uint32_t *__restrict__ counters, int L,
                                                           int last_level) {
    __shared__ WFrame stk_all[kTreeWaves][kMaxDepth];
    WFrame *stk = stk_all[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    const uint32_t base = counters[kCntLevelBase0 + L];
    uint32_t cnt = counters[kCntLevel0 + L];
    if (base >= T.cap) cnt = 0;
    else if (cnt > T.cap - base) cnt = T.cap - base;
    const uint32_t next_base = base + cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[kCntLevelBase0 + L + 1] = next_base;
        // the last level allocates nothing: the workgroup phase continues here
        if (L == last_level) counters[kCntTopAlloc] = next_base;
    }
    const uint32_t wave = (blockIdx.x * kTreeBlock + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * kTreeBlock) >> 6;
    for (uint32_t t = wave; t < cnt; t += nwaves) {
        const uint32_t node = base + t;
        if (T.kind[node] == kKindDead) continue;
        const float px = T.px[node], py = T.py[node], size = T.size[node];
        const uint32_t s = T.start[node], e = T.end[node];
        if (finish_small(T, x, y, node, px, py, size, s, e, stk, counters + kCntError)) continue;
        if (e - s <= kBlockSubMax || L == last_level) {
            if (lane == 0) push_block_root(T, counters, node);
            continue;
        }
        uint32_t b[5];
        wave_split(x, y, px, py, size, s, e, b);
        if (lane == 0) {
            const uint32_t k = child_count(b);
            const uint32_t off = atomicAdd(&counters[kCntLevel0 + L + 1], k);
            const uint32_t first = next_base + off;
            if (first + k > T.cap) {
                // no room on the next level: hand the node to the workgroup phase
                for (uint32_t cc = first; cc < T.cap && cc < first + k; ++cc) T.kind[cc] = kKindDead;
                push_block_root(T, counters, node);
            } else {
                write_children(T, node, px, py, size, b, first, k, nullptr);
            }
        }
    }
}

// One workgroup per subtree root handed over by the level phase: the level
// loop of that subtree with its frontier in LDS.
__global__ __launch_bounds__(kTreeBlock) void k_tree_blocks(TopNodes T, const float *__restrict__ x,
                                                            const float *__restrict__ y,
                                                            uint32_t *__restrict__ counters) {
    __shared__ WFrame stk_all[kTreeWaves][kMaxDepth];
    __shared__ uint32_t front[2][kFrontMax];
    __shared__ uint32_t fcount[2];
    WFrame *stk = stk_all[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t nroots = min(counters[kCntBlockRoots], T.cap);
    for (uint32_t r = blockIdx.x; r < nroots; r += gridDim.x) {
        if (threadIdx.x == 0) {
            front[0][0] = T.broot[r];
            fcount[0] = 1;
            fcount[1] = 0;
        }
        __syncthreads();
        int cur = 0;
        int lvl = 0;
        for (; lvl < kBlockLevelsMax; ++lvl) {
            const uint32_t nf = min(fcount[cur], kFrontMax);
            if (nf == 0) break;
            for (uint32_t i = w; i < nf; i += kTreeWaves) {
                const uint32_t node = front[cur][i];
                const float px = T.px[node], py = T.py[node], size = T.size[node];
                const uint32_t s = T.start[node], e = T.end[node];
                if (finish_small(T, x, y, node, px, py, size, s, e, stk, counters + kCntError)) continue;
                uint32_t b[5];
                wave_split(x, y, px, py, size, s, e, b);
                if (lane == 0) {
                    const uint32_t k = child_count(b);
                    const uint32_t first = atomicAdd(&counters[kCntTopAlloc], k);
                    const uint32_t slot = atomicAdd(&fcount[cur ^ 1], k);
                    if (first + k > T.cap || slot + k > kFrontMax) {
                        atomicOr(&counters[kCntError], kErrFrontier);
                    } else {
                        write_children(T, node, px, py, size, b, first, k, &front[cur ^ 1][slot]);
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) fcount[cur] = 0;
            cur ^= 1;
            __syncthreads();
        }
        if (lvl == kBlockLevelsMax && threadIdx.x == 0) atomicOr(&counters[kCntError], kErrFrontier);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_offsets(TopNodes T, uint32_t *__restrict__ counters,
                                                             uint32_t node_cap) {
    const uint32_t total = min(counters[kCntTopAlloc], T.cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t nn = T.count[0];
        counters[kCntTopTotal] = total;
        counters[kCntNodeTotal] = nn <= node_cap ? nn : node_cap;
        if (nn > node_cap) atomicOr(&counters[kCntError], kErrNodeCap);
    }
    for (uint32_t t = blockIdx.x * kTreeBlock + threadIdx.x; t < total; t += gridDim.x * kTreeBlock) {
        if (T.kind[t] == kKindDead) continue;
        uint32_t o = 0, nd = t;
        int hops = 0;
        while (T.parent[nd] != kNone) {
            const uint32_t p = T.parent[nd];
            if (p >= T.cap || ++hops > 4096) {  // a corrupt chain: flag it, never loop
                atomicOr(&counters[kCntError], kErrFrontier);
                break;
            }
            o += 1;
            for (int q = 0; q < 4; ++q) {
                const uint32_t ch = T.child[4 * p + q];
                if (ch == nd) break;
                if (ch != kNone && ch < T.cap) o += T.count[ch];
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
    for (int hops = 0; p != kNone && p < T.cap && hops < 4096; ++hops) {
        __threadfence();
        uint32_t old = __hip_atomic_fetch_sub(&T.pending2[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old != 1u) return;
        __threadfence();
        float gm = 0.0f, cx = 0.0f, cy = 0.0f;
        uint32_t mask = 0;
        const int order[4] = {0, 2, 1, 3};
        for (int k = 0; k < 4; ++k) {
            const uint32_t ch = T.child[4 * p + order[k]];
            if (ch == kNone || ch >= T.cap) continue;
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
                                                          uint32_t *__restrict__ counters, TreeOut o,
                                                          uint32_t n) {
#pragma clang fp contract(off)
    __shared__ WFrame stk_all[kTreeWaves][kMaxDepth];
    WFrame *stk = stk_all[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    const uint32_t total = counters[kCntTopTotal];
    const uint32_t wave = (blockIdx.x * kTreeBlock + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * kTreeBlock) >> 6;
    for (uint32_t t = wave; t < total; t += nwaves) {
        const uint32_t kind = T.kind[t];
        if (kind != kKindLeaf && kind != kKindWaveSub) continue;
        const uint32_t out = T.offset[t];
        const uint32_t s = T.start[t], e = T.end[t];
        if (e > n || s >= e || (kind == kKindWaveSub && e - s > kWaveSubMax)) {
            if (lane == 0) atomicOr(&counters[kCntError], kErrFrontier);
            continue;
        }
        if (lane == 0 && out + T.count[t] > o.cap) atomicOr(&counters[kCntError], kErrNodeCap);
        float gm, cx, cy;
        if (kind == kKindWaveSub) {
            wave_subtree<true>(x, y, m, T.px[t], T.py[t], T.size[t], s, e - s, out, o, stk,
                               counters + kCntError, gm, cx, cy);
        } else if (lane == 0) {
            // computeLeafMass (quadtree.h:53-67)
            gm = 0.0f; cx = 0.0f; cy = 0.0f;
            for (uint32_t i = s; i < e; ++i) {
                const float mi = m[i];
                gm = gm + mi;
                cx = cx + x[i] * mi;
                cy = cy + y[i] * mi;
            }
            if (gm > 0) { cx = cx / gm; cy = cy / gm; }
            write_node(o, out, cx, cy, gm, T.px[t], T.py[t], T.size[t], 0, s, e, 0);
        }
        if (lane == 0) propagate_mass(T, o, t);
    }
}

uint32_t wave_grid(uint64_t waves) {
    uint64_t blocks = (waves + kTreeWaves - 1) / kTreeWaves;
    if (blocks < 1) blocks = 1;
    if (blocks > kTreeGridMax) blocks = kTreeGridMax;
    return (uint32_t)blocks;
}

}  // namespace

void launch_tree_build(ge_ctx *c, const float *bb_dev) {
    const State &s = c->b.st[c->b.cur];
    TopNodes &T = c->b.top;
    const uint64_t n = c->n;
    hipLaunchKernelGGL(k_tree_init, dim3(1), dim3(kTreeBlock), 0, c->stream, T, bb_dev, c->n, c->b.counters);
    // level L+1 only holds children of nodes above kBlockSubMax particles
    const uint64_t big_children = 4 * (n / (kBlockSubMax + 1)) + 4;
    for (int L = 0; L < kTopLevels; ++L) {
        uint64_t maxn = 1;
        for (int k = 0; k < L && maxn < n; ++k) maxn *= 4;
        if (L > 0 && maxn > big_children) maxn = big_children;
        if (maxn > n) maxn = n;
        hipLaunchKernelGGL(k_tree_level, dim3(wave_grid(maxn)), dim3(kTreeBlock), 0, c->stream, T, s.x, s.y,
                           c->b.counters, L, kTopLevels - 1);
    }
    // roots of workgroup subtrees each hold more than 64 particles
    uint64_t roots = n / (kWaveSubMax + 1) + 1;
    if (roots > 4096) roots = 4096;
    hipLaunchKernelGGL(k_tree_blocks, dim3((uint32_t)roots), dim3(kTreeBlock), 0, c->stream, T, s.x, s.y,
                       c->b.counters);
    uint32_t tg = (uint32_t)((T.cap + kTreeBlock - 1) / kTreeBlock);
    if (tg > 1024) tg = 1024;
    hipLaunchKernelGGL(k_tree_offsets, dim3(tg), dim3(kTreeBlock), 0, c->stream, T, c->b.counters, c->b.node_cap);
    TreeOut o{c->b.nodes, c->b.node_geom, c->b.node_cap, c->debug ? 1 : 0};
    hipLaunchKernelGGL(k_tree_emit, dim3(wave_grid(T.cap)), dim3(kTreeBlock), 0, c->stream, T, s.x, s.y, s.m,
                       c->b.counters, o, c->n);
}

}  // namespace ge
