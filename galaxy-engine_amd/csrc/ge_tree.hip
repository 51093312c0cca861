// ge_tree.hip -- quadtree build reproducing Quadtree / Node::Node /
// Node::subGridMaker (GE/src/Physics/quadtree.cpp:5-107) node for node:
// the same float partition points (std::partition_point's halving search),
// the same leaf rule (quadtree.cpp:16-17), the same DFS pre-order layout with
// `next` skip counts, and the same mass / centre-of-mass rounding order
// (quadtree.h:53-111).
//
// Pre-order without parent links: nodes of nested Morton ranges appear in DFS
// pre-order sorted by (start particle, level).  Every node registers at its
// start particle i: C[i] += 1 and F[i] = min(F[i], level).  After one
// exclusive scan P of C, node (level l, range [a, b)) has pre-order index
// P[a] + (l - F[a]) and next = P[b] - P[a] - (l - F[a]) - 1.  No kernel waits
// on another workgroup: workgroups only meet at launch boundaries.
//
// Device schedule (fixed launch count, no host round trips):
//   1. k_tree_clear: C, F and the root.
//   2. kTopLevels launches of k_tree_level, one wave per pending node of
//      level L.  A node above kBlockSubMax particles is split with a
//      speculative binary search: each round the wave evaluates the predicate
//      at the 63 probes of the next six levels of partition_point's decision
//      tree and follows the decision path, exactly the sequential result.
//   3. k_tree_blocks: one workgroup per subtree of at most kBlockSubMax
//      particles runs that subtree's level loop, frontier in LDS and a
//      workgroup barrier per level instead of a launch per level.
//   A node of at most 64 particles is a "wave subtree": one particle per
//   lane, all of its levels built at once from ballot masks (each partition
//   point is a bit search on a 64-bit mask).
//   4. k_scan_* : the exclusive scan of C.
//   5. k_tree_emit: every node's record at its pre-order index; wave
//      subtrees and leaves also get their masses.
//   6. k_tree_block_mass: one workgroup per workgroup subtree finishes the
//      masses of its internal nodes, deepest level first, from the list of
//      internal nodes k_tree_blocks kept level by level.
//   7. k_tree_mass: one workgroup finishes the rest (the nodes above the
//      workgroup subtrees), deepest first.  Both sum children in the order
//      q = 0, 2, 1, 3 of computeInternalMass.
#include <algorithm>

#include "ge_internal.hpp"
#include "ge_tree_common.hpp"

namespace ge {
namespace {

struct TreeOut {
    NodeRec *nodes;
    float4 *geom;
    uint32_t cap;
    int debug;
};

__device__ __forceinline__ void write_node(const TreeOut &o, uint32_t idx, float cx, float cy, float gm,
                                           float px, float py, float size, uint32_t next, uint32_t s,
                                           uint32_t e, uint32_t mask, bool geom = true) {
    if (idx >= o.cap) return;  // kErrNodeCap is raised by the caller
    NodeRec r;
    r.comx = cx; r.comy = cy; r.mass = gm; r.size_sq = size * size;
    r.next = next; r.start = s; r.end = e; r.child_mask = mask;
    o.nodes[idx] = r;
    if (geom && o.debug) o.geom[idx] = make_float4(px, py, size, 0.0f);
}

struct WLevel {
    uint64_t starts, covered, leafs, qb0, qb1;
};

__device__ __forceinline__ WLevel load_level(const WLevel *lev, int L) {
    WLevel r;
    r.starts = rfl64(lev[L].starts);
    r.covered = rfl64(lev[L].covered);
    r.leafs = rfl64(lev[L].leafs);
    r.qb0 = rfl64(lev[L].qb0);
    r.qb1 = rfl64(lev[L].qb1);
    return r;
}

// end of the node starting at `lane` on a level: the next node start or the
// first particle no longer covered
__device__ __forceinline__ uint32_t node_end(const WLevel &lv, uint32_t lane) {
    const uint64_t stop = (lv.starts | ~lv.covered) & above(lane);
    return stop ? (uint32_t)__builtin_ctzll(stop) : 64u;
}

// lane of the child with quadrant q inside [a, b) on the next level, or 64
__device__ __forceinline__ uint32_t child_lane(const WLevel &nx, uint32_t a, uint32_t b, int q) {
    const uint64_t range = (b >= 64 ? ~0ull : ((1ull << b) - 1ull)) & (~0ull << a);
    uint64_t sel = nx.starts & range;
    sel &= (q & 1) ? nx.qb0 : ~nx.qb0;
    sel &= (q & 2) ? nx.qb1 : ~nx.qb1;
    return sel ? (uint32_t)__builtin_ctzll(sel) : 64u;
}

// Wave subtree: particles [s0, s0 + cnt), cnt <= 64, one per lane, built
// level-parallel.  At each level every lane carries the range and corner of
// the node holding its particle; the node's three partition points are bit
// searches on two ballot masks (x < mid, y < mid).  EMIT=false registers the
// subtree's nodes in the start table; EMIT=true (after the scan) writes them
// and returns the root's mass and centre of mass.  Per level, four 64-bit
// masks kept in LDS describe the nodes: starts, covered particles, leaf
// starts and the two quadrant bits of each start.
template <bool EMIT>
__device__ uint32_t wave_subtree(const float *__restrict__ x, const float *__restrict__ y,
                                 const float *__restrict__ m, float px0, float py0, float size0, uint32_t s0,
                                 uint32_t cnt, uint32_t depth0, const StartTable &st, const TreeOut &o,
                                 WLevel *lev, uint32_t *err, float &rgm, float &rcx, float &rcy,
                                 uint32_t off = 0) {
#pragma clang fp contract(off)
    const uint32_t lane = lane_id();
    const bool have = lane < cnt;
    const float lx = have ? x[s0 + lane - off] : 0.0f;  // x[i - off] is particle i
    const float ly = have ? y[s0 + lane - off] : 0.0f;

    // ---- pass A: top-down, every node of a level at once
    bool act = have;
    uint32_t a = 0, b = cnt;
    float npx = px0, npy = py0, size = size0;
    uint32_t C = 0, F = 0xFFu, leaf_end = 0;
    uint64_t qb0 = 0, qb1 = 0;
    uint32_t total = 0;
    int nlev = 0;
    for (; nlev < kWaveLevels; ++nlev) {
        const uint64_t actm = __ballot(act);
        if (actm == 0) break;
        const bool is_start = act && lane == a;
        const bool leaf = act && leaf_rule(b - a, size, st.leaf_cap);
        const uint64_t startm = __ballot(is_start);
        const uint64_t leafm = __ballot(is_start && leaf);
        total += (uint32_t)__popcll(startm);
        if (is_start) {
            C += 1;
            if (F == 0xFFu) F = (uint32_t)nlev;
            if (leaf) leaf_end = b;
        }
        if (EMIT && lane == 0) {
            WLevel &lv = lev[nlev];
            lv.starts = startm; lv.covered = actm; lv.leafs = leafm; lv.qb0 = qb0; lv.qb1 = qb1;
        }
        const float half = size * 0.5f;
        const bool split = act && !leaf;
        const float midx = npx + half, midy = npy + half;
        const uint64_t MX = __ballot(split && lx < midx);
        const uint64_t MY = __ballot(split && ly < midy);
        uint32_t q = 0;
        if (split) {
            const uint32_t b2 = mask_part_point(MY, a, b);
            const uint32_t b1 = mask_part_point(MX, a, b2);
            const uint32_t b3 = mask_part_point(MX, b2, b);
            q = (lane >= b1 ? 1u : 0u) + (lane >= b2 ? 1u : 0u) + (lane >= b3 ? 1u : 0u);
            const uint32_t na = q == 0 ? a : (q == 1 ? b1 : (q == 2 ? b2 : b3));
            const uint32_t nb = q == 0 ? b1 : (q == 1 ? b2 : (q == 2 ? b3 : b));
            a = na;
            b = nb;
            npx = npx + ((q & 1) ? half : 0.0f);
            npy = npy + ((q & 2) ? half : 0.0f);
        }
        act = split;
        qb0 = __ballot(split && lane == a && (q & 1u));
        qb1 = __ballot(split && lane == a && (q & 2u));
        size = half;
    }
    if (nlev == kWaveLevels && __ballot(act)) {
        if (lane == 0) atomicOr(err, kErrDepth);
    }
    if (!EMIT) {
        if (C > 0) {
            atomicAdd(&st.count[s0 + lane], C);
            atomicMin(&st.level[s0 + lane], depth0 + F);
        }
        return total;
    }

    // after the scan: P = exclusive prefix at this lane's particle, Fg = its
    // shallowest start level (valid where this lane starts a node)
    const uint32_t P = have ? st.count[s0 + lane] : 0u;
    const uint32_t Fg = have ? st.level[s0 + lane] : 0u;

    // ---- pass B (debug geometry): corners top-down from the parent's start lane
    if (o.debug) {
        float gpx = px0, gpy = py0, gsize = size0;
        for (int L = 0; L < nlev; ++L) {
            const WLevel lv = load_level(lev, L);
            const bool is_start = (lv.starts >> lane) & 1ull;
            float ppx = gpx, ppy = gpy;
            if (L > 0) {
                const WLevel pv = load_level(lev, L - 1);
                const uint64_t below = pv.starts & (lane >= 63 ? ~0ull : ((2ull << lane) - 1ull));
                const uint32_t plane = below ? 63u - (uint32_t)__builtin_clzll(below) : lane;
                ppx = shfl_f(gpx, plane);  // every lane takes part in the shuffle
                ppy = shfl_f(gpy, plane);
            }
            if (is_start) {
                if (L > 0) {
                    const uint32_t qq = (uint32_t)((lv.qb0 >> lane) & 1ull) | ((uint32_t)((lv.qb1 >> lane) & 1ull) << 1);
                    gpx = ppx + ((qq & 1) ? gsize : 0.0f);
                    gpy = ppy + ((qq & 2) ? gsize : 0.0f);
                }
                const uint32_t idx = P + (depth0 + (uint32_t)L - Fg);
                if (idx < o.cap) o.geom[idx] = make_float4(gpx, gpy, gsize, 0.0f);
            }
            gsize = gsize * 0.5f;  // size of the next level's nodes
        }
    }

    // ---- leaf masses: computeLeafMass order, every leaf at once
    const float lm = have ? m[s0 + lane] : 0.0f;
    float lgm = 0.0f, lcx = 0.0f, lcy = 0.0f;
    // a leaf's start lane sums its particles lane, lane + 1, ... in order,
    // shuffled in; the loop runs for the longest leaf of the wave
    const uint32_t span = leaf_end > lane ? leaf_end - lane : 0u;
    uint32_t maxspan = span;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxspan = max(maxspan, (uint32_t)__shfl_xor((int)maxspan, o, 64));
    maxspan = __builtin_amdgcn_readfirstlane(maxspan);
    for (uint32_t t = 0; t < maxspan; ++t) {
        const uint32_t src = min(lane + t, 63u);  // every lane takes part in the shuffles
        const float mj = shfl_f(lm, src);
        const float xj = shfl_f(lx, src);
        const float yj = shfl_f(ly, src);
        if (t < span) {
            lgm = lgm + mj;
            lcx = lcx + xj * mj;
            lcy = lcy + yj * mj;
        }
    }
    if (lgm > 0) { lcx = lcx / lgm; lcy = lcy / lgm; }

    // ---- pass C: bottom-up masses (computeInternalMass, q = 0, 2, 1, 3), records
    float M = 0.0f, X = 0.0f, Y = 0.0f;
    float lsize = size0;
    for (int L = 0; L < nlev - 1; ++L) lsize = lsize * 0.5f;
    for (int L = nlev - 1; L >= 0; --L) {
        const WLevel lv = load_level(lev, L);
        const bool is_start = (lv.starts >> lane) & 1ull;
        const bool is_leaf = (lv.leafs >> lane) & 1ull;
        const uint32_t e = node_end(lv, lane);
        float gm = 0.0f, cx = 0.0f, cy = 0.0f;
        uint32_t mask = 0;
        if (L + 1 < nlev) {
            const WLevel nx = load_level(lev, L + 1);
            const int order[4] = {0, 2, 1, 3};
            for (int k = 0; k < 4; ++k) {
                const int qq = order[k];
                const uint32_t cl = child_lane(nx, lane, e, qq);
                // every lane takes part in the shuffles
                const float cm = shfl_f(M, cl & 63u);
                const float ccx = shfl_f(X, cl & 63u);
                const float ccy = shfl_f(Y, cl & 63u);
                if (is_start && !is_leaf && cl < 64u) {
                    mask |= 1u << qq;
                    gm = gm + cm;
                    cx = cx + ccx * cm;
                    cy = cy + ccy * cm;
                }
            }
            if (gm > 0) { cx = cx / gm; cy = cy / gm; }
        }
        if (is_leaf) { gm = lgm; cx = lcx; cy = lcy; }
        if (is_start) {
            const uint32_t depth = depth0 + (uint32_t)L - Fg;
            const uint32_t idx = P + depth;
            const uint32_t next = (st.count[s0 + e] - P) - depth - 1u;
            write_node(o, idx, cx, cy, gm, 0.0f, 0.0f, lsize, next, s0 + lane, s0 + e, mask, false);
            M = gm; X = cx; Y = cy;
        }
        lsize = lsize * 2.0f;
    }
    rgm = readlane_f(M, 0);
    rcx = readlane_f(X, 0);
    rcy = readlane_f(Y, 0);
    return total;
}

// Split a node with the whole wave: subGridMaker's three searches
// (quadtree.cpp:30-65), the y search by the wave, then the two x searches
// by the two half-waves together.
__device__ __forceinline__ void wave_split(const float *__restrict__ x, const float *__restrict__ y, float px,
                                           float py, float size, uint32_t s, uint32_t e, uint32_t b[5],
                                           uint32_t off = 0) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half;
    b[0] = s;
    b[4] = e;
    b[2] = wave_part_point(y, s, e, midy, off);
    // the two x searches at once, one per half-wave
    const bool hi = lane_id() >= 32;
    const uint32_t r = group_part_point<32>(x, hi ? b[2] : s, hi ? e : b[2], midx, off);
    b[1] = readlane_u(r, 0);
    b[3] = readlane_u(r, 32);
}

__device__ __forceinline__ uint32_t child_count(const uint32_t b[5]) {
    uint32_t k = 0;
    for (int q = 0; q < 4; ++q) k += (b[q + 1] > b[q]) ? 1u : 0u;
    return k;
}

// Lane 0: an internal node and its children at indices first.. (q order).
// exact: the node's own prefix tag (kExactTag | P) when its children's
// ranges are Morton-prefix runs (key-run split); child q then gets the tag
// of prefix 4 P + q.
__device__ void write_children(TopNodes &T, uint32_t node, float px, float py, float size, uint32_t depth,
                               const uint32_t b[5], uint32_t first, uint32_t *out_list, uint32_t exact = 0u) {
    T.kind[node] = kKindInternal;
    const float half = size * 0.5f;
    uint32_t c = first, j = 0;
    for (int q = 0; q < 4; ++q) {
        if (b[q + 1] > b[q]) {
            T.px[c] = px + ((q & 1) ? half : 0.0f);
            T.py[c] = py + ((q & 2) ? half : 0.0f);
            T.size[c] = half;
            T.start[c] = b[q];
            T.end[c] = b[q + 1];
            T.depth[c] = depth + 1;
            T.exact[c] = exact ? (kExactTag | (((exact & ~kExactTag) << 2) + (uint32_t)q)) : 0u;
            T.kind[c] = kKindPending;
            T.child[4 * node + q] = c;
            if (out_list) out_list[j++] = c;
            ++c;
        } else {
            T.child[4 * node + q] = kNone;
        }
    }
}

// Leaf or wave subtree: finish the node now (registering its nodes).
// Returns false if the node needs splitting.  Uniform across the wave.
__device__ bool finish_small(TopNodes &T, const StartTable &st, const float *__restrict__ x,
                             const float *__restrict__ y, uint32_t node, float px, float py, float size,
                             uint32_t s, uint32_t e, uint32_t depth, uint32_t *err, uint32_t off = 0) {
    const uint32_t c = e - s;
    if (leaf_rule(c, size, st.leaf_cap)) {
        if (lane_id() == 0) {
            T.kind[node] = kKindLeaf;
            register_start(st, s, depth);
        }
        return true;
    }
    if (c <= kWaveSubMax) {
        const TreeOut none{nullptr, nullptr, 0, 0};
        float d0, d1, d2;
        wave_subtree<false>(x, y, nullptr, px, py, size, s, c, depth, st, none, nullptr, err, d0, d1, d2, off);
        if (lane_id() == 0) T.kind[node] = kKindWaveSub;
        return true;
    }
    return false;
}

__device__ __forceinline__ void push_block_root(TopNodes &T, uint32_t *counters, uint32_t node) {
    const uint32_t i = atomicAdd(&counters[kCntBlockRoots], 1u);
    if (i < T.cap) T.broot[i] = node;
    else atomicOr(&counters[kCntError], kErrFrontier);
}

// Key-run splits: the data a node's split needs when its range is a run of
// equal Morton prefixes (one system)
struct KeySplit {
    const uint64_t *words;  // sorted packed words: key << idx_bits | slot
    uint32_t *start;        // kKeyStartEntries
    uint8_t *dirty;         // kKeyDirtyEntries
    int idx_bits;
    int on;                 // 0: every split searches
};

__device__ __forceinline__ uint64_t word_key(const KeySplit &k, uint32_t i) {
    return (k.words[i] >> k.idx_bits) & ((1ull << kMortonBits) - 1ull);
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_clear(TopNodes T, StartTable st, const float *__restrict__ sys_bb,
                                                           const uint32_t *__restrict__ sys_off, uint32_t nsys,
                                                           uint32_t n, uint32_t *__restrict__ counters, KeySplit ks) {
    const uint32_t i = blockIdx.x * kTreeBlock + threadIdx.x;
    if (i <= n) {
        st.count[i] = 0;
        st.level[i] = 0xFFFFFFFFu;
    }
    if (ks.on) {
        if (i < kKeyStartEntries) ks.start[i] = 0u;
        if (i < kKeyDirtyEntries) ks.dirty[i] = 0u;
    }
    if (i < nsys) {  // the root of system i: Quadtree::root (quadtree.cpp:90-107)
        const float *bb = sys_bb + 8ull * i;
        T.px[i] = bb[0]; T.py[i] = bb[1]; T.size[i] = bb[2];
        T.start[i] = sys_off[i]; T.end[i] = sys_off[i + 1];
        T.kind[i] = kKindPending; T.depth[i] = 0;
        T.exact[i] = ks.on ? kExactTag : 0u;  // prefix 0
    }
    if (blockIdx.x == 0) {
        const uint32_t t = threadIdx.x;
        if (t < kCntSlots - kCntLevel0) counters[kCntLevel0 + t] = 0;
        if (t == 0) {
            counters[kCntError] &= ~(kErrNodeCap | kErrDepth | kErrFrontier);
            counters[kCntBlockRoots] = 0;
            counters[kCntTopAlloc] = 0;
            counters[kCntStamp] += 1u;
        }
        __syncthreads();
        if (t == 0) {
            counters[kCntLevel0 + 0] = nsys;
            counters[kCntLevelBase0 + 0] = 0;
        }
    }
}

// Key runs of the top levels, one thread per sorted particle i (one
// system).  i starts the run of its L-digit prefix for every level L deeper
// than the digits it shares with particle i - 1; it records that for
// L = 1..kKeyLevels.  It also follows its key's path down from the root
// square with subGridMaker's float arithmetic (mid = pos + size / 2, child
// pos = pos + size / 2 on the upper side) and flags the prefix of every
// level whose quadrant test (x < mid.x, y < mid.y) disagrees with its key
// digit.  A node whose range is a prefix run and whose prefix is not flagged
// has every particle's quadrant equal to its key digit, so the predicates are
// partitioned along the sorted range and partition_point returns the
// children's run starts.
__global__ __launch_bounds__(kTreeBlock) void k_tree_keys(KeySplit ks, const float *__restrict__ x,
                                                          const float *__restrict__ y, uint32_t n,
                                                          const float *__restrict__ sys_bb) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kTreeBlock + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = word_key(ks, i);
    int shared = -1;  // leading digits equal to particle i - 1's
    if (i > 0) {
        const uint64_t d = k ^ word_key(ks, i - 1);
        shared = d ? (__clzll(d) - (64 - kMortonBits)) / 2 : kMortonBits / 2;
    }
    for (int L = shared + 1 < 1 ? 1 : shared + 1; L <= kKeyLevels; ++L)
        ks.start[key_start_base(L) + (uint32_t)(k >> (kMortonBits - 2 * L))] = i + 1u;
    const float px0 = x[i], py0 = y[i];
    float px = sys_bb[0], py = sys_bb[1], size = sys_bb[2];
    for (int L = 0; L < kKeyLevels; ++L) {
        const float half = size * 0.5f;
        const uint32_t dig = (uint32_t)(k >> (kMortonBits - 2 - 2 * L)) & 3u;
        const uint32_t qx = (px0 < px + half) ? 0u : 1u;
        const uint32_t qy = (py0 < py + half) ? 0u : 2u;
        if ((qx | qy) != dig)
            ks.dirty[key_dirty_base(L) + (uint32_t)(L ? k >> (kMortonBits - 2 * L) : 0ull)] = 1u;
        px = px + ((dig & 1u) ? half : 0.0f);
        py = py + ((dig & 2u) ? half : 0.0f);
        size = half;
    }
}

// Probe t (0-based) of partition_point's halving search over [first, last)
// when every probe below `guess` tests true and every other false; false if
// that search takes fewer probes.
__device__ __forceinline__ bool guessed_probe(uint32_t first, uint32_t last, uint32_t guess, uint32_t t,
                                              uint32_t &pos) {
    uint32_t f = first, len = last - first;
    for (uint32_t k = 0; len > 0; ++k) {
        const uint32_t half = len >> 1, m = f + half;
        if (k == t) {
            pos = m;
            return true;
        }
        if (m < guess) {
            f = m + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return false;
}

// A prefix run holding a particle whose quadrant disagrees with its key
// digit: the partition points are still the children's run starts b[1..3]
// whenever every probe the three searches would take on that assumption
// (the y search over [b0, b4), the x searches over [b0, b2) and [b2, b4))
// tests as assumed, because the searches then take exactly those paths.
// The wave checks all probes in one round; false means a probe disagrees and
// the node must be searched.
__device__ __forceinline__ bool verify_guess(const float *__restrict__ x, const float *__restrict__ y, float px,
                                             float py, float size, const uint32_t b[5]) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t c = lane_id() + 64u * (uint32_t)r;  // check c of 96: search c / 32, probe c % 32
        const uint32_t sidx = c >> 5, t = c & 31u;
        if (sidx > 2u) continue;
        // search 0: y over [b0, b4) toward b2; 1: x over [b0, b2) toward b1;
        // 2: x over [b2, b4) toward b3
        const uint32_t lo = sidx == 2u ? b[2] : b[0];
        const uint32_t hi = sidx == 1u ? b[2] : b[4];
        const uint32_t g = sidx == 0u ? b[2] : (sidx == 1u ? b[1] : b[3]);
        uint32_t pos;
        if (!guessed_probe(lo, hi, g, t, pos)) continue;
        const bool pred = sidx == 0u ? (y[pos] < midy) : (x[pos] < midx);
        if (pred != (pos < g)) ok = false;
    }
    // ranges hold fewer than 2^28 particles, so no search takes more than 32 probes
    return __ballot(!ok) == 0ull;
}

__global__ __launch_bounds__(kTreeBlock) void k_tree_level(TopNodes T, StartTable st, const float *__restrict__ x,
                                                           const float *__restrict__ y,
                                                           uint32_t *__restrict__ counters, int L,
                                                           int last_level, KeySplit ks) {
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
        // every field read at once: the chain of dependent loads sets a level's time
        const uint32_t kind = T.kind[node], ex = T.exact[node];
        const float px = T.px[node], py = T.py[node], size = T.size[node];
        const uint32_t s = T.start[node], e = T.end[node];
        if (kind == kKindDead) continue;
        if (finish_small(T, st, x, y, node, px, py, size, s, e, (uint32_t)L, counters + kCntError)) continue;
        if (e - s <= kBlockSubMax || L == last_level) {
            if (lane == 0) push_block_root(T, counters, node);
            continue;
        }
        uint32_t b[5];
        // a prefix run whose quadrants all agree with its key digits splits
        // at its children's run starts; any other node searches
        bool keyed = false;
        if (ks.on && L < kKeyLevels && ex) {
            const uint32_t P = ex & ~kExactTag;  // the run's key prefix, L digits
            const uint32_t *cs = ks.start + key_start_base(L + 1) + 4u * P;
            const uint32_t c1 = cs[1], c2 = cs[2], c3 = cs[3];
            b[0] = s;
            b[4] = e;
            b[2] = c2 ? c2 - 1u : (c3 ? c3 - 1u : e);
            b[1] = c1 ? c1 - 1u : b[2];
            b[3] = c3 ? c3 - 1u : e;
            keyed = !ks.dirty[key_dirty_base(L) + P] || verify_guess(x, y, px, py, size, b);
        }
        if (!keyed) wave_split(x, y, px, py, size, s, e, b);
        if (lane == 0) {
            const uint32_t k = child_count(b);
            const uint32_t off = atomicAdd(&counters[kCntLevel0 + L + 1], k);
            const uint32_t first = next_base + off;
            if (first + k > T.cap) {
                // no room on the next level: hand the node to the workgroup phase
                for (uint32_t cc = first; cc < T.cap && cc < first + k; ++cc) T.kind[cc] = kKindDead;
                push_block_root(T, counters, node);
            } else {
                register_start(st, s, (uint32_t)L);
                write_children(T, node, px, py, size, (uint32_t)L, b, first, nullptr, keyed ? ex : 0u);
            }
        }
    }
}

// Workgroup size of the subtree phase: more waves take a level's nodes in
// parallel (its LDS use does not depend on it)
constexpr int kSubBlock = 512;
constexpr int kSubWaves = kSubBlock / 64;
// Nodes split per level of a workgroup subtree (a staged subtree splits at
// most 4096 / 65 per level), and its internal nodes listed for the mass
// pass (a longer list leaves them to the top-node pass).  Both keep the
// workgroup's LDS under a third of the CU's, so three subtrees run per CU.
constexpr uint32_t kSplitMax = 384;
constexpr uint32_t kQuadListMax = 1024;

// One workgroup per subtree root handed over by the level phase: the level
// loop of that subtree with its frontier in LDS.
__global__ __launch_bounds__(kSubBlock) void k_tree_blocks(TopNodes T, StartTable st, const float *__restrict__ x,
                                                            const float *__restrict__ y,
                                                            uint32_t *__restrict__ counters) {
    __shared__ uint32_t front[2][kFrontMax];
    __shared__ uint32_t fcount[2];
    __shared__ uint32_t blist[kQuadListMax];  // internal nodes, level by level
    __shared__ uint32_t bcount, bbase;
    // per node split on a level: its frontier entry, the three inner
    // partition points and the offset of its children among the level's
    // children
    __shared__ uint32_t sfi[kSplitMax], sb[kSplitMax][3], soff[kSplitMax];
    __shared__ uint32_t nsplit, scount, lcount, lbase;
    __shared__ float sx[kBlockSubMax], sy[kBlockSubMax];  // the subtree's coordinates
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t nroots = min(counters[kCntBlockRoots], T.cap);
    for (uint32_t r = blockIdx.x; r < nroots; r += gridDim.x) {
        // a subtree of at most kBlockSubMax particles is staged in LDS, and its
        // partition-point probes read LDS; larger ones (handed over at the
        // last level) read HBM.
        const uint32_t root = T.broot[r];
        const uint32_t rs = T.start[root], re = T.end[root];
        const bool staged = re > rs && re - rs <= kBlockSubMax;
        if (staged) {
            for (uint32_t i = threadIdx.x; i < re - rs; i += kSubBlock) {
                sx[i] = x[rs + i];
                sy[i] = y[rs + i];
            }
        }
        const float *xs = staged ? sx : x;  // xs[i - off] is particle i
        const float *ys = staged ? sy : y;
        const uint32_t off = staged ? rs : 0u;
        if (threadIdx.x == 0) {
            front[0][0] = root;
            fcount[0] = 1;
            fcount[1] = 0;
            bcount = 0;
            lcount = 0;
            nsplit = 0;
        }
        __syncthreads();
        int cur = 0;
        int lvl = 0;
        for (; lvl < kBlockLevelsMax; ++lvl) {
            const uint32_t nf = min(fcount[cur], kFrontMax);
            if (nf == 0) break;
            // A: every node of the level finished or split (one wave per
            // node); split nodes take their children's offsets in LDS
            for (uint32_t i = w; i < nf; i += kSubWaves) {
                const uint32_t node = front[cur][i];
                const float px = T.px[node], py = T.py[node], size = T.size[node];
                const uint32_t s = T.start[node], e = T.end[node], depth = T.depth[node];
                if (finish_small(T, st, xs, ys, node, px, py, size, s, e, depth, counters + kCntError, off))
                    continue;
                uint32_t b[5];
                wave_split(xs, ys, px, py, size, s, e, b, off);
                if (lane == 0) {
                    const uint32_t k = atomicAdd(&nsplit, 1u);
                    if (k < kSplitMax) {
                        sfi[k] = i;
                        sb[k][0] = b[1]; sb[k][1] = b[2]; sb[k][2] = b[3];
                        soff[k] = atomicAdd(&lcount, child_count(b));
                    } else {
                        atomicOr(&counters[kCntError], kErrFrontier);
                    }
                }
            }
            __syncthreads();
            // one global allocation for all children of the level: slots in
            // the node table are only names, so their order is free
            if (threadIdx.x == 0) {
                const uint32_t k = lcount;
                const uint32_t first = k ? atomicAdd(&counters[kCntTopAlloc], k) : 0u;
                const bool ok = k <= kFrontMax && first + k <= T.cap;
                if (!ok) atomicOr(&counters[kCntError], kErrFrontier);
                lbase = ok ? first : kNone;
                fcount[cur ^ 1] = ok ? k : 0u;
                lcount = 0;
                scount = min(nsplit, kSplitMax);
                nsplit = 0;
            }
            __syncthreads();
            // B: the children records, one thread per split node
            const uint32_t base = lbase;
            const uint32_t ns = scount;
            if (base != kNone) {
                for (uint32_t k = threadIdx.x; k < ns; k += kSubBlock) {
                    const uint32_t o = soff[k];
                    const uint32_t node = front[cur][sfi[k]];
                    const float px = T.px[node], py = T.py[node], size = T.size[node];
                    const uint32_t s = T.start[node], e = T.end[node], depth = T.depth[node];
                    const uint32_t b[5] = {s, sb[k][0], sb[k][1], sb[k][2], e};
                    register_start(st, s, depth);
                    write_children(T, node, px, py, size, depth, b, base + o, &front[cur ^ 1][o]);
                    const uint32_t li = atomicAdd(&bcount, 1u);
                    if (li < kQuadListMax) blist[li] = node;
                }
            }
            __syncthreads();
            cur ^= 1;
        }
        if (lvl == kBlockLevelsMax && threadIdx.x == 0) atomicOr(&counters[kCntError], kErrFrontier);
        // hand the internal-node list to the block mass pass (kNone: too long,
        // the last mass pass takes these nodes instead)
        const uint32_t nl = bcount;
        if (threadIdx.x == 0) {
            bbase = nl <= kQuadListMax ? atomicAdd(&counters[kCntBlistAlloc], nl) : kNone;
            const bool fits = nl <= kQuadListMax && bbase + nl <= T.cap;
            T.blist_off[r] = fits ? bbase : 0u;
            T.blist_cnt[r] = fits ? nl : kNone;
            if (!fits) counters[kCntBlistOverflow] = 1u;
        }
        __syncthreads();
        if (nl <= kQuadListMax && bbase + nl <= T.cap)
            for (uint32_t i = threadIdx.x; i < nl; i += kSubBlock) T.blist[bbase + i] = blist[i];
        __syncthreads();
    }
}

// Every top node's record at its pre-order index; leaves and wave subtrees
// complete (with masses), internal nodes without masses yet.
__global__ __launch_bounds__(kTreeBlock) void k_tree_emit(TopNodes T, StartTable st, const float *__restrict__ x,
                                                          const float *__restrict__ y,
                                                          const float *__restrict__ m,
                                                          uint32_t *__restrict__ counters, TreeOut o,
                                                          uint32_t n) {
#pragma clang fp contract(off)
    __shared__ WLevel lev_all[kTreeWaves][kWaveLevels];
    WLevel *lev = lev_all[threadIdx.x >> 6];
    const uint32_t lane = lane_id();
    const uint32_t total = counters[kCntTopTotal];
    const uint32_t wave = (blockIdx.x * kTreeBlock + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * kTreeBlock) >> 6;
    for (uint32_t t = wave; t < total; t += nwaves) {
        const uint32_t kind = T.kind[t];
        if (kind != kKindLeaf && kind != kKindWaveSub && kind != kKindInternal) continue;
        const uint32_t s = T.start[t], e = T.end[t], depth = T.depth[t];
        if (e > n || s >= e || (kind == kKindWaveSub && e - s > kWaveSubMax)) {
            if (lane == 0) atomicOr(&counters[kCntError], kErrFrontier);
            continue;
        }
        const uint32_t Pa = st.count[s], Fa = st.level[s], Pb = st.count[e];
        const uint32_t d = depth - Fa;
        const uint32_t idx = Pa + d;
        const uint32_t next = (Pb - Pa) - d - 1u;
        if (lane == 0) T.offset[t] = idx;
        if (kind == kKindWaveSub) {
            float gm, cx, cy;
            wave_subtree<true>(x, y, m, T.px[t], T.py[t], T.size[t], s, e - s, depth, st, o, lev,
                               counters + kCntError, gm, cx, cy);
        } else {
            float gm = 0.0f, cx = 0.0f, cy = 0.0f;
            uint32_t mask = 0;
            if (kind == kKindLeaf) {
                // computeLeafMass (quadtree.h:53-67): the wave loads up to 64
                // particles at once, the sums run in particle order
                for (uint32_t c0 = s; c0 < e; c0 += 64u) {
                    const uint32_t cnt = min(64u, e - c0);
                    const bool have = lane < cnt;
                    const float lm = have ? m[c0 + lane] : 0.0f;
                    const float lx = have ? x[c0 + lane] : 0.0f;
                    const float ly = have ? y[c0 + lane] : 0.0f;
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const float mj = readlane_f(lm, j);
                        gm = gm + mj;
                        cx = cx + readlane_f(lx, j) * mj;
                        cy = cy + readlane_f(ly, j) * mj;
                    }
                }
                if (gm > 0) { cx = cx / gm; cy = cy / gm; }
            } else {
                for (int q = 0; q < 4; ++q) mask |= (T.child[4 * t + q] != kNone) ? (1u << q) : 0u;
            }
            if (lane == 0) write_node(o, idx, cx, cy, gm, T.px[t], T.py[t], T.size[t], next, s, e, mask);
        }
    }
}

// computeInternalMass (quadtree.h:69-100) of top node t from its children's
// final records: children q = 0, 2, 1, 3 (subGrids[lr][ud] loop order)
__device__ __forceinline__ void internal_mass(const TopNodes &T, const TreeOut &o, uint32_t t) {
#pragma clang fp contract(off)
    float gm = 0.0f, cx = 0.0f, cy = 0.0f;
    const int order[4] = {0, 2, 1, 3};
    for (int k = 0; k < 4; ++k) {
        const uint32_t ch = T.child[4 * t + order[k]];
        if (ch == kNone || ch >= T.cap) continue;
        const uint32_t oc = T.offset[ch];
        if (oc >= o.cap) continue;
        const NodeRec r = o.nodes[oc];
        gm = gm + r.mass;
        cx = cx + r.comx * r.mass;
        cy = cy + r.comy * r.mass;
    }
    if (gm > 0) { cx = cx / gm; cy = cy / gm; }
    const uint32_t ot = T.offset[t];
    if (ot < o.cap) {
        const float sz = T.size[t];
        reinterpret_cast<float4 *>(o.nodes)[2 * ot] = make_float4(cx, cy, gm, sz * sz);
    }
}

// Masses of each workgroup subtree's internal nodes, deepest level first:
// its list holds the levels in order, so each level is one range.
__global__ __launch_bounds__(kTreeBlock) void k_tree_block_mass(TopNodes T, uint32_t *__restrict__ counters,
                                                                TreeOut o) {
    __shared__ uint32_t lstart[kBlockLevelsMax + 2];
    __shared__ uint32_t maxl;
    const uint32_t stamp = counters[kCntStamp];
    const uint32_t nroots = min(counters[kCntBlockRoots], T.cap);
    for (uint32_t r = blockIdx.x; r < nroots; r += gridDim.x) {
        const uint32_t cnt = T.blist_cnt[r];
        if (cnt == kNone || cnt == 0) continue;  // uniform across the workgroup
        const uint32_t base = T.blist_off[r];
        const uint32_t d0 = T.depth[T.blist[base]];
        if (threadIdx.x == 0) maxl = T.depth[T.blist[base + cnt - 1]] - d0;
        for (uint32_t l = threadIdx.x; l < kBlockLevelsMax + 2; l += kTreeBlock) lstart[l] = cnt;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cnt; i += kTreeBlock) {
            const uint32_t l = T.depth[T.blist[base + i]] - d0;
            const uint32_t lp = i == 0 ? kNone : T.depth[T.blist[base + i - 1]] - d0;
            if (l != lp && l <= kBlockLevelsMax) lstart[l] = i;
        }
        __syncthreads();
        const uint32_t ml = maxl;
        if (ml <= kBlockLevelsMax) {
            for (int l = (int)ml; l >= 0; --l) {
                const uint32_t lo = min(lstart[l], cnt), hi = l == (int)ml ? cnt : min(lstart[l + 1], cnt);
                for (uint32_t i = lo + threadIdx.x; i < hi; i += kTreeBlock) {
                    const uint32_t t = T.blist[base + i];
                    internal_mass(T, o, t);
                    T.pending2[t] = stamp;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// Masses of the remaining internal top nodes (above the workgroup subtrees,
// or in subtrees whose list overflowed), deepest first, by one workgroup:
// nodes of one depth are independent; barriers order the depths.  The nodes
// are listed by depth (T.broot, each node's entry in T.slot).  Up to
// kMassLds listed nodes: every child value a node needs from below the list
// (leaves, wave subtrees, workgroup subtrees: final before this pass) is
// loaded up front into the registers of the thread that owns the node, and
// the listed nodes' own results stay in LDS, so a depth costs LDS reads and
// one barrier.  Longer lists take the same sums through global memory.
constexpr int kMassThreads = 1024;
constexpr int kMaxTopDepth = 256;
constexpr int kMassPer = 4;  // list entries per thread
constexpr uint32_t kMassLds = kMassThreads * kMassPer;
constexpr uint32_t kMassExt = 0xFFFFFFFEu;  // child value held in registers

__global__ __launch_bounds__(kMassThreads) void k_tree_mass(TopNodes T, uint32_t *__restrict__ counters, TreeOut o) {
#pragma clang fp contract(off)
    __shared__ uint32_t hist[kMaxTopDepth + 1];
    __shared__ uint32_t cursor[kMaxTopDepth];
    __shared__ uint32_t maxd;
    __shared__ float lm[kMassLds], lx[kMassLds], ly[kMassLds];
    // without an overflowed list, every node left lies above the workgroup
    // subtrees, i.e. among the nodes the level phase allocated
    const uint32_t total = counters[kCntBlistOverflow] ? counters[kCntTopTotal]
                                                       : min(counters[kCntLevelBase0 + kTopLevels], counters[kCntTopTotal]);
    const uint32_t stamp = counters[kCntStamp];
    for (uint32_t d = threadIdx.x; d <= kMaxTopDepth; d += kMassThreads) hist[d] = 0;
    if (threadIdx.x == 0) maxd = 0;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < total; t += kMassThreads) {
        if (T.kind[t] != kKindInternal || T.pending2[t] == stamp) continue;
        const uint32_t d = T.depth[t];
        if (d < kMaxTopDepth) {
            atomicAdd(&hist[d], 1u);
            atomicMax(&maxd, d);
        } else {
            atomicOr(&counters[kCntError], kErrFrontier);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int d = 0; d < kMaxTopDepth; ++d) {
            const uint32_t v = hist[d];
            hist[d] = run;
            cursor[d] = run;
            run += v;
        }
        hist[kMaxTopDepth] = run;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < total; t += kMassThreads) {
        if (T.kind[t] != kKindInternal || T.pending2[t] == stamp) continue;
        const uint32_t d = T.depth[t];
        if (d < kMaxTopDepth) {
            const uint32_t e = atomicAdd(&cursor[d], 1u);
            T.broot[e] = t;
            T.slot[t] = e;
        }
    }
    __syncthreads();
    const uint32_t len = hist[kMaxTopDepth];
    if (len > kMassLds) {
        for (int d = (int)maxd; d >= 0; --d) {
            const uint32_t lo = hist[d], hi = hist[d + 1];
            for (uint32_t i = lo + threadIdx.x; i < hi; i += kMassThreads) internal_mass(T, o, T.broot[i]);
            __syncthreads();
        }
        return;
    }
    // children in the summation order q = 0, 2, 1, 3 of computeInternalMass:
    // a listed node's entry, kMassExt (value in registers) or kNone
    uint32_t ref[kMassPer][4], ot[kMassPer];
    float em[kMassPer][4], ex[kMassPer][4], ey[kMassPer][4], sz[kMassPer];
#pragma unroll
    for (int k = 0; k < kMassPer; ++k) {
        const uint32_t i = threadIdx.x + (uint32_t)k * kMassThreads;
        ot[k] = kNone;
        sz[k] = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ref[k][q] = kNone;
            em[k][q] = ex[k][q] = ey[k][q] = 0.0f;
        }
        if (i >= len) continue;
        const uint32_t t = T.broot[i];
        ot[k] = T.offset[t];
        sz[k] = T.size[t];
        const int order[4] = {0, 2, 1, 3};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t ch = T.child[4 * t + order[q]];
            if (ch == kNone || ch >= T.cap) continue;
            const uint32_t oc = T.offset[ch];
            if (oc >= o.cap) continue;
            if (T.kind[ch] == kKindInternal && T.pending2[ch] != stamp) {
                ref[k][q] = T.slot[ch];
            } else {
                const NodeRec r = o.nodes[oc];
                ref[k][q] = kMassExt;
                em[k][q] = r.mass;
                ex[k][q] = r.comx;
                ey[k][q] = r.comy;
            }
        }
    }
    for (int d = (int)maxd; d >= 0; --d) {
        const uint32_t lo = hist[d], hi = hist[d + 1];
#pragma unroll
        for (int k = 0; k < kMassPer; ++k) {
            const uint32_t i = threadIdx.x + (uint32_t)k * kMassThreads;
            if (i < lo || i >= hi) continue;
            float gm = 0.0f, cx = 0.0f, cy = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t rq = ref[k][q];
                if (rq == kNone) continue;
                const bool ext = rq == kMassExt;
                const uint32_t li = ext ? 0u : rq;  // an LDS entry in range either way
                const float cm = ext ? em[k][q] : lm[li];
                const float ccx = ext ? ex[k][q] : lx[li];
                const float ccy = ext ? ey[k][q] : ly[li];
                gm = gm + cm;
                cx = cx + ccx * cm;
                cy = cy + ccy * cm;
            }
            if (gm > 0) { cx = cx / gm; cy = cy / gm; }
            lm[i] = gm;
            lx[i] = cx;
            ly[i] = cy;
            if (ot[k] < o.cap) reinterpret_cast<float4 *>(o.nodes)[2 * ot[k]] = make_float4(cx, cy, gm, sz[k] * sz[k]);
        }
        __syncthreads();
    }
}

}  // namespace

void launch_tree_build(ge_ctx *c, bool sorted_words) {
    const State &s = c->b.st[c->b.cur];
    TopNodes &T = c->b.top;
    const uint64_t n = c->n;
    const StartTable st{c->b.start_count, c->b.start_level, c->leaf_cap};
    // key-run splits for a single system (a batch's systems search)
    const KeySplit ks{c->b.sort_a, c->b.kstart, c->b.kdirty, c->idx_bits,
                      (sorted_words && c->nsys == 1) ? 1 : 0};
    const uint64_t clear_n = ks.on ? std::max<uint64_t>(n + 1, kKeyStartEntries) : n + 1;
    hipLaunchKernelGGL(k_tree_clear, dim3(blocks_for(clear_n, kTreeBlock)), dim3(kTreeBlock), 0, c->stream, T, st,
                       c->b.sys_bb, c->b.sys_off, c->nsys, c->n, c->b.counters, ks);
    if (ks.on && n > 0)
        hipLaunchKernelGGL(k_tree_keys, dim3(blocks_for(n, kTreeBlock)), dim3(kTreeBlock), 0, c->stream, ks, s.x, s.y,
                           c->n, c->b.sys_bb);
    // level L+1 only holds children of nodes above kBlockSubMax particles
    const uint64_t big_children = 4 * (n / (kBlockSubMax + 1)) + 4;
    for (int L = 0; L < kTopLevels; ++L) {
        uint64_t maxn = c->nsys;
        for (int k = 0; k < L && maxn < n; ++k) maxn *= 4;
        if (L > 0 && maxn > big_children) maxn = big_children;
        if (maxn > n) maxn = n;
        hipLaunchKernelGGL(k_tree_level, dim3(wave_grid(maxn)), dim3(kTreeBlock), 0, c->stream, T, st, s.x, s.y,
                           c->b.counters, L, kTopLevels - 1, ks);
    }
    // roots of workgroup subtrees each hold more than 64 particles
    uint64_t roots = n / (kWaveSubMax + 1) + c->nsys;
    if (roots > 4096) roots = 4096;
    hipLaunchKernelGGL(k_tree_blocks, dim3((uint32_t)roots), dim3(kSubBlock), 0, c->stream, T, st, s.x, s.y,
                       c->b.counters);
    const uint32_t len = c->n + 1;
    const uint32_t parts = (uint32_t)((len + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_sums, dim3(parts), dim3(kScanThreads), 0, c->stream, st.count, len, c->b.scan_part);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanThreads), 0, c->stream, c->b.scan_part, parts, c->b.counters,
                       c->b.node_cap, T.cap);
    hipLaunchKernelGGL(k_scan_apply, dim3(parts), dim3(kScanThreads), 0, c->stream, st.count, len, c->b.scan_part);
    TreeOut o{c->b.nodes, c->b.node_geom, c->b.node_cap, c->debug ? 1 : 0};
    uint64_t top_bound = 8 * (n / (kWaveSubMax + 1)) + 64;
    if (top_bound > T.cap) top_bound = T.cap;
    hipLaunchKernelGGL(k_tree_emit, dim3(wave_grid(top_bound)), dim3(kTreeBlock), 0, c->stream, T, st, s.x, s.y, s.m,
                       c->b.counters, o, c->n);
    hipLaunchKernelGGL(k_tree_block_mass, dim3((uint32_t)roots), dim3(kTreeBlock), 0, c->stream, T, c->b.counters, o);
    hipLaunchKernelGGL(k_tree_mass, dim3(1), dim3(kMassThreads), 0, c->stream, T, c->b.counters, o);
}

}  // namespace ge
