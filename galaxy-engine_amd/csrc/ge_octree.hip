// ge_octree.hip -- octree build reproducing Octree / Node3D::Node3D /
// Node3D::subGridMaker3D (GE/src/Physics/quadtree.cpp:109-292) node for node:
// the same float partition points (z over the node, y over each z half, x
// over each quarter), the same leaf rule (quadtree.cpp:120-121), the same DFS
// pre-order with `next` skip counts (calculateNextNeighbor3D, :255-273), and
// the same mass / centre-of-mass rounding order (computeLeafMass3D,
// quadtree.h:159-174; computeInternalMass3D, quadtree.cpp:228-253, which
// loops subGrids[x][y][z] and so sums octants q = x + 2y + 4z in the order
// 0, 4, 2, 6, 1, 5, 3, 7).
//
// The schedule is the quadtree build's (ge_tree.hip) with eight children per
// node: pre-order indices from the node-start table and its exclusive scan,
// a grid-wide phase of one launch per level for nodes above kBlockSubMax
// particles, one workgroup per subtree below that, one wave per subtree of at
// most 64 particles (three ballot masks, seven bit searches per split), then
// the records, and the internal masses deepest level first.
#include <algorithm>

#include "ge_internal3d.hpp"
#include "ge_tree_common.hpp"

namespace ge {
namespace {

// k-th octant summed by computeInternalMass3D: the three bits of k reversed
// (0, 4, 2, 6, 1, 5, 3, 7)
__device__ __forceinline__ int order3(int k) { return ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1); }

struct TreeOut3 {
    NodeRec3 *nodes;
    float4 *geom;
    uint32_t *mask;
    uint32_t cap;
};

__device__ __forceinline__ void write_node3(const TreeOut3 &o, uint32_t idx, float cx, float cy, float cz, float gm,
                                            float size, uint32_t next, uint32_t s, uint32_t e, uint32_t mask) {
    if (idx >= o.cap) return;  // kErrNodeCap is raised by the scan
    NodeRec3 r;
    r.comx = cx; r.comy = cy; r.comz = cz; r.mass = gm;
    r.size_sq = size * size;
    r.next = next; r.start = s; r.end = e;
    o.nodes[idx] = r;
    o.mask[idx] = mask;
}

struct WLevel3 {
    uint64_t starts, covered, leafs, qb0, qb1, qb2;
};

__device__ __forceinline__ WLevel3 load_level3(const WLevel3 *lev, int L) {
    WLevel3 r;
    r.starts = rfl64(lev[L].starts);
    r.covered = rfl64(lev[L].covered);
    r.leafs = rfl64(lev[L].leafs);
    r.qb0 = rfl64(lev[L].qb0);
    r.qb1 = rfl64(lev[L].qb1);
    r.qb2 = rfl64(lev[L].qb2);
    return r;
}

__device__ __forceinline__ uint32_t node_end3(const WLevel3 &lv, uint32_t lane) {
    const uint64_t stop = (lv.starts | ~lv.covered) & above(lane);
    return stop ? (uint32_t)__builtin_ctzll(stop) : 64u;
}

// lane of the child with octant q inside [a, b) on the next level, or 64
__device__ __forceinline__ uint32_t child_lane3(const WLevel3 &nx, uint32_t a, uint32_t b, int q) {
    const uint64_t range = (b >= 64 ? ~0ull : ((1ull << b) - 1ull)) & (~0ull << a);
    uint64_t sel = nx.starts & range;
    sel &= (q & 1) ? nx.qb0 : ~nx.qb0;
    sel &= (q & 2) ? nx.qb1 : ~nx.qb1;
    sel &= (q & 4) ? nx.qb2 : ~nx.qb2;
    return sel ? (uint32_t)__builtin_ctzll(sel) : 64u;
}

// Wave subtree: particles [s0, s0 + cnt), cnt <= 64, one per lane, all of its
// levels at once (as the quadtree's wave subtree, with three masks).  EMIT =
// false registers its nodes in the start table; EMIT = true writes them.
template <bool EMIT>
__device__ uint32_t wave_subtree3(const float *__restrict__ x, const float *__restrict__ y,
                                  const float *__restrict__ z, const float *__restrict__ m, float px0, float py0,
                                  float pz0, float size0, uint32_t s0, uint32_t cnt, uint32_t depth0,
                                  const StartTable &st, const TreeOut3 &o, WLevel3 *lev, uint32_t *err,
                                  uint32_t off = 0) {
#pragma clang fp contract(off)
    const uint32_t lane = lane_id();
    const bool have = lane < cnt;
    const float lx = have ? x[s0 + lane - off] : 0.0f;  // x[i - off] is particle i
    const float ly = have ? y[s0 + lane - off] : 0.0f;
    const float lz = have ? z[s0 + lane - off] : 0.0f;

    // ---- pass A: top-down, every node of a level at once
    bool act = have;
    uint32_t a = 0, b = cnt;
    float npx = px0, npy = py0, npz = pz0, size = size0;
    uint32_t C = 0, F = 0xFFu, leaf_end = 0;
    uint64_t qb0 = 0, qb1 = 0, qb2 = 0;
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
            WLevel3 &lv = lev[nlev];
            lv.starts = startm; lv.covered = actm; lv.leafs = leafm;
            lv.qb0 = qb0; lv.qb1 = qb1; lv.qb2 = qb2;
        }
        const float half = size * 0.5f;
        const bool split = act && !leaf;
        const float midx = npx + half, midy = npy + half, midz = npz + half;
        const uint64_t MX = __ballot(split && lx < midx);
        const uint64_t MY = __ballot(split && ly < midy);
        const uint64_t MZ = __ballot(split && lz < midz);
        uint32_t q = 0;
        if (split) {
            uint32_t B[9];
            B[0] = a;
            B[8] = b;
            B[4] = mask_part_point(MZ, a, b);
            B[2] = mask_part_point(MY, a, B[4]);
            B[6] = mask_part_point(MY, B[4], b);
            B[1] = mask_part_point(MX, a, B[2]);
            B[3] = mask_part_point(MX, B[2], B[4]);
            B[5] = mask_part_point(MX, B[4], B[6]);
            B[7] = mask_part_point(MX, B[6], b);
            // B is non-decreasing: the lane's child is [last B[k] <= lane,
            // first B[k] > lane) and q counts the boundaries at or below it
            uint32_t na = a, nb = b;
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                const bool le = lane >= B[k];
                q += le ? 1u : 0u;
                na = le ? B[k] : na;
                nb = (!le && B[k] < nb) ? B[k] : nb;
            }
            a = na;
            b = nb;
            npx = npx + ((q & 1) ? half : 0.0f);
            npy = npy + ((q & 2) ? half : 0.0f);
            npz = npz + ((q & 4) ? half : 0.0f);
        }
        act = split;
        qb0 = __ballot(split && lane == a && (q & 1u));
        qb1 = __ballot(split && lane == a && (q & 2u));
        qb2 = __ballot(split && lane == a && (q & 4u));
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

    const uint32_t P = have ? st.count[s0 + lane] : 0u;
    const uint32_t Fg = have ? st.level[s0 + lane] : 0u;

    // ---- pass B: corners top-down from the parent's start lane
    {
        float gpx = px0, gpy = py0, gpz = pz0, gsize = size0;
        for (int L = 0; L < nlev; ++L) {
            const WLevel3 lv = load_level3(lev, L);
            const bool is_start = (lv.starts >> lane) & 1ull;
            float ppx = gpx, ppy = gpy, ppz = gpz;
            if (L > 0) {
                const WLevel3 pv = load_level3(lev, L - 1);
                const uint64_t below = pv.starts & (lane >= 63 ? ~0ull : ((2ull << lane) - 1ull));
                const uint32_t plane = below ? 63u - (uint32_t)__builtin_clzll(below) : lane;
                ppx = shfl_f(gpx, plane);  // every lane takes part in the shuffles
                ppy = shfl_f(gpy, plane);
                ppz = shfl_f(gpz, plane);
            }
            if (is_start) {
                if (L > 0) {
                    const uint32_t qq = (uint32_t)((lv.qb0 >> lane) & 1ull) |
                                        ((uint32_t)((lv.qb1 >> lane) & 1ull) << 1) |
                                        ((uint32_t)((lv.qb2 >> lane) & 1ull) << 2);
                    gpx = ppx + ((qq & 1) ? gsize : 0.0f);
                    gpy = ppy + ((qq & 2) ? gsize : 0.0f);
                    gpz = ppz + ((qq & 4) ? gsize : 0.0f);
                }
                const uint32_t idx = P + (depth0 + (uint32_t)L - Fg);
                if (idx < o.cap) o.geom[idx] = make_float4(gpx, gpy, gpz, gsize);
            }
            gsize = gsize * 0.5f;
        }
    }

    // ---- leaf masses: computeLeafMass3D order, every leaf at once
    const float lm = have ? m[s0 + lane] : 0.0f;
    float lgm = 0.0f, lcx = 0.0f, lcy = 0.0f, lcz = 0.0f;
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
        const float zj = shfl_f(lz, src);
        if (t < span) {
            lgm = lgm + mj;
            lcx = lcx + xj * mj;
            lcy = lcy + yj * mj;
            lcz = lcz + zj * mj;
        }
    }
    if (lgm > 0) { lcx = lcx / lgm; lcy = lcy / lgm; lcz = lcz / lgm; }

    // ---- pass C: bottom-up masses in the order 0,4,2,6,1,5,3,7; records
    float M = 0.0f, X = 0.0f, Y = 0.0f, Z = 0.0f;
    float lsize = size0;
    for (int L = 0; L < nlev - 1; ++L) lsize = lsize * 0.5f;
    for (int L = nlev - 1; L >= 0; --L) {
        const WLevel3 lv = load_level3(lev, L);
        const bool is_start = (lv.starts >> lane) & 1ull;
        const bool is_leaf = (lv.leafs >> lane) & 1ull;
        const uint32_t e = node_end3(lv, lane);
        float gm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        uint32_t mask = 0;
        if (L + 1 < nlev) {
            const WLevel3 nx = load_level3(lev, L + 1);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int qq = order3(k);
                const uint32_t cl = child_lane3(nx, lane, e, qq);
                const float cm = shfl_f(M, cl & 63u);
                const float ccx = shfl_f(X, cl & 63u);
                const float ccy = shfl_f(Y, cl & 63u);
                const float ccz = shfl_f(Z, cl & 63u);
                if (is_start && !is_leaf && cl < 64u) {
                    mask |= 1u << qq;
                    gm = gm + cm;
                    cx = cx + ccx * cm;
                    cy = cy + ccy * cm;
                    cz = cz + ccz * cm;
                }
            }
            if (gm > 0) { cx = cx / gm; cy = cy / gm; cz = cz / gm; }
        }
        if (is_leaf) { gm = lgm; cx = lcx; cy = lcy; cz = lcz; }
        if (is_start) {
            const uint32_t depth = depth0 + (uint32_t)L - Fg;
            const uint32_t idx = P + depth;
            const uint32_t next = (st.count[s0 + e] - P) - depth - 1u;
            write_node3(o, idx, cx, cy, cz, gm, lsize, next, s0 + lane, s0 + e, mask);
            M = gm; X = cx; Y = cy; Z = cz;
        }
        lsize = lsize * 2.0f;
    }
    return total;
}

// subGridMaker3D's seven searches (quadtree.cpp:146-199): z by the whole
// wave, the two y searches by half-waves, the four x searches by
// quarter-waves, each group giving the sequential search's result
__device__ __forceinline__ void wave_split3(const float *__restrict__ x, const float *__restrict__ y,
                                            const float *__restrict__ z, float px, float py, float pz, float size,
                                            uint32_t s, uint32_t e, uint32_t B[9], uint32_t off = 0) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half, midz = pz + half;
    B[0] = s;
    B[8] = e;
    B[4] = wave_part_point(z, s, e, midz, off);
    // the y searches of the two z halves by the two half-waves together
    const uint32_t lane = lane_id();
    const bool h = lane >= 32;
    const uint32_t ry = group_part_point<32>(y, h ? B[4] : s, h ? e : B[4], midy, off);
    B[2] = readlane_u(ry, 0);
    B[6] = readlane_u(ry, 32);
    // the x searches of the four quarters by the four quarter-waves together
    const uint32_t g = lane >> 4;
    const uint32_t lo = g == 0 ? s : (g == 1 ? B[2] : (g == 2 ? B[4] : B[6]));
    const uint32_t hi = g == 0 ? B[2] : (g == 1 ? B[4] : (g == 2 ? B[6] : e));
    const uint32_t rx = group_part_point<16>(x, lo, hi, midx, off);
    B[1] = readlane_u(rx, 0);
    B[3] = readlane_u(rx, 16);
    B[5] = readlane_u(rx, 32);
    B[7] = readlane_u(rx, 48);
}

__device__ __forceinline__ uint32_t child_count8(const uint32_t B[9]) {
    uint32_t k = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) k += (B[q + 1] > B[q]) ? 1u : 0u;
    return k;
}

// Lane 0: an internal node and its children at indices first.. (octant order).
// exact: the node's own prefix tag (kExactTag | P) when its children's
// ranges are Morton-prefix runs (key-run split); child q then gets the tag
// of prefix 8 P + q.
__device__ void write_children3(TopNodes3 &T, uint32_t node, float px, float py, float pz, float size,
                                uint32_t depth, const uint32_t B[9], uint32_t first, uint32_t *out_list,
                                uint32_t exact = 0u) {
    T.kind[node] = kKindInternal;
    const float half = size * 0.5f;
    uint32_t c = first, j = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (B[q + 1] > B[q]) {
            T.px[c] = px + ((q & 1) ? half : 0.0f);
            T.py[c] = py + ((q & 2) ? half : 0.0f);
            T.pz[c] = pz + ((q & 4) ? half : 0.0f);
            T.size[c] = half;
            T.start[c] = B[q];
            T.end[c] = B[q + 1];
            T.depth[c] = depth + 1;
            T.exact[c] = exact ? (kExactTag | (((exact & ~kExactTag) << 3) + (uint32_t)q)) : 0u;
            T.kind[c] = kKindPending;
            T.child[8 * node + q] = c;
            if (out_list) out_list[j++] = c;
            ++c;
        } else {
            T.child[8 * node + q] = kNone;
        }
    }
}

__device__ bool finish_small3(TopNodes3 &T, const StartTable &st, const float *__restrict__ x,
                              const float *__restrict__ y, const float *__restrict__ z, uint32_t node, float px,
                              float py, float pz, float size, uint32_t s, uint32_t e, uint32_t depth,
                              uint32_t *err, uint32_t off = 0) {
    const uint32_t c = e - s;
    if (leaf_rule(c, size, st.leaf_cap)) {
        if (lane_id() == 0) {
            T.kind[node] = kKindLeaf;
            register_start(st, s, depth);
        }
        return true;
    }
    if (c <= kWaveSubMax) {
        const TreeOut3 none{nullptr, nullptr, nullptr, 0};
        wave_subtree3<false>(x, y, z, nullptr, px, py, pz, size, s, c, depth, st, none, nullptr, err, off);
        if (lane_id() == 0) T.kind[node] = kKindWaveSub;
        return true;
    }
    return false;
}

__device__ __forceinline__ void push_block_root3(TopNodes3 &T, uint32_t *counters, uint32_t node) {
    const uint32_t i = atomicAdd(&counters[kCntBlockRoots], 1u);
    if (i < T.cap) T.broot[i] = node;
    else atomicOr(&counters[kCntError], kErrFrontier);
}

// Key-run splits: the data a node's split needs when its range is a run of
// equal Morton prefixes
struct KeySplit3 {
    const uint64_t *keys;  // sorted 63-bit keys
    uint32_t *start;       // kKeyStartEntries3
    uint8_t *dirty;        // kKeyDirtyEntries3
    int on;                // 0: every split searches
};

__global__ __launch_bounds__(kTreeBlock) void k3_tree_clear(TopNodes3 T, StartTable st, const float *__restrict__ bb,
                                                            uint32_t n, uint32_t *__restrict__ counters,
                                                            KeySplit3 ks) {
    const uint32_t i = blockIdx.x * kTreeBlock + threadIdx.x;
    if (i <= n) {
        st.count[i] = 0;
        st.level[i] = 0xFFFFFFFFu;
    }
    if (ks.on) {
        if (i < kKeyStartEntries3) ks.start[i] = 0u;
        if (i < kKeyDirtyEntries3) ks.dirty[i] = 0u;
    }
    if (i == 0) {  // Octree::root (quadtree.cpp:275-292)
        T.px[0] = bb[0]; T.py[0] = bb[1]; T.pz[0] = bb[2]; T.size[0] = bb[3];
        T.start[0] = 0; T.end[0] = n;
        T.kind[0] = kKindPending; T.depth[0] = 0;
        T.exact[0] = ks.on ? kExactTag : 0u;  // prefix 0
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
            counters[kCntLevel0 + 0] = 1;
            counters[kCntLevelBase0 + 0] = 0;
        }
    }
}

// Key runs of the top levels, one thread per sorted particle i: the octree
// counterpart of the quadtree's k_tree_keys.  i starts the run of its
// L-digit prefix for every level L deeper than the digits it shares with
// particle i - 1 (recorded for L = 1..kKeyLevels3), and flags the prefix of
// every level on its key's path whose octant tests (subGridMaker3D's
// x < mid.x, y < mid.y, z < mid.z on the node cube, child corner
// pos + size / 2 on the upper side) disagree with its key digit.
__global__ __launch_bounds__(kTreeBlock) void k3_tree_keys(KeySplit3 ks, const float *__restrict__ x,
                                                           const float *__restrict__ y, const float *__restrict__ z,
                                                           uint32_t n, const float *__restrict__ bb) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * kTreeBlock + threadIdx.x;
    if (i >= n) return;
    constexpr int kDigits = kMortonBits3 / 3;
    const uint64_t k = ks.keys[i];
    int shared = -1;  // leading digits equal to particle i - 1's
    if (i > 0) {
        const uint64_t d = k ^ ks.keys[i - 1];
        shared = d ? (__clzll(d) - (64 - kMortonBits3)) / 3 : kDigits;
    }
    for (int L = shared + 1 < 1 ? 1 : shared + 1; L <= kKeyLevels3; ++L)
        ks.start[key_start_base3(L) + (uint32_t)(k >> (kMortonBits3 - 3 * L))] = i + 1u;
    const float qx0 = x[i], qy0 = y[i], qz0 = z[i];
    float px = bb[0], py = bb[1], pz = bb[2], size = bb[3];
    for (int L = 0; L < kKeyLevels3; ++L) {
        const float half = size * 0.5f;
        const uint32_t dig = (uint32_t)(k >> (kMortonBits3 - 3 - 3 * L)) & 7u;
        const uint32_t q = ((qx0 < px + half) ? 0u : 1u) | ((qy0 < py + half) ? 0u : 2u) | ((qz0 < pz + half) ? 0u : 4u);
        if (q != dig)
            ks.dirty[key_dirty_base3(L) + (uint32_t)(L ? k >> (kMortonBits3 - 3 * L) : 0ull)] = 1u;
        px = px + ((dig & 1u) ? half : 0.0f);
        py = py + ((dig & 2u) ? half : 0.0f);
        pz = pz + ((dig & 4u) ? half : 0.0f);
        size = half;
    }
}

// Probe t (0-based) of partition_point's halving search over [first, last)
// when every probe below `guess` tests true and every other false; false if
// that search takes fewer probes.
__device__ __forceinline__ bool guessed_probe3(uint32_t first, uint32_t last, uint32_t guess, uint32_t t,
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

// A prefix run holding a particle whose octant disagrees with its key digit:
// the partition points are the children's run starts B[1..7] when every
// probe the seven searches would take on that assumption tests as assumed
// (z over [B0, B8) toward B4; y over [B0, B4) toward B2 and [B4, B8) toward
// B6; x over the four quarters toward B1, B3, B5, B7).  The wave checks all
// of them in one round, up to 32 probes per search.
__device__ __forceinline__ bool verify_guess3(const float *__restrict__ x, const float *__restrict__ y,
                                              const float *__restrict__ z, float px, float py, float pz, float size,
                                              const uint32_t B[9]) {
#pragma clang fp contract(off)
    const float half = size * 0.5f;
    const float midx = px + half, midy = py + half, midz = pz + half;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t c = lane_id() + 64u * (uint32_t)r;  // check c of 256: search c / 32, probe c % 32
        const uint32_t sidx = c >> 5, t = c & 31u;
        if (sidx > 6u) continue;
        // search 0: z over [B0, B8) toward B4; 1, 2: y over the z halves
        // toward B2, B6; 3..6: x over the quarters toward B1, B3, B5, B7
        uint32_t lo, hi, g;
        const float *v;
        float mid;
        if (sidx == 0u) {
            lo = B[0]; hi = B[8]; g = B[4]; v = z; mid = midz;
        } else if (sidx <= 2u) {
            // selects, not B[4 h]: an index that varies across lanes would
            // move B to scratch memory for the whole kernel
            const bool h = sidx == 2u;
            lo = h ? B[4] : B[0]; hi = h ? B[8] : B[4]; g = h ? B[6] : B[2]; v = y; mid = midy;
        } else {
            const uint32_t h = sidx - 3u;
            lo = h == 0u ? B[0] : (h == 1u ? B[2] : (h == 2u ? B[4] : B[6]));
            hi = h == 0u ? B[2] : (h == 1u ? B[4] : (h == 2u ? B[6] : B[8]));
            g = h == 0u ? B[1] : (h == 1u ? B[3] : (h == 2u ? B[5] : B[7]));
            v = x; mid = midx;
        }
        uint32_t pos;
        if (!guessed_probe3(lo, hi, g, t, pos)) continue;
        if ((v[pos] < mid) != (pos < g)) ok = false;
    }
    // ranges hold fewer than 2^28 particles, so no search takes more than 32 probes
    return __ballot(!ok) == 0ull;
}

__global__ __launch_bounds__(kTreeBlock) void k3_tree_level(TopNodes3 T, StartTable st, const float *__restrict__ x,
                                                            const float *__restrict__ y, const float *__restrict__ z,
                                                            uint32_t *__restrict__ counters, int L, int last_level,
                                                            KeySplit3 ks) {
    const uint32_t lane = lane_id();
    const uint32_t base = counters[kCntLevelBase0 + L];
    uint32_t cnt = counters[kCntLevel0 + L];
    if (base >= T.cap) cnt = 0;
    else if (cnt > T.cap - base) cnt = T.cap - base;
    const uint32_t next_base = base + cnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[kCntLevelBase0 + L + 1] = next_base;
        if (L == last_level) counters[kCntTopAlloc] = next_base;
    }
    const uint32_t wave = (blockIdx.x * kTreeBlock + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * kTreeBlock) >> 6;
    for (uint32_t t = wave; t < cnt; t += nwaves) {
        const uint32_t node = base + t;
        // every field read at once: the chain of dependent loads sets a level's time
        const uint32_t kind = T.kind[node], ex = T.exact[node];
        const float px = T.px[node], py = T.py[node], pz = T.pz[node], size = T.size[node];
        const uint32_t s = T.start[node], e = T.end[node];
        if (kind == kKindDead) continue;
        if (finish_small3(T, st, x, y, z, node, px, py, pz, size, s, e, (uint32_t)L, counters + kCntError)) continue;
        if (e - s <= kBlockSubMax || L == last_level) {
            if (lane == 0) push_block_root3(T, counters, node);
            continue;
        }
        uint32_t B[9];
        // a prefix run whose octants all agree with its key digits splits at
        // its children's run starts; a run holding a disagreeing particle
        // takes them after one round of probe checks; any other node searches
        bool keyed = false;
        if (ks.on && L < kKeyLevels3 && ex) {
            const uint32_t P = ex & ~kExactTag;  // the run's key prefix, L digits
            const uint32_t *cs = ks.start + key_start_base3(L + 1) + 8u * P;
            B[0] = s;
            B[8] = e;
#pragma unroll
            for (int q = 7; q >= 1; --q) {
                const uint32_t cq = cs[q];
                B[q] = cq ? cq - 1u : B[q + 1];
            }
            keyed = !ks.dirty[key_dirty_base3(L) + P] || verify_guess3(x, y, z, px, py, pz, size, B);
        }
        if (!keyed) wave_split3(x, y, z, px, py, pz, size, s, e, B);
        if (lane == 0) {
            const uint32_t k = child_count8(B);
            const uint32_t off = atomicAdd(&counters[kCntLevel0 + L + 1], k);
            const uint32_t first = next_base + off;
            if (first + k > T.cap) {
                for (uint32_t cc = first; cc < T.cap && cc < first + k; ++cc) T.kind[cc] = kKindDead;
                push_block_root3(T, counters, node);
            } else {
                register_start(st, s, (uint32_t)L);
                write_children3(T, node, px, py, pz, size, (uint32_t)L, B, first, nullptr, keyed ? ex : 0u);
            }
        }
    }
}

// Workgroup size of the subtree phase: more waves take a level's nodes in
// parallel (its LDS use does not depend on it)
constexpr int kSubBlock3 = 1024;
constexpr int kSubWaves3 = kSubBlock3 / 64;

__global__ __launch_bounds__(kSubBlock3) void k3_tree_blocks(TopNodes3 T, StartTable st, const float *__restrict__ x,
                                                             const float *__restrict__ y,
                                                             const float *__restrict__ z,
                                                             uint32_t *__restrict__ counters) {
    __shared__ uint32_t front[2][kFrontMax];
    __shared__ uint32_t fcount[2];
    __shared__ uint32_t blist[kBlockListMax];
    __shared__ uint32_t bcount, bbase;
    __shared__ float sx[kBlockSubMax], sy[kBlockSubMax], sz[kBlockSubMax];  // the subtree's coordinates
    // per frontier node of a level: the offset of its children among the
    // level's children (kNone: not split)
    __shared__ uint32_t soff[kFrontMax];
    __shared__ uint32_t lcount, lbase;
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t nroots = min(counters[kCntBlockRoots], T.cap);
    for (uint32_t r = blockIdx.x; r < nroots; r += gridDim.x) {
        // a subtree of at most kBlockSubMax particles is staged in LDS (the
        // probes of its splits read LDS); larger ones read HBM
        const uint32_t root = T.broot[r];
        const uint32_t rs = T.start[root], re = T.end[root];
        const bool staged = re > rs && re - rs <= kBlockSubMax;
        if (staged) {
            for (uint32_t i = threadIdx.x; i < re - rs; i += kSubBlock3) {
                sx[i] = x[rs + i];
                sy[i] = y[rs + i];
                sz[i] = z[rs + i];
            }
        }
        const float *xs = staged ? sx : x;  // xs[i - off] is particle i
        const float *ys = staged ? sy : y;
        const float *zs = staged ? sz : z;
        const uint32_t off = staged ? rs : 0u;
        if (threadIdx.x == 0) {
            front[0][0] = root;
            fcount[0] = 1;
            fcount[1] = 0;
            bcount = 0;
            lcount = 0;
        }
        __syncthreads();
        int cur = 0;
        int lvl = 0;
        for (; lvl < kBlockLevelsMax; ++lvl) {
            const uint32_t nf = min(fcount[cur], kFrontMax);
            if (nf == 0) break;
            // A: every node of the level finished or split (one wave per
            // node); a split node parks its seven inner partition points in
            // its own child slots and takes its children's offset in LDS
            for (uint32_t i = w; i < nf; i += kSubWaves3) {
                const uint32_t node = front[cur][i];
                const float px = T.px[node], py = T.py[node], pz = T.pz[node], size = T.size[node];
                const uint32_t s = T.start[node], e = T.end[node], depth = T.depth[node];
                if (finish_small3(T, st, xs, ys, zs, node, px, py, pz, size, s, e, depth, counters + kCntError,
                                  off)) {
                    if (lane == 0) soff[i] = kNone;
                    continue;
                }
                uint32_t B[9];
                wave_split3(xs, ys, zs, px, py, pz, size, s, e, B, off);
                if (lane == 0) {
#pragma unroll
                    for (int q = 1; q < 8; ++q) T.child[8 * node + q - 1] = B[q];
                    soff[i] = atomicAdd(&lcount, child_count8(B));
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
            }
            __syncthreads();
            // B: the children records, one thread per split node
            const uint32_t base = lbase;
            if (base != kNone) {
                for (uint32_t i = threadIdx.x; i < nf; i += kSubBlock3) {
                    const uint32_t o = soff[i];
                    if (o == kNone) continue;
                    const uint32_t node = front[cur][i];
                    const float px = T.px[node], py = T.py[node], pz = T.pz[node], size = T.size[node];
                    const uint32_t s = T.start[node], e = T.end[node], depth = T.depth[node];
                    uint32_t B[9];
                    B[0] = s;
                    B[8] = e;
#pragma unroll
                    for (int q = 1; q < 8; ++q) B[q] = T.child[8 * node + q - 1];
                    register_start(st, s, depth);
                    write_children3(T, node, px, py, pz, size, depth, B, base + o, &front[cur ^ 1][o]);
                    const uint32_t li = atomicAdd(&bcount, 1u);
                    if (li < kBlockListMax) blist[li] = node;
                }
            }
            __syncthreads();
            cur ^= 1;
        }
        if (lvl == kBlockLevelsMax && threadIdx.x == 0) atomicOr(&counters[kCntError], kErrFrontier);
        const uint32_t nl = bcount;
        if (threadIdx.x == 0) {
            bbase = nl <= kBlockListMax ? atomicAdd(&counters[kCntBlistAlloc], nl) : kNone;
            const bool fits = nl <= kBlockListMax && bbase + nl <= T.cap;
            T.blist_off[r] = fits ? bbase : 0u;
            T.blist_cnt[r] = fits ? nl : kNone;
            if (!fits) counters[kCntBlistOverflow] = 1u;
        }
        __syncthreads();
        if (nl <= kBlockListMax && bbase + nl <= T.cap)
            for (uint32_t i = threadIdx.x; i < nl; i += kSubBlock3) T.blist[bbase + i] = blist[i];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kTreeBlock) void k3_tree_emit(TopNodes3 T, StartTable st, const float *__restrict__ x,
                                                           const float *__restrict__ y, const float *__restrict__ z,
                                                           const float *__restrict__ m,
                                                           uint32_t *__restrict__ counters, TreeOut3 o, uint32_t n) {
#pragma clang fp contract(off)
    __shared__ WLevel3 lev_all[kTreeWaves][kWaveLevels];
    WLevel3 *lev = lev_all[threadIdx.x >> 6];
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
            wave_subtree3<true>(x, y, z, m, T.px[t], T.py[t], T.pz[t], T.size[t], s, e - s, depth, st, o, lev,
                                counters + kCntError);
        } else {
            float gm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
            uint32_t mask = 0;
            if (kind == kKindLeaf) {
                // the wave loads up to 64 particles at once, the sums run in
                // particle order
                for (uint32_t c0 = s; c0 < e; c0 += 64u) {
                    const uint32_t cnt = min(64u, e - c0);
                    const bool have = lane < cnt;
                    const float lm = have ? m[c0 + lane] : 0.0f;
                    const float lx = have ? x[c0 + lane] : 0.0f;
                    const float ly = have ? y[c0 + lane] : 0.0f;
                    const float lz = have ? z[c0 + lane] : 0.0f;
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const float mj = readlane_f(lm, j);
                        gm = gm + mj;
                        cx = cx + readlane_f(lx, j) * mj;
                        cy = cy + readlane_f(ly, j) * mj;
                        cz = cz + readlane_f(lz, j) * mj;
                    }
                }
                if (gm > 0) { cx = cx / gm; cy = cy / gm; cz = cz / gm; }
            } else {
                for (int q = 0; q < 8; ++q) mask |= (T.child[8 * t + q] != kNone) ? (1u << q) : 0u;
            }
            if (lane == 0) {
                write_node3(o, idx, cx, cy, cz, gm, T.size[t], next, s, e, mask);
                if (idx < o.cap) o.geom[idx] = make_float4(T.px[t], T.py[t], T.pz[t], T.size[t]);
            }
        }
    }
}

// computeInternalMass3D of top node t from its children's final records
__device__ __forceinline__ void internal_mass3(const TopNodes3 &T, const TreeOut3 &o, uint32_t t) {
#pragma clang fp contract(off)
    float gm = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t ch = T.child[8 * t + order3(k)];
        if (ch == kNone || ch >= T.cap) continue;
        const uint32_t oc = T.offset[ch];
        if (oc >= o.cap) continue;
        const float4 r = reinterpret_cast<const float4 *>(o.nodes)[2 * oc];
        gm = gm + r.w;
        cx = cx + r.x * r.w;
        cy = cy + r.y * r.w;
        cz = cz + r.z * r.w;
    }
    if (gm > 0) { cx = cx / gm; cy = cy / gm; cz = cz / gm; }
    const uint32_t ot = T.offset[t];
    if (ot < o.cap) reinterpret_cast<float4 *>(o.nodes)[2 * ot] = make_float4(cx, cy, cz, gm);
}

__global__ __launch_bounds__(kTreeBlock) void k3_tree_block_mass(TopNodes3 T, uint32_t *__restrict__ counters,
                                                                 TreeOut3 o) {
    __shared__ uint32_t lstart[kBlockLevelsMax + 2];
    __shared__ uint32_t maxl;
    const uint32_t stamp = counters[kCntStamp];
    const uint32_t nroots = min(counters[kCntBlockRoots], T.cap);
    for (uint32_t r = blockIdx.x; r < nroots; r += gridDim.x) {
        const uint32_t cnt = T.blist_cnt[r];
        if (cnt == kNone || cnt == 0) continue;
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
                    internal_mass3(T, o, t);
                    T.pending2[t] = stamp;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

constexpr int kMassThreads = 1024;
constexpr int kMaxTopDepth = 256;

__global__ __launch_bounds__(kMassThreads) void k3_tree_mass(TopNodes3 T, uint32_t *__restrict__ counters,
                                                             TreeOut3 o) {
#pragma clang fp contract(off)
    __shared__ uint32_t hist[kMaxTopDepth + 1];
    __shared__ uint32_t cursor[kMaxTopDepth];
    __shared__ uint32_t maxd;
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
        if (d < kMaxTopDepth) T.broot[atomicAdd(&cursor[d], 1u)] = t;
    }
    __syncthreads();
    for (int d = (int)maxd; d >= 0; --d) {
        const uint32_t lo = hist[d], hi = hist[d + 1];
        for (uint32_t i = lo + threadIdx.x; i < hi; i += kMassThreads) internal_mass3(T, o, T.broot[i]);
        __syncthreads();
    }
}

}  // namespace

void launch_octree_build(ge3_ctx *c, const float *bb_device, bool sorted_keys) {
    const State3 &s = c->b.st[c->b.cur];
    TopNodes3 &T = c->b.top;
    const uint64_t n = c->n;
    const StartTable st{c->b.start_count, c->b.start_level, c->leaf_cap};
    const KeySplit3 ks{c->b.sk_a, c->b.kstart, c->b.kdirty, sorted_keys ? 1 : 0};
    const uint64_t clear_n = ks.on ? std::max<uint64_t>(n + 1, kKeyStartEntries3) : n + 1;
    hipLaunchKernelGGL(k3_tree_clear, dim3(blocks_for(clear_n, kTreeBlock)), dim3(kTreeBlock), 0, c->stream, T, st,
                       bb_device, c->n, c->b.counters, ks);
    if (ks.on && n > 0)
        hipLaunchKernelGGL(k3_tree_keys, dim3(blocks_for(n, kTreeBlock)), dim3(kTreeBlock), 0, c->stream, ks, s.x, s.y,
                           s.z, c->n, bb_device);
    // level L+1 only holds children of nodes above kBlockSubMax particles
    const uint64_t big_children = 8 * (n / (kBlockSubMax + 1)) + 8;
    for (int L = 0; L < kTopLevels; ++L) {
        uint64_t maxn = 1;
        for (int k = 0; k < L && maxn < n; ++k) maxn *= 8;
        if (L > 0 && maxn > big_children) maxn = big_children;
        if (maxn > n) maxn = n;
        hipLaunchKernelGGL(k3_tree_level, dim3(wave_grid(maxn)), dim3(kTreeBlock), 0, c->stream, T, st, s.x, s.y,
                           s.z, c->b.counters, L, kTopLevels - 1, ks);
    }
    uint64_t roots = n / (kWaveSubMax + 1) + 1;
    if (roots > 4096) roots = 4096;
    hipLaunchKernelGGL(k3_tree_blocks, dim3((uint32_t)roots), dim3(kSubBlock3), 0, c->stream, T, st, s.x, s.y, s.z,
                       c->b.counters);
    const uint32_t len = c->n + 1;
    const uint32_t parts = (uint32_t)((len + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_sums, dim3(parts), dim3(kScanThreads), 0, c->stream, st.count, len, c->b.scan_part);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanThreads), 0, c->stream, c->b.scan_part, parts, c->b.counters,
                       c->b.node_cap, T.cap);
    hipLaunchKernelGGL(k_scan_apply, dim3(parts), dim3(kScanThreads), 0, c->stream, st.count, len, c->b.scan_part);
    TreeOut3 o{c->b.nodes, c->b.node_geom, c->b.node_mask, c->b.node_cap};
    uint64_t top_bound = 16 * (n / (kWaveSubMax + 1)) + 64;
    if (top_bound > T.cap) top_bound = T.cap;
    hipLaunchKernelGGL(k3_tree_emit, dim3(wave_grid(top_bound)), dim3(kTreeBlock), 0, c->stream, T, st, s.x, s.y, s.z,
                       s.m, c->b.counters, o, c->n);
    hipLaunchKernelGGL(k3_tree_block_mass, dim3((uint32_t)roots), dim3(kTreeBlock), 0, c->stream, T, c->b.counters, o);
    hipLaunchKernelGGL(k3_tree_mass, dim3(1), dim3(kMassThreads), 0, c->stream, T, c->b.counters, o);
}

}  // namespace ge
